// wfa_engine.h — host-side internals shared by the C ABI (wfa_api.cpp)
// and the BiWFA driver (wfa_biwfa.cpp): penalties, the supported
// configurations, the engine object and launch-parameter helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstring>
#include <map>
#include <mutex>

#include "wfa_device.h"

namespace wfahost {

using wfa::KParams;

// ---- penalties: WavefrontPenalties::new_* (src/penalties.rs:122-296) ----
struct Internal {
  int metric, match, x, o1, e1, o2, e2;
};

inline int make_internal(const wfa_penalties_t& u, Internal* p) {
  p->metric = u.distance_metric;
  switch (u.distance_metric) {
    case WFA_METRIC_INDEL:
      *p = {u.distance_metric, 0, -1, 1, -1, -1, -1};
      return WFA_OK;
    case WFA_METRIC_EDIT:
      *p = {u.distance_metric, 0, 1, 1, -1, -1, -1};
      return WFA_OK;
    case WFA_METRIC_GAP_LINEAR:
      if (u.match > 0 || u.mismatch <= 0 || u.gap_opening1 <= 0) return WFA_ERR_INVALID_PENALTIES;
      if (u.match < 0)
        *p = {u.distance_metric, u.match, 2 * u.mismatch - 2 * u.match,
              2 * u.gap_opening1 - u.match, -1, -1, -1};
      else
        *p = {u.distance_metric, 0, u.mismatch, u.gap_opening1, -1, -1, -1};
      return WFA_OK;
    case WFA_METRIC_GAP_AFFINE:
      if (u.match > 0 || u.mismatch <= 0 || u.gap_opening1 < 0 || u.gap_extension1 <= 0)
        return WFA_ERR_INVALID_PENALTIES;
      if (u.match < 0)
        *p = {u.distance_metric, u.match, 2 * u.mismatch - 2 * u.match, 2 * u.gap_opening1,
              2 * u.gap_extension1 - u.match, -1, -1};
      else
        *p = {u.distance_metric, 0, u.mismatch, u.gap_opening1, u.gap_extension1, -1, -1};
      return WFA_OK;
    case WFA_METRIC_GAP_AFFINE_2P:
      if (u.match > 0 || u.mismatch <= 0 || u.gap_opening1 < 0 || u.gap_extension1 <= 0 ||
          u.gap_opening2 < 0 || u.gap_extension2 <= 0)
        return WFA_ERR_INVALID_PENALTIES;
      if (u.match < 0)
        *p = {u.distance_metric, u.match, 2 * u.mismatch - 2 * u.match, 2 * u.gap_opening1,
              2 * u.gap_extension1 - u.match, 2 * u.gap_opening2,
              2 * u.gap_extension2 - u.match};
      else
        *p = {u.distance_metric, 0, u.mismatch, u.gap_opening1, u.gap_extension1,
              u.gap_opening2, u.gap_extension2};
      return WFA_OK;
    default:
      return WFA_ERR_INVALID_PENALTIES;
  }
}

// max_score_scope (src/components.rs:435-483)
inline int score_scope(const Internal& p) {
  switch (p.metric) {
    case WFA_METRIC_INDEL:
    case WFA_METRIC_EDIT: return 2;
    case WFA_METRIC_GAP_LINEAR: return std::max(p.x, p.o1) + 1;
    case WFA_METRIC_GAP_AFFINE: return std::max(p.o1 + p.e1, p.x) + 1;
    default: return std::max(std::max(p.o1 + p.e1, p.o2 + p.e2), p.x) + 1;
  }
}

// Which configurations the device path runs. The rest return
// WFA_ERR_UNSUPPORTED rather than silently falling back to anything.
inline int check_supported(const wfa_attr_t& a) {
  if (a.span != WFA_SPAN_END2END && a.span != WFA_SPAN_ENDSFREE) return WFA_ERR_INVALID_ARGUMENT;
  if (a.span == WFA_SPAN_ENDSFREE && !a.extension &&
      (a.pattern_begin_free < 0 || a.pattern_end_free < 0 || a.text_begin_free < 0 ||
       a.text_end_free < 0))
    return WFA_ERR_INVALID_ARGUMENT;
  if (a.heuristic < WFA_HEURISTIC_NONE || a.heuristic > WFA_HEURISTIC_BANDED_ADAPTIVE)
    return WFA_ERR_INVALID_ARGUMENT;
  if (a.alignment_scope != WFA_SCOPE_SCORE && a.alignment_scope != WFA_SCOPE_ALIGNMENT)
    return WFA_ERR_INVALID_ARGUMENT;
  if (a.memory_mode < WFA_MEMORY_HIGH || a.memory_mode > WFA_MEMORY_ULTRALOW)
    return WFA_ERR_INVALID_ARGUMENT;
  if (a.compat != WFA_COMPAT_C && a.compat != WFA_COMPAT_CORRECTED) return WFA_ERR_INVALID_ARGUMENT;
  // BiWFA: end-to-end only, and no Z-drop (C refuses it since 7110ee3, and
  // the reference's driver refuses X-drop and Z-drop,
  // src/bin/align_benchmark.rs:529-560); X-drop sub-alignments run, as in
  // the reference's aligner (tests/xdrop_steps.rs:57-92)
  if (a.memory_mode == WFA_MEMORY_ULTRALOW &&
      (a.span != WFA_SPAN_END2END || a.heuristic == WFA_HEURISTIC_ZDROP))
    return WFA_ERR_UNSUPPORTED;
  return WFA_OK;
}

inline int pow2ceil(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

#define HIPCHK(x)                                   \
  do {                                              \
    hipError_t _e = (x);                            \
    if (_e != hipSuccess) return WFA_ERR_DEVICE;    \
  } while (0)

}  // namespace wfahost

// =====================================================================
// Engine
// =====================================================================
struct wfa_engine_t {
  int device = 0;
  int num_cus = 0;
  size_t lds_max = 0;
  int32_t* d_counters = nullptr;  // [0] queue head, [1] retries, [2] scratch, [3] next pool chunk
  int32_t* d_lists[2] = {nullptr, nullptr};
  size_t list_cap = 0;
  int32_t* h_pinned = nullptr;
  // scratch
  int32_t* d_hist = nullptr;
  size_t hist_bytes = 0;
  int4* d_hdr = nullptr;
  size_t hdr_bytes = 0;
  int32_t* d_gring = nullptr;
  size_t gring_bytes = 0;
  uint8_t* d_chain = nullptr;
  size_t chain_bytes = 0;
  int32_t* d_tab = nullptr;  // history pool: chunk ids per slot
  size_t tab_bytes = 0;
  int32_t* d_dom = nullptr;  // corrected ends-free dominance records per slot
  size_t dom_bytes = 0;
  int4* d_ztr = nullptr;     // Compat::C Z-drop trimmed ranges per slot and score
  size_t ztr_bytes = 0;
  uint8_t* d_zsign = nullptr;  // Compat::C Z-drop offset signs per history cell
  size_t zsign_bytes = 0;
  void* d_order = nullptr;   // longest-first work order: keys, sorted keys, values, order, temp
  size_t order_bytes = 0;
  int last_retries = 0;
  // device scratch one launch may take (wavefront rings, CIGAR history);
  // `auto_budget`: recomputed before each call from the device's free
  // memory (refresh_budget) until wfa_engine_set_scratch_budget fixes it
  int64_t scratch_budget = (int64_t)48 << 30;
  bool auto_budget = true;
  int64_t long_budget = 0;  // the last long-read budget and when it was taken
  std::chrono::steady_clock::time_point long_budget_at;
  // most workgroups of a long-read launch (0 = as many as fit): a serving
  // loop leaves room on every CU for the short requests' launches
  int long_grid_cap = 0;
  // long reads: the first pass's history pool takes the whole budget (a
  // batch engine, allocated once and reused); a serving loop's engines size
  // it by the slots' average and leave the rest to retries
  bool long_pool_whole_budget = true;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  float first_ms = 0.f, all_ms = 0.f;
  // BiWFA scratch (wfa_biwfa.cpp), grown on demand
  static constexpr int kBiBuffers = 12;
  void* bi_buf[kBiBuffers] = {};
  size_t bi_cap[kBiBuffers] = {};
  // No-allocation mode (wfa_align_batch_noalloc): every device buffer is a
  // piece of one fixed device array of `arena_cap` bytes, handed out by a
  // bump pointer: the work queue counters at its start, then one call's
  // pair lists and work order, then (from `arena_launch`, again for every
  // launch of the call) that launch's rings, history and tables. Host
  // reads of the counters go through `h_pinned` (then ordinary memory).
  uint8_t* arena = nullptr;
  size_t arena_cap = 0;
  // the arena is the engine's own allocation (a serving loop's engines:
  // one allocation up front, none while launches of other engines run)
  bool own_arena = false;
  // phase-clock builds: the short-read kernel's cycle sums of the last call
  unsigned long long* d_phase = nullptr;
  // Launch shapes of long-read launches (pairs over 1024 bases, rings in
  // HBM, no heuristic) by (kind, mode, length bucket): {waves per pair,
  // workgroups per CU}; from wfa_engine_autotune or
  // wfa_engine_set_launch_shape, else the fixed plan
  std::map<int, std::pair<int, int>> shapes;
  size_t arena_top = 0;
  size_t arena_launch = 0;
  // longest pair a call may bring (0: no limit), WFA_ERR_LIMIT beyond it
  int max_len_limit = 0;
  std::mutex mu;
};

namespace wfahost {

// Device scratch grows geometrically (to at most twice the old size and
// 1 GiB beyond the need), so a batch engine reallocates a buffer only a
// few times. hipFree waits for the whole device: engines that must not
// stall other streams' work (a serving loop's) take their buffers from a
// fixed arena instead (egrow).
inline int grow(void** ptr, size_t* have, size_t need) {
  if (*have >= need) return WFA_OK;
  size_t cap = std::max(need, std::min(*have * 2, need + ((size_t)1 << 30)));
  if (*ptr) hipFree(*ptr);
  *ptr = nullptr;
  *have = 0;
  if (hipMalloc(ptr, cap) != hipSuccess) {
    (void)hipGetLastError();
    cap = need;
    if (hipMalloc(ptr, cap) != hipSuccess) {
      (void)hipGetLastError();
      *ptr = nullptr;
      return WFA_ERR_OUT_OF_MEMORY;
    }
  }
  *have = cap;
  return WFA_OK;
}

// grow() or, in no-allocation mode, the next `need` bytes of the arena
// (256-byte aligned); the arena never grows: WFA_ERR_OUT_OF_MEMORY when the
// piece does not fit.
inline int egrow(wfa_engine_t* e, void** ptr, size_t* have, size_t need) {
  if (!e->arena) return grow(ptr, have, need);
  const size_t at = (e->arena_top + 255) & ~(size_t)255;
  if (at > e->arena_cap || need > e->arena_cap - at) return WFA_ERR_OUT_OF_MEMORY;
  *ptr = e->arena + at;
  *have = need;
  e->arena_top = at + need;
  return WFA_OK;
}

// Device bytes this engine holds in its scratch buffers
inline int64_t held_bytes(const wfa_engine_t* e) {
  if (e->arena) return 0;
  int64_t h = (int64_t)(e->hist_bytes + e->hdr_bytes + e->gring_bytes + e->chain_bytes +
                        e->tab_bytes + e->dom_bytes + e->ztr_bytes + e->zsign_bytes +
                        e->order_bytes) +
              (int64_t)e->list_cap * 8;
  for (size_t c : e->bi_cap) h += (int64_t)c;
  return h;
}

// The automatic scratch budget of a call. Short reads: 48 GiB. Long reads
// (`long_reads`: sequences over 1024 bases, and BiWFA): half of what this
// engine could reach (the device's free memory and what it already
// holds), between 48 GiB (or 90% of that reach, when less is free) and
// 128 GiB: on an idle 288 GB device a long-read launch keeps twice the
// workgroups' rows and history a fixed 48 GiB allowed, and engines that
// start in a process already holding memory get less. The device query
// (hipMemGetInfo, about a millisecond) runs at most once a second.
inline void refresh_budget(wfa_engine_t* e, bool long_reads) {
  if (!e->auto_budget || e->arena) return;
  constexpr int64_t kShort = (int64_t)48 << 30;
  if (!long_reads) {
    e->scratch_budget = kShort;
    return;
  }
  const auto now = std::chrono::steady_clock::now();
  if (e->long_budget > 0 && now - e->long_budget_at < std::chrono::seconds(1)) {
    e->scratch_budget = e->long_budget;
    return;
  }
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
    (void)hipGetLastError();
    e->scratch_budget = kShort;
    return;
  }
  const int64_t reach = (int64_t)free_b + held_bytes(e);
  int64_t b = std::min<int64_t>(reach / 2, (int64_t)128 << 30);
  b = std::max<int64_t>(b, std::min<int64_t>(kShort, reach / 10 * 9));
  e->long_budget = std::max<int64_t>(b, (int64_t)64 << 20);
  e->long_budget_at = now;
  e->scratch_budget = e->long_budget;
}

inline int kind_of(int metric) {
  if (metric == WFA_METRIC_GAP_AFFINE) return wfa::KIND_AFFINE;
  if (metric == WFA_METRIC_GAP_AFFINE_2P) return wfa::KIND_AFFINE2P;
  if (metric == WFA_METRIC_GAP_LINEAR) return wfa::KIND_LINEAR;
  return wfa::KIND_EDIT;
}

// Build kernel parameters shared by every launch of one batch.
inline void base_params(const Internal& ip, const wfa_attr_t& attr, const wfa_batch_t& b, KParams* p) {
  std::memset(p, 0, sizeof(*p));
  p->metric = ip.metric;
  p->x = ip.x;
  p->o1 = ip.o1;
  p->e1 = ip.e1;
  p->o2 = ip.o2;
  p->e2 = ip.e2;
  p->oe1 = ip.o1 + ip.e1;
  p->oe2 = ip.o2 + ip.e2;
  p->scope = score_scope(ip);
  const int kind = kind_of(ip.metric);
  const bool gaps = kind == wfa::KIND_AFFINE || kind == wfa::KIND_AFFINE2P;
  p->ringM = kind == wfa::KIND_EDIT ? 2 : p->scope;
  p->ringE1 = gaps ? ip.e1 + 1 : 0;
  p->ringE2 = kind == wfa::KIND_AFFINE2P ? ip.e2 + 1 : 0;
  // score-only and medium/low memory keep only the wavefronts in scope
  // (src/aligner.rs:855-866): no score-indexed slot limit
  p->modular = attr.alignment_scope == WFA_SCOPE_SCORE || attr.memory_mode != WFA_MEMORY_HIGH;
  p->max_steps = attr.max_alignment_steps;
  p->endsfree = attr.span == WFA_SPAN_ENDSFREE;
  p->extension = p->endsfree && attr.extension;
  p->pb_free = attr.pattern_begin_free;
  p->pe_free = attr.pattern_end_free;
  p->tb_free = attr.text_begin_free;
  p->te_free = attr.text_end_free;
  p->pair_cfg = (const int4*)b.pair_config;
  p->u_match = attr.penalties.match;
  p->u_x = attr.penalties.mismatch;
  p->u_o1 = attr.penalties.gap_opening1;
  p->u_e1 = attr.penalties.gap_extension1;
  p->u_o2 = attr.penalties.gap_opening2;
  p->u_e2 = attr.penalties.gap_extension2;
  p->heuristic = attr.heuristic;
  p->hp0 = attr.heuristic_p[0];
  p->hp1 = attr.heuristic_p[1];
  p->hp2 = attr.heuristic_p[2];
  p->fmt = b.seq_format;
  p->seq = b.seq_data;
  p->poff = b.pattern_offset;
  p->plen = b.pattern_length;
  p->toff = b.text_offset;
  p->tlen = b.text_length;
  p->score = b.score;
  p->status = b.status;
  p->cig = b.cigar_ops;
  p->cig_off = b.cigar_offset;
  p->cig_len = b.cigar_length;
  p->end_pos = b.end_position;
  p->compat = attr.compat;
  p->low = attr.memory_mode == WFA_MEMORY_LOW;
}

// BiWFA batches (memory mode ULTRALOW): wfa_biwfa.cpp
int biwfa_align_batch(wfa_engine_t* e, const wfa_attr_t* attr, const wfa_batch_t* b,
                      hipStream_t st, int first, int count);

// wfa_engine_align_batch's work: validate, then align under the engine's
// lock (wfa_api.cpp)
int align_batch_locked(wfa_engine_t* e, const wfa_attr_t* attr, const wfa_batch_t* b,
                       hipStream_t st);

inline int read_counter(wfa_engine_t* e, int idx, hipStream_t st, int32_t* out) {
  HIPCHK(hipMemcpyAsync(e->h_pinned, e->d_counters + idx, 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  *out = e->h_pinned[0];
  return WFA_OK;
}

}  // namespace wfahost
