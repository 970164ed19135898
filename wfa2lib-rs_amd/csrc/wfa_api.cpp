// wfa_api.cpp — host side of the C ABI (include/wfa_mi355x.h).
//
// The aligner object mirrors the reference's WavefrontAligner<N>
// (src/aligner.rs:146-211): it holds the configuration the setters change
// and the last result (score, status, CIGAR). Alignment runs on the GPU:
// each call stages its pair in device memory and launches the same
// persistent kernel the batch engine uses, with one work item. The engine
// owns the device work queue and per-slot scratch, sizes the launch from
// the occupancy the kernel's LDS use allows on the device, and re-runs the
// (rare) pairs that outgrow the first launch's per-pair storage.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "wfa_device.h"
#include "wfa_engine.h"

using namespace wfahost;
using wfa::KParams;

namespace {

struct LaunchCfg {
  int kind;
  int mode;  // 0 score only, 1 high memory CIGAR, 2 medium/low memory CIGAR
  bool cig;
  bool lds_ring;
  bool i16;
  bool shortk;  // the short-read kernel (wfa_short.h)
  int short_rc;  // its ring class
  int short_gsh; // its score step, 2^short_gsh
  int nw;       // waves per pair
  int wcap;
  int seq_words;
  int ring_rows;
  size_t lds_bytes;
  int grid;
  int hdr_cap;
  int64_t hist_cap;  // history elements per slot, on average
  int chunk;         // history pool: elements per chunk
  int pool_chunks;
  int tab_cap;       // chunks one slot may hold
  int chain_cap;
  int dom_stride;    // corrected ends-free dominance record per slot (ints)
};


// Launches whose pairs may end by Z-drop with a medium/low-memory CIGAR in
// Compat::C (a heuristic launch: the heuristic may be per pair) keep every
// score's trimmed ranges for that backtrace.
bool zdrop_ranges(const KParams& p, int mode) {
  return mode == 2 && p.heuristic != WFA_HEURISTIC_NONE && p.compat == WFA_COMPAT_C;
}

// The length bucket of a launch: ceil(log2(longest sequence))
int length_bucket(int max_len) {
  int b = 0;
  while ((1 << b) < max_len && b < 30) ++b;
  return b;
}

int shape_key(int kind, int mode, int max_len) {
  return (kind * 4 + mode) * 64 + length_bucket(max_len);
}

// Waves per pair a wide kernel is built for (mode, 16-bit rows)
bool wide_waves_built(int nw, int mode, bool i16) {
  return nw == wfa::kWideWaves || (i16 && mode != 1 && (nw == 4 || nw == 16));
}

int ring_rows_of(const KParams& p, int kind) {
  if (kind == wfa::KIND_EDIT || kind == wfa::KIND_LINEAR) return p.ringM;
  if (kind == wfa::KIND_AFFINE) return p.ringM + 2 * p.ringE1;
  return p.ringM + 2 * p.ringE1 + 2 * p.ringE2;
}

// Choose storage for a launch over pairs of length <= max_len. `wide`
// selects the full-width configuration that cannot overflow the ring.
int plan(wfa_engine_t* e, const KParams& p, int kind, int mode, int fmt, int max_len, int npairs,
         int pass, LaunchCfg* c) {
  const bool cig = mode != 0;
  c->kind = kind;
  c->mode = mode;
  c->cig = cig;
  c->ring_rows = ring_rows_of(p, kind);
  // ASCII pairs are staged 2-bit when every byte is A/C/G/T and byte-wise
  // otherwise; when byte-wise staging of the longest pair would not fit in
  // LDS, LDS is sized for 2-bit staging and a pair that needs bytes comes
  // back for retry (and ends with WFA_STATUS_OOM)
  const int per = fmt == WFA_SEQ_PACKED2BITS ||
                          (size_t)2 * (((max_len + 7) / 8 + 3) & ~1) * 8 > e->lds_max / 2
                      ? 32
                      : 8;
  // even, so the per-row int4 metadata after the two sequences is 16-byte aligned
  c->seq_words = ((max_len + per - 1) / per + 2 + 1) & ~1;
  // the widest possible computed range: [-(plen+scope+1), tlen+scope+1]
  const int full = pow2ceil(2 * max_len + 2 * p.scope + 4);
  // Short reads start narrow and re-run the few pairs that outgrow it; long
  // reads (whose wavefronts widen by up to two diagonals per score) start at
  // full width.
  const bool long_reads = max_len > 1024;
  c->wcap = pass == 0 && !long_reads ? std::min(256, full) : full;
  const size_t seq_meta = (size_t)2 * c->seq_words * 8 + (size_t)c->ring_rows * 16;
  // 16-bit offsets for pairs of at most 8000 bases (the kernel re-runs any
  // pair whose score passes 16000 with 32-bit offsets)
  // (a Compat::C Z-drop end of a medium/low CIGAR follows the winners of
  // cells holding values below zero, whose order 16 bits do not keep)
  c->i16 = max_len <= 8000 && pass < 2 && !zdrop_ranges(p, mode);
  const size_t ring_bytes = (size_t)c->ring_rows * c->wcap * (c->i16 ? 2 : 4);
  c->lds_ring = seq_meta + ring_bytes <= 64 * 1024;
  // rings in HBM are wide: several waves share each pair's rows (heuristic
  // cut-offs need the single-wave kernel); their cross-wave reductions take
  // 16 ints per wave of LDS
  c->nw = !c->lds_ring && !wfa::heur_kernel(p) ? wfa::kWideWaves : 1;
  // The wide kernels are bound by the rings' HBM traffic (every score reads
  // and writes whole rows): 16-bit rows for pairs of at most 16000 bases
  // halve it (offsets stay below 2^15; a NULL row value rises by at most
  // one per score, and pairs whose score passes 16000 re-run with 32 bits)
  if (!c->lds_ring) c->i16 = c->nw > 1 && max_len <= 16000 && pass < 2 && !zdrop_ranges(p, mode);
  // a tuned shape (wfa_engine_autotune) for long reads: waves per pair and
  // workgroups per CU
  int tuned_per_cu = 0;
  if (long_reads && !c->lds_ring && c->nw > 1) {
    auto it = e->shapes.find(shape_key(kind, mode, max_len));
    if (it != e->shapes.end()) {
      if (wide_waves_built(it->second.first, mode, c->i16)) c->nw = it->second.first;
      tuned_per_cu = it->second.second;
    }
  }
  const size_t red_bytes = c->lds_ring ? 0 : (size_t)64 * c->nw;
  c->lds_bytes = seq_meta + (c->lds_ring ? ring_bytes : red_bytes);
  if (c->lds_bytes > e->lds_max) return WFA_ERR_UNSUPPORTED;
  // First pass of gap-affine end-to-end alignments without a heuristic,
  // score only or with high-memory CIGARs: the short-read kernel (a fixed
  // 448-diagonal window per pair, 16-bit rows, a per-slot history of packed
  // rows); pairs that leave its window come back for the general kernel.
  // Its rings hold a fixed number of rows (the ring class), enough for the
  // lookbacks of most penalty sets once scores step by their common power
  // of two.
  c->short_rc = wfa::short_ring_class(p.x, p.oe1, p.e1, &c->short_gsh);
  c->shortk = pass == 0 && kind == wfa::KIND_AFFINE && (mode == 0 || mode == 1) &&
              p.heuristic == WFA_HEURISTIC_NONE && !p.endsfree && max_len <= 8000 &&
              c->short_rc >= 0;
  int per_cu;
  if (c->shortk) {
    c->seq_words = ((max_len + 31) / 32 + 2 + 1) & ~1;
    c->lds_bytes = (size_t)2 * c->seq_words * 8 + wfa::short_lds_row_bytes_host(c->short_rc);
    if (c->lds_bytes > e->lds_max) c->shortk = false;
  }
  if (c->shortk) {
    per_cu = wfa::short_blocks_per_cu(mode, fmt == WFA_SEQ_PACKED2BITS, c->short_rc,
                                      c->lds_bytes);
    int grid = std::max(1, std::min(e->num_cus * per_cu, npairs));
    c->lds_ring = true;
    c->i16 = true;
    c->nw = 1;
    c->wcap = wfa::kShortWindow;
    c->dom_stride = 0;
    c->chunk = c->pool_chunks = c->tab_cap = c->chain_cap = 0;
    c->hist_cap = 0;
    c->hdr_cap = 0;
    if (cig) {
      // rows for the first 1024 scores of the kernel's score step (a pair
      // scoring more goes to the general kernel): 16-bit cells over the
      // window, one header per row
      c->hdr_cap = 1024;
      const int64_t per_slot = (int64_t)c->hdr_cap * (wfa::kShortWindow * 2 + 4);
      if ((int64_t)grid * per_slot > e->scratch_budget)
        grid = (int)std::max<int64_t>(1, e->scratch_budget / per_slot);
    }
    c->grid = grid;
    return WFA_OK;
  }
  c->seq_words = ((max_len + per - 1) / per + 2 + 1) & ~1;
  c->lds_bytes = seq_meta + (c->lds_ring ? ring_bytes : red_bytes);
  per_cu = wfa::max_blocks_per_cu(kind, mode, c->lds_ring, c->i16, fmt == WFA_SEQ_PACKED2BITS,
                                  wfa::heur_kernel(p), c->nw, c->lds_bytes);
  if (tuned_per_cu > 0) per_cu = std::min(per_cu, tuned_per_cu);
  int grid = e->num_cus * per_cu;
  if (long_reads && e->long_grid_cap > 0) grid = std::min(grid, e->long_grid_cap);
  grid = std::max(1, std::min(grid, npairs));
  c->grid = grid;
  int64_t per_slot = c->lds_ring ? 0 : (int64_t)c->ring_rows * c->wcap * 4;
  // corrected ends-free: one pair's dominance record (components x the
  // diagonals of [-(plen+scope+2), tlen+scope+2])
  c->dom_stride = wfa::corrected_endsfree(p)
                      ? wfa::kind_components(kind) * (2 * max_len + 2 * p.scope + 8)
                      : 0;
  per_slot += (int64_t)c->dom_stride * 4;
  if (cig) {
    // worst case: every score up to the slot bound, every row full width
    const int64_t nwf_bound = (int64_t)std::max(2 * max_len, 1000) * 64 + 16;
    if (pass == 0 && !long_reads) {
      c->hdr_cap = 2048;
      c->hist_cap = 128 * 1024;
    } else {
      // long reads: scores up to twice the longest sequence and 48 MiB of
      // history per slot first, x4 per later pass
      const int g = long_reads ? pass : pass - 1;
      c->hdr_cap = (int)std::min<int64_t>(nwf_bound, (int64_t)std::max(2048, 2 * max_len) << (2 * g));
      c->hist_cap = std::min<int64_t>((int64_t)c->hdr_cap * c->wcap,
                                      (long_reads ? (int64_t)12 << 20 : (int64_t)128 << 10) << (2 * g));
    }
    // medium/low memory keeps one flag byte per cell instead of an offset:
    // the same bytes hold four times the cells
    if (mode == 2) c->hist_cap *= 4;
    // The history is one pool of chunks shared by the launch's slots
    // (hist_cap elements per slot on average): a pair takes chunks as its
    // rows need them, so one pair can hold far more than the average. A
    // row (at most wcap cells) never spans chunks.
    c->chunk = (int)std::max<int64_t>(16 * (int64_t)c->wcap, 1 << 16);
    c->tab_cap = (int)std::min<int64_t>((int64_t)c->hdr_cap * c->wcap / c->chunk + 2, 1 << 20);
    per_slot += c->hist_cap * (mode == 2 ? 1 : 4) + (int64_t)c->hdr_cap * 16 +
                (int64_t)c->tab_cap * 4 + (mode == 2 ? 4 * (int64_t)max_len + 64 : 0);
    // Compat::C Z-drop ends of medium/low CIGARs: each score's trimmed
    // ranges, and a sign byte beside every flag byte of the pool
    if (zdrop_ranges(p, mode)) per_slot += (int64_t)c->hdr_cap * 48 + c->hist_cap;
  } else {
    c->hdr_cap = 0;
    c->hist_cap = 0;
    c->chunk = c->pool_chunks = c->tab_cap = 0;
  }
  // keep the scratch of one launch within the engine's budget (48 GiB of
  // the 288 GB of HBM unless a serving loop splits it between engines)
  const int64_t budget = e->scratch_budget;
  if (per_slot > 0 && (int64_t)c->grid * per_slot > budget)
    c->grid = (int)std::max<int64_t>(1, budget / per_slot);
  if (cig) {
    int64_t elems = (int64_t)c->grid * c->hist_cap;
    // what the budget leaves after the rings, headers and chunk tables
    // bytes per pool element (and its sign byte when kept)
    const int64_t cell = mode == 2 ? (zdrop_ranges(p, mode) ? 2 : 1) : 4;
    const int64_t rest = budget - (int64_t)c->grid * (per_slot - c->hist_cap * cell);
    // long reads, or the pairs that outgrew the first pass: the pool takes
    // all of it; never more than it (at least one chunk)
    if (pass > 0 || (long_reads && e->long_pool_whole_budget)) elems = std::max(elems, rest / cell);
    elems = std::min(elems, std::max<int64_t>(rest / cell, c->chunk));
    elems = std::min<int64_t>(elems, (int64_t)c->chunk * INT32_MAX);
    // whole chunks, never more than the budget leaves (at least one per slot)
    const int64_t fit = std::max<int64_t>(rest / cell / c->chunk, 0);
    c->pool_chunks = (int)std::max<int64_t>(
        c->grid, std::min<int64_t>((elems + c->chunk - 1) / c->chunk, fit));
  }
  // op chain of one pair: at most every base consumed by a gap op plus one
  // gap-closing mismatch per gap
  c->chain_cap = 4 * max_len + 64;
  return WFA_OK;
}

int run_launch(wfa_engine_t* e, KParams p, const LaunchCfg& c, int num_items,
               const int32_t* list, int32_t* retry_out, hipStream_t st) {
  p.num_pairs = num_items;
  p.pair_list = list;
  p.counters = e->d_counters;
  p.retry_list = retry_out;
  p.wcap = c.wcap;
  p.seq_words = c.seq_words;
  p.nrows_meta = c.ring_rows;
  p.ring_rows = c.ring_rows;
  if (e->arena) {
    // no-allocation mode: this launch's buffers start again where the
    // call's lists end (the previous launch on the stream has finished
    // with them by the time this one runs)
    e->arena_top = e->arena_launch;
    e->gring_bytes = e->hist_bytes = e->hdr_bytes = e->tab_bytes = e->chain_bytes =
        e->dom_bytes = e->ztr_bytes = e->zsign_bytes = 0;
  }
  if (!c.lds_ring) {
    int r = egrow(e, (void**)&e->d_gring, &e->gring_bytes,
                  (size_t)c.grid * c.ring_rows * c.wcap * 4);
    if (r) return r;
    p.gring = e->d_gring;
  }
  if (c.shortk) {
    // the short-read kernel: per slot hdr_cap history rows of the window's
    // 16-bit cells and hdr_cap int32 headers
    if (c.cig) {
      int r = egrow(e, (void**)&e->d_hist, &e->hist_bytes,
                    (size_t)c.grid * c.hdr_cap * wfa::kShortWindow * 2);
      if (r) return r;
      r = egrow(e, (void**)&e->d_hdr, &e->hdr_bytes, (size_t)c.grid * c.hdr_cap * 4);
      if (r) return r;
      p.hist = e->d_hist;
      p.hdr = e->d_hdr;
      p.hdr_cap = c.hdr_cap;
    }
    p.sh_gsh = c.short_gsh;
    p.phase = e->d_phase;
    HIPCHK(wfa::launch_reset(e->d_counters, st));
    HIPCHK(wfa::launch_short(p, c.mode, c.short_rc, c.grid, c.lds_bytes, st));
    return WFA_OK;
  }
  if (c.cig) {
    const size_t cell = c.mode == 2 ? 1 : 4;
    int r = egrow(e, (void**)&e->d_hist, &e->hist_bytes, (size_t)c.pool_chunks * c.chunk * cell);
    if (r) return r;
    r = egrow(e, (void**)&e->d_hdr, &e->hdr_bytes, (size_t)c.grid * c.hdr_cap * sizeof(int4));
    if (r) return r;
    r = egrow(e, (void**)&e->d_tab, &e->tab_bytes, (size_t)c.grid * c.tab_cap * 4);
    if (r) return r;
    p.hist = e->d_hist;
    p.hist_cap = 0;
    p.chunk = c.chunk;
    p.pool_chunks = c.pool_chunks;
    p.chunk_tab = e->d_tab;
    p.tab_cap = c.tab_cap;
    p.hdr = e->d_hdr;
    p.hdr_cap = c.hdr_cap;
    if (c.mode == 2) {
      p.chain_cap = c.chain_cap;
      r = egrow(e, (void**)&e->d_chain, &e->chain_bytes, (size_t)c.grid * c.chain_cap);
      if (r) return r;
      p.chain = e->d_chain;
    }
  }
  p.ztr = nullptr;
  p.zsign = nullptr;
  if (c.cig && zdrop_ranges(p, c.mode)) {
    int r = egrow(e, (void**)&e->d_ztr, &e->ztr_bytes, (size_t)c.grid * c.hdr_cap * 48);
    if (r) return r;
    r = egrow(e, (void**)&e->d_zsign, &e->zsign_bytes, (size_t)c.pool_chunks * c.chunk);
    if (r) return r;
    p.ztr = e->d_ztr;
    p.zsign = e->d_zsign;
  }
  p.dom_stride = c.dom_stride;
  if (c.dom_stride > 0) {
    int r = egrow(e, (void**)&e->d_dom, &e->dom_bytes, (size_t)c.grid * c.dom_stride * 4);
    if (r) return r;
    p.dom = e->d_dom;
  }
  HIPCHK(wfa::launch_reset(e->d_counters, st));
  HIPCHK(wfa::launch_align(p, c.kind, c.mode, c.lds_ring, c.i16, c.nw, c.grid, c.lds_bytes,
                           st));
  return WFA_OK;
}

int ensure_lists(wfa_engine_t* e, int n) {
  if (e->arena) {
    // no-allocation mode: two lists of n entries from the arena
    size_t have = 0;
    for (auto& l : e->d_lists) {
      int r = egrow(e, (void**)&l, &have, (size_t)std::max(n, 1) * 4);
      if (r) return r;
    }
    e->list_cap = n;
    return WFA_OK;
  }
  if ((size_t)n <= e->list_cap) return WFA_OK;
  const size_t cap = std::max<size_t>(n, 1024);
  size_t have[2] = {e->list_cap * 4, e->list_cap * 4};
  for (int i = 0; i < 2; ++i) {
    int r = grow((void**)&e->d_lists[i], &have[i], cap * 4);
    if (r) {
      e->list_cap = 0;
      return r;
    }
  }
  e->list_cap = std::min(have[0], have[1]) / 4;
  return WFA_OK;
}


int validate_batch(const wfa_attr_t* attr, const wfa_batch_t* b) {
  if (!attr || !b) return WFA_ERR_INVALID_ARGUMENT;
  if (b->num_pairs < 0) return WFA_ERR_INVALID_ARGUMENT;
  if (b->seq_format != WFA_SEQ_ASCII && b->seq_format != WFA_SEQ_PACKED2BITS)
    return WFA_ERR_INVALID_ARGUMENT;
  if (b->num_pairs > 0 && (!b->seq_data || !b->pattern_offset || !b->pattern_length ||
                           !b->text_offset || !b->text_length || !b->score || !b->status))
    return WFA_ERR_INVALID_ARGUMENT;
  if (attr->alignment_scope == WFA_SCOPE_ALIGNMENT &&
      (!b->cigar_ops || !b->cigar_offset || !b->cigar_length) && b->num_pairs > 0)
    return WFA_ERR_INVALID_ARGUMENT;
  return WFA_OK;
}

int align_batch_run(wfa_engine_t* e, const wfa_attr_t* attr, const wfa_batch_t* b,
                    hipStream_t st, int first, int count);

// The passes after a first pass that left `nretry` pairs in d_lists[0]:
// ever wider launches over the pairs that asked for retry, then, after
// six, the out-of-memory status for what still does not fit.
int retry_passes(wfa_engine_t* e, KParams p, int kind, int mode, const wfa_batch_t* b,
                 int max_len, hipStream_t st, int32_t nretry) {
  int cur = 0;
  LaunchCfg c;
  for (int pass = 1; nretry > 0; ++pass) {
    if (pass > 6) {
      // what still does not fit the scratch budget gets the reference's
      // out-of-memory status (AlignStatus::Oom, src/aligner.rs:120-135)
      HIPCHK(wfa::launch_mark_oom(e->d_lists[cur], nretry, b->score, b->status, b->cigar_length,
                                  b->end_position, st));
      break;
    }
    int r = plan(e, p, kind, mode, b->seq_format, max_len, nretry, pass, &c);
    if (r) return r;
    r = run_launch(e, p, c, nretry, e->d_lists[cur], e->d_lists[cur ^ 1], st);
    if (r) return r;
    cur ^= 1;
    r = read_counter(e, 1, st, &nretry);
    if (r) return r;
  }
  return WFA_OK;
}

// One batch (see align_batch_run). A WFA_DEVICE_CHECKS build counts the
// out-of-range ring and history indexes of the call's launches and fails
// the call when there were any.
int align_batch_impl(wfa_engine_t* e, const wfa_attr_t* attr, const wfa_batch_t* b,
                     hipStream_t st, int first, int count) {
#ifdef WFA_PHASE_CLOCKS
  HIPCHK(hipSetDevice(e->device));
  if (!e->d_phase) HIPCHK(hipMalloc((void**)&e->d_phase, 5 * sizeof(unsigned long long)));
  HIPCHK(hipMemsetAsync(e->d_phase, 0, 5 * sizeof(unsigned long long), st));
#endif
#ifdef WFA_DEVICE_CHECKS
  HIPCHK(hipSetDevice(e->device));
  HIPCHK(hipMemsetAsync(e->d_counters + wfa::kCheckCounter, 0, 4, st));
  int r = align_batch_run(e, attr, b, st, first, count);
  if (r) return r;
  int32_t bad = 0;
  r = read_counter(e, wfa::kCheckCounter, st, &bad);
  if (r) return r;
  if (bad) {
    std::fprintf(stderr, "wfa_mi355x: %d ring/history index(es) out of range\n", bad);
    return WFA_ERR_DEVICE;
  }
  return WFA_OK;
#else
  return align_batch_run(e, attr, b, st, first, count);
#endif
}

// One batch: first pass over all pairs (or `first..first+count` when
// `per_pair`), then wider passes over the pairs that asked for retry.
int align_batch_run(wfa_engine_t* e, const wfa_attr_t* attr, const wfa_batch_t* b,
                    hipStream_t st, int first, int count) {
  Internal ip;
  int r = make_internal(attr->penalties, &ip);
  if (r) return r;
  r = check_supported(*attr);
  if (r) return r;
  if (count == 0) return WFA_OK;
  HIPCHK(hipSetDevice(e->device));
  if (attr->memory_mode == WFA_MEMORY_ULTRALOW) {
    if (b->pair_config) return WFA_ERR_UNSUPPORTED;
    refresh_budget(e, true);
    return biwfa_align_batch(e, attr, b, st, first, count);
  }
  KParams p;
  base_params(ip, *attr, *b, &p);
  const int kind = kind_of(ip.metric);
  const bool cig = attr->alignment_scope == WFA_SCOPE_ALIGNMENT;
  const int mode = !cig ? 0 : attr->memory_mode == WFA_MEMORY_HIGH ? 1 : 2;
  int max_len = b->max_length;
  if (max_len <= 0) {
    HIPCHK(wfa::launch_max_length(b->pattern_length + first, b->text_length + first, count,
                                  e->d_counters + 2, st));
    r = read_counter(e, 2, st, &max_len);
    if (r) return r;
  }
  if (e->max_len_limit > 0 && max_len > e->max_len_limit) return WFA_ERR_LIMIT;
  // no-allocation mode: the call's lists and work order follow the
  // counters; the launches' buffers follow them, within what is left
  if (e->arena) {
    e->arena_top = 256;
    e->order_bytes = 0;
  }
  r = ensure_lists(e, count);
  if (r) return r;
  // Long reads: hand the pairs out longest first, so the longest ones do
  // not start last and stretch the launch (a device radix sort of the
  // lengths; short reads are uniform enough to take in order)
  const int32_t* order = nullptr;
  if (max_len > 1024 && count > 1) {
    const size_t tb = wfa::lpt_temp_bytes(count);
    const size_t nb = ((size_t)count * 4 + 255) & ~(size_t)255;
    r = egrow(e, &e->d_order, &e->order_bytes, 4 * nb + tb);
    if (r) return r;
    uint8_t* ob = (uint8_t*)e->d_order;
    int32_t* ord = (int32_t*)(ob + 3 * nb);
    HIPCHK(wfa::launch_lpt_order(b->pattern_length, b->text_length, first, count,
                                 (uint32_t*)ob, (uint32_t*)(ob + nb), (int32_t*)(ob + 2 * nb),
                                 ord, ob + 4 * nb, tb, st));
    order = ord;
  }
  if (e->arena) {
    // what the launches may take: the rest of the arena, less room for the
    // 256-byte alignment of their buffers
    e->arena_launch = (e->arena_top + 255) & ~(size_t)255;
    e->scratch_budget = (int64_t)(e->arena_cap - e->arena_launch) - (64 << 10);
  } else {
    refresh_budget(e, max_len > 1024);
  }
  LaunchCfg c;
  r = plan(e, p, kind, mode, b->seq_format, max_len, count, 0, &c);
  if (r) return r;
  p.pair_base = first;
  // (no-allocation mode keeps no timing events: creating them allocates)
  const bool timed = e->ev[0] != nullptr;
  if (timed) HIPCHK(hipEventRecord(e->ev[0], st));
  r = run_launch(e, p, c, count, order, e->d_lists[0], st);
  if (r) return r;
  if (timed) HIPCHK(hipEventRecord(e->ev[1], st));
  int32_t nretry = 0;
  r = read_counter(e, 1, st, &nretry);
  if (r) return r;
  e->last_retries = nretry;
  r = retry_passes(e, p, kind, mode, b, max_len, st, nretry);
  if (r) return r;
  if (!timed) {
    HIPCHK(hipStreamSynchronize(st));
    return WFA_OK;
  }
  HIPCHK(hipEventRecord(e->ev[2], st));
  HIPCHK(hipEventSynchronize(e->ev[2]));
  HIPCHK(hipEventElapsedTime(&e->first_ms, e->ev[0], e->ev[1]));
  HIPCHK(hipEventElapsedTime(&e->all_ms, e->ev[0], e->ev[2]));
  return WFA_OK;
}

// 0 score only, 1 high-memory CIGAR, 2 medium/low-memory CIGAR
int mode_of(const wfa_attr_t& a) {
  if (a.alignment_scope != WFA_SCOPE_ALIGNMENT) return 0;
  return a.memory_mode == WFA_MEMORY_HIGH ? 1 : 2;
}

// The autotuner: time candidate launch shapes of a long-read launch on a
// strided sample of the batch (longest first, as the batch runs) and keep
// the fastest for the batch's (kind, mode, length bucket).
int autotune_impl(wfa_engine_t* e, const wfa_attr_t* attr, const wfa_batch_t* b, hipStream_t st,
                  int sample_pairs, wfa_tune_report_t* rep) {
  Internal ip;
  int r = make_internal(attr->penalties, &ip);
  if (r) return r;
  r = check_supported(*attr);
  if (r) return r;
  if (b->num_pairs == 0 || attr->memory_mode == WFA_MEMORY_ULTRALOW) return WFA_OK;
  HIPCHK(hipSetDevice(e->device));
  KParams p;
  base_params(ip, *attr, *b, &p);
  // heuristic launches run the single-wave kernels: nothing to tune
  if (wfa::heur_kernel(p)) return WFA_OK;
  const int kind = kind_of(ip.metric);
  const int mode = mode_of(*attr);
  const int n = b->num_pairs;
  int max_len = b->max_length;
  if (max_len <= 0) {
    HIPCHK(wfa::launch_max_length(b->pattern_length, b->text_length, n, e->d_counters + 2, st));
    r = read_counter(e, 2, st, &max_len);
    if (r) return r;
  }
  // short reads start on the narrow kernels: no shape to choose
  if (max_len <= 1024) return WFA_OK;
  // the sample: enough pairs that the widest candidate grid (4 workgroups
  // per CU) runs 16 pairs per slot, so a trial measures the shape's rate
  // more than its tail
  const int m = std::max(1, std::min(n, sample_pairs > 0 ? sample_pairs : e->num_cus * 64));
  std::vector<int32_t> pl(n), tl(n), idx(m);
  HIPCHK(hipMemcpy(pl.data(), b->pattern_length, (size_t)n * 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(tl.data(), b->text_length, (size_t)n * 4, hipMemcpyDeviceToHost));
  // the m longest pairs, longest first: a long-read batch's cost is
  // dominated by its longest pairs (cells grow with the square of the
  // score), and they are what every slot takes first
  std::vector<int32_t> all(n);
  for (int i = 0; i < n; ++i) all[i] = i;
  std::stable_sort(all.begin(), all.end(), [&](int a, int c) {
    return (int64_t)pl[a] + tl[a] > (int64_t)pl[c] + tl[c];
  });
  std::copy(all.begin(), all.begin() + m, idx.begin());
  r = ensure_lists(e, m);
  if (r) return r;
  r = egrow(e, &e->d_order, &e->order_bytes, (size_t)m * 4);
  if (r) return r;
  int32_t* sample = (int32_t*)e->d_order;
  HIPCHK(hipMemcpy(sample, idx.data(), (size_t)m * 4, hipMemcpyHostToDevice));
  p.pair_base = 0;
  refresh_budget(e, true);
  const int key = shape_key(kind, mode, max_len);
  const auto saved = e->shapes.find(key);
  const bool had = saved != e->shapes.end();
  const std::pair<int, int> old = had ? saved->second : std::make_pair(0, 0);
  // the fixed plan's shape
  e->shapes.erase(key);
  LaunchCfg c;
  r = plan(e, p, kind, mode, b->seq_format, max_len, m, 0, &c);
  if (r) return r;
  rep->default_waves_per_pair = c.nw;
  rep->default_groups_per_cu = std::max(1, c.grid / e->num_cus);
  rep->sample_pairs = m;
  if (c.lds_ring || c.nw == 1) {
    if (had) e->shapes[key] = old;
    return WFA_OK;
  }
  hipEvent_t ev[2] = {nullptr, nullptr};
  HIPCHK(hipEventCreate(&ev[0]));
  HIPCHK(hipEventCreate(&ev[1]));
  float best_ms = 0.f;
  std::pair<int, int> best(0, 0);
  bool warmed = false;
  // The candidates split the same resident waves differently: 4, 8 or 16
  // waves per pair, each at as many workgroups per CU as occupancy allows
  // (a shape with fewer leaves wave slots idle, and measured slower on
  // every long-read workload tried)
  for (int nw : {4, wfa::kWideWaves, 16}) {
    e->shapes[key] = std::make_pair(nw, 16);
    if (plan(e, p, kind, mode, b->seq_format, max_len, m, 0, &c) != WFA_OK || c.nw != nw)
      continue;
    const int g = std::max(1, c.grid / e->num_cus);
    float ms = 0.f;
    for (int rep_i = warmed ? 1 : 0; rep_i < 2; ++rep_i) {
      // the first pass over the sample and its retry passes, as a batch
      // runs: a shape whose slots exhaust the history pool pays for it
      r = hipEventRecord(ev[0], st) == hipSuccess ? WFA_OK : WFA_ERR_DEVICE;
      if (!r) r = run_launch(e, p, c, m, sample, e->d_lists[0], st);
      int32_t nretry = 0;
      if (!r) r = read_counter(e, 1, st, &nretry);
      if (!r) r = retry_passes(e, p, kind, mode, b, max_len, st, nretry);
      if (!r && hipEventRecord(ev[1], st) != hipSuccess) r = WFA_ERR_DEVICE;
      if (!r && hipEventSynchronize(ev[1]) != hipSuccess) r = WFA_ERR_DEVICE;
      if (!r && hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess) r = WFA_ERR_DEVICE;
      if (r) break;
    }
    if (r) break;
    warmed = true;
    wfa_tune_trial_t& t = rep->trials[rep->num_trials++];
    t.waves_per_pair = c.nw;
    t.groups_per_cu = g;
    t.ms = ms;
    if (best.first == 0 || ms < best_ms) {
      best_ms = ms;
      best = std::make_pair(c.nw, g);
    }
  }
  hipEventDestroy(ev[0]);
  hipEventDestroy(ev[1]);
  if (r || best.first == 0) {
    if (had) e->shapes[key] = old;
    else e->shapes.erase(key);
    return r;
  }
  e->shapes[key] = best;
  rep->waves_per_pair = best.first;
  rep->groups_per_cu = best.second;
  return WFA_OK;
}

}  // namespace

namespace wfahost {

int align_batch_locked(wfa_engine_t* e, const wfa_attr_t* attr, const wfa_batch_t* b,
                       hipStream_t st) {
  int r = validate_batch(attr, b);
  if (r) return r;
  std::lock_guard<std::mutex> g(e->mu);
  return align_batch_impl(e, attr, b, st, 0, b->num_pairs);
}

}  // namespace wfahost

extern "C" {

wfa_engine_t* wfa_engine_new(int32_t device, int32_t* err) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    if (err) *err = WFA_ERR_NO_DEVICE;
    return nullptr;
  }
  if (device < 0 || device >= n) {
    if (err) *err = WFA_ERR_INVALID_ARGUMENT;
    return nullptr;
  }
  auto* e = new wfa_engine_t();
  e->device = device;
  hipDeviceProp_t prop;
  if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess) {
    delete e;
    if (err) *err = WFA_ERR_DEVICE;
    return nullptr;
  }
  e->num_cus = prop.multiProcessorCount;
  e->lds_max = prop.sharedMemPerBlock;
  for (auto& ev : e->ev) {
    if (hipEventCreate(&ev) != hipSuccess) {
      delete e;
      if (err) *err = WFA_ERR_DEVICE;
      return nullptr;
    }
  }
  if (hipMalloc((void**)&e->d_counters, 64) != hipSuccess ||
      hipHostMalloc((void**)&e->h_pinned, 64, hipHostMallocDefault) != hipSuccess) {
    delete e;
    if (err) *err = WFA_ERR_OUT_OF_MEMORY;
    return nullptr;
  }
  if (err) *err = WFA_OK;
  return e;
}

void wfa_engine_free(wfa_engine_t* e) {
  if (!e) return;
  hipSetDevice(e->device);
  hipDeviceSynchronize();
  hipFree(e->d_counters);
  if (e->d_phase) hipFree(e->d_phase);
  if (e->own_arena) {
    hipFree(e->arena);
  } else if (!e->arena) {
    void* bufs[] = {e->d_lists[0], e->d_lists[1], e->d_hist, e->d_hdr, e->d_gring,
                    e->d_chain, e->d_tab, e->d_dom, e->d_ztr, e->d_zsign, e->d_order};
    for (void* b : bufs) hipFree(b);
  }
  for (void* b : e->bi_buf) hipFree(b);
  hipHostFree(e->h_pinned);
  for (auto ev : e->ev) hipEventDestroy(ev);
  delete e;
}

int32_t wfa_engine_align_batch(wfa_engine_t* e, const wfa_attr_t* attr, const wfa_batch_t* b,
                               void* stream) {
  if (!e) return WFA_ERR_INVALID_ARGUMENT;
  int r = validate_batch(attr, b);
  if (r) return r;
  std::lock_guard<std::mutex> g(e->mu);
  return align_batch_impl(e, attr, b, (hipStream_t)stream, 0, b->num_pairs);
}

int32_t wfa_engine_align_batch_per_pair(wfa_engine_t* e, const wfa_attr_t* attr,
                                        const wfa_batch_t* b, void* stream) {
  if (!e) return WFA_ERR_INVALID_ARGUMENT;
  int r = validate_batch(attr, b);
  if (r) return r;
  std::lock_guard<std::mutex> g(e->mu);
  wfa_batch_t one = *b;
  for (int i = 0; i < b->num_pairs; ++i) {
    r = align_batch_impl(e, attr, &one, (hipStream_t)stream, i, 1);
    if (r) return r;
  }
  return WFA_OK;
}

int32_t wfa_engine_autotune(wfa_engine_t* e, const wfa_attr_t* attr, const wfa_batch_t* b,
                            void* stream, int32_t sample_pairs, wfa_tune_report_t* report) {
  if (!e || e->arena) return WFA_ERR_INVALID_ARGUMENT;
  int r = validate_batch(attr, b);
  if (r) return r;
  wfa_tune_report_t local;
  wfa_tune_report_t* rep = report ? report : &local;
  std::memset(rep, 0, sizeof(*rep));
  std::lock_guard<std::mutex> g(e->mu);
  return autotune_impl(e, attr, b, (hipStream_t)stream, sample_pairs, rep);
}

int32_t wfa_engine_set_launch_shape(wfa_engine_t* e, const wfa_attr_t* attr, int32_t max_length,
                                    int32_t waves_per_pair, int32_t groups_per_cu) {
  if (!e || !attr || max_length <= 0) return WFA_ERR_INVALID_ARGUMENT;
  const bool clear = waves_per_pair == 0 && groups_per_cu == 0;
  if (!clear && ((waves_per_pair != 4 && waves_per_pair != wfa::kWideWaves &&
                  waves_per_pair != 16) ||
                 groups_per_cu < 1 || groups_per_cu > 16))
    return WFA_ERR_INVALID_ARGUMENT;
  Internal ip;
  int r = make_internal(attr->penalties, &ip);
  if (r) return r;
  std::lock_guard<std::mutex> g(e->mu);
  const int key = shape_key(kind_of(ip.metric), mode_of(*attr), max_length);
  if (clear) e->shapes.erase(key);
  else e->shapes[key] = std::make_pair(waves_per_pair, groups_per_cu);
  return WFA_OK;
}

int32_t wfa_engine_get_launch_shape(wfa_engine_t* e, const wfa_attr_t* attr, int32_t max_length,
                                    int32_t* waves_per_pair, int32_t* groups_per_cu) {
  if (!e || !attr || max_length <= 0 || !waves_per_pair || !groups_per_cu)
    return WFA_ERR_INVALID_ARGUMENT;
  Internal ip;
  int r = make_internal(attr->penalties, &ip);
  if (r) return r;
  std::lock_guard<std::mutex> g(e->mu);
  const auto it = e->shapes.find(shape_key(kind_of(ip.metric), mode_of(*attr), max_length));
  *waves_per_pair = it == e->shapes.end() ? 0 : it->second.first;
  *groups_per_cu = it == e->shapes.end() ? 0 : it->second.second;
  return WFA_OK;
}

int32_t wfa_engine_last_phase_cycles(wfa_engine_t* e, uint64_t* out) {
  if (!e || !out) return WFA_ERR_INVALID_ARGUMENT;
#ifdef WFA_PHASE_CLOCKS
  std::lock_guard<std::mutex> g(e->mu);
  unsigned long long v[5] = {0, 0, 0, 0, 0};
  if (e->d_phase) {
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipMemcpy(v, e->d_phase, sizeof(v), hipMemcpyDeviceToHost));
  }
  for (int i = 0; i < 5; ++i) out[i] = v[i];
  return WFA_OK;
#else
  (void)out;
  return WFA_ERR_UNSUPPORTED;
#endif
}

int32_t wfa_engine_set_scratch_budget(wfa_engine_t* e, int64_t bytes) {
  if (!e || bytes < ((int64_t)64 << 20)) return WFA_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> g(e->mu);
  e->scratch_budget = bytes;
  e->auto_budget = false;
  return WFA_OK;
}

int32_t wfa_engine_last_retries(const wfa_engine_t* e) { return e ? e->last_retries : 0; }

int32_t wfa_engine_last_kernel_ms(const wfa_engine_t* e, float* first_ms, float* all_ms) {
  if (!e) return WFA_ERR_INVALID_ARGUMENT;
  if (first_ms) *first_ms = e->first_ms;
  if (all_ms) *all_ms = e->all_ms;
  return WFA_OK;
}

const char* wfa_version(void) { return "wfa-mi355x 0.1.0 (gfx950)"; }

void wfa_attr_default(wfa_attr_t* a, int32_t metric) {
  std::memset(a, 0, sizeof(*a));
  a->penalties.distance_metric = metric;
  switch (metric) {
    case WFA_METRIC_GAP_LINEAR:
      a->penalties.match = 0; a->penalties.mismatch = 4; a->penalties.gap_opening1 = 2;
      break;
    case WFA_METRIC_GAP_AFFINE:
      a->penalties.match = 0; a->penalties.mismatch = 4;
      a->penalties.gap_opening1 = 6; a->penalties.gap_extension1 = 2;
      break;
    case WFA_METRIC_GAP_AFFINE_2P:
      a->penalties.match = 0; a->penalties.mismatch = 4;
      a->penalties.gap_opening1 = 6; a->penalties.gap_extension1 = 2;
      a->penalties.gap_opening2 = 24; a->penalties.gap_extension2 = 1;
      break;
    default:
      break;
  }
  a->alignment_scope = WFA_SCOPE_SCORE;
  a->memory_mode = WFA_MEMORY_HIGH;
  a->span = WFA_SPAN_END2END;
  a->heuristic = WFA_HEURISTIC_NONE;
  a->max_alignment_steps = INT32_MAX;
}

int32_t wfa_penalties_check(const wfa_penalties_t* p) {
  if (!p) return WFA_ERR_INVALID_ARGUMENT;
  Internal ip;
  return make_internal(*p, &ip);
}

}  // extern "C"

// =====================================================================
// Aligner: WavefrontAligner<N> over the engine, one pair per call
// =====================================================================
struct wfa_aligner_t {
  wfa_attr_t attr;
  wfa_engine_t* engine = nullptr;
  int err = WFA_OK;
  int status = WFA_STATUS_OK;
  int32_t score = 0;
  std::vector<uint8_t> cigar;
  int32_t cigar_score = INT32_MIN;
  int32_t end_pos[3] = {-1, INT32_MAX, INT32_MIN / 2};
  // set_max_memory_compact's value (stored only: no buffer here compacts)
  uint64_t max_memory_compact = uint64_t(256) << 20;
  // device staging for one pair: [pattern | text | cigar], plus arrays
  uint8_t* d_buf = nullptr;
  size_t d_cap = 0;
  int64_t* d_i64 = nullptr;  // poff, toff, coff
  int32_t* d_i32 = nullptr;  // plen, tlen, score, status, clen, end position (3)
  uint8_t* h_buf = nullptr;  // pinned
  size_t h_cap = 0;
};

namespace {

// cigar_wfa_score (src/aligner.rs:2700-2719) of ops[0, n): the classic
// score (Cigar::score_*, src/cigar.rs:118-202) on the user penalties, in
// the WFA scale over the bases the ops span.
int32_t wfa_score_of_ops(const wfa_penalties_t& u, const uint8_t* ops, int n) {
  int v = 0, h = 0, classic = 0, edits = 0;
  uint8_t last = 0;
  for (int i = 0; i < n; ++i) {
    const uint8_t op = ops[i];
    v += op != 'I';
    h += op != 'D';
    edits += op != 'M';
    if (op == 'M') classic -= u.match;
    else if (op == 'X') classic -= u.mismatch;
    else if (u.distance_metric == WFA_METRIC_GAP_LINEAR) classic -= u.gap_opening1;
    else if (u.distance_metric == WFA_METRIC_GAP_AFFINE)
      classic -= u.gap_extension1 + (last == op ? 0 : u.gap_opening1);
    last = op;
  }
  if (u.distance_metric == WFA_METRIC_EDIT || u.distance_metric == WFA_METRIC_INDEL) return edits;
  if (u.distance_metric == WFA_METRIC_GAP_AFFINE_2P) {
    // runs, each gap run priced by the cheaper piece
    classic = 0;
    for (int i = 0; i < n;) {
      int j = i;
      while (j < n && ops[j] == ops[i]) ++j;
      const int len = j - i;
      if (ops[i] == 'M') classic -= u.match * len;
      else if (ops[i] == 'X') classic -= u.mismatch * len;
      else
        classic -= std::min(u.gap_opening1 + u.gap_extension1 * len,
                            u.gap_opening2 + u.gap_extension2 * len);
      i = j;
    }
  }
  return u.match < 0 ? -u.match * (v + h) - 2 * classic : -classic;
}

int aligner_run(wfa_aligner_t* a, const uint8_t* P, int plen, const uint8_t* T, int tlen,
                int fmt) {
  a->cigar.clear();
  a->cigar_score = INT32_MIN;
  a->status = WFA_STATUS_OK;
  a->end_pos[0] = -1;
  a->end_pos[1] = INT32_MAX;
  a->end_pos[2] = INT32_MIN / 2;
  if (plen < 0 || tlen < 0 || (plen && !P) || (tlen && !T)) return WFA_ERR_INVALID_ARGUMENT;
  int r = check_supported(a->attr);
  if (r) return r;
  const size_t pb = fmt == WFA_SEQ_PACKED2BITS ? (size_t)(plen + 3) / 4 : (size_t)plen;
  const size_t tb = fmt == WFA_SEQ_PACKED2BITS ? (size_t)(tlen + 3) / 4 : (size_t)tlen;
  const bool cig = a->attr.alignment_scope == WFA_SCOPE_ALIGNMENT;
  const size_t need = pb + tb + (cig ? (size_t)plen + tlen : 0) + 64;
  if (hipSetDevice(a->engine->device) != hipSuccess) return WFA_ERR_DEVICE;
  if (need > a->d_cap) {
    if (a->d_buf) hipFree(a->d_buf);
    a->d_buf = nullptr;
    a->d_cap = 0;
    if (hipMalloc((void**)&a->d_buf, need) != hipSuccess) return WFA_ERR_OUT_OF_MEMORY;
    a->d_cap = need;
  }
  if (need > a->h_cap) {
    if (a->h_buf) hipHostFree(a->h_buf);
    a->h_buf = nullptr;
    a->h_cap = 0;
    if (hipHostMalloc((void**)&a->h_buf, need, hipHostMallocDefault) != hipSuccess)
      return WFA_ERR_OUT_OF_MEMORY;
    a->h_cap = need;
  }
  if (!a->d_i64) {
    if (hipMalloc((void**)&a->d_i64, 64) != hipSuccess) return WFA_ERR_OUT_OF_MEMORY;
    if (hipMalloc((void**)&a->d_i32, 64) != hipSuccess) return WFA_ERR_OUT_OF_MEMORY;
  }
  if (pb) std::memcpy(a->h_buf, P, pb);
  if (tb) std::memcpy(a->h_buf + pb, T, tb);
  int64_t offs[3] = {0, (int64_t)pb, (int64_t)(pb + tb)};
  int32_t lens[2] = {plen, tlen};
  hipStream_t st = nullptr;
  if (pb + tb && hipMemcpyAsync(a->d_buf, a->h_buf, pb + tb, hipMemcpyHostToDevice, st) != hipSuccess)
    return WFA_ERR_DEVICE;
  if (hipMemcpyAsync(a->d_i64, offs, sizeof(offs), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(a->d_i32, lens, sizeof(lens), hipMemcpyHostToDevice, st) != hipSuccess)
    return WFA_ERR_DEVICE;
  wfa_batch_t b;
  std::memset(&b, 0, sizeof(b));
  b.num_pairs = 1;
  b.seq_format = fmt;
  b.max_length = std::max(plen, tlen);
  b.seq_data = a->d_buf;
  b.pattern_offset = a->d_i64;
  b.text_offset = a->d_i64 + 1;
  b.pattern_length = a->d_i32;
  b.text_length = a->d_i32 + 1;
  b.score = a->d_i32 + 2;
  b.status = a->d_i32 + 3;
  if (cig) {
    b.cigar_ops = a->d_buf;
    b.cigar_offset = a->d_i64 + 2;
    b.cigar_length = a->d_i32 + 4;
  }
  const bool biwfa = a->attr.memory_mode == WFA_MEMORY_ULTRALOW;
  if (!biwfa) b.end_position = a->d_i32 + 5;
  {
    std::lock_guard<std::mutex> g(a->engine->mu);
    r = align_batch_impl(a->engine, &a->attr, &b, st, 0, 1);
  }
  if (r) return r;
  int32_t res[6];
  if (hipMemcpy(res, a->d_i32 + 2, sizeof(res), hipMemcpyDeviceToHost) != hipSuccess)
    return WFA_ERR_DEVICE;
  a->score = res[0];
  a->status = res[1];
  if (!biwfa) std::memcpy(a->end_pos, res + 3, sizeof(a->end_pos));
  if (cig && res[2] > 0) {
    a->cigar.resize(res[2]);
    if (hipMemcpy(a->cigar.data(), a->d_buf + pb + tb, res[2], hipMemcpyDeviceToHost) !=
        hipSuccess)
      return WFA_ERR_DEVICE;
    a->cigar_score = a->score;
    // a max-trimmed CIGAR (extension, or a Z-drop end) scores its kept
    // part, a free leading gap excluded (maxtrim_cigar, src/aligner.rs:1545-1583)
    const int m = a->attr.penalties.distance_metric;
    const bool ef = a->attr.span == WFA_SPAN_ENDSFREE;
    const bool ext = ef && a->attr.extension;
    const bool zdrop_end =
        a->attr.heuristic == WFA_HEURISTIC_ZDROP && a->status == WFA_STATUS_PARTIAL;
    if ((ext || zdrop_end) && m != WFA_METRIC_EDIT && m != WFA_METRIC_INDEL) {
      int b = 0, bh = 0;
      const int n = (int)a->cigar.size();
      if (ef && !ext) {
        while (bh < a->attr.text_begin_free && b < n && a->cigar[b] == 'I') { ++b; ++bh; }
        for (int bv = 0; bh == 0 && bv < a->attr.pattern_begin_free && b < n &&
                         a->cigar[b] == 'D';
             ++bv)
          ++b;
      }
      a->cigar_score = wfa_score_of_ops(a->attr.penalties, a->cigar.data() + b, n - b);
    }
  }
  return WFA_OK;
}

}  // namespace

extern "C" {

wfa_aligner_t* wfa_aligner_new(const wfa_attr_t* attr, int32_t device, int32_t* err) {
  if (!attr) {
    if (err) *err = WFA_ERR_INVALID_ARGUMENT;
    return nullptr;
  }
  Internal ip;
  int r = make_internal(attr->penalties, &ip);
  if (r) {
    if (err) *err = r;
    return nullptr;
  }
  int32_t e = 0;
  wfa_engine_t* eng = wfa_engine_new(device, &e);
  if (!eng) {
    if (err) *err = e;
    return nullptr;
  }
  auto* a = new wfa_aligner_t();
  a->attr = *attr;
  a->engine = eng;
  if (err) *err = WFA_OK;
  return a;
}

void wfa_aligner_free(wfa_aligner_t* a) {
  if (!a) return;
  hipSetDevice(a->engine->device);
  hipFree(a->d_buf);
  hipFree(a->d_i64);
  hipFree(a->d_i32);
  hipHostFree(a->h_buf);
  wfa_engine_free(a->engine);
  delete a;
}

int32_t wfa_aligner_set_alignment_scope(wfa_aligner_t* a, int32_t scope) {
  if (!a || (scope != WFA_SCOPE_SCORE && scope != WFA_SCOPE_ALIGNMENT)) return WFA_ERR_INVALID_ARGUMENT;
  a->attr.alignment_scope = scope;
  return WFA_OK;
}
int32_t wfa_aligner_set_memory_mode(wfa_aligner_t* a, int32_t mode) {
  if (!a || mode < WFA_MEMORY_HIGH || mode > WFA_MEMORY_ULTRALOW) return WFA_ERR_INVALID_ARGUMENT;
  a->attr.memory_mode = mode;
  return WFA_OK;
}
int32_t wfa_aligner_set_compat(wfa_aligner_t* a, int32_t compat) {
  if (!a || (compat != WFA_COMPAT_C && compat != WFA_COMPAT_CORRECTED))
    return WFA_ERR_INVALID_ARGUMENT;
  a->attr.compat = compat;
  return WFA_OK;
}
int32_t wfa_aligner_set_heuristic(wfa_aligner_t* a, int32_t h, int32_t p0, int32_t p1,
                                  int32_t p2) {
  if (!a || h < WFA_HEURISTIC_NONE || h > WFA_HEURISTIC_BANDED_ADAPTIVE) return WFA_ERR_INVALID_ARGUMENT;
  a->attr.heuristic = h;
  a->attr.heuristic_p[0] = p0;
  a->attr.heuristic_p[1] = p1;
  a->attr.heuristic_p[2] = p2;
  return WFA_OK;
}
int32_t wfa_aligner_set_alignment_free_ends(wfa_aligner_t* a, int32_t pb, int32_t pe, int32_t tb,
                                            int32_t te) {
  if (!a) return WFA_ERR_INVALID_ARGUMENT;
  a->attr.span = WFA_SPAN_ENDSFREE;
  a->attr.extension = 0;
  a->attr.pattern_begin_free = pb;
  a->attr.pattern_end_free = pe;
  a->attr.text_begin_free = tb;
  a->attr.text_end_free = te;
  return WFA_OK;
}
int32_t wfa_aligner_set_alignment_extension(wfa_aligner_t* a) {
  if (!a) return WFA_ERR_INVALID_ARGUMENT;
  a->attr.span = WFA_SPAN_ENDSFREE;
  a->attr.extension = 1;
  return WFA_OK;
}
int32_t wfa_aligner_set_alignment_end_to_end(wfa_aligner_t* a) {
  if (!a) return WFA_ERR_INVALID_ARGUMENT;
  a->attr.span = WFA_SPAN_END2END;
  a->attr.extension = 0;
  a->attr.pattern_begin_free = a->attr.pattern_end_free = 0;
  a->attr.text_begin_free = a->attr.text_end_free = 0;
  return WFA_OK;
}
int32_t wfa_aligner_set_max_alignment_steps(wfa_aligner_t* a, int32_t max_steps) {
  if (!a) return WFA_ERR_INVALID_ARGUMENT;
  a->attr.max_alignment_steps = max_steps;
  return WFA_OK;
}

int32_t wfa_aligner_set_max_memory_compact(wfa_aligner_t* a, uint64_t bytes) {
  if (!a) return WFA_ERR_INVALID_ARGUMENT;
  a->max_memory_compact = bytes;
  return WFA_OK;
}

uint64_t wfa_aligner_max_memory_compact(const wfa_aligner_t* a) {
  return a ? a->max_memory_compact : 0;
}

int32_t wfa_align_end2end(wfa_aligner_t* a, const uint8_t* P, int32_t plen, const uint8_t* T,
                          int32_t tlen) {
  if (!a) return INT32_MIN;
  // align_end2end is end-to-end whatever the span (src/aligner.rs:627-636)
  wfa_attr_t saved = a->attr;
  a->attr.span = WFA_SPAN_END2END;
  a->attr.extension = 0;
  a->err = aligner_run(a, P, plen, T, tlen, WFA_SEQ_ASCII);
  a->attr = saved;
  return a->err ? INT32_MIN : a->score;
}

int32_t wfa_align_lambda_classes(wfa_aligner_t* a, const uint8_t* P, int32_t plen,
                                 const uint8_t* T, int32_t tlen, const uint8_t* pattern_class,
                                 const uint8_t* text_class) {
  if (!a) return INT32_MIN;
  if (!pattern_class || !text_class || plen < 0 || tlen < 0 || (plen > 0 && !P) ||
      (tlen > 0 && !T)) {
    a->err = WFA_ERR_INVALID_ARGUMENT;
    return INT32_MIN;
  }
  // pattern[i] matches text[j] iff their classes are equal: align the class
  // sequences, whose bytes the extension compares
  std::vector<uint8_t> pc((size_t)plen), tc((size_t)tlen);
  for (int32_t i = 0; i < plen; ++i) pc[i] = pattern_class[P[i]];
  for (int32_t j = 0; j < tlen; ++j) tc[j] = text_class[T[j]];
  return wfa_align_end2end(a, pc.data(), plen, tc.data(), tlen);
}

int32_t wfa_align_endsfree(wfa_aligner_t* a, const uint8_t* P, int32_t plen, const uint8_t* T,
                           int32_t tlen) {
  if (!a) return INT32_MIN;
  // align_endsfree runs the ends-free loop whatever the span; an aligner
  // never given free ends has them all zero (src/aligner.rs:707-718)
  wfa_attr_t saved = a->attr;
  if (a->attr.span != WFA_SPAN_ENDSFREE) {
    a->attr.span = WFA_SPAN_ENDSFREE;
    a->attr.extension = 0;
    a->attr.pattern_begin_free = a->attr.pattern_end_free = 0;
    a->attr.text_begin_free = a->attr.text_end_free = 0;
  }
  a->err = aligner_run(a, P, plen, T, tlen, WFA_SEQ_ASCII);
  a->attr = saved;
  return a->err ? INT32_MIN : a->score;
}

// BiWFA through the batch path: ultralow memory, end-to-end
static int32_t biwfa_call(wfa_aligner_t* a, const uint8_t* P, int32_t plen, const uint8_t* T,
                          int32_t tlen, int32_t scope) {
  if (!a) return INT32_MIN;
  wfa_attr_t saved = a->attr;
  a->attr.memory_mode = WFA_MEMORY_ULTRALOW;
  a->attr.alignment_scope = scope;
  a->attr.span = WFA_SPAN_END2END;
  a->attr.extension = 0;
  a->err = aligner_run(a, P, plen, T, tlen, WFA_SEQ_ASCII);
  a->attr = saved;
  return a->err ? INT32_MIN : a->score;
}

int32_t wfa_align_biwfa(wfa_aligner_t* a, const uint8_t* P, int32_t plen, const uint8_t* T,
                        int32_t tlen) {
  return biwfa_call(a, P, plen, T, tlen, WFA_SCOPE_ALIGNMENT);
}

int32_t wfa_align_biwfa_score(wfa_aligner_t* a, const uint8_t* P, int32_t plen,
                              const uint8_t* T, int32_t tlen) {
  return biwfa_call(a, P, plen, T, tlen, WFA_SCOPE_SCORE);
}

int32_t wfa_align_packed2bits(wfa_aligner_t* a, const uint8_t* P, int32_t plen, const uint8_t* T,
                              int32_t tlen) {
  if (!a) return INT32_MIN;
  wfa_attr_t saved = a->attr;
  a->attr.span = WFA_SPAN_END2END;
  a->attr.extension = 0;
  a->err = aligner_run(a, P, plen, T, tlen, WFA_SEQ_PACKED2BITS);
  a->attr = saved;
  return a->err ? INT32_MIN : a->score;
}

int32_t wfa_aligner_status(const wfa_aligner_t* a) { return a ? a->status : WFA_STATUS_OK; }
int32_t wfa_aligner_cigar(const wfa_aligner_t* a, const uint8_t** ops) {
  if (!a || a->cigar.empty()) {
    if (ops) *ops = nullptr;
    return 0;
  }
  if (ops) *ops = a->cigar.data();
  return (int32_t)a->cigar.size();
}
int32_t wfa_aligner_cigar_score(const wfa_aligner_t* a) { return a ? a->cigar_score : INT32_MIN; }
void wfa_aligner_end_position(const wfa_aligner_t* a, int32_t* out) {
  if (!a || !out) return;
  std::memcpy(out, a->end_pos, sizeof(a->end_pos));
}
int32_t wfa_aligner_error(const wfa_aligner_t* a) { return a ? a->err : WFA_ERR_INVALID_ARGUMENT; }

}  // extern "C"
