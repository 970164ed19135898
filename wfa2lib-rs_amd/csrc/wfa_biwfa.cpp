// wfa_biwfa.cpp — host driver of BiWFA batches (memory mode ULTRALOW).
//
// align_biwfa (src/aligner.rs:2563-2642) recurses depth-first; here the
// recursion runs breadth-first over every pair of the batch at once: each
// level launches one breakpoint search per pending sub-problem, the results
// split them into the next level, and all base cases run in one launch at
// the end. Every sub-problem carries an interval of a 64-bit key space
// (children split their parent's interval), so the depth-first order of the
// CIGAR pieces, and of failures, is the order of the keys.
//
// align_biwfa_score (src/aligner.rs:2650-2690) is one breakpoint search per
// pair.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "wfa_engine.h"

namespace wfahost {
namespace {

using wfa::BiJob;
using wfa::BiParams;
using wfa::BiPiece;
using wfa::BiResult;

enum {
  B_JOBS, B_RESULTS, B_RING, B_HDR, B_HIST, B_FRAG, B_FRAG_OFF, B_FRAG_LEN, B_FRAG_STATUS,
  B_FRAG_SCORE, B_PIECES, B_PAIRS
};

// BIALIGN_* thresholds (src/bialign.rs:511-520)
constexpr int kFallbackMinScore = 250;
constexpr int kFallbackMinLength = 100;
constexpr int kRecoveryMinScore = 500;
constexpr size_t kRingBudget = size_t(16) << 30;   // bytes of rings per launch
constexpr size_t kHistBudget = size_t(8) << 30;    // bytes of base-case history per launch

template <class T>
int dev(wfa_engine_t* e, int idx, size_t count, T** out) {
  int r = grow(&e->bi_buf[idx], &e->bi_cap[idx], std::max<size_t>(count, 1) * sizeof(T));
  *out = (T*)e->bi_buf[idx];
  return r;
}

struct Node {
  BiJob job;
  int score_remaining;
  uint64_t key_lo, key_hi;  // depth-first order interval
};

// A piece of a pair's CIGAR in depth-first order
struct Leaf {
  uint64_t key;
  int frag;      // base-case fragment index, or -1 for a gap run
  int op, len;   // gap run
};

struct PairState {
  bool failed = false;
  int fail_status = 0;
  uint64_t fail_key = ~0ull;
  int fail_frag = -1;  // a failing base case's fragment stays in the CIGAR
  int score = 0;
  std::vector<Leaf> leaves;
};

int ceil_pow2(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

int nc_of(int kind) { return wfa::kind_components(kind); }

// Run one kind of BiWFA launch over `jobs`, re-running jobs that asked for
// more room (longer LDS sequences, a wider ring, more history) with larger
// configurations. `which`: 0 breakpoint search, 1 base case.
int run_jobs(wfa_engine_t* e, const KParams& kp, int kind, bool heur, int which,
             std::vector<BiJob>& jobs, std::vector<BiResult>* results,
             std::vector<int64_t>* frag_off, uint8_t* d_frag, std::vector<int32_t>* frag_len,
             std::vector<int32_t>* frag_status, std::vector<int32_t>* frag_score,
             hipStream_t st) {
  const int n = (int)jobs.size();
  if (n == 0) return WFA_OK;
  const int NC = nc_of(kind);
  const int dirs = which == 0 ? 2 : 1;
  // Largest sub-problems first (the launch's workgroups take them from one
  // queue in this order), so the longest searches do not start last and
  // stretch the launch
  std::vector<int> todo(n);
  for (int i = 0; i < n; ++i) todo[i] = i;
  auto size_of = [&](int i) { return (int64_t)(jobs[i].pe - jobs[i].pb) + (jobs[i].te - jobs[i].tb); };
  std::stable_sort(todo.begin(), todo.end(), [&](int a, int b) { return size_of(a) > size_of(b); });
  if (results) results->assign(n, BiResult{});
  if (frag_len) {
    frag_len->assign(n, 0);
    frag_status->assign(n, wfa::BI_RETRY);
    frag_score->assign(n, 0);
  }
  for (int pass = 0; !todo.empty(); ++pass) {
    if (pass > 3) return WFA_ERR_OUT_OF_MEMORY;
    const int m = (int)todo.size();
    int max_len = 0, max_sum = 0;
    for (int i : todo) {
      const BiJob& j = jobs[i];
      max_len = std::max({max_len, j.pe - j.pb, j.te - j.tb});
      max_sum = std::max(max_sum, (j.pe - j.pb) + (j.te - j.tb));
    }
    BiParams bp;
    std::memset(&bp, 0, sizeof(bp));
    bp.k = kp;
    // sequences: 2-bit first; later passes make room for bytes (a
    // sub-problem with other symbols than A/C/G/T) where LDS allows
    const size_t meta_bytes = (size_t)dirs * NC * kp.scope * 16;
    auto words_for = [&](int per) { return ((max_len + per - 1) / per + 2 + 1) & ~1; };
    bp.seq_words = words_for(32);
    if (pass > 0 && (size_t)2 * words_for(8) * 8 + meta_bytes <= e->lds_max)
      bp.seq_words = words_for(8);
    // breakpoint launches: 16 ints per wave for cross-wave reductions
    const size_t lds = (size_t)2 * bp.seq_words * 8 + meta_bytes +
                       (which == 0 ? (size_t)64 * wfa::kWideWaves : 0);
    if (lds > e->lds_max) return WFA_ERR_UNSUPPORTED;
    int grid = e->num_cus * std::max(1, wfa::biwfa_blocks_per_cu(kind, which, heur, lds));
    grid = std::min(grid, m);
    const size_t ring_budget = std::min<size_t>(kRingBudget, (size_t)e->scratch_budget / 2);
    // Rows: no row is wider than the sub-problem's diagonal range. A first
    // breakpoint pass takes the widest power of two whose rings fit the
    // budget for the whole grid (at least 16384 diagonals): a sub-problem
    // whose range outgrows its rows starts again at full width, after the
    // work it did. Base cases start at 4096. Later passes take the full
    // width.
    const int full = ceil_pow2(max_sum + 2 * kp.scope + 4);
    int first_cap = 4096;
    if (which == 0) {
      const size_t per_diag = (size_t)dirs * NC * kp.scope * 4 * grid;
      first_cap = 16384;
      while (first_cap < full && (size_t)first_cap * 2 * per_diag <= ring_budget) first_cap *= 2;
    }
    bp.wcap = pass == 0 ? std::min(full, first_cap) : full;
    const size_t ring_slot = (size_t)dirs * NC * kp.scope * bp.wcap * 4;
    const size_t hist_budget = std::min<size_t>(kHistBudget, (size_t)e->scratch_budget / 4);
    grid = (int)std::max<size_t>(1, std::min<size_t>(grid, ring_budget / ring_slot));
    if (which == 1) {
      bp.hdr_cap = pass == 0 ? 1024 : 1024 << (2 * pass);
      bp.hist_cap = pass == 0 ? (int64_t(1) << 20) : (int64_t(1) << (20 + 2 * pass));
      const size_t per_slot = (size_t)bp.hdr_cap * NC * 16 + (size_t)bp.hist_cap * 4;
      grid = (int)std::max<size_t>(1, std::min<size_t>(grid, hist_budget / per_slot));
    }
    BiJob* d_jobs;
    int r = dev(e, B_JOBS, m, &d_jobs);
    if (r) return r;
    std::vector<BiJob> sub(m);
    for (int i = 0; i < m; ++i) sub[i] = jobs[todo[i]];
    HIPCHK(hipMemcpyAsync(d_jobs, sub.data(), m * sizeof(BiJob), hipMemcpyHostToDevice, st));
    int32_t* d_ring;
    r = dev(e, B_RING, ring_slot / 4 * grid, &d_ring);
    if (r) return r;
    bp.jobs = d_jobs;
    bp.njobs = m;
    bp.counters = e->d_counters;
    bp.ring = d_ring;
    BiResult* d_res = nullptr;
    int64_t* d_foff = nullptr;
    int32_t *d_flen = nullptr, *d_fst = nullptr, *d_fsc = nullptr;
    if (which == 0) {
      r = dev(e, B_RESULTS, m, &d_res);
      if (r) return r;
      bp.results = d_res;
    } else {
      int4* d_hdr;
      int32_t* d_hist;
      r = dev(e, B_HDR, (size_t)bp.hdr_cap * NC * grid, &d_hdr);
      if (!r) r = dev(e, B_HIST, (size_t)bp.hist_cap * grid, &d_hist);
      if (!r) r = dev(e, B_FRAG_OFF, m, &d_foff);
      if (!r) r = dev(e, B_FRAG_LEN, m, &d_flen);
      if (!r) r = dev(e, B_FRAG_STATUS, m, &d_fst);
      if (!r) r = dev(e, B_FRAG_SCORE, m, &d_fsc);
      if (r) return r;
      std::vector<int64_t> offs(m);
      for (int i = 0; i < m; ++i) offs[i] = (*frag_off)[todo[i]];
      HIPCHK(hipMemcpyAsync(d_foff, offs.data(), m * 8, hipMemcpyHostToDevice, st));
      bp.hdr = d_hdr;
      bp.hist = d_hist;
      bp.frag = d_frag;
      bp.frag_off = d_foff;
      bp.frag_len = d_flen;
      bp.frag_status = d_fst;
      bp.frag_score = d_fsc;
    }
    HIPCHK(wfa::launch_reset(e->d_counters, st));
    HIPCHK(wfa::launch_biwfa(kind, which, heur, bp, grid, lds, st));
    std::vector<int> again;
    if (which == 0) {
      std::vector<BiResult> res(m);
      HIPCHK(hipMemcpyAsync(res.data(), d_res, m * sizeof(BiResult), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      for (int i = 0; i < m; ++i) {
        if (res[i].status == wfa::BI_RETRY) again.push_back(todo[i]);
        else (*results)[todo[i]] = res[i];
      }
    } else {
      std::vector<int32_t> len(m), stt(m), sc(m);
      HIPCHK(hipMemcpyAsync(len.data(), d_flen, m * 4, hipMemcpyDeviceToHost, st));
      HIPCHK(hipMemcpyAsync(stt.data(), d_fst, m * 4, hipMemcpyDeviceToHost, st));
      HIPCHK(hipMemcpyAsync(sc.data(), d_fsc, m * 4, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      for (int i = 0; i < m; ++i) {
        if (stt[i] == wfa::BI_RETRY) {
          again.push_back(todo[i]);
        } else {
          (*frag_len)[todo[i]] = len[i];
          (*frag_status)[todo[i]] = stt[i];
          (*frag_score)[todo[i]] = sc[i];
        }
      }
    }
    todo.swap(again);
  }
  return WFA_OK;
}

}  // namespace

int biwfa_align_batch(wfa_engine_t* e, const wfa_attr_t* attr, const wfa_batch_t* b,
                      hipStream_t st, int first, int count) {
  Internal ip;
  int r = make_internal(attr->penalties, &ip);
  if (r) return r;
  KParams kp;
  base_params(ip, *attr, *b, &kp);
  const int kind = kind_of(ip.metric);
  const bool heur = attr->heuristic != WFA_HEURISTIC_NONE;
  const bool cig = attr->alignment_scope == WFA_SCOPE_ALIGNMENT;
  HIPCHK(hipEventRecord(e->ev[0], st));
  // lengths (and CIGAR offsets) of the pairs, on the host
  std::vector<int32_t> plen(count), tlen(count);
  std::vector<int64_t> coff(count);
  HIPCHK(hipMemcpyAsync(plen.data(), b->pattern_length + first, count * 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(tlen.data(), b->text_length + first, count * 4, hipMemcpyDeviceToHost, st));
  if (cig)
    HIPCHK(hipMemcpyAsync(coff.data(), b->cigar_offset + first, count * 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  std::vector<int32_t> score(count, 0), status(count, WFA_STATUS_OK), clen(count, 0);

  if (!cig) {
    // align_biwfa_score: the first breakpoint, or the end reached first
    std::vector<BiJob> jobs(count);
    for (int i = 0; i < count; ++i)
      jobs[i] = BiJob{first + i, 0, plen[i], 0, tlen[i], wfa::C_M, wfa::C_M, 0};
    std::vector<BiResult> res;
    r = run_jobs(e, kp, kind, heur, 0, jobs, &res, nullptr, nullptr, nullptr, nullptr, nullptr, st);
    if (r) return r;
    for (int i = 0; i < count; ++i) {
      const BiResult& x = res[i];
      score[i] = x.status == wfa::BI_OK ? x.score : x.status == wfa::BI_END_REACHED ? x.end_score : 0;
      status[i] = (x.status == wfa::BI_OK || x.status == wfa::BI_END_REACHED)
                      ? WFA_STATUS_COMPLETED
                      : x.status == wfa::BI_MAX_STEPS ? WFA_STATUS_MAX_STEPS_REACHED
                                                      : WFA_STATUS_END_UNREACHABLE;
    }
  } else {
    std::vector<PairState> ps(count);
    std::vector<Node> frontier, bases;
    const uint64_t kTop = uint64_t(1) << 62;
    for (int i = 0; i < count; ++i) {
      Node nd;
      nd.job = BiJob{first + i, 0, plen[i], 0, tlen[i], wfa::C_M, wfa::C_M, 0};
      // short sequences go straight to unidirectional WFA, at the top level only
      nd.score_remaining = std::max(plen[i], tlen[i]) <= kFallbackMinLength ? 0 : INT32_MAX;
      nd.key_lo = 0;
      nd.key_hi = kTop;
      frontier.push_back(nd);
    }
    std::vector<int> root_score(count, 0);
    auto fail = [&](const Node& nd, int s, int frag) {
      PairState& p = ps[nd.job.pair - first];
      if (nd.key_lo < p.fail_key) {
        p.failed = true;
        p.fail_status = s;
        p.fail_key = nd.key_lo;
        p.fail_frag = frag;
      }
    };
    while (!frontier.empty()) {
      std::vector<Node> search;
      for (const Node& nd : frontier) {
        const BiJob& j = nd.job;
        const int pl = j.pe - j.pb, tl = j.te - j.tb;
        PairState& p = ps[j.pair - first];
        if (tl == 0) {
          if (pl > 0) p.leaves.push_back(Leaf{nd.key_lo, -1, 'D', pl});
        } else if (pl == 0) {
          p.leaves.push_back(Leaf{nd.key_lo, -1, 'I', tl});
        } else if (nd.score_remaining <= kFallbackMinScore) {
          bases.push_back(nd);
        } else {
          search.push_back(nd);
        }
      }
      std::vector<BiJob> jobs(search.size());
      for (size_t i = 0; i < search.size(); ++i) jobs[i] = search[i].job;
      std::vector<BiResult> res;
      r = run_jobs(e, kp, kind, heur, 0, jobs, &res, nullptr, nullptr, nullptr, nullptr, nullptr,
                   st);
      if (r) return r;
      std::vector<Node> next;
      for (size_t i = 0; i < search.size(); ++i) {
        const Node& nd = search[i];
        const BiResult& x = res[i];
        if (x.status == wfa::BI_OK) {
          if (nd.key_hi - nd.key_lo < 2) return WFA_ERR_UNSUPPORTED;  // deeper than 62 levels
          const uint64_t mid = nd.key_lo + (nd.key_hi - nd.key_lo) / 2;
          const int bp_h = x.off_fwd, bp_v = x.off_fwd - x.k_fwd;
          const BiJob& j = nd.job;
          Node l, rr;
          l.job = BiJob{j.pair, j.pb, j.pb + bp_v, j.tb, j.tb + bp_h, j.cbeg, x.comp, 0};
          l.score_remaining = x.score_fwd;
          l.key_lo = nd.key_lo;
          l.key_hi = mid;
          rr.job = BiJob{j.pair, j.pb + bp_v, j.pe, j.tb + bp_h, j.te, x.comp, j.cend, 0};
          rr.score_remaining = x.score_rev;
          rr.key_lo = mid;
          rr.key_hi = nd.key_hi;
          next.push_back(l);
          next.push_back(rr);
          if (nd.key_lo == 0 && nd.key_hi == kTop) root_score[j.pair - first] = x.score;
        } else if (x.status == wfa::BI_END_REACHED) {
          // one direction reached the end first: unidirectional WFA if cheap
          // (wavefront_bialign_find_breakpoint_exception)
          if (x.end_score <= kRecoveryMinScore) bases.push_back(nd);
          else fail(nd, wfa::BI_UNREACHABLE, -1);
        } else {
          fail(nd, x.status, -1);
        }
      }
      frontier.swap(next);
    }
    // base cases
    std::vector<BiJob> bjobs(bases.size());
    std::vector<int64_t> foff(bases.size());
    int64_t ftotal = 0;
    for (size_t i = 0; i < bases.size(); ++i) {
      bjobs[i] = bases[i].job;
      foff[i] = ftotal;
      ftotal += (bjobs[i].pe - bjobs[i].pb) + (bjobs[i].te - bjobs[i].tb);
    }
    uint8_t* d_frag;
    r = dev(e, B_FRAG, (size_t)ftotal, &d_frag);
    if (r) return r;
    std::vector<int32_t> flen, fst, fsc;
    r = run_jobs(e, kp, kind, heur, 1, bjobs, nullptr, &foff, d_frag, &flen, &fst, &fsc, st);
    if (r) return r;
    for (size_t i = 0; i < bases.size(); ++i) {
      const Node& nd = bases[i];
      PairState& p = ps[nd.job.pair - first];
      p.leaves.push_back(Leaf{nd.key_lo, (int)i, 0, flen[i]});
      if (fst[i] != wfa::BI_OK) fail(nd, wfa::BI_UNATTAINABLE, (int)i);
      if (nd.key_lo == 0 && nd.key_hi == kTop) root_score[nd.job.pair - first] = fsc[i];
    }
    // assemble each pair's CIGAR in depth-first order: everything before
    // the first failure, and a failing base case's own fragment
    std::vector<BiPiece> pieces;
    std::vector<int32_t> rescore;
    for (int i = 0; i < count; ++i) {
      PairState& p = ps[i];
      std::sort(p.leaves.begin(), p.leaves.end(),
                [](const Leaf& a, const Leaf& c) { return a.key < c.key; });
      int64_t pos = 0;
      for (const Leaf& lf : p.leaves) {
        const bool keep = !p.failed || lf.key < p.fail_key || (lf.frag >= 0 && lf.frag == p.fail_frag);
        if (!keep || lf.len <= 0) continue;
        pieces.push_back(BiPiece{coff[i] + pos, lf.frag >= 0 ? foff[lf.frag] : -1, lf.len, lf.op});
        pos += lf.len;
      }
      clen[i] = (int32_t)pos;
      if (p.failed) {
        status[i] = p.fail_status == wfa::BI_MAX_STEPS ? WFA_STATUS_MAX_STEPS_REACHED
                                                       : WFA_STATUS_END_UNREACHABLE;
        score[i] = 0;
      } else {
        status[i] = WFA_STATUS_COMPLETED;
        score[i] = root_score[i];
        // an empty sequence, or a heuristic, takes the CIGAR's own score
        if (plen[i] == 0 || tlen[i] == 0 || (heur && pos > 0)) rescore.push_back(first + i);
      }
    }
    if (!pieces.empty()) {
      BiPiece* d_pieces;
      r = dev(e, B_PIECES, pieces.size(), &d_pieces);
      if (r) return r;
      HIPCHK(hipMemcpyAsync(d_pieces, pieces.data(), pieces.size() * sizeof(BiPiece),
                            hipMemcpyHostToDevice, st));
      HIPCHK(wfa::launch_biwfa_gather(d_pieces, (int)pieces.size(), d_frag, b->cigar_ops, st));
    }
    HIPCHK(hipMemcpyAsync(b->cigar_length + first, clen.data(), count * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b->score + first, score.data(), count * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b->status + first, status.data(), count * 4, hipMemcpyHostToDevice, st));
    if (!rescore.empty()) {
      int32_t* d_pairs;
      r = dev(e, B_PAIRS, rescore.size(), &d_pairs);
      if (r) return r;
      HIPCHK(hipMemcpyAsync(d_pairs, rescore.data(), rescore.size() * 4, hipMemcpyHostToDevice, st));
      HIPCHK(wfa::launch_biwfa_rescore(kp, d_pairs, (int)rescore.size(), st));
    }
    HIPCHK(hipStreamSynchronize(st));
    HIPCHK(hipEventRecord(e->ev[1], st));
    HIPCHK(hipEventSynchronize(e->ev[1]));
    HIPCHK(hipEventElapsedTime(&e->all_ms, e->ev[0], e->ev[1]));
    e->first_ms = e->all_ms;
    e->last_retries = 0;
    return WFA_OK;
  }
  HIPCHK(hipMemcpyAsync(b->score + first, score.data(), count * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(b->status + first, status.data(), count * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipEventRecord(e->ev[1], st));
  HIPCHK(hipEventSynchronize(e->ev[1]));
  HIPCHK(hipEventElapsedTime(&e->all_ms, e->ev[0], e->ev[1]));
  e->first_ms = e->all_ms;
  e->last_retries = 0;
  return WFA_OK;
}

}  // namespace wfahost
