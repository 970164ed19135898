// wfa_short.h — the short-read kernel: gap-affine, end-to-end, no
// heuristic, score only or high-memory CIGAR, sequences of A/C/G/T.
//
// Semantics are those of align_pair (the unialign loop, src/aligner.rs:
// 638-698, with compute_affine_step, :2006-2193, and the Singletrack
// backtrace, src/backtrace.rs:751-858); the storage is laid out so that a
// step costs as few scalar instructions as possible (the general kernel is
// bound by its per-score scalar bookkeeping on rows of a few dozen
// diagonals):
//
// * Each pair gets a fixed window of 256 diagonals [base, base + 256),
//   centred on the diagonals its alignment must cross. Every ring row is an
//   LDS row over the whole window plus one NULL pad cell on each side, all
//   NULL when the pair starts. A step writes its three output rows over the
//   64-diagonal chunks it computes, OFFSET_NULL outside their trimmed
//   ranges, so a row's cells outside its range read as NULL with no range
//   test: the reference NULL-fills exactly those cells before it reads them
//   (src/wavefront.rs:422-457).
// * The chunks a step computes cover [lo - 1, hi + 1] of the union of every
//   row computed so far. That holds the reference's computed range (the
//   union of the inputs' ranges, widened by one, src/aligner.rs:2045-2062);
//   a cell outside the latter has only NULL inputs, so it comes out below
//   zero, outside the trimmed ranges, and is stored as NULL — as it would
//   be read in the reference. The union only grows, so no cell outside it
//   has been written since the pair started.
// * Which ring rows hold a wavefront (not absent, not trimmed empty) is one
//   64-bit scalar mask; an absent input row is read from an all-NULL row
//   instead, and a step whose inputs are all absent is a null step
//   (src/aligner.rs:2031-2043) that touches no vector registers at all.
// * In CIGAR scope the extended M row of every score goes to the slot's
//   history in HBM at a fixed stride (cell k of score s at s * 256 + k -
//   base, 16-bit), with the computed chunks' first and last diagonal packed
//   in one int32 header per score; the Singletrack backtrace reads it
//   through WinView.
//
// Offsets are held in 16 bits as in the general kernel's short-read rings
// (sequences of at most 8000 bases, scores up to 16000; NULL-derived values
// saturate at -16384 and only their sign matters). A pair whose union of
// diagonals would leave its window, whose score passes the history's rows,
// or whose sequences hold other bytes than A/C/G/T is handed back through
// the retry list to the general kernel, which aligns it from the start.
#pragma once

#include "wfa_align.h"

namespace wfa {

constexpr int SH_WIN = 256;              // diagonals of a pair's window
constexpr int SH_CHUNKS = SH_WIN / 64;   // most 64-diagonal chunks a step computes
constexpr int SH_ROW = SH_WIN + 8;       // LDS cells per row: pad, window, pad (16-byte rows)
constexpr int SH_NULL = -16384;          // OFFSET_NULL as held in 16 bits
constexpr int SH_HDR_EMPTY = 0x00000001; // header of an absent M row: lo 1, hi 0

static __device__ __forceinline__ int16_t sh16(int v) { return (int16_t)imax(v, SH_NULL); }

// One slot's high-memory history as the Singletrack backtrace reads it
// (st_m_offset, src/backtrace.rs:645-663): M[s] at diagonal k, if score s
// has a row covering k and the cell is not NULL.
struct WinView {
  const int16_t* cells;  // row s at cells + s * SH_WIN, diagonal k at k - base
  const int32_t* hdr;    // per score: first diagonal (low 16 bits), last (high 16 bits)
  int base, nscores, nwf;
  __device__ __forceinline__ bool m(int s, int k, int* out) const {
    if (s < 0 || s >= nscores || s >= nwf) return false;
    const int h = hdr[s];
    const int lo = (int)(int16_t)(h & 0xffff), hi = h >> 16;
    if (k < lo || k > hi) return false;
    const int v = cells[(int64_t)s * SH_WIN + (k - base)];
    if (v < 0) return false;
    *out = v;
    return true;
  }
};

static __device__ __forceinline__ int32_t sh_hdr(int lo, int hi) {
  return (int32_t)(((uint32_t)hi << 16) | ((uint32_t)lo & 0xffffu));
}

// Record the first and last set lane of chunk ballot `b` (chunk starting at
// diagonal kb) into [lo, hi]; chunks come in ascending order.
static __device__ __forceinline__ void sh_track(int& lo, int& hi, uint64_t b, int kb) {
  if (b) {
    if (lo == INT32_MAX) lo = kb + __builtin_ctzll(b);
    hi = kb + 63 - __builtin_clzll(b);
  }
}

// Bases [pos, pos + 16) of a 2-bit sequence staged as 32-bit words
// (16 bases each, base j at bits [2j, 2j + 2)): one funnel shift of two
// consecutive words.
static __device__ __forceinline__ uint32_t sh_w16(const uint32_t* s, uint32_t pos) {
  const uint32_t q = pos >> 4;
  return __builtin_amdgcn_alignbit(s[q + 1], s[q], (pos & 15u) * 2u);
}

// Matching prefix length of pattern[v..] and text[h..], bounded by the
// sequence ends (extend_matches_packed_end2end's per-diagonal LCP,
// src/extend.rs:15-38): 16 bases per XOR of funnel-shifted words.
static __device__ __forceinline__ int sh_lcp(const uint32_t* P, const uint32_t* T, uint32_t v,
                                             uint32_t h, int rem) {
  int n = 0;
  while (rem > 0) {
    const uint32_t x = sh_w16(P, v + n) ^ sh_w16(T, h + n);
    // matching bases before the first difference (16 when none)
    const int c = x ? (int)(__builtin_ctz(x) >> 1) : 16;
    if (c >= rem) { n += rem; break; }
    n += c;
    if (c < 16) break;
    rem -= 16;
  }
  return n;
}

template <int MODE>
static __device__ void short_pair(const KParams& p, int pair, int plen, int tlen,
                                  const uint64_t* sP, const uint64_t* sT, int16_t* rows,
                                  int16_t* hcells, int32_t* hhdr) {
  constexpr bool CIG = MODE == 1;
  const int lane = threadIdx.x;
  const int k_end = uniform(tlen - plen);
  const int ak = k_end < 0 ? -k_end : k_end;
  auto retry = [&]() {
    if (lane == 0) {
      const int slot = atomicAdd(&p.counters[1], 1);
      p.retry_list[slot] = pair;
      p.status[pair] = ST_RETRY;
    }
  };
  if (ak > SH_WIN - 64 || plen > 8000 || tlen > 8000 || (CIG && p.hdr_cap < 1)) {
    retry();
    return;
  }
  // the window centred on the diagonals between 0 and k_end
  const int base = uniform(imin(0, k_end) - (SH_WIN - ak) / 2);
  const int ringM = p.ringM, ringE = p.ringE1;
  const int rI1 = ringM, rD1 = ringM + ringE, rNull = ringM + 2 * ringE;
  // every row (and the NULL row) NULL over the window and its pads
  {
    const int n16 = (rNull + 1) * SH_ROW / 8;
    const int nn = (int)(((uint32_t)(uint16_t)SH_NULL << 16) | (uint16_t)SH_NULL);
    int4* r4 = reinterpret_cast<int4*>(rows);
    for (int i = lane; i < n16; i += 64) r4[i] = make_int4(nn, nn, nn, nn);
  }
  const uint32_t* P32 = reinterpret_cast<const uint32_t*>(sP);
  const uint32_t* T32 = reinterpret_cast<const uint32_t*>(sT);
  // row r's cell of diagonal k
  auto rowp = [&](int r) { return rows + r * SH_ROW + 1 - base; };
  const int nwf = uniform(high_mem_nwf(p, plen, tlen));
  int status = WFA_STATUS_OK, result = 0, score = 0, nulls = 0;
  bool overflow = false;
  __syncthreads();

  // score 0: M0 = {k 0: offset 0} (init_alignment, src/aligner.rs:846-1000),
  // extended
  {
    // (the same in every lane: keep it scalar, so the loop state stays
    // scalar)
    const int off = uniform(lcp<2>(sP, sT, 0, 0, plen, tlen));
    if (lane == 0) {
      rowp(0)[0] = sh16(off);
      if (CIG) hcells[0 - base] = sh16(off);
    }
    if (k_end == 0 && off >= tlen) status = WFA_STATUS_END_REACHED;
  }
  __syncthreads();

  // Which earlier rows hold a wavefront (not absent, not trimmed empty), by
  // age: bit d of agM is M[s - 1 - d] at the top of step s, likewise agI
  // and agD for I1 and D1; every step shifts them by one.
  uint64_t agM = 1;  // M0
  uint32_t agI = 0, agD = 0;
  const uint64_t needM = (1ull << (p.x - 1)) | (1ull << (p.oe1 - 1));
  const uint32_t needE = 1u << (p.e1 - 1);
  int ulo = 0, uhi = 0;   // union of the diagonals of every row so far
  // ring rows as LDS byte offsets from `rows`: score s's M row at
  // (s mod ringM) * RB, its I1 / D1 rows at iB / dB + (s mod ringE) * RB
  constexpr int RB = SH_ROW * 2;
  const int ringMB = ringM * RB, ringEB = ringE * RB;
  const int xB = p.x * RB, oeB = p.oe1 * RB, eB = p.e1 * RB;
  const int iB = rI1 * RB, dB = rD1 * RB, nullB = rNull * RB;
  int sMB = 0, sEB = 0;   // (s mod ringM) * RB, (s mod ringE) * RB
  auto back = [](int cur, int d, int ring) { return cur >= d ? cur - d : cur - d + ring; };
  auto cell = [&](int byte_off) {
    return reinterpret_cast<int16_t*>(reinterpret_cast<char*>(rows) + byte_off) + 1 - base;
  };
  // CIGAR headers of 64 consecutive scores, lane i holding score
  // (s & ~63) + i, stored once the block is full (and at the end)
  int hbuf = lane == 0 ? sh_hdr(0, 0) : 0;
  // the first score at which a per-score limit applies (one compare per score)
  int limit = imin(16001, p.max_steps);
  if (!p.modular) limit = imin(limit, nwf);
  if (CIG) limit = imin(limit, p.hdr_cap);
  limit = uniform(limit);
  // Every exit of the loop is a uniform break that sets the outcome; the
  // loop-carried state is read back as scalar at the top of each step (the
  // same in every lane; the compiler cannot tell across the per-lane
  // extension, and would otherwise run the whole loop under exec masks).
  for (int step = status == WFA_STATUS_OK ? 1 : 0; step; ) {
    score = uniform(score) + 1;
    ulo = uniform(ulo);
    uhi = uniform(uhi);
    nulls = uniform(nulls);
    agM = ((uint64_t)(uint32_t)uniform((int)(agM >> 32)) << 32) |
          (uint32_t)uniform((int)(uint32_t)agM);
    agI = (uint32_t)uniform((int)agI);
    agD = (uint32_t)uniform((int)agD);
    sMB = uniform(sMB) + RB;
    sEB = uniform(sEB) + RB;
    if (sMB == ringMB) sMB = 0;
    if (sEB == ringEB) sEB = 0;
    const int s = score;
    bool last = false;  // max steps: this step is computed, then the loop ends
    if (s >= limit) {
      if (!p.modular && s >= nwf) { status = WFA_STATUS_MAX_STEPS_REACHED; result = s; break; }
      if (CIG && s >= p.hdr_cap) { overflow = true; break; }
      if (s > 16000) { overflow = true; break; }
      last = s >= p.max_steps;
    }
    // inputs M[s-x], M[s-o-e], I1/D1[s-e]; a lookback below score 0 finds
    // its age bit clear
    if (!(agM & needM) && !((agI | agD) & needE)) {
      // null step (src/aligner.rs:2031-2043): the outputs are absent
      nulls += 1;
      agM <<= 1;
      agI <<= 1;
      agD <<= 1;
      if (CIG) {
        hbuf = lane == (s & 63) ? SH_HDR_EMPTY : hbuf;
        if ((s & 63) == 63) hhdr[(s & ~63) + lane] = hbuf;
      }
      if (last) { status = WFA_STATUS_MAX_STEPS_REACHED; result = s; break; }
      if (nulls > p.scope) { status = WFA_STATUS_END_UNREACHABLE; result = s; break; }
      continue;
    }
    nulls = 0;
    const int eb = back(sEB, eB, ringEB);
    const int16_t* Px = cell((agM >> (p.x - 1)) & 1 ? back(sMB, xB, ringMB) : nullB);
    const int16_t* Po = cell((agM >> (p.oe1 - 1)) & 1 ? back(sMB, oeB, ringMB) : nullB);
    const int16_t* Pi = cell(agI & needE ? iB + eb : nullB);
    const int16_t* Pd = cell(agD & needE ? dB + eb : nullB);
    const int clo = ulo - 1, chi = uhi + 1;
    if (clo < base || chi > base + SH_WIN - 1) { overflow = true; break; }
    const int nch = ((chi - clo) >> 6) + 1;
    // first chunk's diagonal: the chunks stay inside the window
    const int a = imin(clo, base + SH_WIN - 64 * nch);

    // ---- compute_affine_idm (src/compute/affine.rs:318-351) over the chunks,
    // trims by ballot (trim_ends, src/compute/mod.rs:14-49) ----
    int mv[SH_CHUNKS], iv[SH_CHUNKS], dv[SH_CHUNKS];
    int m_lo = INT32_MAX, m_hi = INT32_MIN, i_lo = INT32_MAX, i_hi = INT32_MIN;
    int d_lo = INT32_MAX, d_hi = INT32_MIN;
#pragma unroll
    for (int j = 0; j < SH_CHUNKS; ++j) {
      if (j < nch) {
        const int k = a + 64 * j + lane;
        const int ins = imax(Po[k - 1], Pi[k - 1]) + 1;
        const int del = imax(Po[k + 1], Pd[k + 1]);
        const int m = imax(Px[k] + 1, imax(ins, del));
        // in bounds: offset h <= tlen and h - k <= plen, both unsigned
        auto vcol = [&](int h) {
          return (uint32_t)h <= (uint32_t)tlen ? (uint32_t)(h - k) : 0xffffffffu;
        };
        const bool vm = vcol(m) <= (uint32_t)plen;
        mv[j] = vm ? m : WFA_NULL;
        iv[j] = ins;
        dv[j] = del;
        sh_track(m_lo, m_hi, ballot(vm), a + 64 * j);
        sh_track(i_lo, i_hi, ballot(vcol(ins) <= (uint32_t)plen), a + 64 * j);
        sh_track(d_lo, d_hi, ballot(vcol(del) <= (uint32_t)plen), a + 64 * j);
      }
    }
    const bool mne = m_lo != INT32_MAX, ine = i_lo != INT32_MAX, dne = d_lo != INT32_MAX;
    // I1 and D1 keep their cells inside their trimmed ranges (out-of-bounds
    // ones included), NULL elsewhere; M is NULL wherever it is out of bounds
    const Guard gi = ine ? Guard{i_lo, i_hi - i_lo} : Guard{INT32_MIN / 2, 0};
    const Guard gd = dne ? Guard{d_lo, d_hi - d_lo} : Guard{INT32_MIN / 2, 0};
    int16_t* Wm = cell(sMB);
    int16_t* Wi = cell(iB + sEB);
    int16_t* Wd = cell(dB + sEB);
#pragma unroll
    for (int j = 0; j < SH_CHUNKS; ++j) {
      if (j < nch) {
        const int k = a + 64 * j + lane;
        Wi[k] = (uint32_t)(k - gi.first) <= (uint32_t)gi.span ? sh16(iv[j]) : (int16_t)SH_NULL;
        Wd[k] = (uint32_t)(k - gd.first) <= (uint32_t)gd.span ? sh16(dv[j]) : (int16_t)SH_NULL;
      }
    }
    agM = (agM << 1) | (uint64_t)mne;
    agI = (agI << 1) | (uint32_t)ine;
    agD = (agD << 1) | (uint32_t)dne;
    if (mne) { ulo = imin(ulo, m_lo); uhi = imax(uhi, m_hi); }
    if (ine) { ulo = imin(ulo, i_lo); uhi = imax(uhi, i_hi); }
    if (dne) { ulo = imin(ulo, d_lo); uhi = imax(uhi, d_hi); }
    if (last) {
      // the row is computed but not extended (the loop stops first)
#pragma unroll
      for (int j = 0; j < SH_CHUNKS; ++j)
        if (j < nch) Wm[a + 64 * j + lane] = sh16(mv[j]);
      status = WFA_STATUS_MAX_STEPS_REACHED;
      result = s;
      break;
    }
    // ---- extend M[s] (extend_matches_packed_end2end, src/extend.rs:47-69;
    // the reference extends at the top of the next iteration and nothing
    // reads M[s] in between), keep it, test the end cell
    // (termination_end2end, src/termination.rs:14-23) ----
    bool reached = false;
    int16_t* hrow = CIG ? hcells + (int64_t)s * SH_WIN - base : nullptr;
#pragma unroll
    for (int j = 0; j < SH_CHUNKS; ++j) {
      if (j < nch) {
        const int k = a + 64 * j + lane;
        int off = mv[j];
        if (off >= 0)
          off += sh_lcp(P32, T32, (uint32_t)(off - k), (uint32_t)off, imin(plen - (off - k), tlen - off));
        const int16_t o16 = sh16(off);
        Wm[k] = o16;
        if (CIG) hrow[k] = o16;
        reached |= k == k_end && off >= tlen;
      }
    }
    if (CIG) {
      hbuf = lane == (s & 63) ? sh_hdr(a, a + 64 * nch - 1) : hbuf;
      if ((s & 63) == 63) hhdr[(s & ~63) + lane] = hbuf;
    }
    // the step's rows are written before the next step reads them: one
    // wave's LDS accesses complete in order, so only the compiler must keep
    // them in program order
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    if (ballot(reached)) { status = WFA_STATUS_END_REACHED; result = s; break; }
  }

  if (overflow) {
    retry();
    return;
  }
  // the headers of the last, partial block
  if (CIG && lane <= (score & 63)) hhdr[(score & ~63) + lane] = hbuf;
  int out_status = status;
  int ncig = 0;
  if (status == WFA_STATUS_END_REACHED) {
    out_status = WFA_STATUS_COMPLETED;
    if (CIG) {
      __threadfence_block();
      uint8_t* out = p.cig + p.cig_off[pair];
      const int cap = plen + tlen;
      ncig = singletrack<2>(p, WinView{hcells, hhdr, base, score + 1, nwf}, sP, sT, plen, tlen,
                            score, k_end, tlen, out, cap);
      // move the ops from the tail of the region to its head
      const int gap = cap - ncig;
      if (gap > 0) {
        for (int c = 0; c < ncig; c += 64) {
          const int i = c + lane;
          uint8_t v = 0;
          if (i < ncig) v = out[gap + i];
          __builtin_amdgcn_wave_barrier();
          if (i < ncig) out[i] = v;
          __builtin_amdgcn_wave_barrier();
        }
      }
    }
  }
  if (lane == 0) {
    p.score[pair] = result;
    p.status[pair] = out_status;
    if (CIG) p.cig_len[pair] = ncig;
    if (p.end_pos) {
      // the end cell, or unset (src/aligner.rs:903-909)
      const bool end = status == WFA_STATUS_END_REACHED;
      p.end_pos[3 * pair] = end ? score : -1;
      p.end_pos[3 * pair + 1] = end ? k_end : INT32_MAX;
      p.end_pos[3 * pair + 2] = end ? tlen : WFA_NULL;
    }
  }
}

// One wave per pair, persistent over the launch's work queue. The slot's
// history: p.hdr_cap rows of SH_WIN 16-bit cells in p.hist and p.hdr_cap
// packed headers in p.hdr, both per workgroup.
template <int MODE, bool PACKED>
__global__ void __launch_bounds__(64) wfa_short_kernel(KParams p) {
  extern __shared__ uint64_t smem[];
  uint64_t* sP = smem;
  uint64_t* sT = smem + p.seq_words;
  int16_t* rows = reinterpret_cast<int16_t*>(sT + p.seq_words);
  int16_t* hcells =
      MODE == 1 ? reinterpret_cast<int16_t*>(p.hist) + (int64_t)blockIdx.x * p.hdr_cap * SH_WIN
                : nullptr;
  int32_t* hhdr =
      MODE == 1 ? reinterpret_cast<int32_t*>(p.hdr) + (int64_t)blockIdx.x * p.hdr_cap : nullptr;
  for (;;) {
    int item = 0;
    if (threadIdx.x == 0) item = atomicAdd(&p.counters[0], 1);
    item = uniform(__shfl(item, 0));
    if (item >= p.num_pairs) break;
    // pair-level values are uniform: keep them in scalar registers
    const int pair = uniform(p.pair_list ? p.pair_list[item] : p.pair_base + item);
    const int plen = uniform(p.plen[pair]), tlen = uniform(p.tlen[pair]);
    const int bits = stage_sequences<1>(p, pair, plen, tlen, sP, sT, nullptr);
    if (bits != 2) {
      // does not fit, or not A/C/G/T: the general kernel takes it
      if (threadIdx.x == 0) {
        const int slot = atomicAdd(&p.counters[1], 1);
        p.retry_list[slot] = pair;
        p.status[pair] = ST_RETRY;
      }
      continue;
    }
    short_pair<MODE>(p, pair, plen, tlen, sP, sT, rows, hcells, hhdr);
    __syncthreads();
  }
}

// LDS bytes of the short kernel's rows (the ring rows and the NULL row)
static inline size_t short_lds_row_bytes(int ring_rows) {
  return (size_t)(ring_rows + 1) * SH_ROW * 2;
}

}  // namespace wfa
