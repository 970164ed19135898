// wfa_window.h — the short-read fast path: gap-affine, end-to-end, no
// heuristic, 2-bit sequences, score only or high-memory CIGAR.
//
// Semantics are those of align_pair (the unialign loop, src/aligner.rs:
// 638-698, with compute_affine_step, :2006-2193); only the storage differs.
// Each pair's diagonals live in a fixed window [base, base + 256) of four
// 64-lane chunks. A wavefront row is stored in LDS over the chunks its step
// computed (its cells inside the row's trimmed range, OFFSET_NULL elsewhere
// in those chunks) with a per-row mask of those chunks; a NULL cell pads
// each end of the window. The compute reads a neighbour at k-1 or k+1 with
// a fixed offset and keeps it only if its chunk is in the row's mask (lane
// 0 of a k-1 read, and lane 63 of a k+1 read, look at the next chunk's
// bit) — a select on a uniform mask instead of a bounds test; a row missing
// from the ring (a null step, or a score below zero) reads as NULL. The
// three output rows of a step stay in registers through the trims and the
// extension and are written once, over the computed chunks only.
// Offsets are held in 16 bits as in the general kernel (pairs of at most
// 8000 bases, scores up to 16000).
//
// A pair whose wavefront would leave its window, or that the window cannot
// hold, is handed back through the retry list to the general kernel, which
// computes it from the start.
#pragma once

#include "wfa_align.h"

namespace wfa {

constexpr int WIN_CHUNKS = 4;
constexpr int WIN = 64 * WIN_CHUNKS;  // diagonals per window
constexpr int WROW = WIN + 2;         // cells per row: the window and two NULL pads
constexpr int16_t NULL16 = -16384;    // OFFSET_NULL as held in 16 bits

static __device__ __forceinline__ int16_t sat16(int v) { return (int16_t)imax(v, NULL16); }

template <int MODE>
static __device__ void align_window(const KParams& p, int pair, int plen, int tlen,
                                    const uint64_t* sP, const uint64_t* sT, int16_t* rows,
                                    int2* meta, int* vchunks, const Hist& H, int64_t hcap) {
  constexpr bool CIG = MODE == 1;
  const int lane = threadIdx.x;
  const int k_end = tlen - plen;
  const int ak = k_end < 0 ? -k_end : k_end;
  auto retry = [&]() {
    if (lane == 0) {
      const int slot = atomicAdd(&p.counters[1], 1);
      p.retry_list[slot] = pair;
      p.status[pair] = ST_RETRY;
    }
  };
  if (ak > WIN - 64 || plen > 8000 || tlen > 8000) { retry(); return; }
  // the window centred on the diagonals the alignment must cross
  const int base = imin(0, k_end) - (WIN - ak) / 2;
  const int ringM = p.ringM, ringE = p.ringE1;
  const int rI1 = ringM, rD1 = ringM + ringE;
  const int nrows = ringM + 2 * ringE;
  // row r's cell for diagonal k: rows[r * WROW + (k - base + 1)]
  auto row_ptr = [&](int r) { return rows + r * WROW + 1 - base; };
  // Write row r over the computed chunks [clo, chi]: v inside [lo, hi],
  // NULL elsewhere; the row's chunk mask becomes those chunks.
  auto write_row = [&](int r, const int (&v)[WIN_CHUNKS], int clo, int chi, int lo, int hi) {
    int16_t* P = row_ptr(r);
#pragma unroll
    for (int c = 0; c < WIN_CHUNKS; ++c) {
      const int k = base + 64 * c + lane;
      if (c >= clo && c <= chi) P[k] = k >= lo && k <= hi ? sat16(v[c]) : NULL16;
    }
    if (lane == 0) vchunks[r] = ((2 << chi) - 1) & ~((1 << clo) - 1);
  };
  const int M_ABSENT = INT32_MAX;
  for (int r = lane; r < nrows; r += 64) meta[r] = make_int2(M_ABSENT, M_ABSENT - 1);
  __syncthreads();

  const int nwf = high_mem_nwf(p, plen, tlen);
  const int hist_lo = -(plen + p.scope + 2), hist_hi = tlen + p.scope + 2;
  int64_t hpos = 0;
  int nulls = 0, score = 0, status = WFA_STATUS_OK, result = 0;
  bool overflow = false;

  // registers of the current step's output rows, one cell per chunk
  int mv[WIN_CHUNKS], iv[WIN_CHUNKS], dv[WIN_CHUNKS];

  // Extend mv over [clo, chi] chunks (cells >= 0), keep the row's history
  // over [lo, hi], write the M row; returns whether the end cell is reached.
  auto finish_m = [&](int s, int lo, int hi, int clo, int chi, int m_lo, int m_hi) -> bool {
    bool reached = false;
#pragma unroll
    for (int c = 0; c < WIN_CHUNKS; ++c) {
      if (c < clo || c > chi) continue;
      const int k = base + 64 * c + lane;
      int off = mv[c];
      if (off >= 0) {
        off += lcp<2>(sP, sT, off - k, off, plen, tlen);
        mv[c] = off;
      }
      if (k == k_end && off >= tlen) reached = true;
      if (CIG && k >= lo && k <= hi) H.cells[hpos + (k - lo)] = off < 0 ? WFA_NULL : off;
    }
    write_row(s % ringM, mv, clo, chi, m_lo, m_hi);
    if (CIG) {
      if (lane == 0) H.hdr[s] = make_int4(lo, hi, (int)hpos, 0);
      if (lo <= hi) hpos += hi - lo + 1;
    }
    return ballot(reached) != 0 && m_lo <= k_end && k_end <= m_hi;
  };

  // ---- score 0: M0 = {k = 0: offset 0} ----
  if (CIG && (H.hdr_cap < 1 || hcap < WIN)) overflow = true;
  if (!overflow) {
    const int c0 = (0 - base) >> 6;
#pragma unroll
    for (int c = 0; c < WIN_CHUNKS; ++c) mv[c] = base + 64 * c + lane == 0 ? 0 : WFA_NULL;
    if (lane == 0) meta[0] = make_int2(0, 0);
    if (finish_m(0, 0, 0, c0, c0, 0, 0)) { status = WFA_STATUS_END_REACHED; result = 0; }
    __syncthreads();
  }

  while (status == WFA_STATUS_OK && !overflow) {
    score += 1;
    const int s = score;
    if (!p.modular && s >= nwf) { status = WFA_STATUS_MAX_STEPS_REACHED; result = s; break; }
    if (CIG && (s >= H.hdr_cap || hpos + WIN > hcap)) { overflow = true; break; }
    if (s > 16000) { overflow = true; break; }
    // ---- inputs: M[s-x], M[s-o-e], I1/D1[s-e] ----
    int2 mx = make_int2(1, 0), mo = mx, i1 = mx, d1 = mx;
    int rx = 0, ro = 0, re = 0;
    if (s - p.x >= 0) { rx = (s - p.x) % ringM; mx = meta[rx]; }
    if (s - p.oe1 >= 0) { ro = (s - p.oe1) % ringM; mo = meta[ro]; }
    if (s - p.e1 >= 0) { re = (s - p.e1) % ringE; i1 = meta[rI1 + re]; d1 = meta[rD1 + re]; }
    const bool lx = uniform(mx.x) <= uniform(mx.y), lo_ = uniform(mo.x) <= uniform(mo.y);
    const bool li = uniform(i1.x) <= uniform(i1.y), ld = uniform(d1.x) <= uniform(d1.y);
    const int rowM = s % ringM, rowE = s % ringE;
    __syncthreads();  // every lane has read the metadata before it is rewritten
    if (!(lx || lo_ || li || ld)) {
      // null step (src/aligner.rs:2031-2043)
      nulls += 1;
      if (lane == 0) {
        meta[rowM] = make_int2(M_ABSENT, M_ABSENT - 1);
        meta[rI1 + rowE] = make_int2(M_ABSENT, M_ABSENT - 1);
        meta[rD1 + rowE] = make_int2(M_ABSENT, M_ABSENT - 1);
        if (CIG) H.hdr[s] = make_int4(1, 0, 0, 0);
      }
      __syncthreads();
      if (s >= p.max_steps) { status = WFA_STATUS_MAX_STEPS_REACHED; result = s; break; }
      if (nulls > p.scope) { status = WFA_STATUS_END_UNREACHABLE; result = s; break; }
      continue;
    }
    nulls = 0;
    int lo = INT32_MAX, hi = INT32_MIN;
    if (lx) { lo = imin(lo, uniform(mx.x)); hi = imax(hi, uniform(mx.y)); }
    if (lo_) { lo = imin(lo, uniform(mo.x) - 1); hi = imax(hi, uniform(mo.y) + 1); }
    if (li) { lo = imin(lo, uniform(i1.x) + 1); hi = imax(hi, uniform(i1.y) + 1); }
    if (ld) { lo = imin(lo, uniform(d1.x) - 1); hi = imax(hi, uniform(d1.y) - 1); }
    lo = imax(lo, hist_lo + 1);
    hi = imin(hi, hist_hi - 1);
    // the range must stay inside the window; reads reach one cell further,
    // into the pads
    if (lo < base || hi > base + WIN - 1) { overflow = true; break; }
    const int clo = (lo - base) >> 6, chi = (hi - base) >> 6;
    const int16_t* Px = row_ptr(rx);
    const int16_t* Po = row_ptr(ro);
    const int16_t* Pi = row_ptr(rI1 + re);
    const int16_t* Pd = row_ptr(rD1 + re);
    // chunk masks of the inputs (0: absent)
    const int vx = lx ? uniform(vchunks[rx]) : 0, vo = lo_ ? uniform(vchunks[ro]) : 0;
    const int vi = li ? uniform(vchunks[rI1 + re]) : 0, vd = ld ? uniform(vchunks[rD1 + re]) : 0;
    int m_lo = INT32_MAX, m_hi = INT32_MIN, i_lo = INT32_MAX, i_hi = INT32_MIN;
    int d_lo = INT32_MAX, d_hi = INT32_MIN;
#pragma unroll
    for (int c = 0; c < WIN_CHUNKS; ++c) {
      mv[c] = WFA_NULL; iv[c] = WFA_NULL; dv[c] = WFA_NULL;
      if (c < clo || c > chi) continue;
      const int k = base + 64 * c + lane;
      const bool in = k >= lo && k <= hi;
      // a k-1 read of lane 0 falls in chunk c-1, a k+1 read of lane 63 in c+1
      auto at = [&](int mask) { return (mask >> c) & 1; };
      auto below = [&](int mask) { return lane == 0 ? (c > 0 ? (mask >> (c - 1)) & 1 : 0) : at(mask); };
      auto above = [&](int mask) {
        return lane == 63 ? (c < WIN_CHUNKS - 1 ? (mask >> (c + 1)) & 1 : 0) : at(mask);
      };
      const int om = below(vo) ? Po[k - 1] : WFA_NULL;
      const int op = above(vo) ? Po[k + 1] : WFA_NULL;
      const int ie = below(vi) ? Pi[k - 1] : WFA_NULL;
      const int de = above(vd) ? Pd[k + 1] : WFA_NULL;
      const int ms = at(vx) ? Px[k] : WFA_NULL;
      const int ins = imax(om, ie) + 1;
      const int del = imax(op, de);
      int m = imax(ms + 1, imax(ins, del));
      const bool vm = in && in_bounds(m, k, plen, tlen);
      const bool vi = in && in_bounds(ins, k, plen, tlen);
      const bool vd = in && in_bounds(del, k, plen, tlen);
      mv[c] = vm ? m : WFA_NULL;
      iv[c] = ins;
      dv[c] = del;
      uint64_t b = ballot(vm);
      if (b) {
        if (m_lo == INT32_MAX) m_lo = 64 * c + __builtin_ctzll(b);
        m_hi = 64 * c + 63 - __builtin_clzll(b);
      }
      b = ballot(vi);
      if (b) {
        if (i_lo == INT32_MAX) i_lo = 64 * c + __builtin_ctzll(b);
        i_hi = 64 * c + 63 - __builtin_clzll(b);
      }
      b = ballot(vd);
      if (b) {
        if (d_lo == INT32_MAX) d_lo = 64 * c + __builtin_ctzll(b);
        d_hi = 64 * c + 63 - __builtin_clzll(b);
      }
    }
    // trimmed ranges (trim_ends, src/compute/mod.rs:14-49), as diagonals
    auto fix = [&](int& l, int& h) {
      if (l == INT32_MAX) { l = 1; h = 0; }
      else { l += base; h += base; }
    };
    fix(m_lo, m_hi); fix(i_lo, i_hi); fix(d_lo, d_hi);
    __syncthreads();  // the step's reads are done before its rows are written
    write_row(rI1 + rowE, iv, clo, chi, i_lo, i_hi);
    write_row(rD1 + rowE, dv, clo, chi, d_lo, d_hi);
    if (lane == 0) {
      meta[rowM] = make_int2(m_lo, m_hi);
      meta[rI1 + rowE] = make_int2(i_lo, i_hi);
      meta[rD1 + rowE] = make_int2(d_lo, d_hi);
    }
    if (s >= p.max_steps) {
      // the row is computed but not extended (the loop stops first)
      write_row(rowM, mv, clo, chi, m_lo, m_hi);
      status = WFA_STATUS_MAX_STEPS_REACHED;
      result = s;
      __syncthreads();
      break;
    }
    // ---- extend M[s] (the reference extends at the top of the next
    // iteration; nothing reads M[s] in between) ----
    const bool reached = finish_m(s, lo, hi, clo, chi, m_lo, m_hi);
    __syncthreads();
    if (reached) { status = WFA_STATUS_END_REACHED; result = s; break; }
  }

  if (overflow) { retry(); return; }
  int out_status = status;
  int ncig = 0;
  if (status == WFA_STATUS_END_REACHED) {
    out_status = WFA_STATUS_COMPLETED;
    if (CIG) {
      __threadfence_block();
      uint8_t* out = p.cig + p.cig_off[pair];
      const int cap = plen + tlen;
      ncig = singletrack<2>(p, H, score + 1, nwf, sP, sT, plen, tlen, score, k_end, tlen, out, cap);
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

template <int MODE, bool PACKED>
__global__ void __launch_bounds__(64, 5) wfa_window_kernel(KParams p) {
  extern __shared__ uint64_t smem[];
  const int lane = threadIdx.x;
  uint64_t* sP = smem;
  uint64_t* sT = smem + p.seq_words;
  int2* meta = (int2*)(sT + p.seq_words);
  const int nrows = p.ringM + 2 * p.ringE1;
  int* vchunks = (int*)(meta + nrows);
  int16_t* rows = (int16_t*)(vchunks + nrows);
  // the pads are never written by a step: NULL once for the launch
  for (int r = lane; r < nrows; r += 64) {
    rows[r * WROW] = NULL16;
    rows[r * WROW + WROW - 1] = NULL16;
    vchunks[r] = 0;
  }
  Hist H;
  H.hdr = MODE == 1 ? p.hdr + (int64_t)blockIdx.x * p.hdr_cap : nullptr;
  H.cells = MODE == 1 ? p.hist + (int64_t)blockIdx.x * p.hist_cap : nullptr;
  H.bytes = nullptr;
  H.chain = nullptr;
  H.chain_cap = 0;
  H.hdr_cap = p.hdr_cap;
  H.chunk = 0;  // one contiguous history per slot
  __syncthreads();
  for (;;) {
    int item = 0;
    if (lane == 0) item = atomicAdd(&p.counters[0], 1);
    item = uniform(__shfl(item, 0));
    if (item >= p.num_pairs) break;
    // pair-level values are uniform: keep them in scalar registers
    const int pair = uniform(p.pair_list ? p.pair_list[item] : p.pair_base + item);
    const int plen = uniform(p.plen[pair]), tlen = uniform(p.tlen[pair]);
    const int bits = stage_sequences<1>(p, pair, plen, tlen, sP, sT, nullptr);
    if (bits != 2) {
      // does not fit, or not A/C/G/T: the general kernel takes it
      if (lane == 0) {
        const int slot = atomicAdd(&p.counters[1], 1);
        p.retry_list[slot] = pair;
        p.status[pair] = ST_RETRY;
      }
      continue;
    }
    align_window<MODE>(p, pair, plen, tlen, sP, sT, rows, meta, vchunks, H, p.hist_cap);
    __syncthreads();
  }
}

// LDS bytes of the window kernel's rows and metadata
static inline size_t window_lds_bytes(int ring_rows) {
  return (size_t)ring_rows * (8 + 4 + WROW * 2);
}

}  // namespace wfa
