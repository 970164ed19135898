// wfa_biwfa.h — BiWFA (bidirectional WFA) on gfx950.
//
// The divide-and-conquer recursion of align_biwfa (src/aligner.rs:2563-2950)
// runs level by level from the host: every level, one launch searches the
// breakpoints of all pending sub-problems of all pairs (one workgroup per
// sub-problem: one wave with a heuristic, else kWideWaves waves sharing
// every row's chunks; forward and reverse searches interleaved as in
// bialign_find_breakpoint, src/aligner.rs:2958-3120), and one launch aligns
// the base cases (bialign_base, src/aligner.rs:3127-3202) with full storage
// and the component backtrace (backtrace_affine_c, src/backtrace.rs:492-635).
//
// Sub-problem sequences are staged into LDS (2-bit when every base is
// A/C/G/T, else bytes). The reverse search reads the forward sequences
// backwards. Wavefront rows live in a per-slot HBM ring of `scope` rows per
// component and direction (the reference's modular storage); their limits
// and initialized ranges live in LDS.
#pragma once

#include "wfa_align.h"

namespace wfa {

// Stage bases [start, start + len) of one stored sequence as 64-bit words
// (2-bit: base j at bits 2j; bytes: byte j at bits 8j), zero-padded to
// `words` words.
template <int NW>
static __device__ void stage_range(const uint8_t* src, int fmt, int start, int len, int bits,
                                   uint64_t* dst, int words) {
  const int lane = threadIdx.x;  // all 64 * NW threads
  constexpr int NT = 64 * NW;
  if (fmt == WFA_SEQ_PACKED2BITS) {
    const int last_byte = len > 0 ? (start + len - 1) >> 2 : -1;
    for (int w = lane; w < words; w += NT) {
      uint64_t out = 0;
      const int b0 = start + 32 * w;  // first base of the word
      const int nb = imin(32, len - 32 * w);
      if (nb > 0) {
        const int fb = b0 >> 2, sh = 2 * (b0 & 3);
        uint64_t lo = 0;
        for (int j = 0; j < 8; ++j)
          if (fb + j <= last_byte) lo |= (uint64_t)src[fb + j] << (8 * j);
        const uint64_t hi = (sh && fb + 8 <= last_byte) ? (uint64_t)src[fb + 8] : 0;
        out = sh ? (lo >> sh) | (hi << (64 - sh)) : lo;
        if (nb < 32) out &= (1ull << (2 * nb)) - 1;
      }
      dst[w] = out;
    }
  } else if (bits == 2) {
    for (int w = lane; w < words; w += NT) {
      uint64_t out = 0;
      const int nb = imin(32, len - 32 * w);
      for (int j = 0; j < nb; ++j) {
        const uint8_t c = src[start + 32 * w + j];
        const uint64_t code = c == 'A' ? 0 : c == 'C' ? 1 : c == 'G' ? 2 : 3;
        out |= code << (2 * j);
      }
      dst[w] = out;
    }
  } else {
    for (int w = lane; w < words; w += NT) {
      uint64_t out = 0;
      const int nb = imin(8, len - 8 * w);
      for (int j = 0; j < nb; ++j) out |= (uint64_t)src[start + 8 * w + j] << (8 * j);
      dst[w] = out;
    }
  }
}

// Whether bases [start, start + len) of an ASCII sequence are all A/C/G/T.
// With NW > 1 waves the answer is combined through red[0..NW).
template <int NW>
static __device__ bool all_acgt(const uint8_t* src, int start, int len, int* red) {
  bool bad = false;
  for (int i = threadIdx.x; i < len; i += 64 * NW) {
    const uint8_t c = src[start + i];
    bad |= !(c == 'A' || c == 'C' || c == 'G' || c == 'T');
  }
  bool any_bad = ballot(bad) != 0;
  if (NW > 1) {
    __syncthreads();  // earlier readers of red are done
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = any_bad;
    __syncthreads();
    for (int w = 0; w < NW; ++w) any_bad |= red[w] != 0;
  }
  return !any_bad;
}

// Stage one sub-problem's pattern and text. Returns the symbol width, or 0
// when they do not fit in `seq_words` words each.
template <int NW>
static __device__ int stage_job(const BiParams& bp, const BiJob& J, uint64_t* sP, uint64_t* sT,
                                int* red) {
  const KParams& p = bp.k;
  const uint8_t* P = p.seq + p.poff[J.pair];
  const uint8_t* T = p.seq + p.toff[J.pair];
  const int plen = J.pe - J.pb, tlen = J.te - J.tb;
  int bits = 2;
  if (p.fmt == WFA_SEQ_ASCII && !(all_acgt<NW>(P, J.pb, plen, red) && all_acgt<NW>(T, J.tb, tlen, red)))
    bits = 8;
  const int per = 64 / bits;
  const int pw = (plen + per - 1) / per + 2, tw = (tlen + per - 1) / per + 2;
  if (pw > bp.seq_words || tw > bp.seq_words) return 0;
  stage_range<NW>(P, p.fmt, J.pb, plen, bits, sP, pw);
  stage_range<NW>(T, p.fmt, J.tb, tlen, bits, sT, tw);
  __syncthreads();
  return bits;
}

// ---------------------------------------------------------------------
// One unidirectional wavefront aligner over a staged sub-problem: the
// forward or reverse search of a breakpoint, or (HISTORY) the base case,
// which also keeps every row it computes.
// ---------------------------------------------------------------------
// With NW > 1 waves (no heuristic, no history) the waves take every row's
// 64-diagonal chunks round-robin and combine trims and maxima through `red`
// (LDS): [0, 10 NW) trims, [10 NW, 11 NW) maxima, [11 NW, 14 NW) overlap
// scans.
template <int KIND, int BITS, bool HEUR, bool HISTORY, int NW = 1>
struct WfDir {
  static constexpr int NC = kind_components(KIND);
  static constexpr int kNW = NW;
  static constexpr int NT = 64 * NW;
  static_assert(NW == 1 || (!HEUR && !HISTORY), "several waves: no heuristic, no history");
  int* red;
  const KParams* p;
  Rows<int32_t> R;  // ring rows: component c, score s -> row c * scope + s % scope
  int scope;
  const uint64_t* sP;
  const uint64_t* sT;
  int plen, tlen;
  bool rev;         // the reverse search: sequences read backwards
  int heur, band_lo, band_hi;
  HeurState hs;
  int hist_lo, hist_hi;
  int nulls;
  int status, end_score, end_k, end_off;
  bool overflow;
  // HISTORY: headers hdr[s * NC + c] = {init lo, init hi, cell position, 0}
  int4* hdr;
  int hdr_cap;      // scores
  int32_t* cells;
  int64_t cap, hpos;
  // searches without a heuristic or a history extend M as they compute it
  // (step); the largest antidiagonal of the step's extended cells
  static constexpr bool FUSED = !HEUR && !HISTORY;
  int step_amax;

  __device__ int row(int c, int s) const { return c * scope + s % scope; }
  __device__ int cell(int c, int s, int k) const {
    return R.at(row(c, s) * R.wcap + (k & R.mask));
  }

  __device__ void set_hdr_empty(int s, int c) {
    if (HISTORY && threadIdx.x == 0) hdr[s * NC + c] = make_int4(1, 0, 0, 0);
  }

  // Keep row (c, s) over [lo, hi] in the history (cells as computed).
  __device__ void store(int c, int s, int lo, int hi) {
    if (!HISTORY) return;
    if (lo > hi) { set_hdr_empty(s, c); return; }
    if (hpos + (hi - lo + 1) > cap) { overflow = true; return; }
    for (int k = lo + (int)threadIdx.x; k <= hi; k += 64) cells[hpos + (k - lo)] = cell(c, s, k);
    if (threadIdx.x == 0) hdr[s * NC + c] = make_int4(lo, hi, (int)hpos, 0);
    hpos += hi - lo + 1;
  }

  // init_alignment_component (src/aligner.rs:3205-3266): cell (k = 0,
  // offset 0) in component `comp` at score 0; the end position starts at
  // (0, 0, 0).
  __device__ void init(int comp) {
    const int lane = threadIdx.x;
    for (int r = lane; r < NC * scope; r += NT) R.meta[r] = meta_absent();
    __syncthreads();
    if (lane == 0) {
      R.set(row(comp, 0) * R.wcap, 0);
      R.meta[row(comp, 0)] = make_int4(0, 0, 0, 0);
    }
    __syncthreads();
    hs.reset(HeurCfg{heur, p->hp0, p->hp1, p->hp2});
    hist_lo = -(plen + scope + 2);
    hist_hi = tlen + scope + 2;
    nulls = 0;
    status = WFA_STATUS_OK;
    end_score = 0; end_k = 0; end_off = 0;
    overflow = false;
    hpos = 0;
    if (HISTORY) {
      for (int c = 0; c < NC; ++c) {
        if (c == comp) store(c, 0, 0, 0);
        else set_hdr_empty(0, c);
      }
    }
  }

  __device__ void set_absent(int s) {
    __syncthreads();
    if (threadIdx.x == 0)
      for (int c = 0; c < NC; ++c) R.meta[row(c, s)] = meta_absent();
    for (int c = 0; c < NC; ++c) set_hdr_empty(s, c);
    __syncthreads();
  }

  // compute_step (src/aligner.rs:1769-1865, 2006-2193, 2201-2553) on the
  // ring; the caller checks `overflow`.
  __device__ void compute(int s) {
    const int lane = threadIdx.x;
    if (HISTORY && s >= hdr_cap) { overflow = true; return; }
    int lo = INT32_MAX, hi = INT32_MIN;
    bool any = false;
    // inputs: (limits, initialized range) of the rows each step reads
    struct In { int row, l, h, gl, gh; };
    auto load = [&](int c, int sc) -> In {
      In in{0, 1, 0, 1, 0};
      if (sc < 0) return in;
      in.row = row(c, sc);
      const int4 m = R.meta[in.row];
      in.l = uniform(m.x); in.h = uniform(m.y);
      in.gl = uniform(m.z); in.gh = uniform(m.w);
      if (in.l > in.h) { in.gl = 1; in.gh = 0; }
      return in;
    };
    if constexpr (KIND == KIND_EDIT) {
      In x = load(C_M, s - 1);
      __syncthreads();
      if (x.l > x.h) {
        nulls += 1;
        set_absent(s);
        return;
      }
      nulls = 0;
      // the two cells beyond each end of the previous row are NULL for
      // this step (src/aligner.rs:1796-1806)
      x.gl = x.l; x.gh = x.h;
      lo = imax(x.l - 1, hist_lo + 1);
      hi = imin(x.h + 1, hist_hi - 1);
      if (lo > hi) { set_absent(s); return; }
      if (hi - lo + 1 > R.wcap) { overflow = true; return; }
      const int rM = row(C_M, s);
      int m_lo = INT32_MAX, m_hi = INT32_MIN;
      for (int kb = lo + 64 * (lane >> 6); kb <= hi; kb += NT) {
        const int k = kb + (lane & 63);
        bool vm = false;
        if (k <= hi) {
          const int ins = rd(R, x.row, x.gl, x.gh, k - 1);
          const int del = rd(R, x.row, x.gl, x.gh, k + 1);
          int m;
          if (p->metric == WFA_METRIC_INDEL) {
            m = imax(del, ins + 1);
          } else {
            const int mis = rd(R, x.row, x.gl, x.gh, k);
            m = imax(del, imax(ins, mis) + 1);
          }
          vm = in_bounds(m, k, plen, tlen);
          R.set(rM * R.wcap + (k & R.mask), vm ? m : WFA_NULL);
        }
        const uint64_t b = ballot(vm);
        if (b) {
          if (m_lo == INT32_MAX) m_lo = kb + __builtin_ctzll(b);
          m_hi = kb + 63 - __builtin_clzll(b);
        }
      }
      if (NW > 1) {
        if ((lane & 63) == 0) { red[2 * (lane >> 6)] = m_lo; red[2 * (lane >> 6) + 1] = m_hi; }
        __syncthreads();
        m_lo = INT32_MAX; m_hi = INT32_MIN;
        for (int w = 0; w < NW; ++w) { m_lo = imin(m_lo, red[2 * w]); m_hi = imax(m_hi, red[2 * w + 1]); }
      }
      if (m_lo == INT32_MAX) { m_lo = 1; m_hi = 0; }
      __syncthreads();
      if (lane == 0) R.meta[rM] = make_int4(m_lo, m_hi, m_lo, m_hi);
      __syncthreads();
    } else {
      // gap-affine kinds, and gap-linear (compute_linear_step,
      // src/aligner.rs:1866-2002): M from M[s-x] and M[s-indel], whose ends
      // are NULL-filled only beyond their initialized ranges
      constexpr bool GAPS = KIND != KIND_LINEAR;
      const bool two = KIND == KIND_AFFINE2P;
      In mx = load(C_M, s - p->x), mo = load(C_M, s - (GAPS ? p->oe1 : p->o1));
      In i1{0, 1, 0, 1, 0}, d1{0, 1, 0, 1, 0};
      if (GAPS) {
        i1 = load(C_I1, s - p->e1);
        d1 = load(C_D1, s - p->e1);
      }
      In mo2{0, 1, 0, 1, 0}, i2{0, 1, 0, 1, 0}, d2{0, 1, 0, 1, 0};
      if (two) {
        mo2 = load(C_M, s - p->oe2);
        i2 = load(C_I2, s - p->e2);
        d2 = load(C_D2, s - p->e2);
      }
      __syncthreads();
      // Missing and empty inputs read as the null wavefront; with a
      // heuristic their bounds (1, -1) enter the range union
      // (src/aligner.rs:2045-2062, 2265-2289), which without one is a no-op.
      auto bound = [&](const In& in, int dlo, int dhi) {
        int l = in.l, h = in.h;
        if (l <= h) any = true;
        else if (!HEUR) return;
        else { l = 1; h = -1; }
        lo = imin(lo, l + dlo);
        hi = imax(hi, h + dhi);
      };
      bound(mx, 0, 0);
      bound(mo, -1, 1);
      if (GAPS) {
        bound(i1, 1, 1);
        bound(d1, -1, -1);
      }
      if (two) {
        bound(mo2, -1, 1);
        bound(i2, 1, 1);
        bound(d2, -1, -1);
      }
      if (!any) {
        nulls += 1;
        set_absent(s);
        return;
      }
      nulls = 0;
      if (heur == WFA_HEURISTIC_BANDED_STATIC || heur == WFA_HEURISTIC_BANDED_ADAPTIVE) {
        lo = imax(lo, band_lo);
        hi = imin(hi, band_hi);
      }
      lo = imax(lo, hist_lo + 1);
      hi = imin(hi, hist_hi - 1);
      if (lo > hi) { set_absent(s); return; }
      if (hi - lo + 1 > R.wcap) { overflow = true; return; }
      const int rM = row(C_M, s);
      const int rI1 = GAPS ? row(C_I1, s) : 0, rD1 = GAPS ? row(C_D1, s) : 0;
      const int rI2 = two ? row(C_I2, s) : 0, rD2 = two ? row(C_D2, s) : 0;
      int lim[5][2];
      for (int c = 0; c < 5; ++c) { lim[c][0] = INT32_MAX; lim[c][1] = INT32_MIN; }
      int ext_max = 0;  // (FUSED) antidiagonals of the extended M cells
      for (int kb = lo + 64 * (lane >> 6); kb <= hi; kb += NT) {
        const int k = kb + (lane & 63);
        bool v[5] = {false, false, false, false, false};
        if (k <= hi) {
          int iv = rd(R, mo.row, mo.gl, mo.gh, k - 1);
          int dv = rd(R, mo.row, mo.gl, mo.gh, k + 1);
          if (GAPS) {
            iv = imax(iv, rd(R, i1.row, i1.gl, i1.gh, k - 1));
            dv = imax(dv, rd(R, d1.row, d1.gl, d1.gh, k + 1));
          }
          iv += 1;
          const int mis = rd(R, mx.row, mx.gl, mx.gh, k) + 1;
          int ins = iv, del = dv;
          if (two) {
            const int iv2 =
                imax(rd(R, mo2.row, mo2.gl, mo2.gh, k - 1), rd(R, i2.row, i2.gl, i2.gh, k - 1)) + 1;
            const int dv2 =
                imax(rd(R, mo2.row, mo2.gl, mo2.gh, k + 1), rd(R, d2.row, d2.gl, d2.gh, k + 1));
            R.set(rI2 * R.wcap + (k & R.mask), iv2);
            R.set(rD2 * R.wcap + (k & R.mask), dv2);
            v[C_I2] = in_bounds(iv2, k, plen, tlen);
            v[C_D2] = in_bounds(dv2, k, plen, tlen);
            ins = imax(ins, iv2);
            del = imax(del, dv2);
          }
          int m = imax(mis, imax(ins, del));
          v[C_M] = in_bounds(m, k, plen, tlen);
          if (!v[C_M]) {
            m = WFA_NULL;
          } else if constexpr (FUSED) {
            const int vv = m - k;
            m += rev ? backward_matches<BITS>(sP, sT, plen - vv, tlen - m)
                     : lcp<BITS>(sP, sT, vv, m, plen, tlen);
            ext_max = imax(ext_max, 2 * m - k);
          }
          R.set(rM * R.wcap + (k & R.mask), m);
          if (GAPS) {
            R.set(rI1 * R.wcap + (k & R.mask), iv);
            R.set(rD1 * R.wcap + (k & R.mask), dv);
            v[C_I1] = in_bounds(iv, k, plen, tlen);
            v[C_D1] = in_bounds(dv, k, plen, tlen);
          }
        }
        for (int c = 0; c < NC; ++c) {
          const uint64_t b = ballot(v[c]);
          if (b) {
            if (lim[c][0] == INT32_MAX) lim[c][0] = kb + __builtin_ctzll(b);
            lim[c][1] = kb + 63 - __builtin_clzll(b);
          }
        }
      }
      int amax = FUSED ? uniform(wave_max(ext_max)) : 0;
      if (NW > 1) {
        if ((lane & 63) == 0) {
          for (int c = 0; c < NC; ++c) {
            red[10 * (lane >> 6) + 2 * c] = lim[c][0];
            red[10 * (lane >> 6) + 2 * c + 1] = lim[c][1];
          }
          if (FUSED) red[10 * NW + (lane >> 6)] = amax;
        }
        __syncthreads();
        for (int c = 0; c < NC; ++c) {
          lim[c][0] = INT32_MAX; lim[c][1] = INT32_MIN;
          for (int w = 0; w < NW; ++w) {
            lim[c][0] = imin(lim[c][0], red[10 * w + 2 * c]);
            lim[c][1] = imax(lim[c][1], red[10 * w + 2 * c + 1]);
          }
        }
        if (FUSED)
          for (int w = 0; w < NW; ++w) amax = imax(amax, red[10 * NW + w]);
      }
      step_amax = amax;
      for (int c = 0; c < NC; ++c)
        if (lim[c][0] == INT32_MAX) { lim[c][0] = 1; lim[c][1] = 0; }
      __syncthreads();
      if (lane == 0)
        for (int c = 0; c < NC; ++c)
          R.meta[row(c, s)] = make_int4(lim[c][0], lim[c][1], lim[c][0], lim[c][1]);
      // the gap components are kept as computed; M after its extension
      for (int c = 1; c < NC; ++c) store(c, s, lim[c][0], lim[c][1]);
      __syncthreads();
    }
  }

  // One score of a search: compute(s), then extend(s); the caller checks
  // `overflow`. Edit and indel searches without a heuristic or a history
  // extend each M cell as they compute it, and write the row once: the
  // extension of a cell reads only that cell and the sequences, and the
  // trims test the unextended offsets, which the extension keeps inside
  // the sequences.
  __device__ bool step(int s, int comp_end, int* max_ak) {
    if constexpr (KIND == KIND_EDIT && FUSED) {
      return edit_step(s, comp_end, max_ak);
    } else if constexpr (FUSED) {
      // compute(s) extended M already; what extend(s) does after that
      step_amax = 0;
      compute(s);
      if (overflow) return false;
      *max_ak = 0;
      const int4 mm = R.meta[row(C_M, s)];
      if (uniform(mm.x) == META_ABSENT) {
        if (nulls > scope) {
          status = WFA_STATUS_END_UNREACHABLE;
          return true;
        }
        return false;
      }
      *max_ak = step_amax;
      return end_reached(s, comp_end);
    } else {
      compute(s);
      if (overflow) return false;
      return extend(s, comp_end, max_ak);
    }
  }

  // check_component_termination (src/aligner.rs:3269-3302) at score s
  __device__ bool end_reached(int s, int comp_end) {
    const int4 ce = R.meta[row(comp_end, s)];
    const int clo = uniform(ce.x), chi = uniform(ce.y);
    const int k_end = tlen - plen;
    if (clo != META_ABSENT && k_end >= clo && k_end <= chi &&
        uniform(cell(comp_end, s, k_end)) >= tlen) {
      status = WFA_STATUS_END_REACHED;
      end_score = s; end_k = k_end; end_off = tlen;
      return true;
    }
    return false;
  }

  // compute(s) and extend(s) of the edit/indel step in one pass
  // (src/aligner.rs:1769-1865, 3516-3554).
  __device__ bool edit_step(int s, int comp_end, int* max_ak) {
    const int lane = threadIdx.x;
    *max_ak = 0;
    const int xrow = row(C_M, s - 1);
    const int4 xm = R.meta[xrow];
    const int xl = uniform(xm.x), xh = uniform(xm.y);
    __syncthreads();
    if (xl > xh) {
      nulls += 1;
      set_absent(s);
      return nulls > scope ? (status = WFA_STATUS_END_UNREACHABLE, true) : false;
    }
    nulls = 0;
    // the two cells beyond each end of the previous row are NULL for this
    // step (src/aligner.rs:1796-1806): reads are guarded by [xl, xh]
    const int lo = imax(xl - 1, hist_lo + 1);
    const int hi = imin(xh + 1, hist_hi - 1);
    if (lo > hi) {
      set_absent(s);
      return false;
    }
    if (hi - lo + 1 > R.wcap) { overflow = true; return false; }
    const int rM = row(C_M, s);
    int m_lo = INT32_MAX, m_hi = INT32_MIN, mx = 0;
    for (int kb = lo + 64 * (lane >> 6); kb <= hi; kb += NT) {
      const int k = kb + (lane & 63);
      bool vm = false;
      if (k <= hi) {
        const int ins = rd(R, xrow, xl, xh, k - 1);
        const int del = rd(R, xrow, xl, xh, k + 1);
        int m;
        if (p->metric == WFA_METRIC_INDEL) {
          m = imax(del, ins + 1);
        } else {
          const int mis = rd(R, xrow, xl, xh, k);
          m = imax(del, imax(ins, mis) + 1);
        }
        vm = in_bounds(m, k, plen, tlen);
        if (vm) {
          const int v = m - k;
          m += rev ? backward_matches<BITS>(sP, sT, plen - v, tlen - m)
                   : lcp<BITS>(sP, sT, v, m, plen, tlen);
          mx = imax(mx, 2 * m - k);
        }
        R.set(rM * R.wcap + (k & R.mask), vm ? m : WFA_NULL);
      }
      const uint64_t b = ballot(vm);
      if (b) {
        if (m_lo == INT32_MAX) m_lo = kb + __builtin_ctzll(b);
        m_hi = kb + 63 - __builtin_clzll(b);
      }
    }
    int amax = uniform(wave_max(mx));
    if (NW > 1) {
      if ((lane & 63) == 0) {
        red[2 * (lane >> 6)] = m_lo;
        red[2 * (lane >> 6) + 1] = m_hi;
        red[10 * NW + (lane >> 6)] = amax;
      }
      __syncthreads();
      m_lo = INT32_MAX; m_hi = INT32_MIN;
      for (int w = 0; w < NW; ++w) {
        m_lo = imin(m_lo, red[2 * w]);
        m_hi = imax(m_hi, red[2 * w + 1]);
        amax = imax(amax, red[10 * NW + w]);
      }
    }
    if (m_lo == INT32_MAX) { m_lo = 1; m_hi = 0; }
    __syncthreads();
    if (lane == 0) R.meta[rM] = make_int4(m_lo, m_hi, m_lo, m_hi);
    __syncthreads();
    *max_ak = amax;
    // check_component_termination (src/aligner.rs:3269-3302)
    const int k_end = tlen - plen;
    if (comp_end == C_M && k_end >= m_lo && k_end <= m_hi &&
        uniform(cell(C_M, s, k_end)) >= tlen) {
      status = WFA_STATUS_END_REACHED;
      end_score = s; end_k = k_end; end_off = tlen;
      return true;
    }
    return false;
  }

  // extend_biwfa_step (src/aligner.rs:3516-3554). Returns true when the
  // search is done (end reached on `comp_end`, or unreachable); *max_ak is
  // the largest antidiagonal 2*offset - k of an extended cell (0 if none).
  __device__ bool extend(int s, int comp_end, int* max_ak) {
    const int lane = threadIdx.x;
    *max_ak = 0;
    const int rM = row(C_M, s);
    const int4 mm = R.meta[rM];
    const int lo = uniform(mm.x), hi = uniform(mm.y);
    if (lo == META_ABSENT) {
      if (nulls > scope) {
        status = WFA_STATUS_END_UNREACHABLE;
        return true;
      }
      return false;
    }
    int mx = 0;
    for (int kb = lo + 64 * (lane >> 6); kb <= hi; kb += NT) {
      const int k = kb + (lane & 63);
      if (k <= hi) {
        const int idx = rM * R.wcap + (k & R.mask);
        int off = R.at(idx);
        if (off >= 0) {
          const int v = off - k;
          off += rev ? backward_matches<BITS>(sP, sT, plen - v, tlen - off)
                     : lcp<BITS>(sP, sT, v, off, plen, tlen);
          R.set(idx, off);
          mx = imax(mx, 2 * off - k);
        }
      }
    }
    int amax = uniform(wave_max(mx));
    if (NW > 1) {
      if ((lane & 63) == 0) red[10 * NW + (lane >> 6)] = amax;
      __syncthreads();
      for (int w = 0; w < NW; ++w) amax = imax(amax, red[10 * NW + w]);
    }
    *max_ak = amax;
    __syncthreads();
    store(C_M, s, lo, hi);
    if (end_reached(s, comp_end)) return true;
    if (HEUR && heur != WFA_HEURISTIC_NONE) cutoff(s, lo, hi, mm);
    return false;
  }

  // apply_heuristic_cutoff (src/aligner.rs:1644-1766) on M[s]
  __device__ void cutoff(int s, int lo, int hi, int4 mm) {
    const int rM = row(C_M, s);
    int nlo = lo, nhi = hi;
    if (!heuristic_limits(*p, HeurCfg{heur, p->hp0, p->hp1, p->hp2}, band_lo, band_hi, hs, R,
                          rM, lo, hi, s, plen, tlen, 0, 0, &nlo, &nhi))
      return;
    if (nlo == lo && nhi == hi) return;
    const bool empty = nlo > nhi;
    __syncthreads();
    if (threadIdx.x == 0) {
      R.meta[rM] = empty ? make_int4(nlo, nhi, 1, 0) : make_int4(nlo, nhi, mm.z, mm.w);
      if (HISTORY && empty) hdr[s * NC + C_M] = make_int4(1, 0, 0, 0);
      for (int c = 1; c < NC; ++c) {
        const int r = row(c, s);
        const int4 g = R.meta[r];
        if (g.x > g.y) continue;
        const int l = imax(g.x, nlo), h = imin(g.y, nhi);
        if (l > h) {
          R.meta[r] = make_int4(l, h, 1, 0);
          if (HISTORY) hdr[s * NC + c] = make_int4(1, 0, 0, 0);
        } else {
          R.meta[r] = make_int4(l, h, g.z, g.w);
        }
      }
    }
    if (HISTORY && KIND == KIND_EDIT && !empty) {
      // the next edit step nulls the two cells beyond each end of M[s] in
      // place (src/aligner.rs:1796-1806); the backtrace sees them
      const int4 hd = hdr[s];
      const int lane = threadIdx.x;
      if (lane < 4) {
        const int k = lane < 2 ? nlo - 2 + lane : nhi + lane - 1;
        if (k >= hd.x && k <= hd.y) cells[hd.z + (k - hd.x)] = WFA_NULL;
      }
    }
    __syncthreads();
  }
};

// ---------------------------------------------------------------------
// Breakpoint search
// ---------------------------------------------------------------------
struct Bp {
  int score, score_fwd, score_rev, k_fwd, off_fwd, comp;
};

static __device__ __forceinline__ int k_inverse(int k, int plen, int tlen) {
  return (tlen - plen) - k;
}

// breakpoint_m2m / breakpoint_indel2indel (src/bialign.rs:79-227) with
// scan_overlap (:244-295): the first diagonal of row (c, s0) of `a`, in
// ascending k, whose text position and that of row (c, s1) of `b` (in b's
// reversed diagonals) add up to the text length; an indel breakpoint also
// needs the forward cell inside the sequences.
template <class D>
static __device__ void bi_scan(const D& a, const D& b, int c, int s0, int s1, int go, bool fwd,
                               Bp& bp) {
  if (s0 + s1 - go >= bp.score) return;
  const int plen = a.plen, tlen = a.tlen;
  const int4 ma = a.R.meta[a.row(c, s0)], mb = b.R.meta[b.row(c, s1)];
  const int alo = uniform(ma.x), ahi = uniform(ma.y), blo = uniform(mb.x), bhi = uniform(mb.y);
  if (alo > ahi || blo > bhi) return;
  const int lo1 = k_inverse(bhi, plen, tlen), hi1 = k_inverse(blo, plen, tlen);
  if (hi1 < alo || ahi < lo1) return;
  const int lo = imax(alo, lo1), hi = imin(ahi, hi1);
  constexpr int NW = D::kNW;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // every wave finds the first hit of its chunks; the lowest one wins
  int hk = INT32_MAX, ho0 = 0, ho1 = 0;
  for (int kb = lo + 64 * wave; kb <= hi; kb += 64 * NW) {
    const int k0 = kb + lane;
    bool ok = false;
    int o0 = 0, o1 = 0;
    if (k0 <= hi) {
      const int k1 = k_inverse(k0, plen, tlen);
      o0 = a.cell(c, s0, k0);
      o1 = b.cell(c, s1, k1);
      ok = o0 + o1 >= tlen;
      if (ok && c != C_M) {
        const int ck = fwd ? k0 : k1, ch = fwd ? o0 : o1;
        ok = !(ch - ck > plen || ch > tlen);
      }
    }
    const uint64_t m = ballot(ok);
    if (m) {
      const int j = __builtin_ctzll(m);
      hk = kb + j;
      ho0 = uniform(__shfl(o0, j));
      ho1 = uniform(__shfl(o1, j));
      break;
    }
  }
  if (NW > 1) {
    int* x = a.red + 11 * NW;
    if (lane == 0) { x[3 * wave] = hk; x[3 * wave + 1] = ho0; x[3 * wave + 2] = ho1; }
    __syncthreads();
    hk = INT32_MAX;
    for (int w = 0; w < NW; ++w)
      if (x[3 * w] < hk) { hk = x[3 * w]; ho0 = x[3 * w + 1]; ho1 = x[3 * w + 2]; }
    __syncthreads();  // the next scan reuses x
  }
  if (hk == INT32_MAX) return;
  if (fwd) {
    bp.score_fwd = s0; bp.score_rev = s1; bp.k_fwd = hk; bp.off_fwd = ho0;
  } else {
    bp.score_fwd = s1; bp.score_rev = s0;
    bp.k_fwd = k_inverse(hk, plen, tlen); bp.off_fwd = ho1;
  }
  bp.score = s0 + s1 - go;
  bp.comp = c;
}

// bialign_overlap (src/bialign.rs:362-508): a at s0 against b's last
// `scope` scores from s1 down; D2, I2, D1, I1, then M at each.
template <int KIND, class D>
static __device__ void bi_overlap(const D& a, const D& b, int s0, int s1, bool fwd, Bp& bp) {
  const KParams& p = *a.p;
  const int4 m0 = a.R.meta[a.row(C_M, s0)];
  if (uniform(m0.x) > uniform(m0.y)) return;  // absent or empty M
  for (int i = 0; i < b.scope; ++i) {
    const int si = s1 - i;
    if (si < 0) break;
    if (KIND == KIND_AFFINE2P && s0 + si - p.o2 < bp.score) {
      bi_scan(a, b, C_D2, s0, si, p.o2, fwd, bp);
      bi_scan(a, b, C_I2, s0, si, p.o2, fwd, bp);
    }
    if ((KIND == KIND_AFFINE || KIND == KIND_AFFINE2P) && s0 + si - p.o1 < bp.score) {
      bi_scan(a, b, C_D1, s0, si, p.o1, fwd, bp);
      bi_scan(a, b, C_I1, s0, si, p.o1, fwd, bp);
    }
    if (s0 + si >= bp.score) continue;
    bi_scan(a, b, C_M, s0, si, 0, fwd, bp);
  }
}

// bialign_find_breakpoint (src/aligner.rs:2958-3120) for one sub-problem.
template <int KIND, int BITS, bool HEUR, int NW>
static __device__ BiResult find_breakpoint(const BiParams& bp, const BiJob& J, uint64_t* sP,
                                           uint64_t* sT, int4* metaF, int4* metaR,
                                           int32_t* ringF, int32_t* ringR, int* red) {
  using Dir = WfDir<KIND, BITS, HEUR, false, NW>;
  const KParams& p = bp.k;
  const int plen = J.pe - J.pb, tlen = J.te - J.tb;
  Dir f, r;
  auto setup = [&](Dir& d, int4* meta, int32_t* ring, bool rev) {
    d.p = &p;
    d.R.ring = ring; d.R.meta = meta; d.R.wcap = bp.wcap; d.R.mask = bp.wcap - 1;
    d.R.cap = kind_components(KIND) * p.scope * bp.wcap;
    d.R.chk = bp.counters;
    d.scope = p.scope;
    d.sP = sP; d.sT = sT; d.plen = plen; d.tlen = tlen; d.rev = rev;
    d.heur = HEUR ? p.heuristic : WFA_HEURISTIC_NONE;
    d.hdr = nullptr; d.cells = nullptr; d.hdr_cap = 0; d.cap = 0;
    d.red = red;
  };
  setup(f, metaF, ringF, false);
  setup(r, metaR, ringR, true);
  // wavefront_bialign_set_subsidiary_band: the band in each search's own
  // diagonals (src/aligner.rs:2866-2874)
  const int shift = J.tb - J.pb, diag = tlen - plen;
  f.band_lo = p.hp0 - shift; f.band_hi = p.hp1 - shift;
  r.band_lo = diag - f.band_hi; r.band_hi = diag - f.band_lo;
  f.init(J.cbeg);
  r.init(J.cend);
  const int fwd_end = J.cend, rev_end = J.cbeg;
  BiResult res;
  res.status = BI_OK;
  res.score = res.score_fwd = res.score_rev = res.k_fwd = res.off_fwd = res.end_score = 0;
  res.comp = C_M;
  auto quit = [&](const Dir& d) {
    if (d.status == WFA_STATUS_END_REACHED) {
      res.status = BI_END_REACHED;
      res.end_score = d.end_score;
    } else {
      res.status = BI_UNREACHABLE;
    }
    return res;
  };
  auto retry = [&]() {
    res.status = BI_RETRY;
    return res;
  };
  const int gap_open = KIND == KIND_AFFINE ? p.o1 : KIND == KIND_AFFINE2P ? imax(p.o1, p.o2) : 0;
  const int max_ad = plen + tlen - 1;
  // Both searches together never need more than twice the score of the
  // all-gap alignment (plus the null steps that end a search): a bound that
  // keeps every wavefront finite whatever the input.
  const int64_t pen_max = imax(imax(p.x, KIND == KIND_LINEAR ? p.o1 : p.oe1),
                               KIND == KIND_AFFINE2P ? p.oe2 : 0);
  const int64_t step_bound = 4 * ((int64_t)(plen + tlen + 2) * imax((int)pen_max, 1)) + 4 * p.scope;
  int s_fwd = 0, s_rev = 0, ak = 0, fmax = 0, rmax = 0;
  auto runaway = [&]() { return (int64_t)s_fwd + s_rev > step_bound; };
  bool last_fwd = false;
  if (f.extend(0, fwd_end, &fmax)) return quit(f);
  if (r.extend(0, rev_end, &rmax)) return quit(r);
  while (fmax + rmax < max_ad) {
    s_fwd += 1;
    bool done = f.step(s_fwd, fwd_end, &ak);
    if (f.overflow) return retry();
    fmax = imax(fmax, ak);
    last_fwd = true;
    if (done) return quit(f);
    if (fmax + rmax >= max_ad) break;
    s_rev += 1;
    done = r.step(s_rev, rev_end, &ak);
    if (r.overflow) return retry();
    rmax = imax(rmax, ak);
    last_fwd = false;
    if (done) return quit(r);
    if (s_fwd + s_rev >= p.max_steps) { res.status = BI_MAX_STEPS; return res; }
    if (runaway()) { res.status = BI_UNREACHABLE; return res; }
  }
  Bp b;
  b.score = INT32_MAX;
  b.score_fwd = b.score_rev = b.k_fwd = b.off_fwd = 0;
  b.comp = C_M;
  const int scope = p.scope;
  for (;;) {
    if (last_fwd) {
      const int min_rev = s_rev > scope - 1 ? s_rev - (scope - 1) : 0;
      if (s_fwd + min_rev - gap_open >= b.score) break;
      bi_overlap<KIND>(f, r, s_fwd, s_rev, true, b);
      s_rev += 1;
      const bool rdone = r.step(s_rev, rev_end, &ak);
      if (r.overflow) return retry();
      if (rdone) return quit(r);
    }
    const int min_fwd = s_fwd > scope - 1 ? s_fwd - (scope - 1) : 0;
    if (min_fwd + s_rev - gap_open >= b.score) break;
    bi_overlap<KIND>(r, f, s_rev, s_fwd, false, b);
    s_fwd += 1;
    const bool fdone = f.step(s_fwd, fwd_end, &ak);
    if (f.overflow) return retry();
    if (fdone) return quit(f);
    if (s_fwd + s_rev >= p.max_steps) { res.status = BI_MAX_STEPS; return res; }
    if (runaway()) { res.status = BI_UNREACHABLE; return res; }
    last_fwd = true;
  }
  res.score = b.score;
  res.score_fwd = b.score_fwd;
  res.score_rev = b.score_rev;
  res.k_fwd = b.k_fwd;
  res.off_fwd = b.off_fwd;
  res.comp = b.comp;
  return res;
}

template <int KIND, bool HEUR, int NW>
__global__ void __launch_bounds__(64 * NW) biwfa_breakpoint_kernel(BiParams bp) {
  extern __shared__ uint64_t smem[];
  const int lane = threadIdx.x;
  constexpr int NC = kind_components(KIND);
  uint64_t* sP = smem;
  uint64_t* sT = smem + bp.seq_words;
  int4* metaF = (int4*)(sT + bp.seq_words);
  int4* metaR = metaF + NC * bp.k.scope;
  const int64_t dir_ints = (int64_t)NC * bp.k.scope * bp.wcap;
  int32_t* ringF = bp.ring + (int64_t)blockIdx.x * 2 * dir_ints;
  int32_t* ringR = ringF + dir_ints;
  // cross-wave reductions: 16 ints per wave after the metadata
  int* red = (int*)(metaR + NC * bp.k.scope);
  for (;;) {
    int item = 0;
    if (NW == 1) {
      if (lane == 0) item = atomicAdd(&bp.counters[0], 1);
      item = uniform(__shfl(item, 0));
    } else {
      if (lane == 0) red[15 * NW] = atomicAdd(&bp.counters[0], 1);
      __syncthreads();
      item = uniform(red[15 * NW]);
      __syncthreads();
    }
    if (item >= bp.njobs) break;
    const BiJob J = bp.jobs[item];
    const int bits = stage_job<NW>(bp, J, sP, sT, red);
    BiResult res;
    if (bits == 0) {
      res.status = BI_RETRY;
      res.score = res.score_fwd = res.score_rev = res.k_fwd = res.off_fwd = res.end_score = 0;
      res.comp = 0;
    } else if (bits == 2) {
      res = find_breakpoint<KIND, 2, HEUR, NW>(bp, J, sP, sT, metaF, metaR, ringF, ringR, red);
    } else {
      res = find_breakpoint<KIND, 8, HEUR, NW>(bp, J, sP, sT, metaF, metaR, ringF, ringR, red);
    }
    if (lane == 0) bp.results[item] = res;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------
// Base case: unidirectional WFA between components with every row kept,
// then the component backtrace.
// ---------------------------------------------------------------------
#define BT_I1_OPEN 1
#define BT_I1_EXT 2
#define BT_I2_OPEN 3
#define BT_I2_EXT 4
#define BT_D1_OPEN 5
#define BT_D1_EXT 6
#define BT_D2_OPEN 7
#define BT_D2_EXT 8
#define BT_M 9

// backtrace_affine_c (src/backtrace.rs:492-635): gap-affine / 2-piece
// CIGAR from the end cell in component `comp_end`, every predecessor read
// from its source row (bt_* helpers, src/backtrace.rs:45-449), the largest
// packed (offset << 4 | type) winning. Writes ops backwards ending at
// out[cap-1]; returns their number.
template <int KIND>
static __device__ int backtrace_classic(const KParams& p, const int4* hdr, const int32_t* cells,
                                        int nscores, int plen, int tlen, int score, int k,
                                        int offset, int comp_end, uint8_t* out, int cap) {
  constexpr int NC = KIND == KIND_AFFINE ? 3 : 5;
  const int64_t NEG = WFA_NULL;
  // the cell of row (c, s) at kk if inside its initialized range
  auto get = [&](int c, int s, int kk, int* v) -> bool {
    if (s < 0 || s >= nscores) return false;
    const int4 h = hdr[s * NC + c];
    if (kk < h.x || kk > h.y) return false;
    *v = cells[h.z + (kk - h.x)];
    return true;
  };
  auto ins = [&](int c, int s, int kk, int64_t t) -> int64_t {
    int v;
    return get(c, s, kk - 1, &v) ? ((int64_t)(v + 1) * 16 + t) : NEG;
  };
  auto del = [&](int c, int s, int kk, int64_t t) -> int64_t {
    int v;
    return get(c, s, kk + 1, &v) ? ((int64_t)v * 16 + t) : NEG;
  };
  auto mx2 = [](int64_t a, int64_t b) { return a > b ? a : b; };
  int pos = cap - 1;
  int matrix = comp_end;
  int v = offset - k, h = offset;
  if (comp_end == C_M) {
    write_ops(out, pos, 'D', plen - v);
    write_ops(out, pos, 'I', tlen - h);
  }
  while (v > 0 && h > 0 && score > 0) {
    const int mis = score - p.x;
    const int go1 = score - p.o1 - p.e1, ge1 = score - p.e1;
    const int go2 = score - p.o2 - p.e2, ge2 = score - p.e2;
    int64_t best;
    if (matrix == C_M) {
      int mv;
      const int64_t misms = get(C_M, mis, k, &mv) ? ((int64_t)(mv + 1) * 16 + BT_M) : NEG;
      const int64_t i1 = mx2(ins(C_M, go1, k, BT_I1_OPEN), ins(C_I1, ge1, k, BT_I1_EXT));
      const int64_t d1 = mx2(del(C_M, go1, k, BT_D1_OPEN), del(C_D1, ge1, k, BT_D1_EXT));
      if (KIND == KIND_AFFINE2P) {
        const int64_t i2 = mx2(ins(C_M, go2, k, BT_I2_OPEN), ins(C_I2, ge2, k, BT_I2_EXT));
        const int64_t d2 = mx2(del(C_M, go2, k, BT_D2_OPEN), del(C_D2, ge2, k, BT_D2_EXT));
        best = mx2(misms, mx2(mx2(i1, i2), mx2(d1, d2)));
      } else {
        best = mx2(misms, mx2(i1, d1));
      }
    } else if (matrix == C_I1) {
      best = mx2(ins(C_M, go1, k, BT_I1_OPEN), ins(C_I1, ge1, k, BT_I1_EXT));
    } else if (matrix == C_D1) {
      best = mx2(del(C_M, go1, k, BT_D1_OPEN), del(C_D1, ge1, k, BT_D1_EXT));
    } else if (matrix == C_I2) {
      best = mx2(ins(C_M, go2, k, BT_I2_OPEN), ins(C_I2, ge2, k, BT_I2_EXT));
    } else {
      best = mx2(del(C_M, go2, k, BT_D2_OPEN), del(C_D2, ge2, k, BT_D2_EXT));
    }
    if (best < 0) break;
    if (matrix == C_M) {
      const int moff = (int)(best >> 4);
      const int nm = offset - moff;
      // write_matches of a negative count drops the last ops written
      if (nm < 0) pos = imin(pos - nm, cap - 1);
      else write_ops(out, pos, 'M', nm);
      offset = moff;
      v = offset - k; h = offset;
      if (v <= 0 || h <= 0) break;
    }
    const int t = (int)(best & 0xF);
    switch (t) {
      case BT_M: score = mis; matrix = C_M; break;
      case BT_I1_OPEN: case BT_D1_OPEN: score = go1; matrix = C_M; break;
      case BT_I2_OPEN: case BT_D2_OPEN: score = go2; matrix = C_M; break;
      case BT_I1_EXT: score = ge1; matrix = C_I1; break;
      case BT_I2_EXT: score = ge2; matrix = C_I2; break;
      case BT_D1_EXT: score = ge1; matrix = C_D1; break;
      default: score = ge2; matrix = C_D2; break;
    }
    if (t == BT_M) {
      write_ops(out, pos, 'X', 1);
      offset -= 1;
    } else if (t == BT_I1_OPEN || t == BT_I1_EXT || t == BT_I2_OPEN || t == BT_I2_EXT) {
      write_ops(out, pos, 'I', 1);
      k -= 1;
      offset -= 1;
    } else {
      write_ops(out, pos, 'D', 1);
      k += 1;
    }
    v = offset - k; h = offset;
  }
  if (matrix == C_M) {
    if (v > 0 && h > 0) {
      const int nm = imin(v, h);
      write_ops(out, pos, 'M', nm);
      v -= nm; h -= nm;
    }
    write_ops(out, pos, 'D', v > 0 ? v : 0);
    write_ops(out, pos, 'I', h > 0 ? h : 0);
  }
  return cap - 1 - pos;
}

// bialign_base (src/aligner.rs:3127-3202) for one sub-problem.
template <int KIND, int BITS, bool HEUR>
static __device__ void base_case(const BiParams& bp, int item, const BiJob& J, uint64_t* sP,
                                 uint64_t* sT, int4* meta, int32_t* ring, int4* hdr,
                                 int32_t* cells) {
  using Dir = WfDir<KIND, BITS, HEUR, true>;
  const KParams& p = bp.k;
  const int lane = threadIdx.x;
  const int plen = J.pe - J.pb, tlen = J.te - J.tb;
  Dir d;
  d.p = &p;
  d.R.ring = ring; d.R.meta = meta; d.R.wcap = bp.wcap; d.R.mask = bp.wcap - 1;
  d.R.cap = kind_components(KIND) * p.scope * bp.wcap;
  d.R.chk = bp.counters;
  d.scope = p.scope;
  d.sP = sP; d.sT = sT; d.plen = plen; d.tlen = tlen; d.rev = false;
  d.heur = HEUR ? p.heuristic : WFA_HEURISTIC_NONE;
  d.band_lo = p.hp0 - (J.tb - J.pb);
  d.band_hi = p.hp1 - (J.tb - J.pb);
  d.hdr = hdr; d.hdr_cap = bp.hdr_cap; d.cells = cells; d.cap = bp.hist_cap;
  d.init(J.cbeg);
  const int nwf = high_mem_nwf(p, plen, tlen);
  int score = 0, ak = 0;
  bool ok = !d.overflow;
  while (ok) {
    if (d.extend(score, J.cend, &ak)) break;
    if (d.overflow) { ok = false; break; }
    score += 1;
    if (score >= nwf) { d.status = WFA_STATUS_MAX_STEPS_REACHED; break; }
    d.compute(score);
    if (d.overflow) { ok = false; break; }
  }
  if (d.overflow) ok = false;
  uint8_t* out = bp.frag + bp.frag_off[item];
  const int cap = plen + tlen;
  int n = 0, st = BI_UNATTAINABLE;
  if (ok) {
    __threadfence_block();
    __syncthreads();
    int from_score = score, from_k = d.end_k, from_off = d.end_off;
    bool trace = true;
    if (d.status == WFA_STATUS_END_REACHED) {
      st = BI_OK;
    } else if (d.status == WFA_STATUS_END_UNREACHABLE) {
      // the default end position (0, 0, 0): C backtraces from it and
      // max-trims (src/aligner.rs:3178-3190)
      from_score = d.end_score >= 0 ? d.end_score : score;
    } else {
      trace = false;  // max steps: the CIGAR is cleared
    }
    if (trace) {
      if constexpr (KIND == KIND_EDIT || KIND == KIND_LINEAR) {
        // one contiguous history per sub-problem: chunk 0 of size 0
        Hist H;
        H.hdr = HBuf<int4>{hdr, (int64_t)bp.hdr_cap * kind_components(KIND), bp.counters};
        H.cells = HBuf<int32_t>{cells, bp.hist_cap, bp.counters};
        H.chunk = 0;
        n = backtrace_linear<BITS>(p, H, score + 1, plen, tlen, from_score, from_k, from_off,
                                   out, cap);
      } else {
        n = backtrace_classic<KIND>(p, hdr, cells, score + 1, plen, tlen, from_score, from_k,
                                    from_off, J.cend, out, cap);
      }
      // move the ops to the head of the fragment
      const int gap = cap - n;
      if (gap > 0) {
        for (int c = 0; c < n; c += 64) {
          const int i = c + lane;
          uint8_t v = 0;
          if (i < n) v = out[gap + i];
          __builtin_amdgcn_wave_barrier();
          if (i < n) out[i] = v;
          __builtin_amdgcn_wave_barrier();
        }
      }
      if (st != BI_OK && KIND != KIND_EDIT) {
        __threadfence_block();
        bool trimmed = false;
        n = maxtrim_cigar(p, out, n, 0, 0, &trimmed);
      }
    }
  } else {
    st = BI_RETRY;
  }
  if (lane == 0) {
    bp.frag_len[item] = n;
    bp.frag_status[item] = st;
    bp.frag_score[item] = score;
  }
}

template <int KIND, bool HEUR>
__global__ void __launch_bounds__(64) biwfa_base_kernel(BiParams bp) {
  extern __shared__ uint64_t smem[];
  const int lane = threadIdx.x;
  constexpr int NC = kind_components(KIND);
  uint64_t* sP = smem;
  uint64_t* sT = smem + bp.seq_words;
  int4* meta =(int4*)(sT + bp.seq_words);
  int32_t* ring = bp.ring + (int64_t)blockIdx.x * NC * bp.k.scope * bp.wcap;
  int4* hdr = bp.hdr + (int64_t)blockIdx.x * bp.hdr_cap * NC;
  int32_t* cells = bp.hist + (int64_t)blockIdx.x * bp.hist_cap;
  for (;;) {
    int item = 0;
    if (lane == 0) item = atomicAdd(&bp.counters[0], 1);
    item = uniform(__shfl(item, 0));
    if (item >= bp.njobs) break;
    const BiJob J = bp.jobs[item];
    const int bits = stage_job<1>(bp, J, sP, sT, nullptr);
    if (bits == 0) {
      if (lane == 0) {
        bp.frag_len[item] = 0;
        bp.frag_status[item] = BI_RETRY;
        bp.frag_score[item] = 0;
      }
    } else if (bits == 2) {
      base_case<KIND, 2, HEUR>(bp, item, J, sP, sT, meta, ring, hdr, cells);
    } else {
      base_case<KIND, 8, HEUR>(bp, item, J, sP, sT, meta, ring, hdr, cells);
    }
    __syncthreads();
  }
}

template <int K>
static KernelFn2 pick_biwfa_kind(int which, bool heur) {
  if (which == 0)
    return heur ? biwfa_breakpoint_kernel<K, true, 1> : biwfa_breakpoint_kernel<K, false, kWideWaves>;
  return heur ? biwfa_base_kernel<K, true> : biwfa_base_kernel<K, false>;
}

}  // namespace wfa
