// wfa_align.h — gfx950 wavefront-alignment kernels (included by one
// translation unit per distance-metric kind, which instantiates them).
//
// One 64-lane wavefront aligns one pattern/text pair at a time and pulls
// the next pair from a device work queue (a persistent launch). Within a
// pair, the diagonals k of a wavefront row sit across the lanes (lane i
// handles k = lo + i + 64*j); the rows of the last `scope` scores live in
// an LDS ring indexed by (k & (wcap-1)), read through a range guard that
// returns OFFSET_NULL outside a row's [lo, hi] (the reference fills those
// halo cells with OFFSET_NULL before every read, src/wavefront.rs:422-457).
// Extension compares 32 bases per 64-bit XOR of 2-bit packed sequences held
// in LDS; trims reduce per-chunk ballots; the termination cell is read by
// the lane that owns it. In CIGAR scope the extended M rows go to a
// per-slot history in HBM, which the same wavefront walks backwards with
// the Singletrack backtrace (src/backtrace.rs:751-858) — gap-length search
// parallel over lanes — writing one byte per CIGAR op.
//
// Semantics follow the reference's unialign loop (src/aligner.rs:638-698)
// and compute steps (src/aligner.rs:1769-1865, 2006-2193, 2201-2553) bit
// for bit; the kernels never approximate.

#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "wfa_device.h"

namespace wfa {

#define WFA_NULL (INT32_MIN / 2)  // OFFSET_NULL, src/offset.rs:15
#define META_ABSENT 0x7fffffff   // row never computed (a null step)
static __device__ __forceinline__ int4 meta_absent() {
  return make_int4(META_ABSENT, META_ABSENT - 1, META_ABSENT, META_ABSENT - 1);
}

static __device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }
static __device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }

static __device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

static __device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }
static __device__ __forceinline__ int wave_min(int v) {
  for (int d = 32; d >= 1; d >>= 1) v = imin(v, __shfl_xor(v, d));
  return v;
}
static __device__ __forceinline__ int wave_max(int v) {
  for (int d = 32; d >= 1; d >>= 1) v = imax(v, __shfl_xor(v, d));
  return v;
}
static __device__ __forceinline__ int64_t wave_max64(int64_t v) {
  for (int d = 32; d >= 1; d >>= 1) {
    const int64_t o = __shfl_xor(v, d);
    v = v > o ? v : o;
  }
  return v;
}

// In-bounds test of the kernels and trim_ends (src/compute/affine.rs:338-344,
// src/compute/mod.rs:14-49).
static __device__ __forceinline__ bool in_bounds(int off, int k, int plen, int tlen) {
  return (uint32_t)off <= (uint32_t)tlen && (uint32_t)(off - k) <= (uint32_t)plen;
}

// ---------------------------------------------------------------------
// Sequences staged in LDS as little-endian 64-bit words.
//   2-bit mode: base j at bits [2j, 2j+2) (src/sequences.rs:55-74 order)
//   byte mode:  byte j at bits [8j, 8j+8)
// ---------------------------------------------------------------------
template <int BITS>
static __device__ __forceinline__ uint64_t load_fwd(const uint64_t* s, int pos) {
  constexpr int PER = 64 / BITS;
  const int q = pos / PER;
  const int sh = (pos % PER) * BITS;
  const uint64_t lo = s[q];
  const uint64_t hi = s[q + 1];
  return sh ? (lo >> sh) | (hi << (64 - sh)) : lo;
}

// Matching prefix length of pattern[v..] and text[h..], bounded by the
// sequence ends (src/extend.rs:15-38).
template <int BITS>
static __device__ __forceinline__ int lcp(const uint64_t* P, const uint64_t* T, int v, int h,
                                          int plen, int tlen) {
  constexpr int PER = 64 / BITS;
  int rem = imin(plen - v, tlen - h);
  int n = 0;
  while (rem > 0) {
    const uint64_t x = load_fwd<BITS>(P, v + n) ^ load_fwd<BITS>(T, h + n);
    const int c = x ? (int)(__builtin_ctzll(x) / BITS) : PER;
    if (c >= rem) { n += rem; break; }
    n += c;
    if (c < PER) break;
    rem -= PER;
  }
  return n;
}

// Matching suffix length of pattern[..v) and text[..h), at most min(v, h)
// (st_backward_matches, src/backtrace.rs:668-686).
template <int BITS>
static __device__ __forceinline__ uint64_t load_back(const uint64_t* s, int end) {
  constexpr int PER = 64 / BITS;
  if (end >= PER) return load_fwd<BITS>(s, end - PER);
  // fewer than PER symbols before `end`: align them to the top bits
  const uint64_t w = s[0];
  return end > 0 ? (w << ((PER - end) * BITS)) : 0;
}
template <int BITS>
static __device__ __forceinline__ int backward_matches(const uint64_t* P, const uint64_t* T, int v,
                                                       int h) {
  constexpr int PER = 64 / BITS;
  int rem = imin(v, h);
  int n = 0;
  while (rem > 0) {
    const uint64_t x = load_back<BITS>(P, v - n) ^ load_back<BITS>(T, h - n);
    const int c = x ? (int)(__builtin_clzll(x) / BITS) : PER;
    if (c >= rem) { n += rem; break; }
    n += c;
    if (c < PER) break;
    rem -= PER;
  }
  return n;
}

// ---------------------------------------------------------------------
// Row store: ring of rows, each `wcap` cells, row r at ring + r*wcap.
// ---------------------------------------------------------------------
// Offsets are held as RT: int32, or int16 for pairs short enough. Only
// the sign of an offset derived from OFFSET_NULL ever matters (a valid
// offset, >= 0, beats it in every maximum; M is re-bounded to NULL; the
// extension skips it), so int16 keeps valid offsets exactly and saturates
// the NULL-derived ones at -16384, where adding the few +1s of later steps
// keeps them negative.
// Debug builds (WFA_DEVICE_CHECKS: wfa2lib-rs_amd/build.py --checks) test
// every ring and history index against its buffer. An index outside it is
// counted in the engine's counters[kCheckCounter] and replaced by 0 (the
// access stays inside the buffer), and the host reports a call with a
// nonzero count as WFA_ERR_DEVICE. Release builds carry no test.
template <typename T>
static __device__ __forceinline__ int64_t checked_index(int64_t i, int64_t n, int32_t* chk) {
#ifdef WFA_DEVICE_CHECKS
  if ((uint64_t)i >= (uint64_t)n) {
    atomicAdd(chk + kCheckCounter, 1);
    return 0;
  }
#else
  (void)n;
  (void)chk;
#endif
  return i;
}

// A history buffer (pool, headers, signs): element access, tested against
// the buffer's size in WFA_DEVICE_CHECKS builds.
template <typename T>
struct HBuf {
  T* p = nullptr;
  int64_t n = 0;
  int32_t* chk = nullptr;
  __device__ __forceinline__ T& operator[](int64_t i) const {
    return p[checked_index<T>(i, n, chk)];
  }
  __device__ __forceinline__ explicit operator bool() const { return p != nullptr; }
};

template <typename RT>
struct Rows {
  RT* ring;       // LDS or global
  // per row {lo, hi, init_min, init_max}: the limits the next steps use and
  // the initialized range reads see (wider than [lo, hi] after a heuristic
  // cut, whose cells stay readable, src/wavefront.rs:459-462);
  // lo == META_ABSENT: never computed
  int4* meta;
  int mask;       // wcap - 1
  int wcap;
  int cap = 0;            // cells in the ring (tested in WFA_DEVICE_CHECKS builds)
  int32_t* chk = nullptr;  // the launch's counters
  __device__ __forceinline__ int at(int idx) const {
    return (int)ring[checked_index<RT>(idx, cap, chk)];
  }
  __device__ __forceinline__ void set(int idx, int v) const {
    const int64_t i = checked_index<RT>(idx, cap, chk);
    if constexpr (sizeof(RT) == 2) ring[i] = (RT)imax(v, -16384);
    else ring[i] = v;
  }
};

template <typename RT>
static __device__ __forceinline__ int rd(const Rows<RT>& R, int row, int lo, int hi, int k) {
  return (k >= lo && k <= hi) ? R.at(row * R.wcap + (k & R.mask)) : WFA_NULL;
}

// A row's readable diagonals as {first, count - 1}: k is readable iff
// (unsigned)(k - first) <= span. An empty range becomes a first diagonal
// no pair reaches, with span 0.
struct Guard {
  int first, span;
};

static __device__ __forceinline__ Guard make_guard(int lo, int hi) {
  return lo <= hi ? Guard{lo, hi - lo} : Guard{INT32_MIN / 2, 0};
}

// Branch-free guarded read: the ring index is always inside the ring, so
// the read happens for every lane and a select applies the range (the
// reference's NULL halo, src/wavefront.rs:422-457).
template <typename RT>
static __device__ __forceinline__ int rdg(const Rows<RT>& R, int rowbase, Guard g, int k) {
  const int v = R.at(rowbase + (k & R.mask));
  return (uint32_t)(k - g.first) <= (uint32_t)g.span ? v : WFA_NULL;
}

// ---------------------------------------------------------------------
// Per-pair result codes internal to the kernel
// ---------------------------------------------------------------------
#define ST_RETRY (-1)

struct PairCtx {
  int pair;
  int plen, tlen;
  int bits;  // 2 or 8
};

// Stage pair `pair`'s sequences into LDS. Returns the symbol width (2 or 8)
// or 0 if they do not fit (caller retries with a larger configuration).
template <int NW>
static __device__ int stage_sequences(const KParams& p, int pair, int plen, int tlen,
                                      uint64_t* sP, uint64_t* sT, int* red) {
  const int lane = threadIdx.x;  // all 64 * NW threads of the workgroup
  constexpr int NT = 64 * NW;
  const uint8_t* P = p.seq + p.poff[pair];
  const uint8_t* T = p.seq + p.toff[pair];
  int bits = 2;
  if (p.fmt == WFA_SEQ_ASCII) {
    // 2-bit only if every byte is one of A,C,G,T (case-sensitive, as the
    // reference compares raw bytes)
    bool bad = false;
    for (int i = lane; i < plen; i += NT) {
      const uint8_t c = P[i];
      bad |= !(c == 'A' || c == 'C' || c == 'G' || c == 'T');
    }
    for (int i = lane; i < tlen; i += NT) {
      const uint8_t c = T[i];
      bad |= !(c == 'A' || c == 'C' || c == 'G' || c == 'T');
    }
    bool any_bad = ballot(bad) != 0;
    if (NW > 1) {
      if ((lane & 63) == 0) red[lane >> 6] = any_bad;
      __syncthreads();
      for (int w = 0; w < NW; ++w) any_bad |= red[w] != 0;
    }
    bits = any_bad ? 8 : 2;
  }
  const int per = 64 / bits;
  const int pw = (plen + per - 1) / per + 2;
  const int tw = (tlen + per - 1) / per + 2;
  if (pw > p.seq_words || tw > p.seq_words) return 0;
  // zero the words first (padding words are read but masked off)
  for (int i = lane; i < pw; i += NT) sP[i] = 0;
  for (int i = lane; i < tw; i += NT) sT[i] = 0;
  __syncthreads();
  uint8_t* bP = (uint8_t*)sP;
  uint8_t* bT = (uint8_t*)sT;
  if (p.fmt == WFA_SEQ_PACKED2BITS) {
    const int nb = (plen + 3) >> 2, mb = (tlen + 3) >> 2;
    for (int i = lane; i < nb; i += NT) bP[i] = P[i];
    for (int i = lane; i < mb; i += NT) bT[i] = T[i];
  } else if (bits == 8) {
    for (int i = lane; i < plen; i += NT) bP[i] = P[i];
    for (int i = lane; i < tlen; i += NT) bT[i] = T[i];
  } else {
    // pack ASCII ACGT → 2 bits (A0 C1 G2 T3), 4 bases per byte
    auto code = [](uint8_t c) -> uint32_t { return c == 'A' ? 0u : c == 'C' ? 1u : c == 'G' ? 2u : 3u; };
    const int nb = (plen + 3) >> 2, mb = (tlen + 3) >> 2;
    for (int i = lane; i < nb; i += NT) {
      uint32_t b = 0;
      for (int j = 0; j < 4 && 4 * i + j < plen; ++j) b |= code(P[4 * i + j]) << (2 * j);
      bP[i] = (uint8_t)b;
    }
    for (int i = lane; i < mb; i += NT) {
      uint32_t b = 0;
      for (int j = 0; j < 4 && 4 * i + j < tlen; ++j) b |= code(T[4 * i + j]) << (2 * j);
      bT[i] = (uint8_t)b;
    }
  }
  __syncthreads();
  return bits;
}

// ---------------------------------------------------------------------
// CIGAR history: per slot, headers hdr[s] = {lo, hi, offset in chunk,
// chunk id}; row s holds the cells k = lo..hi at element
// chunk id * chunk + offset + (k - lo) of the pool: extended M offsets
// (high memory) or winner flags (medium/low memory).
// ---------------------------------------------------------------------
struct Hist {
  HBuf<int4> hdr;
  HBuf<int32_t> cells;   // high memory: the pool as extended M offsets
  HBuf<uint8_t> bytes;   // medium/low memory: the pool as winner flags per (score, k)
  uint8_t* chain;   // medium/low memory: op chain scratch
  int chain_cap;
  HBuf<uint8_t> signs;   // Compat::C Z-drop: per pool cell, the signs of its offsets (or null)
  int hdr_cap;
  int32_t* tab;     // this slot's chunk ids
  int tab_cap;
  int chunk;        // elements per chunk
  int pool_chunks;
  int32_t* next;    // the pool's next free chunk
};

static __device__ __forceinline__ int64_t hist_row(const Hist& H, const int4& h) {
  return (int64_t)h.w * H.chunk + h.z;
}

// Winner flags of one (score, k) in medium/low memory. Bits 0-2: the
// source of M (0 I1, 1 I2, 2 D1, 3 D2, 4 mismatch; for edit and gap-linear
// 0 ins, 1 del, 2 mismatch; 7 a begin-free cell, where a chain starts; 6
// a cell set NULL when M was widened to a begin-free cell, whose chain is
// empty); bits 3-6: I1, D1, I2, D2 took their extension input.
#define FL_M_SEED 7
#define FL_M_NULLED 6
#define FL_I1_EXT 8
#define FL_D1_EXT 16
#define FL_I2_EXT 32
#define FL_D2_EXT 64

// st_m_offset (src/backtrace.rs:645-663) over the stored rows.
static __device__ __forceinline__ bool st_m(const Hist& H, int nscores, int nwf, int s, int k,
                                            int* out) {
  if (s < 0 || s >= nwf || s >= nscores) return false;
  const int4 h = H.hdr[s];
  if (k < h.x || k > h.y) return false;
  const int v = H.cells[hist_row(H, h) + (k - h.x)];
  if (v == WFA_NULL) return false;
  *out = v;
  return true;
}

// The stored M rows of a CIGAR history pool as the Singletrack backtrace
// reads them (st_m_offset over the slot's headers and pool chunks).
struct PoolView {
  const Hist& H;
  int nscores, nwf;
  __device__ __forceinline__ bool m(int s, int k, int* out) const {
    return st_m(H, nscores, nwf, s, k, out);
  }
};

// write `n` copies of op backwards ending at out[pos] (pos decreases)
static __device__ __forceinline__ void write_ops(uint8_t* out, int& pos, uint8_t op, int n) {
  for (int i = threadIdx.x; i < n; i += 64)
    if (pos - i >= 0) out[pos - i] = op;  // never below the pair's region
  pos -= n > 0 ? n : 0;
}

// Singletrack backtrace (src/backtrace.rs:751-858), end-to-end, from the
// end cell (score, k, offset). Writes ops backwards into out[0..cap) and
// returns the number of ops, which end at out[cap-1].
// `hv.m(s, k, &offset)` reads the stored M offset of score s at diagonal k
// (false when the score has no row, k is outside it, or the cell is NULL).
template <int BITS, class HV>
static __device__ int singletrack(const KParams& p, const HV& hv, const uint64_t* sP,
                                  const uint64_t* sT, int plen, int tlen, int score, int k,
                                  int offset, uint8_t* out, int cap) {
  const int lane = threadIdx.x;
  const bool two = p.metric == WFA_METRIC_GAP_AFFINE_2P;
  int pos = cap - 1;
  int offset_orig = offset;
  write_ops(out, pos, 'D', plen - (offset - k));
  write_ops(out, pos, 'I', tlen - offset);
  bool in_m = true;
  int l = 0;
  while (score != 0) {
    if (in_m) {
      const int v = offset - k, h = offset;
      const int nm = backward_matches<BITS>(sP, sT, v, h);
      offset_orig = offset - nm;
      const int mis = score - p.x;
      int mo;
      if (hv.m(mis, k, &mo) && mo == offset_orig - 1) {
        write_ops(out, pos, 'M', nm);
        write_ops(out, pos, 'X', 1);
        offset = offset_orig - 1;
        score = mis;
      } else {
        in_m = false;
        l = 0;
      }
    } else {
      // lanes test gap lengths l+1 .. l+64 at once; the lowest length whose
      // test succeeds (or ends the search) decides, in the reference's order:
      // piece 1 insertion, piece 1 deletion, piece 2 insertion, deletion.
      const int L = l + 1 + lane;
      const int ind1 = score - p.o1 - L * p.e1;
      const int ind2 = two ? score - p.o2 - L * p.e2 : -1;
      int m1i = 0, m1d = 0, m2i = 0, m2d = 0;
      const bool r1i = hv.m(ind1, k - L, &m1i) && m1i + L >= offset_orig &&
                       m1i + L <= offset;
      const bool r1d = hv.m(ind1, k + L, &m1d) && m1d >= offset_orig &&
                       m1d <= offset;
      bool r2i = false, r2d = false;
      if (two) {
        r2i = hv.m(ind2, k - L, &m2i) && m2i + L >= offset_orig &&
              m2i + L <= offset;
        r2d = hv.m(ind2, k + L, &m2d) && m2d >= offset_orig && m2d <= offset;
      }
      const bool stop = r1i || r1d || r2i || r2d || (ind1 < 0 && ind2 < 0);
      const uint64_t mask = ballot(stop);
      if (!mask) {
        l += 64;
        continue;
      }
      const int j = __builtin_ctzll(mask);
      const int Lj = uniform(__shfl(L, j));
      const int f = (r1i ? 1 : 0) | (r1d ? 2 : 0) | (r2i ? 4 : 0) | (r2d ? 8 : 0);
      const int fj = uniform(__shfl(f, j));
      if (fj & 1) {
        const int m = uniform(__shfl(m1i, j));
        write_ops(out, pos, 'M', offset - (m + Lj));
        write_ops(out, pos, 'I', Lj);
        k -= Lj; offset = m; score = uniform(__shfl(ind1, j)); in_m = true;
      } else if (fj & 2) {
        const int m = uniform(__shfl(m1d, j));
        write_ops(out, pos, 'M', offset - m);
        write_ops(out, pos, 'D', Lj);
        k += Lj; offset = m; score = uniform(__shfl(ind1, j)); in_m = true;
      } else if (fj & 4) {
        const int m = uniform(__shfl(m2i, j));
        write_ops(out, pos, 'M', offset - (m + Lj));
        write_ops(out, pos, 'I', Lj);
        k -= Lj; offset = m; score = uniform(__shfl(ind2, j)); in_m = true;
      } else if (fj & 8) {
        const int m = uniform(__shfl(m2d, j));
        write_ops(out, pos, 'M', offset - m);
        write_ops(out, pos, 'D', Lj);
        k += Lj; offset = m; score = uniform(__shfl(ind2, j)); in_m = true;
      } else {
        break;  // no source: only at an ends-free start
      }
    }
  }
  int v = offset - k, h = offset;
  const int nm = imin(v, h);
  v -= nm; h -= nm;
  write_ops(out, pos, 'M', nm);
  write_ops(out, pos, 'D', v);
  write_ops(out, pos, 'I', h);
  return cap - 1 - pos;
}

// backtrace_linear (src/backtrace.rs:132-258) for edit / indel, high memory.
template <int BITS>
static __device__ int backtrace_linear(const KParams& p, const Hist& H, int nscores, int plen,
                                       int tlen, int score, int k, int offset, uint8_t* out,
                                       int cap) {
  const bool indel = p.metric == WFA_METRIC_INDEL;
  int pos = cap - 1;
  int v = offset - k, h = offset;
  if (v < plen) write_ops(out, pos, 'D', plen - v);
  if (h < tlen) write_ops(out, pos, 'I', tlen - h);
  // packed (offset << 4) | type; NULL when absent
  auto cell = [&](int s, int kk, int lo_adj, int* val) -> bool {
    (void)lo_adj;
    if (s < 0 || s >= nscores) return false;
    const int4 hd = H.hdr[s];
    if (kk < hd.x || kk > hd.y) return false;
    *val = H.cells[hist_row(H, hd) + (kk - hd.x)];
    return true;
  };
  while (v > 0 && h > 0 && score > 0) {
    const int mis_s = score - p.x, gap_s = score - p.o1;
    int64_t mis = WFA_NULL, ins = WFA_NULL, del = WFA_NULL;
    int c;
    // bt_misms: init_min <= k <= init_max
    if (!indel && cell(mis_s, k, 0, &c)) mis = (int64_t)(c + 1) * 16 + 9;
    // bt_ins1_open: init_min < k && k-1 <= init_max  -> cell k-1 in range
    if (cell(gap_s, k - 1, 0, &c)) ins = (int64_t)(c + 1) * 16 + 1;
    // bt_del1_open: init_min <= k+1 && k < init_max -> cell k+1 in range
    if (cell(gap_s, k + 1, 0, &c)) del = (int64_t)c * 16 + 5;
    int64_t mx = ins > del ? ins : del;
    if (mis > mx) mx = mis;
    if (mx < 0) break;
    const int moff = (int)(mx >> 4);
    const int nmatch = offset - moff;
    if (nmatch < 0) pos = imin(pos - nmatch, cap - 1);
    else write_ops(out, pos, 'M', nmatch);
    offset = moff;
    v = offset - k; h = offset;
    if (v <= 0 || h <= 0) break;
    switch ((int)(mx & 0xF)) {
      case 9: score = mis_s; write_ops(out, pos, 'X', 1); offset -= 1; break;
      case 1: score = gap_s; write_ops(out, pos, 'I', 1); k -= 1; offset -= 1; break;
      default: score = gap_s; write_ops(out, pos, 'D', 1); k += 1; break;
    }
    v = offset - k; h = offset;
  }
  if (v > 0 && h > 0) {
    const int nm = imin(v, h);
    write_ops(out, pos, 'M', nm);
    v -= nm; h -= nm;
  }
  write_ops(out, pos, 'D', v > 0 ? v : 0);
  write_ops(out, pos, 'I', h > 0 ? h : 0);
  return cap - 1 - pos;
}

// maxtrim_cigar (src/aligner.rs:1545-1583) with Cigar::maxtrim_gap_linear
// / _affine / _affine2p (src/cigar.rs:207-419) on the user penalties: keep
// the maximal-scoring prefix, a free leading gap skipped and put back; an
// alignment whose best prefix scores 0 is cleared. Returns the new length.
static __device__ int maxtrim_cigar(const KParams& p, const uint8_t* ops, int n, int pb_free,
                                    int tb_free, bool* trimmed) {
  const int lane = threadIdx.x;
  // leading free gap
  int b = 0;
  int bh = 0, bv = 0;
  while (bh < tb_free && b < n && ops[b] == 'I') { ++b; ++bh; }
  if (bh == 0)
    while (bv < pb_free && b < n && ops[b] == 'D') { ++b; ++bv; }
  *trimmed = false;
  if (b >= n) { *trimmed = true; return 0; }
  const int mscore = p.u_match != 0 ? p.u_match : -1;
  const bool two = p.metric == WFA_METRIC_GAP_AFFINE_2P;
  const bool lin = p.metric == WFA_METRIC_GAP_LINEAR;
  int max_score = 0, max_off = b, score = 0, last = 0, len = 0;
  auto run_cost = [&](int op, int l) -> int {
    if (op == 'M') return mscore * l;
    if (op == 'X') return p.u_x * l;
    const int s1 = p.u_o1 + p.u_e1 * l, s2 = p.u_o2 + p.u_e2 * l;
    return s1 < s2 ? s1 : s2;
  };
  for (int c0 = b; c0 < n; c0 += 64) {
    const int mine = c0 + lane < n ? ops[c0 + lane] : 0;
    const int cnt = imin(64, n - c0);
    for (int j = 0; j < cnt; ++j) {
      const int op = uniform(__shfl(mine, j));
      const int i = c0 + j;
      if (two) {
        if (op != last && last != 0) {
          score -= run_cost(last, len);
          len = 0;
          if (max_score < score) { max_score = score; max_off = i - 1; }
        }
        last = op;
        ++len;
      } else {
        if (op == 'M') score -= mscore;
        else if (op == 'X') score -= p.u_x;
        else if (lin) score -= p.u_o1;  // gap_opening1 is the indel penalty
        else score -= p.u_e1 + (last == op ? 0 : p.u_o1);
        last = op;
        if (max_score < score) { max_score = score; max_off = i; }
      }
    }
  }
  if (two) {
    score -= run_cost(last, len);
    if (max_score < score) { max_score = score; max_off = n - 1; }
  }
  *trimmed = max_off != n - 1;
  if (max_score == 0) return 0;  // Cigar::clear
  return max_off + 1;            // the skipped leading gap stays in front
}

static __device__ __forceinline__ int wave_sum(int v) {
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
  return v;
}

// rescore_heuristic_cigar (src/aligner.rs:1605-1640) with cigar_wfa_score
// (src/aligner.rs:2700-2719): with a heuristic the returned score of a
// gap-affine/2p alignment is its CIGAR's score on the user penalties
// (ends-free: a leading free gap and the trailing free ends not counted),
// in the internal scale; a BiWFA alignment of any gap metric returns its
// CIGAR's score too. `ops` holds the n ops at the head of the region.
static __device__ int rescore_cigar(const KParams& p, const uint8_t* ops, int n, bool ef,
                                    int pb_free, int tb_free, int trailing) {
  const int lane = threadIdx.x;
  int b = 0, e = n;
  if (ef) {
    // end_offset.saturating_sub(trailing as usize): a negative count (an
    // end beyond the sequences, after free begin gaps longer than them)
    // wraps to a huge one and leaves nothing
    e = trailing < 0 ? 0 : imax(0, e - trailing);
    const int op = n > 0 ? ops[0] : 0;
    const int fr = op == 'I' ? tb_free : op == 'D' ? pb_free : 0;
    while (b < fr && b < e && ops[b] == op) ++b;
  }
  if (e <= b) return 0;
  int classic = 0, vh = 0;
  if (p.metric == WFA_METRIC_GAP_LINEAR) {
    // Cigar::score_gap_linear (src/cigar.rs:126-136): every gap op costs
    // the indel penalty
    for (int i = b + lane; i < e; i += 64) {
      const int op = ops[i];
      classic -= op == 'M' ? p.u_match : op == 'X' ? p.u_x : p.u_o1;
      vh += (op != 'I') + (op != 'D');
    }
    classic = wave_sum(classic);
    vh = wave_sum(vh);
  } else if (p.metric == WFA_METRIC_GAP_AFFINE) {
    // Cigar::score_gap_affine (src/cigar.rs:139-169): each op on its own,
    // a gap op opening unless it follows the same op
    for (int i = b + lane; i < e; i += 64) {
      const int op = ops[i];
      const int last = i > b ? ops[i - 1] : 0;
      classic -= op == 'M' ? p.u_match
               : op == 'X' ? p.u_x
                           : p.u_e1 + (last == op ? 0 : p.u_o1);
      vh += (op != 'I') + (op != 'D');
    }
    classic = wave_sum(classic);
    vh = wave_sum(vh);
  } else {
    // Cigar::score_gap_affine2p (src/cigar.rs:171-202): runs, each gap run
    // priced by the cheaper of the two gap models
    auto run_cost = [&](int op, int l) -> int {
      if (op == 'M') return p.u_match * l;
      if (op == 'X') return p.u_x * l;
      const int s1 = p.u_o1 + p.u_e1 * l, s2 = p.u_o2 + p.u_e2 * l;
      return s1 < s2 ? s1 : s2;
    };
    int last = 0, len = 0;
    for (int c0 = b; c0 < e; c0 += 64) {
      const int mine = c0 + lane < e ? ops[c0 + lane] : 0;
      vh += c0 + lane < e ? (mine != 'I') + (mine != 'D') : 0;
      const int cnt = imin(64, e - c0);
      for (int j = 0; j < cnt; ++j) {
        const int op = uniform(__shfl(mine, j));
        if (op != last && last != 0) {
          classic -= run_cost(last, len);
          len = 0;
        }
        last = op;
        ++len;
      }
    }
    classic -= run_cost(last, len);
    vh = wave_sum(vh);
  }
  return p.u_match < 0 ? -p.u_match * vh - 2 * classic : -classic;
}

// Medium/low memory: the chain of operations the reference's piggyback
// kernels push (src/compute/piggyback.rs:195-528), recovered by following
// the stored winner flags from the end cell back to M at score 0, then
// unpacked exactly as unpack_piggyback (src/aligner.rs:232-311).
//
// `ztr` (Z-drop ends of Compat::C, medium/low memory, else null): the chain
// starts at M cell (score, k) of the stop score whatever that cell holds
// (backtrace_piggyback, src/aligner.rs:1443-1460): a cell the step did not
// compute has the zeroed chain, and a cell whose chain reads a cell outside
// that component's trimmed range at its score (ztr, per score {M lo, hi, I1
// lo, hi}, {D1 lo, hi, I2 lo, hi}, {D2 lo, hi}) took that cell's zeroed
// chain (init_ends, src/wavefront.rs:422-457): both lead to the first
// backtrace block, the alignment's start (0, 0).
template <int KIND, int BITS>
static __device__ int piggyback_backtrace(const KParams& p, const Hist& H, int nscores,
                                          const uint64_t* sP, const uint64_t* sT, int plen,
                                          int tlen, int score, int k, int offset, uint8_t* out,
                                          int cap, const int4* ztr = nullptr) {
  const int lane = threadIdx.x;
  enum { CM = 0, CI1, CD1, CI2, CD2 };
  int comp = CM, s = score, kk = k;
  int pos = H.chain_cap - 1;
  bool bad = false;
  bool root00 = false;  // the chain reached a zeroed cell: it starts at (0, 0)
  // whether cell kk of component c at score s lies in that component's
  // trimmed range
  auto in_trim = [&](int c, int s_, int k_) -> bool {
    const int4 a = ztr[3 * s_ + (c == CM || c == CI1 ? 0 : c == CD1 || c == CI2 ? 1 : 2)];
    const int l = c == CM || c == CD1 || c == CD2 ? a.x : a.z;
    const int h = c == CM || c == CD1 || c == CD2 ? a.y : a.w;
    return k_ >= l && k_ <= h;
  };
  // (an M cell's I/D source at the same score was computed in the same
  // step, not read through a range: the test applies to earlier scores)
  int last_s = score;  // the score of the cell whose operation came last
  for (;;) {
    const bool first = pos == H.chain_cap - 1;
    const int succ_s = last_s;
    if (ztr && s != succ_s && (s < 0 || !in_trim(comp, s, kk))) { root00 = true; break; }
    last_s = s;
    if (comp == CM && s == 0) break;
    if (s < 0 || s >= nscores || pos < 0) { bad = true; break; }
    const int4 hd = H.hdr[s];
    if (kk < hd.x || kk > hd.y) {
      // the stop score's cell outside the cells its step computed
      if (ztr && first) { root00 = true; break; }
      bad = true;
      break;
    }
    const int f = H.bytes[hist_row(H, hd) + (kk - hd.x)];
    // a begin-free cell added during compute starts its chain
    if (comp == CM && (f & 7) == FL_M_SEED) break;
    // a cell nulled when M was widened holds the zeroed chain
    // (set_null_cell, src/wavefront.rs:400-408): only a Z-drop end reads one
    if (comp == CM && (f & 7) == FL_M_NULLED) {
      if (ztr) root00 = true;
      else bad = true;
      break;
    }
    // (Compat::C Z-drop ends) whether this cell's packed CIGAR went to the
    // backtrace buffer before the chain read it on: its wavefront offloaded
    // at its score (occupancy at the limit; edit every 16 scores), its
    // offset is not negative (offload_selective, src/compute/piggyback.rs:
    // 535-556), and the next cell of the chain is of a later score (an M
    // cell takes its I/D source's chain in the same step) or is the stop
    // score's cell itself
    bool offl = false;
    if (ztr && (first || succ_s != s)) {
      const int pre = (ztr[3 * s + 2].z >> (5 * comp)) & 31;
      const bool at = KIND == KIND_EDIT ? s % 16 == 0 : KIND == KIND_LINEAR ? pre >= 16 : pre >= 15;
      offl = at && ((H.signs[hist_row(H, hd) + (kk - hd.x)] >> comp) & 1);
    }
    uint8_t op;
    if (KIND == KIND_EDIT || KIND == KIND_LINEAR) {
      // one M component: a mismatch came from M[s - x], a gap from
      // M[s - indel] (edit and indel: both 1)
      const int w = f & 7;
      op = w == 0 ? 'I' : w == 1 ? 'D' : 'X';
      kk += w == 0 ? -1 : w == 1 ? 1 : 0;
      s -= KIND == KIND_EDIT ? 1 : w == 2 ? p.x : p.o1;
    } else if (comp == CM) {
      op = 'X';
      const int w = f & 7;
      if (w == 4) s -= p.x;
      else comp = w == 0 ? CI1 : w == 1 ? CI2 : w == 2 ? CD1 : CD2;
    } else if (comp == CI1 || comp == CI2) {
      op = 'I';
      const bool ext = f & (comp == CI1 ? FL_I1_EXT : FL_I2_EXT);
      if (ext) s -= comp == CI1 ? p.e1 : p.e2;
      else { s -= comp == CI1 ? p.oe1 : p.oe2; comp = CM; }
      kk -= 1;
    } else {
      op = 'D';
      const bool ext = f & (comp == CD1 ? FL_D1_EXT : FL_D2_EXT);
      if (ext) s -= comp == CD1 ? p.e1 : p.e2;
      else { s -= comp == CD1 ? p.oe1 : p.oe2; comp = CM; }
      kk += 1;
    }
    if (lane == 0) H.chain[pos] = op | (offl ? 0x80 : 0);
    --pos;
  }
  if (bad) return -1;
  if (ztr && lane == 0) {
    // The reference packs a cell's operations into 32 bits, 16 at most: a
    // chain that grows past 16 since its last offload keeps only the last
    // 16, and an offload starts a new block once the chain holds at least
    // the mask's count (src/pcigar.rs, offload_selective): replay that
    // window forward over the chain and drop what it lost (0 bytes).
    const bool aff = KIND == KIND_AFFINE || KIND == KIND_AFFINE2P;
    const int need = p.low ? (aff ? 15 : 16) : 9;
    int cur = 0, wstart = pos + 1;
    for (int i = pos + 1; i < H.chain_cap; ++i) {
      const uint8_t c = H.chain[i];
      ++cur;
      if (cur > 16) {
        H.chain[wstart++] = 0;
        cur = 16;
      }
      if ((c & 0x80) && cur >= need) {
        cur = 0;
        wstart = i + 1;
      }
      H.chain[i] = c & 0x7f;
    }
  }
  __threadfence_block();
  // the chain starts at the M cell of score 0 it reached: (0,0) or an
  // ends-free seed (h,0) at k = h, (0,v) at k = -v
  const int bv = root00 || kk >= 0 ? 0 : -kk, bh = !root00 && kk > 0 ? kk : 0;
  const int nchain = H.chain_cap - 1 - pos;
  const uint8_t* ch = H.chain + pos + 1;
  const bool affine = KIND == KIND_AFFINE || KIND == KIND_AFFINE2P;
  // unpack: begin at (0, 0); re-extend matches before each op (affine: only
  // while in M); drop the X that closes a gap; pad to the end cell and the
  // sequence ends
  int n = 0, v = bv, h = bh;
  bool in_m = true;
  auto put_run = [&](uint8_t c, int len) {
    for (int i = lane; i < len; i += 64)
      if (n + i < cap) out[n + i] = c;
    n += len > 0 ? len : 0;
  };
  put_run('I', bh);
  put_run('D', bv);
  for (int b = 0; b < nchain; b += 64) {
    const int c = (b + lane < nchain) ? ch[b + lane] : 0;
    const int cnt = imin(64, nchain - b);
    for (int j = 0; j < cnt; ++j) {
      const int op = uniform(__shfl(c, j));
      if (op == 0) continue;  // lost from its packed CIGAR (above)
      if (in_m || !affine) {
        const int m = lcp<BITS>(sP, sT, v, h, plen, tlen);
        put_run('M', m);
        v += m;
        h += m;
      }
      if (affine && !in_m && op == 'X') { in_m = true; continue; }
      put_run((uint8_t)op, 1);
      if (op == 'D') { v += 1; in_m = false; }
      else if (op == 'I') { h += 1; in_m = false; }
      else { v += 1; h += 1; in_m = true; }
    }
  }
  const int ev = offset - k, eh = offset;
  const int nm = imin(ev - v, eh - h);
  put_run('M', nm);
  v += nm;
  h += nm;
  put_run('I', tlen - h);
  put_run('D', plen - v);
  return n;
}

// num_wavefronts of a freshly created aligner in high memory
// (compute_dimensions, src/components.rs:435-504, sized for
// max(1000x1000, plen x tlen)).
static __device__ int high_mem_nwf(const KParams& p, int plen, int tlen) {
  auto dims = [&](int pl, int tl) -> int {
    const int ad = pl > tl ? pl - tl : tl - pl;
    const int mn = imin(pl, tl);
    switch (p.metric) {
      case WFA_METRIC_INDEL: return pl + tl + 1;
      case WFA_METRIC_EDIT: return imax(pl, tl) + 1;
      case WFA_METRIC_GAP_LINEAR: return mn * p.x + p.o1 * ad + 1;
      case WFA_METRIC_GAP_AFFINE: return mn * p.x + p.o1 + ad * p.e1 + 1;
      default: {
        const int i1 = p.o1 + ad * p.e1, i2 = p.o2 + ad * p.e2;
        return mn * p.x + imin(i1, i2) + 1;
      }
    }
  };
  return imax(dims(1000, 1000), dims(plen, tlen));
}

// ---------------------------------------------------------------------
// Heuristic cut-off decision (heuristic_cutoff, src/heuristic.rs:167-525,
// without Z-drop) for an extended M row with limits [lo, hi] at score s:
// the new limits in *nlo / *nhi. Returns false when the strategy does not
// cut at this step. `band_lo/hi` are the banded strategies' (min_k, max_k)
// in this aligner's diagonals.
// ---------------------------------------------------------------------
// A pair's heuristic: its kind and parameters (HeuristicStrategy,
// src/heuristic.rs), launch-wide or per pair.
struct HeurCfg {
  int kind, p0, p1, p2;
};

static __device__ __forceinline__ HeurCfg launch_heuristic(const KParams& p) {
  return HeurCfg{p.heuristic, p.hp0, p.hp1, p.hp2};
}

struct HeurState {
  int steps_wait, max_sw;
  bool have_max;
  // Z-drop: the best cell (k, offset), the score it was reached at, and the
  // last score with no drop
  int best_k, best_off, max_wf, zlast;
  // HeuristicState::reset (src/heuristic.rs:122-143)
  __device__ void reset(const HeurCfg& hc) {
    const int heur = hc.kind;
    steps_wait = heur == WFA_HEURISTIC_WFADAPTIVE || heur == WFA_HEURISTIC_BANDED_ADAPTIVE
                     ? hc.p2
                     : heur == WFA_HEURISTIC_XDROP || heur == WFA_HEURISTIC_ZDROP ? hc.p1 : 0;
    max_sw = 0;
    have_max = false;
    best_k = 0;
    best_off = WFA_NULL;
    max_wf = 0;
    zlast = 0;
  }
};

// Z-drop (zdrop_cutoff, src/heuristic.rs:449-478, with compute_sw_scores,
// :491-525) on the extended M row [lo, hi] at score s: every p1 steps the
// row's best Smith-Waterman cell (the first k with the largest score) is
// compared with the best seen; true when the best has dropped by more than
// p0 for at least the largest single penalty (scope - 1) scores. The row's
// limits are not changed.
template <typename RT>
static __device__ bool zdrop_check(const KParams& p, const HeurCfg& hc, HeurState& hs,
                                   const Rows<RT>& R, int rowM, int lo, int hi, int s, int pb_free,
                                   int tb_free) {
  const int lane = threadIdx.x & 63;
  if (lo > hi) return false;
  hs.steps_wait -= 1;
  if (hs.steps_wait > 0) return false;
  const int swg = p.u_match < 0 ? -p.u_match : 1;
  int64_t best = INT64_MIN;  // (sw << 32) | (~biased k): max sw, first k
  int boff = 0;
  for (int kb = lo; kb <= hi; kb += 64) {
    const int k = kb + lane;
    if (k <= hi) {
      const int off = R.at(rowM * R.wcap + (k & R.mask));
      if (off >= 0) {
        const int fv = k < 0 ? imin(-k, pb_free) : 0;
        const int fh = k > 0 ? imin(k, tb_free) : 0;
        const int sw = (swg * ((off - k - fv) + (off - fh)) - s) / 2;
        const int64_t key = ((int64_t)sw << 32) | (uint32_t)(0x7fffffffu - (uint32_t)k);
        if (key > best) {
          best = key;
          boff = off;
        }
      }
    }
  }
  const int64_t wb = wave_max64(best);
  // no valid cell: the reference's (INT32_MIN, k 0, offset 0)
  int cmax = INT32_MIN, ck = 0, coff = 0;
  if (wb != INT64_MIN) {
    const uint64_t holder = ballot(best == wb);
    coff = uniform(__shfl(boff, (int)__builtin_ctzll(holder)));
    cmax = (int)(wb >> 32);
    ck = (int)(0x7fffffffu - (uint32_t)(wb & 0xffffffffu));
  }
  if (!hs.have_max || cmax > hs.max_sw) {
    hs.have_max = true;
    hs.max_sw = cmax;
    hs.max_wf = s;
    hs.best_k = ck;
    hs.best_off = coff;
    hs.zlast = s;
  } else if ((int64_t)hs.max_sw - cmax <= hc.p0) {
    hs.zlast = s;
  } else if (s - hs.zlast >= p.scope - 1) {
    return true;
  }
  hs.steps_wait = hc.p1;
  return false;
}

template <typename RT>
static __device__ bool heuristic_limits(const KParams& p, const HeurCfg& hc, int band_lo,
                                        int band_hi, HeurState& hs, const Rows<RT>& R, int rowM,
                                        int lo, int hi, int s, int plen, int tlen, int pb_free,
                                        int tb_free, int* out_lo, int* out_hi) {
  const int lane = threadIdx.x;
  const int heur = hc.kind;
  // banded-static inside its band changes nothing and skips the step
  // counter (src/aligner.rs:1645-1654)
  if (heur == WFA_HEURISTIC_BANDED_STATIC && lo >= band_lo && hi <= band_hi) return false;
  if (lo > hi) return false;
  hs.steps_wait -= 1;
  int nlo = lo, nhi = hi;
  auto cell = [&](int k) { return R.at(rowM * R.wcap + (k & R.mask)); };
  // wf_distance_end2end (src/heuristic.rs:150-160)
  auto dist = [&](int off, int k) -> int {
    return off < 0 ? INT32_MAX : imax(plen - (off - k), tlen - off);
  };
  if (heur == WFA_HEURISTIC_WFADAPTIVE) {
    // src/heuristic.rs:258-346
    if (hs.steps_wait > 0 || hi - lo + 1 < hc.p0) return false;
    int md = imax(plen, tlen);
    for (int kb = lo; kb <= hi; kb += 64) {
      const int k = kb + lane;
      if (k <= hi) md = imin(md, dist(cell(k), k));
    }
    md = wave_min(md);
    int fg = INT32_MAX, lg = INT32_MIN;
    for (int kb = lo; kb <= hi; kb += 64) {
      const int k = kb + lane;
      const bool good = k <= hi && (int64_t)dist(cell(k), k) - md <= (int64_t)hc.p1;
      const uint64_t b = ballot(good);
      if (b) {
        if (fg == INT32_MAX) fg = kb + __builtin_ctzll(b);
        lg = kb + 63 - __builtin_clzll(b);
      }
    }
    // the low end moves up to the first close diagonal, stopping at the
    // end diagonal; then the high end down, stopping there or at the new lo
    const int ak = tlen - plen;
    const int top = imin(ak, hi);
    nlo = top <= lo ? lo : imin(fg, top);
    const int bottom = imax(ak, nlo);
    nhi = hi <= bottom ? hi : imax(lg, bottom);
    hs.steps_wait = hc.p2;
  } else if (heur == WFA_HEURISTIC_BANDED_STATIC) {
    // src/heuristic.rs:400-407
    nlo = imax(lo, band_lo);
    nhi = imin(hi, band_hi);
  } else if (heur == WFA_HEURISTIC_BANDED_ADAPTIVE) {
    // src/heuristic.rs:355-394
    if (hs.steps_wait > 0) return false;
    const int len = hi - lo + 1;
    if (len < 4) return false;
    const int maxlen = band_hi - band_lo + 1;
    if (len > maxlen) {
      const int leeway = (len - maxlen) / 2, q = len / 4;
      const int d0 = dist(uniform(cell(lo)), lo);
      const int d1 = dist(uniform(cell(lo + q)), lo + q);
      const int d2 = dist(uniform(cell(lo + 2 * q)), lo + 2 * q);
      const int d3 = dist(uniform(cell(hi)), hi);
      int nl = lo;
      if (d0 > d3) nl += leeway;
      if (d1 > d2) nl += leeway;
      nlo = imax(nl, lo);
      nhi = imin(nl + maxlen - 1, hi);
    }
    hs.steps_wait = hc.p2;
  } else if (heur == WFA_HEURISTIC_XDROP) {
    // src/heuristic.rs:410-448, 491-525: the Smith-Waterman-like score of
    // each cell (free begin gaps not counted), cells more than `xdrop`
    // below the best seen so far cut from both ends
    if (hs.steps_wait > 0) return false;
    const int swg = p.u_match < 0 ? -p.u_match : 1;
    int64_t best = INT64_MIN;  // (sw << 32) | (~biased k): max sw, first k
    int fg = INT32_MAX, lg = INT32_MIN;
    for (int kb = lo; kb <= hi; kb += 64) {
      const int k = kb + lane;
      bool good = false;
      if (k <= hi) {
        const int off = cell(k);
        if (off >= 0) {
          const int fv = k < 0 ? imin(-k, pb_free) : 0;
          const int fh = k > 0 ? imin(k, tb_free) : 0;
          const int sw = (swg * ((off - k - fv) + (off - fh)) - s) / 2;
          const int64_t key = ((int64_t)sw << 32) | (uint32_t)(0x7fffffffu - (uint32_t)k);
          best = best > key ? best : key;
          good = hs.have_max && hs.max_sw - sw < hc.p0;
        }
      }
      const uint64_t b = ballot(good);
      if (b) {
        if (fg == INT32_MAX) fg = kb + __builtin_ctzll(b);
        lg = kb + 63 - __builtin_clzll(b);
      }
    }
    best = wave_max64(best);
    const int cmax = (int)(best >> 32);
    if (hs.have_max) {
      nlo = fg != INT32_MAX ? fg : hi + 1;
      nhi = lg != INT32_MIN ? lg : hi;
      if (cmax > hs.max_sw) hs.max_sw = cmax;
    } else {
      hs.max_sw = cmax;
      hs.have_max = true;
    }
    hs.steps_wait = hc.p1;
  }
  *out_lo = nlo;
  *out_hi = nhi;
  return true;
}

// ---------------------------------------------------------------------
// The per-pair forward pass + backtrace.
// ---------------------------------------------------------------------
// One pair per workgroup of NW waves. With NW > 1 the waves split every
// row's 64-diagonal chunks round-robin and combine their per-chunk results
// (trims, termination, the first ends-free hit) through `red` in LDS; one
// thread writes metadata and results, and wave 0 alone runs the backtrace.
template <int KIND, int MODE, int BITS, typename RT, bool HEUR, int U, int NW, bool BFREE>
static __device__ void align_pair(const KParams& p, int pair, int plen, int tlen, uint64_t* sP,
                                  uint64_t* sT, const Rows<RT>& R, const Hist& H, int* red,
                                  int& owned) {
  // MODE 0: score only; 1: high memory (M rows kept); 2: medium/low memory
  // (winner flags kept)
  constexpr bool CIG = MODE != 0;
  constexpr bool MED = MODE == 2;
  static_assert(NW == 1 || (!HEUR && U == 1), "several waves per pair: no heuristics, U = 1");
  constexpr int NT = 64 * NW;
  // gap-affine kinds keep I/D components; edit and gap-linear only M
  constexpr bool GAPS = KIND == KIND_AFFINE || KIND == KIND_AFFINE2P;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = NW == 1 ? 0 : tid >> 6;
  const bool two = KIND == KIND_AFFINE2P;
  // ring rows: M: [0, ringM), I1: [ringM, ringM+ringE1), D1 next, then I2, D2
  const int ringM = p.ringM, ringE1 = p.ringE1, ringE2 = p.ringE2;
  const int rI1 = ringM, rD1 = ringM + ringE1, rI2 = ringM + 2 * ringE1,
            rD2 = ringM + 2 * ringE1 + ringE2;
  const int nrows = !GAPS ? ringM : two ? rD2 + ringE2 : rD1 + ringE1;
  for (int r = tid; r < nrows; r += NT) R.meta[r] = meta_absent();
  // ends-free limits of this pair (extension: free ends up to the full
  // lengths, no begin gaps; src/aligner.rs:713-718)
  // span and heuristic: the launch's, or this pair's within the launch's
  // envelope (ends-free only in an ends-free launch, a heuristic only in a
  // heuristic launch)
  bool ef = p.endsfree, ext = p.extension;
  int4 fe = make_int4(p.pb_free, p.pe_free, p.tb_free, p.te_free);
  HeurCfg hc = launch_heuristic(p);
  if (p.pair_cfg) {
    const int4 c0 = p.pair_cfg[3 * pair], c1 = p.pair_cfg[3 * pair + 1];
    const int4 c2 = p.pair_cfg[3 * pair + 2];
    ef = p.endsfree && uniform(c0.x) == WFA_SPAN_ENDSFREE;
    ext = ef && uniform(c0.y) != 0;
    fe = make_int4(imax(uniform(c0.z), 0), imax(uniform(c0.w), 0), imax(uniform(c1.x), 0),
                   imax(uniform(c1.y), 0));
    hc = HeurCfg{uniform(c1.z), uniform(c1.w), uniform(c2.x), uniform(c2.y)};
    if (p.heuristic == WFA_HEURISTIC_NONE || hc.kind < WFA_HEURISTIC_NONE ||
        hc.kind > WFA_HEURISTIC_BANDED_ADAPTIVE)
      hc.kind = WFA_HEURISTIC_NONE;
  }
  const int pb_free = ef && !ext ? fe.x : 0;
  const int tb_free = ef && !ext ? fe.z : 0;
  const int pe_free = !ef ? 0 : ext ? plen : fe.y;
  const int te_free = !ef ? 0 : ext ? tlen : fe.w;
  // with a match reward the begin-free cells come during compute
  // (the single-wave kernels only: those launches take them)
  const bool seedable = HEUR && p.u_match < 0 && ef && (pb_free > 0 || tb_free > 0);
  // X-drop and Z-drop leave free leading gaps out of the Smith-Waterman
  // score only with match == 0 (DropParams::new, src/heuristic.rs:92-107)
  const int sw_pb = p.u_match == 0 ? pb_free : 0, sw_tb = p.u_match == 0 ? tb_free : 0;
  // M0: k = 0, offset 0, plus (ends-free, match == 0) the begin-free cells
  // (h, 0) at k = h and (0, v) at k = -v (init_alignment,
  // src/aligner.rs:938-998)
  const int lo0 = p.u_match == 0 ? -pb_free : 0, hi0 = p.u_match == 0 ? tb_free : 0;
  bool overflow = hi0 - lo0 + 1 > R.wcap;
  if (!overflow) {
    for (int k = lo0 + tid; k <= hi0; k += NT) R.set(k & R.mask, k > 0 ? k : 0);
  }
  __syncthreads();
  if (tid == 0) R.meta[0] = make_int4(lo0, hi0, lo0, hi0);
  __syncthreads();

  const int nwf = high_mem_nwf(p, plen, tlen);
  const int k_end = tlen - plen;
  int end_k = k_end, end_off = tlen;
  const int hist_lo = -(plen + p.scope + 2), hist_hi = tlen + p.scope + 2;
  // Z-drop end of a medium/low-memory CIGAR in Compat::C: the backtrace
  // starts at the best cell's diagonal of the stop score's M wavefront
  // (backtrace_zdrop_end, src/aligner.rs:1393-1406), empty when that
  // wavefront's allocation does not hold it. Kept for it: the allocated
  // diagonals [zalo, zahi] (output_alloc_range, src/aligner.rs:3318-3342:
  // every computed range widened by scope + 1) and each score's trimmed
  // ranges (ztr).
  const bool zc = MED && HEUR && hc.kind == WFA_HEURISTIC_ZDROP &&
                  p.compat == WFA_COMPAT_C && p.ztr != nullptr;
  int4* const ztr = zc ? p.ztr + (int64_t)blockIdx.x * p.hdr_cap * 3 : nullptr;
  const int zpad = p.scope + 1;
  int zalo = imin(0, lo0 - zpad), zahi = imax(0, hi0 + zpad);
  // occ: the components' occupancies before offloading (5 bits each, M,
  // I1, D1, I2, D2; the offload rules of src/compute/piggyback.rs:576-633)
  auto ztr_put = [&](int s_, int ml, int mh, int il, int ih, int dl, int dh, int i2l, int i2h,
                     int d2l, int d2h, int occ) {
    if (zc && tid == 0 && s_ < p.hdr_cap) {
      ztr[3 * s_] = make_int4(ml, mh, il, ih);
      ztr[3 * s_ + 1] = make_int4(dl, dh, i2l, i2h);
      ztr[3 * s_ + 2] = make_int4(d2l, d2h, occ, 0);
    }
  };
  ztr_put(0, lo0, hi0, 1, 0, 1, 0, 1, 0, 1, 0, 0);
  // the occupancy a component of score s_ passes on (after its offload),
  // 0 for a missing or null wavefront (occupancy helpers,
  // src/compute/piggyback.rs:591-620)
  auto zlive = [&](int s_, int c) -> bool {
    if (s_ < 0) return false;
    const int4 a = ztr[3 * s_ + (c == 0 || c == 1 ? 0 : c == 2 || c == 3 ? 1 : 2)];
    const bool lohi = c == 0 || c == 2 || c == 4;
    return (lohi ? a.x : a.z) <= (lohi ? a.y : a.w);
  };
  auto zocc = [&](int s_, int c) -> int {
    if (!zlive(s_, c)) return 0;
    const int pre = (ztr[3 * s_ + 2].z >> (5 * c)) & 31;
    if (KIND == KIND_LINEAR) return pre >= 16 ? (p.low ? 15 : 8) : pre;
    return pre >= 15 ? (p.low ? 14 : 8) : pre;
  };
  // CIGAR history rows go to the slot's pool chunks in order: chunk `cur`
  // (id `cid`) is filled up to `cpos`; a row that does not fit moves to the
  // slot's next chunk, taking a new one from the pool when the slot owns no
  // more (it keeps them for its next pairs). False: the pool is exhausted.
  int cur = -1, cpos = H.chunk, cid = 0;
  auto halloc = [&](int n, int4* hd_out, int lo_, int hi_) -> bool {
    if (cpos + n > H.chunk) {
      cur += 1;
      if (cur >= H.tab_cap) return false;
      if (cur >= owned) {
        if (tid == 0) H.tab[cur] = atomicAdd(H.next, 1);
        __syncthreads();
        if (uniform(H.tab[cur]) >= H.pool_chunks) return false;
        owned = cur + 1;
      }
      cid = uniform(H.tab[cur]);
      cpos = 0;
    }
    *hd_out = make_int4(lo_, hi_, cpos, cid);
    cpos += n;
    return true;
  };
  int nulls = 0;
  int score = 0;
  int status = WFA_STATUS_OK;
  int result = 0;

  // ---- Compat::Corrected ends-free with a match reward
  // (src/aligner.rs:1111-1360; the heuristic kernels, one wave): a cell
  // reaching a free end at score s with `left` bases to skip completes an
  // alignment of total s + -match * left; the best candidate ends the
  // search once the score reaches its total. Band pruning against the
  // candidate's total (no heuristic or banded-static only) and begin-free
  // dominance pruning drop cells that cannot do better.
  const bool corrected = HEUR && NW == 1 && ef && corrected_endsfree(p);
  const bool cprune =
      corrected && (hc.kind == WFA_HEURISTIC_NONE || hc.kind == WFA_HEURISTIC_BANDED_STATIC);
  const int mr = -p.u_match;
  bool ext_found = false;  // the last extension's best free-end cell
  int ext_left = 0, ext_k = 0, ext_off = 0;
  bool cand_valid = false;
  int cand_total = 0, cand_s = 0, cand_k = 0, cand_off = 0;
  int end_s = -1;  // the end cell's score when it is a candidate's
  // dominance record: per component and diagonal the furthest offset seen
  // (src/aligner.rs:925-945: begin-free gaps, at most 4 for gap-affine)
  constexpr int NCOMP = kind_components(KIND);
  const int bfree = imax(pb_free, tb_free);
  // (begin-free gaps longer than their sequence put seeds beyond the
  // record's diagonals [hist_lo, hist_hi]: the record stays off)
  const bool dom_on = cprune && bfree > 0 && (KIND == KIND_LINEAR || bfree <= 4) &&
                      pb_free <= plen && tb_free <= tlen;
  const int dlen = hist_hi - hist_lo + 1;
  int32_t* const dom = dom_on ? p.dom + (int64_t)blockIdx.x * p.dom_stride : nullptr;
  // the record is sized from the batch's stated longest sequence: a longer
  // pair has no room in it and goes back for retry (the wider passes plan
  // the same record, so it ends with WFA_STATUS_OOM)
  if (dom_on && NCOMP * dlen > p.dom_stride) overflow = true;
  if (dom_on && !overflow) {
    for (int i = tid; i < NCOMP * dlen; i += NT) dom[i] = -1;
    __syncthreads();
  }

  // extend + terminate + (CIGAR) record M[score] over its rows; `lo/hi`
  // is the computed (pre-trim) range; returns true if the end is reached.
  auto extend_row = [&](int s, int row, int lo, int hi, int4 hrow) -> bool {
    bool reached = false;
    uint64_t hit = 0;  // ends-free: first chunk with a terminating diagonal
    int hit_kb = 0;
    // corrected: the fewest bases left, then the lowest k
    // (extend_matches_packed_endsfree_best, src/extend.rs:179-232)
    int64_t cbest = INT64_MAX;
    int cboff = 0;
    const int64_t base = MODE == 1 ? hist_row(H, hrow) : 0;
    // U chunks at a time: their offsets are read before the first one's
    // extension is written back
    auto ext_chunk = [&](int kb, int off) {
      const int k = kb + lane;
      bool term = false;
      if (k <= hi) {
        const int cell = row * R.wcap + (k & R.mask);
        if (off >= 0) {
          off += lcp<BITS>(sP, sT, off - k, off, plen, tlen);
          R.set(cell, off);
        }
        if (!ef) {
          if (k == k_end && off >= tlen) reached = true;
        } else {
          // termination_endsfree (src/termination.rs:31-59)
          const int v = off - k;
          term = off >= 0 &&
                 ((off >= tlen && plen - v <= pe_free) || (v >= plen && tlen - off <= te_free));
          if (corrected && term) {
            // endsfree_left (src/termination.rs:65-86)
            int left = INT32_MAX;
            if (off >= tlen && plen - v <= pe_free) left = plen - v;
            if (v >= plen && tlen - off <= te_free) left = imin(left, tlen - off);
            const int64_t key = ((int64_t)left << 32) | (uint32_t)(k + (1 << 30));
            if (key < cbest) {
              cbest = key;
              cboff = off;
            }
          }
        }
        if (MODE == 1) H.cells[base + (k - lo)] = off < 0 ? WFA_NULL : off;
      }
      if (ef && !corrected && !hit) {
        // the lowest terminating diagonal wins (src/extend.rs:133-166)
        hit = ballot(term);
        hit_kb = kb;
        if (hit) {
          const int j = __builtin_ctzll(hit);
          end_k = hit_kb + j;
          end_off = uniform(__shfl(off, j));
        }
      }
    };
    if (NW == 1) {
      int kb = lo;
      for (; kb + 64 * (U - 1) <= hi; kb += 64 * U) {
        int off[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int k = kb + 64 * u + lane;
          off[u] = k <= hi ? R.at(row * R.wcap + (k & R.mask)) : WFA_NULL;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) ext_chunk(kb + 64 * u, off[u]);
      }
      for (; kb <= hi; kb += 64) {
        const int k = kb + lane;
        ext_chunk(kb, k <= hi ? R.at(row * R.wcap + (k & R.mask)) : WFA_NULL);
      }
    } else {
      for (int kb = lo + 64 * wave; kb <= hi; kb += NT) {
        const int k = kb + lane;
        ext_chunk(kb, k <= hi ? R.at(row * R.wcap + (k & R.mask)) : WFA_NULL);
      }
    }
    if (ef) reached = hit != 0;
    if (corrected) {
      const int64_t w = ~wave_max64(~cbest);  // the wave's minimum key
      ext_found = w != INT64_MAX;
      if (ext_found) {
        const uint64_t holder = ballot(cbest == w);
        ext_off = uniform(__shfl(cboff, (int)__builtin_ctzll(holder)));
        ext_left = (int)(w >> 32);
        ext_k = (int)(uint32_t)(w & 0xffffffffu) - (1 << 30);
      }
    }
    if (MODE == 1 && tid == 0) H.hdr[s] = hrow;
    bool any_reached = ballot(reached) != 0;
    if (NW > 1) {
      // every wave's first hit is its lowest diagonal: the block's is the
      // lowest of those
      int* x = red + 10 * NW;
      if (lane == 0) {
        x[3 * wave] = any_reached;
        x[3 * wave + 1] = hit ? end_k : INT32_MAX;
        x[3 * wave + 2] = end_off;
      }
      __syncthreads();
      any_reached = false;
      int bk = INT32_MAX, boff = 0;
      for (int w = 0; w < NW; ++w) {
        any_reached |= x[3 * w] != 0;
        if (x[3 * w + 1] < bk) {
          bk = x[3 * w + 1];
          boff = x[3 * w + 2];
        }
      }
      if (ef && bk != INT32_MAX) {
        end_k = bk;
        end_off = boff;
      }
    }
    return any_reached;
  };

  // ---- heuristic cut-off of M[s] after its extension (apply_heuristic_cutoff,
  // src/aligner.rs:1644-1766; heuristic_cutoff, src/heuristic.rs:167-525).
  // The cut narrows M's limits; its cells stay readable within the
  // initialized range (an emptied M reads as the null wavefront) and the
  // gap components of the same score take M's new limits.
  const int heur = HEUR ? hc.kind : WFA_HEURISTIC_NONE;
  HeurState hs;
  hs.reset(hc);
  // Returns true when Z-drop ends the alignment.
  auto cutoff = [&](int s, int rowM, int rowE, int rowE2) -> bool {
    const int4 mm = R.meta[rowM];
    const int lo = uniform(mm.x), hi = uniform(mm.y);
    if (hc.kind == WFA_HEURISTIC_ZDROP)
      return zdrop_check(p, hc, hs, R, rowM, lo, hi, s, sw_pb, sw_tb);
    int nlo = lo, nhi = hi;
    if (!heuristic_limits(p, hc, hc.p0, hc.p1, hs, R, rowM, lo, hi, s, plen, tlen, sw_pb,
                          sw_tb, &nlo, &nhi))
      return false;
    if (nlo == lo && nhi == hi) return false;
    const bool empty = nlo > nhi;
    __syncthreads();  // every lane has read M's cells and metadata
    if (tid == 0) {
      // an emptied M is the null wavefront with an empty initialized range
      R.meta[rowM] = empty ? make_int4(nlo, nhi, 1, 0) : make_int4(nlo, nhi, mm.z, mm.w);
      if (GAPS) {
        auto equate = [&](int row) {
          const int4 g = R.meta[row];
          if (g.x > g.y) return;
          const int l = imax(g.x, nlo), h = imin(g.y, nhi);
          R.meta[row] = l > h ? make_int4(l, h, 1, 0) : make_int4(l, h, g.z, g.w);
        };
        equate(rI1 + rowE);
        equate(rD1 + rowE);
        if (two) {
          equate(rI2 + rowE2);
          equate(rD2 + rowE2);
        }
      }
      if (MODE == 1 && empty) H.hdr[s] = make_int4(1, 0, 0, 0);
    }
    if (MODE == 1 && KIND == KIND_EDIT && !empty) {
      // the next edit step nulls the two cells beyond each end of M[s] in
      // place (src/aligner.rs:1796-1806); inside the initialized range the
      // backtrace sees them
      const int4 hd = H.hdr[s];
      if (lane < 4) {
        const int k = lane < 2 ? nlo - 2 + lane : nhi + lane - 1;
        if (k >= hd.x && k <= hd.y) H.cells[hist_row(H, hd) + (k - hd.x)] = WFA_NULL;
      }
    }
    __syncthreads();
    return false;
  };
  // Z-drop ended the alignment: the end position is the best cell seen
  // (src/aligner.rs:1690-1700); the alignment is partial
  bool zdropped = false;

  // Corrected: every component at score s keeps only the diagonals within
  // (total - s) / -match of k_end, the others NULLed (endsfree_prune,
  // src/aligner.rs:1203-1245, with trim_nulled, src/wavefront.rs:459-490);
  // `hist`: M's stored CIGAR row is NULLed too (high memory, after the
  // extension)
  auto band_prune = [&](int s, int total, int rowM, int rowE, int rowE2, bool hist) {
    const int r = (total - s) / mr;
    const int blo = k_end - r, bhi = k_end + r;
    auto trim = [&](int row, bool is_m) {
      const int4 mm = R.meta[row];
      const int l = uniform(mm.x), h = uniform(mm.y);
      if (l > h || (blo <= l && bhi >= h)) return;
      const int nl = imax(blo, l), nh = imin(bhi, h);
      int4 hd = make_int4(1, 0, 0, 0);
      if (MODE == 1 && is_m && hist) hd = H.hdr[s];
      for (int k = l + tid; k <= h; k += NT) {
        if (k >= nl && k <= nh) continue;
        R.set(row * R.wcap + (k & R.mask), WFA_NULL);
        if (MODE == 1 && k >= hd.x && k <= hd.y) H.cells[hist_row(H, hd) + (k - hd.x)] = WFA_NULL;
      }
      __syncthreads();
      if (tid == 0) R.meta[row] = make_int4(nl, nh, mm.z, mm.w);
      __syncthreads();
    };
    trim(rowM, true);
    if (GAPS) {
      trim(rI1 + rowE, false);
      trim(rD1 + rowE, false);
    }
    if (two) {
      trim(rI2 + rowE2, false);
      trim(rD2 + rowE2, false);
    }
  };
  // Corrected: M's extended offsets enter the dominance record
  // (endsfree_dominance_update, src/aligner.rs:1314-1330)
  auto dom_update = [&](int rowM) {
    const int4 mm = R.meta[rowM];
    const int l = uniform(mm.x), h = uniform(mm.y);
    int32_t* best = dom - hist_lo;
    for (int k = l + tid; k <= h; k += NT) {
      const int o = R.at(rowM * R.wcap + (k & R.mask));
      if (o > best[k]) best[k] = o;
    }
    __syncthreads();
  };
  auto cand_end = [&]() {
    status = WFA_STATUS_END_REACHED;
    result = cand_total;
    end_s = cand_s;
    end_k = cand_k;
    end_off = cand_off;
  };
  // Corrected: the extension of M[s] and its end test
  // (extend_endsfree_corrected, src/aligner.rs:1111-1200); true when the
  // alignment ends here
  auto corrected_step = [&](int s, int rowM, int rowE, int rowE2, int elo, int ehi,
                            int4 hrow) -> bool {
    if (cprune && cand_valid && cand_total > s) band_prune(s, cand_total, rowM, rowE, rowE2, false);
    extend_row(s, rowM, elo, ehi, hrow);
    __syncthreads();
    if (ext_found) {
      const int total = s + mr * ext_left;
      if (!cand_valid || total < cand_total) {
        cand_valid = true;
        cand_total = total;
        cand_s = s;
        cand_k = ext_k;
        cand_off = ext_off;
        if (cprune && total > s) band_prune(s, total, rowM, rowE, rowE2, true);
      }
    }
    if (dom_on) dom_update(rowM);
    if (cand_valid && cand_total <= s) {
      cand_end();
      return true;
    }
    if (heur != WFA_HEURISTIC_NONE && cutoff(s, rowM, rowE, rowE2)) {
      if (cand_valid) {
        cand_end();
      } else {
        status = WFA_STATUS_END_UNREACHABLE;
        result = s;
        zdropped = true;
      }
      return true;
    }
    return false;
  };

  // score 0
  {
    int4 h0 = make_int4(lo0, hi0, 0, 0);
    if (CIG && !overflow && (H.hdr_cap < 1 || !halloc(hi0 - lo0 + 1, &h0, lo0, hi0)))
      overflow = true;
    else if (MED && tid == 0) H.hdr[0] = h0;
    if (corrected) {
      if (!overflow) corrected_step(0, 0, 0, 0, lo0, hi0, h0);
    } else if (!overflow && extend_row(0, 0, lo0, hi0, h0)) {
      status = WFA_STATUS_END_REACHED;
      result = 0;
    }
    __syncthreads();
    if (!corrected && !overflow && status == WFA_STATUS_OK && heur != WFA_HEURISTIC_NONE &&
        cutoff(0, 0, 0, 0)) {
      status = WFA_STATUS_END_UNREACHABLE;
      result = 0;
      zdropped = true;
    }
  }
  __syncthreads();

  // s mod ringM / ringE1 / ringE2, kept as the score advances
  int sM = 0, sE1 = 0, sE2 = 0;
  // (s - d) mod ring for 0 <= d < ring
  auto back = [](int cur, int d, int ring) { return cur >= d ? cur - d : cur - d + ring; };
  // the first score at which one of the per-score limits below applies
  // (one compare per score: the kernel is bound by scalar instructions)
  int limit = INT32_MAX;
  if (!p.modular) limit = imin(limit, nwf);
  if (CIG) limit = imin(limit, H.hdr_cap);
  if (sizeof(RT) == 2) limit = imin(limit, 16001);
  while (status == WFA_STATUS_OK && !overflow) {
    score += 1;
    sM = sM + 1 == ringM ? 0 : sM + 1;
    if (GAPS) sE1 = sE1 + 1 == ringE1 ? 0 : sE1 + 1;
    if (two) sE2 = sE2 + 1 == ringE2 ? 0 : sE2 + 1;
    if (score >= limit) {
      if (!p.modular && score >= nwf) { status = WFA_STATUS_MAX_STEPS_REACHED; result = score; break; }
      if (CIG && score >= H.hdr_cap) { overflow = true; break; }
      // int16 rows hold offsets up to tlen + score: leave long runs to int32
      if (sizeof(RT) == 2 && score > 16000) { overflow = true; break; }
    }
    const int s = score;
    // ---- input rows ----
    int lo = INT32_MAX, hi = INT32_MIN;
    // limits (ml/mh, il/ih, ...) and initialized ranges (gx..., the read guards)
    int ml_x = 1, mh_x = 0, ml_o = 1, mh_o = 0, il = 1, ih = 0, dl = 1, dh = 0;
    int ml_o2 = 1, mh_o2 = 0, il2 = 1, ih2 = 0, dl2 = 1, dh2 = 0;
    int gl_x = 1, gh_x = 0, gl_o = 1, gh_o = 0, gil = 1, gih = 0, gdl = 1, gdh = 0;
    int gl_o2 = 1, gh_o2 = 0, gil2 = 1, gih2 = 0, gdl2 = 1, gdh2 = 0;
    int row_x = 0, row_o = 0, row_e = 0, row_o2 = 0, row_e2 = 0;
    bool any = false;
    auto load = [&](int row, int& l, int& h, int& gl, int& gh) {
      const int4 m = R.meta[row];
      l = uniform(m.x); h = uniform(m.y);
      gl = uniform(m.z); gh = uniform(m.w);
    };
    if (KIND == KIND_EDIT) {
      row_x = back(sM, 1, ringM);
      load(row_x, ml_x, mh_x, gl_x, gh_x);
      // the edit step nulls prev's cells next to [lo, hi] before reading
      // them (src/aligner.rs:1795-1808): reads outside [lo, hi] are NULL
      gl_x = ml_x; gh_x = mh_x;
      if (ml_x != META_ABSENT && ml_x <= mh_x) {
        any = true;
        lo = ml_x - 1;
        hi = mh_x + 1;
      }
    } else {
      // gap-linear (compute_linear_step, src/aligner.rs:1866-2002): M at
      // s - x and s - indel; its inputs' ends are NULL-filled only beyond
      // their initialized ranges, as the affine inputs'
      // An input below score 0 maps to a ring slot this pair has not
      // written yet (every lookback is shorter than its ring), whose
      // metadata is still absent: it reads as the null wavefront, so the
      // loads need no score checks (the kernel is bound by scalar
      // instructions per score).
      const int go = GAPS ? p.oe1 : p.o1;
      row_x = back(sM, p.x, ringM);
      load(row_x, ml_x, mh_x, gl_x, gh_x);
      row_o = back(sM, go, ringM);
      load(row_o, ml_o, mh_o, gl_o, gh_o);
      if (GAPS) {
        row_e = back(sE1, p.e1, ringE1);
        load(rI1 + row_e, il, ih, gil, gih);
        load(rD1 + row_e, dl, dh, gdl, gdh);
      }
      if (two) {
        row_o2 = back(sM, p.oe2, ringM);
        load(row_o2, ml_o2, mh_o2, gl_o2, gh_o2);
        row_e2 = back(sE2, p.e2, ringE2);
        load(rI2 + row_e2, il2, ih2, gil2, gih2);
        load(rD2 + row_e2, dl2, dh2, gdl2, gdh2);
      }
      // an input whose limits are empty is the null wavefront; so are its
      // reads. Only a heuristic cut or a Corrected band prune leaves a row
      // readable beyond empty limits: in the other kernels every row's
      // readable range equals its limits already.
      if (HEUR) {
        auto ng = [](int l, int h, int& gl, int& gh) { if (l > h) { gl = 1; gh = 0; } };
        ng(ml_x, mh_x, gl_x, gh_x); ng(ml_o, mh_o, gl_o, gh_o); ng(il, ih, gil, gih);
        ng(dl, dh, gdl, gdh); ng(ml_o2, mh_o2, gl_o2, gh_o2); ng(il2, ih2, gil2, gih2);
        ng(dl2, dh2, gdl2, gdh2);
      }
      // Absent rows (META_ABSENT, hi = META_ABSENT-1), missing inputs and
      // trimmed- or cut-empty rows all have lo > hi and read as the null
      // wavefront, whose bounds (1, -1) still enter the range union
      // (src/aligner.rs:2045-2062, 2265-2289). Without a heuristic the extra
      // cells only see NULL and are trimmed away, so those kernels leave the
      // null bounds out; after a cut they can read the cut cells, and with
      // begin-free cells to add the computed range decides which stay.
      auto bound = [&](int l, int h, int dlo, int dhi) {
        if (l <= h) any = true;
        else if (!HEUR) return;
        else { l = 1; h = -1; }
        lo = imin(lo, l + dlo);
        hi = imax(hi, h + dhi);
      };
      bound(ml_x, mh_x, 0, 0);
      bound(ml_o, mh_o, -1, 1);
      if (GAPS) {
        bound(il, ih, 1, 1);
        bound(dl, dh, -1, -1);
      }
      if (two) {
        bound(ml_o2, mh_o2, -1, 1);
        bound(il2, ih2, 1, 1);
        bound(dl2, dh2, -1, -1);
      }
      // clamp_to_band (src/aligner.rs:3308-3316); an emptied range takes
      // the no-output path below
      if (heur == WFA_HEURISTIC_BANDED_STATIC || heur == WFA_HEURISTIC_BANDED_ADAPTIVE) {
        lo = imax(lo, hc.p0);
        hi = imin(hi, hc.p1);
      }
    }
    // affine2p with no second-piece input runs the gap-affine kernel
    // (src/aligner.rs:2261-2264)
    const bool use2 = two && (ml_o2 <= mh_o2 || il2 <= ih2 || dl2 <= dh2);
    const int rowM = sM;
    const int rowE = sE1;
    const int rowE2 = two ? sE2 : 0;
    // ends-free with a match reward: the begin-free cells (h = d, 0) and
    // (0, v = d) are reached at score d * -match (endsfree_begin_k,
    // src/aligner.rs:3369-3386); -1 when this score has none
    int seed_k = -1;
    if (KIND != KIND_EDIT && seedable && s % (-p.u_match) == 0) {
      const int d = s / (-p.u_match);
      if (tb_free >= d || pb_free >= d) seed_k = d;
    }
    // a begin-free score whose inputs are all null: M holds just the
    // begin-free cells, and the step still counts as null
    // (endsfree_null_m, src/aligner.rs:3439-3465)
    const bool seed_only = !any && seed_k >= 0;
    __syncthreads();  // every lane has read the metadata before it is rewritten
    if (!any && !seed_only) {
      // null step (src/aligner.rs:2031-2043, 1778-1782)
      nulls += 1;
      if (tid == 0) {
        R.meta[rowM] = meta_absent();
        if (GAPS) {
          R.meta[rI1 + rowE] = meta_absent();
          R.meta[rD1 + rowE] = meta_absent();
        }
        if (two) {
          R.meta[rI2 + rowE2] = meta_absent();
          R.meta[rD2 + rowE2] = meta_absent();
        }
        if (CIG) H.hdr[s] = make_int4(1, 0, 0, 0);
      }
      ztr_put(s, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 0);
      __syncthreads();
      if (s >= p.max_steps) { status = WFA_STATUS_MAX_STEPS_REACHED; result = s; break; }
      // Corrected: with a candidate, null steps do not end the search
      if (corrected && cand_valid) {
        if (cand_total <= s) { cand_end(); break; }
        continue;
      }
      if (nulls > p.scope) { status = WFA_STATUS_END_UNREACHABLE; result = s; break; }
      continue;
    }
    if (seed_only) {
      nulls += 1;
    } else {
      nulls = 0;
      lo = imax(lo, hist_lo + 1);
      hi = imin(hi, hist_hi - 1);
    }
    if (!seed_only && lo > hi) {
      // a band (or the historic limits) left nothing to compute: no outputs
      // at this score (src/aligner.rs:2058-2081)
      if (tid == 0) {
        R.meta[rowM] = meta_absent();
        if (GAPS) {
          R.meta[rI1 + rowE] = meta_absent();
          R.meta[rD1 + rowE] = meta_absent();
        }
        if (two) {
          R.meta[rI2 + rowE2] = meta_absent();
          R.meta[rD2 + rowE2] = meta_absent();
        }
        if (CIG) H.hdr[s] = make_int4(1, 0, 0, 0);
      }
      ztr_put(s, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 0);
      __syncthreads();
      if (s >= p.max_steps) { status = WFA_STATUS_MAX_STEPS_REACHED; result = s; break; }
      if (corrected && cand_valid && cand_total <= s) { cand_end(); break; }
      continue;
    }
    // [lo, hi]: the cells the recurrences compute; [wlo, whi]: the row,
    // widened over NULL cells to reach the begin-free cells (endsfree_init,
    // src/aligner.rs:3388-3419): the text cell replaces a computed cell
    // that reaches no further, and is lost below the computed range; the
    // pattern cell likewise above it
    // The seeding state lives in one scalar across the recurrences' loop
    // (register pressure): seeds = -1, or d | SEED_ONLY | SEED_TEXT |
    // SEED_PATTERN.
    constexpr int SEED_ONLY = 1 << 28, SEED_TEXT = 1 << 29, SEED_PATTERN = 1 << 30;
    int seeds = -1;
    if (seed_k >= 0) {
      const int d = seed_k;
      bool tseed, pseed;
      if (seed_only) {
        tseed = tb_free >= d;
        pseed = pb_free >= d;
        lo = 1;
        hi = 0;
      } else {
        tseed = tb_free >= d && d >= lo;
        pseed = pb_free >= d && -d <= (tseed ? imax(hi, d) : hi);
      }
      seeds = d | (seed_only ? SEED_ONLY : 0) | (tseed ? SEED_TEXT : 0) |
              (pseed ? SEED_PATTERN : 0);
    }
    // the row's range from the computed range and the seeds
    auto widened = [&](int& wl, int& wh) {
      wl = lo;
      wh = hi;
      if (seeds < 0) return;
      const int d = seeds & (SEED_ONLY - 1);
      const bool ts = seeds & SEED_TEXT, ps = seeds & SEED_PATTERN;
      if (seeds & SEED_ONLY) {
        wl = ps ? -d : d;
        wh = ts ? d : -d;
      } else {
        if (ts) wh = imax(hi, d);
        if (ps) wl = imin(lo, -d);
      }
    };
    int wlo, whi;
    widened(wlo, whi);
    if (zc) {
      // the wavefronts of this score are allocated over the computed range,
      // widened to both begin-free diagonals of a begin-free score
      // (output_alloc_range_at, endsfree_null_m, src/aligner.rs:3318-3465)
      int al = wlo, ah = whi;
      if (seeds >= 0 && !(seeds & SEED_ONLY)) {
        al = imin(lo, -seed_k);
        ah = imax(hi, seed_k);
      }
      zalo = imin(zalo, al - zpad);
      zahi = imax(zahi, ah + zpad);
    }
    if (whi - wlo + 1 > R.wcap) { overflow = true; break; }
    int4 hrow = make_int4(wlo, whi, 0, 0);
    if (MED && !halloc(whi - wlo + 1, &hrow, wlo, whi)) { overflow = true; break; }
    // winner flags of diagonal k at fbase + (k - lo)
    const int64_t fbase = MED ? hist_row(H, hrow) + (lo - wlo) : 0;
    // Several waves, end to end, high memory: the extended M row goes to
    // the history as computed, over the computed range (its cells outside
    // the trimmed range are NULL, and the backtrace reads NULL and missing
    // cells alike)
    const bool hfused = MODE == 1 && NW > 1 && !ef;
    if (hfused && !halloc(whi - wlo + 1, &hrow, wlo, whi)) { overflow = true; break; }
    const int64_t hbase = hfused ? hist_row(H, hrow) - wlo : 0;

    // Several waves, end to end: each lane extends the M cell it computes,
    // and the row is written once (the extension of a cell reads only that
    // cell and the sequences; the trims test the unextended offsets, which
    // the extension keeps within the sequences). The end test is the end
    // diagonal's cell.
    constexpr bool FUSE = NW > 1;
    const bool fuse = FUSE && !ef;
    bool freach = false;
    // ---- compute all chunks; trims by ballot ----
    // U chunks at a time: their input reads are all issued before the
    // first chunk's outputs are written (outputs go to other ring rows), so
    // a wave keeps U chunks of loads in flight on wide HBM-resident rows.
    int m_lo = INT32_MAX, m_hi = INT32_MIN, i_lo = INT32_MAX, i_hi = INT32_MIN;
    int d_lo = INT32_MAX, d_hi = INT32_MIN, i2_lo = INT32_MAX, i2_hi = INT32_MIN;
    int d2_lo = INT32_MAX, d2_hi = INT32_MIN;
    struct In {
      int a, b, c, d, e, f, g, h, i;
    };
    // input rows: base index and readable range
    const int bx = row_x * R.wcap, bo = row_o * R.wcap, bi1 = (rI1 + row_e) * R.wcap,
              bd1 = (rD1 + row_e) * R.wcap, bo2 = row_o2 * R.wcap,
              bi2 = (rI2 + row_e2) * R.wcap, bd2 = (rD2 + row_e2) * R.wcap;
    const Guard g_x = make_guard(gl_x, gh_x), g_o = make_guard(gl_o, gh_o),
                g_i1 = make_guard(gil, gih), g_d1 = make_guard(gdl, gdh),
                g_o2 = make_guard(gl_o2, gh_o2), g_i2 = make_guard(gil2, gih2),
                g_d2 = make_guard(gdl2, gdh2);
    // BFREE: branch-free reads (every lane reads, a select applies the
    // range); otherwise reads outside a row's range are skipped, which the
    // wide CIGAR kernels measure faster (their register pressure would
    // serialize nine unconditional HBM loads)
    auto gather = [&](int k, In& in) {
      if constexpr (!BFREE) {
        if (k > hi) return;
        if (KIND == KIND_EDIT) {
          in.a = rd(R, row_x, gl_x, gh_x, k - 1);
          in.b = rd(R, row_x, gl_x, gh_x, k + 1);
          if (p.metric != WFA_METRIC_INDEL) in.c = rd(R, row_x, gl_x, gh_x, k);
        } else {
          in.a = rd(R, row_o, gl_o, gh_o, k - 1);
          if (GAPS) in.b = rd(R, rI1 + row_e, gil, gih, k - 1);
          in.c = rd(R, row_o, gl_o, gh_o, k + 1);
          if (GAPS) in.d = rd(R, rD1 + row_e, gdl, gdh, k + 1);
          in.e = rd(R, row_x, gl_x, gh_x, k);
          if (two) {
            in.f = rd(R, row_o2, gl_o2, gh_o2, k - 1);
            in.g = rd(R, rI2 + row_e2, gil2, gih2, k - 1);
            in.h = rd(R, row_o2, gl_o2, gh_o2, k + 1);
            in.i = rd(R, rD2 + row_e2, gdl2, gdh2, k + 1);
          }
        }
        return;
      }
      if (KIND == KIND_EDIT) {
        in.a = rdg(R, bx, g_x, k - 1);
        in.b = rdg(R, bx, g_x, k + 1);
        if (p.metric != WFA_METRIC_INDEL) in.c = rdg(R, bx, g_x, k);
      } else {
        in.a = rdg(R, bo, g_o, k - 1);
        if (GAPS) in.b = rdg(R, bi1, g_i1, k - 1);
        in.c = rdg(R, bo, g_o, k + 1);
        if (GAPS) in.d = rdg(R, bd1, g_d1, k + 1);
        in.e = rdg(R, bx, g_x, k);
        if (two) {
          in.f = rdg(R, bo2, g_o2, k - 1);
          in.g = rdg(R, bi2, g_i2, k - 1);
          in.h = rdg(R, bo2, g_o2, k + 1);
          in.i = rdg(R, bd2, g_d2, k + 1);
        }
      }
    };
    auto chunk = [&](int kb, const In& in) {
      const int k = kb + lane;
      bool vm = false, vi = false, vd = false, vi2 = false, vd2 = false;
      if (k <= hi) {
        int m;
        int fl = 0;
        int gsg = 0;  // signs of the gap components' offsets
        if (KIND == KIND_EDIT) {
          const int ins = in.a;
          const int del = in.b;
          if (p.metric == WFA_METRIC_INDEL) {
            m = imax(del, ins + 1);
            if (MED) fl = m == del ? 1 : 0;
          } else {
            const int mis = in.c;
            m = imax(del, imax(ins, mis) + 1);
            // piggyback priority (src/compute/piggyback.rs:218-224):
            // mismatch, then deletion, then insertion
            if (MED) fl = m == mis + 1 ? 2 : m == del ? 1 : 0;
          }
          if (!in_bounds(m, k, plen, tlen)) m = WFA_NULL;
        } else if (KIND == KIND_LINEAR) {
          // compute_linear_idm (src/compute/linear.rs:20-131); the same
          // piggyback priority as edit (src/compute/piggyback.rs:279-302)
          const int ins = in.a + 1, del = in.c, mis = in.e + 1;
          m = imax(del, imax(ins, mis));
          if (MED) fl = m == mis ? 2 : m == del ? 1 : 0;
          if (!in_bounds(m, k, plen, tlen)) m = WFA_NULL;
        } else {
          const int io = in.a, ie = in.b;
          const int dop = in.c, de = in.d;
          const int i1 = imax(io, ie) + 1;
          const int d1 = imax(dop, de);
          const int mis = in.e + 1;
          // a gap cell takes its extension input on ties
          // (src/compute/piggyback.rs:146-160)
          if (MED) fl |= (ie >= io ? FL_I1_EXT : 0) | (de >= dop ? FL_D1_EXT : 0);
          int ins = i1, del = d1;
          int i2 = WFA_NULL, d2 = WFA_NULL;
          if (two) {
            const int io2 = in.f, ie2 = in.g;
            const int do2 = in.h, de2 = in.i;
            i2 = imax(io2, ie2) + 1;
            d2 = imax(do2, de2);
            // with no second-piece input the reference runs the gap-affine
            // kernel (src/aligner.rs:2261-2264): I2/D2 never win
            if (zc && !use2) i2 = d2 = INT32_MIN;
            if (MED) fl |= (ie2 >= io2 ? FL_I2_EXT : 0) | (de2 >= do2 ? FL_D2_EXT : 0);
            R.set((rI2 + rowE2) * R.wcap + (k & R.mask), i2);
            R.set((rD2 + rowE2) * R.wcap + (k & R.mask), d2);
            vi2 = in_bounds(i2, k, plen, tlen);
            vd2 = in_bounds(d2, k, plen, tlen);
            ins = imax(ins, i2);
            del = imax(del, d2);
          }
          m = imax(mis, imax(ins, del));
          if (MED) {
            // M's source, a later one winning ties (src/compute/piggyback.rs
            // 383-391, 515-525): I1, I2, D1, D2, then the mismatch
            int w = 0;
            if (two && m == i2) w = 1;
            if (m == d1) w = 2;
            if (two && m == d2) w = 3;
            if (m == mis) w = 4;
            fl |= w;
          }
          if (!in_bounds(m, k, plen, tlen)) m = WFA_NULL;
          R.set((rI1 + rowE) * R.wcap + (k & R.mask), i1);
          R.set((rD1 + rowE) * R.wcap + (k & R.mask), d1);
          vi = in_bounds(i1, k, plen, tlen);
          vd = in_bounds(d1, k, plen, tlen);
          gsg = (i1 >= 0) << 1 | (d1 >= 0) << 2 | (i2 >= 0) << 3 | (d2 >= 0) << 4;
        }
        vm = m != WFA_NULL;
        if (fuse && vm) {
          // the extension of M[s] fused into its computation (a valid cell
          // is in the sequences, so m >= 0)
          m += lcp<BITS>(sP, sT, m - k, m, plen, tlen);
          if (k == k_end && m >= tlen) freach = true;
        }
        R.set(rowM * R.wcap + (k & R.mask), m);
        if (MODE == 1 && FUSE && fuse) H.cells[hbase + k] = m;
        if (MED) H.bytes[fbase + (k - lo)] = (uint8_t)fl;
        if (MED && zc) {
          // the offsets' signs (offload_selective moves only offsets >= 0)
          H.signs[fbase + (k - lo)] = (uint8_t)((m >= 0) | gsg);
        }
      }
      uint64_t b = ballot(vm);
      if (b) {
        if (m_lo == INT32_MAX) m_lo = kb + __builtin_ctzll(b);
        m_hi = kb + 63 - __builtin_clzll(b);
      }
      if (GAPS) {
        b = ballot(vi);
        if (b) {
          if (i_lo == INT32_MAX) i_lo = kb + __builtin_ctzll(b);
          i_hi = kb + 63 - __builtin_clzll(b);
        }
        b = ballot(vd);
        if (b) {
          if (d_lo == INT32_MAX) d_lo = kb + __builtin_ctzll(b);
          d_hi = kb + 63 - __builtin_clzll(b);
        }
        if (two) {
          b = ballot(vi2);
          if (b) {
            if (i2_lo == INT32_MAX) i2_lo = kb + __builtin_ctzll(b);
            i2_hi = kb + 63 - __builtin_clzll(b);
          }
          b = ballot(vd2);
          if (b) {
            if (d2_lo == INT32_MAX) d2_lo = kb + __builtin_ctzll(b);
            d2_hi = kb + 63 - __builtin_clzll(b);
          }
        }
      }
    };
    if (NW == 1) {
      int kb = lo;
      for (; kb + 64 * (U - 1) <= hi; kb += 64 * U) {
        In in[U];
#pragma unroll
        for (int u = 0; u < U; ++u) gather(kb + 64 * u + lane, in[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) chunk(kb + 64 * u, in[u]);
      }
      for (; kb <= hi; kb += 64) {
        In in;
        gather(kb + lane, in);
        chunk(kb, in);
      }
    } else {
      for (int kb = lo + 64 * wave; kb <= hi; kb += NT) {
        In in;
        gather(kb + lane, in);
        chunk(kb, in);
      }
      // every wave saw its own chunks: combine the trims (and, fused, the
      // end test)
      const bool wreach = FUSE && ballot(freach) != 0;
      if (lane == 0) {
        int* x = red + 10 * wave;
        x[0] = m_lo; x[1] = m_hi; x[2] = i_lo; x[3] = i_hi; x[4] = d_lo;
        x[5] = d_hi; x[6] = i2_lo; x[7] = i2_hi; x[8] = d2_lo; x[9] = d2_hi;
        if (FUSE) red[10 * NW + 3 * wave] = wreach;
      }
      __syncthreads();
      if (FUSE) {
        freach = false;
        for (int w = 0; w < NW; ++w) freach |= red[10 * NW + 3 * w] != 0;
      }
      m_lo = i_lo = d_lo = i2_lo = d2_lo = INT32_MAX;
      m_hi = i_hi = d_hi = i2_hi = d2_hi = INT32_MIN;
      for (int w = 0; w < NW; ++w) {
        const int* x = red + 10 * w;
        m_lo = imin(m_lo, x[0]); m_hi = imax(m_hi, x[1]);
        i_lo = imin(i_lo, x[2]); i_hi = imax(i_hi, x[3]);
        d_lo = imin(d_lo, x[4]); d_hi = imax(d_hi, x[5]);
        i2_lo = imin(i2_lo, x[6]); i2_hi = imax(i2_hi, x[7]);
        d2_lo = imin(d2_lo, x[8]); d2_hi = imax(d2_hi, x[9]);
      }
    }
    if (KIND != KIND_EDIT && seeds >= 0) {
      const int seed_k = seeds & (SEED_ONLY - 1);
      const bool seed_only = seeds & SEED_ONLY;
      const bool tseed = seeds & SEED_TEXT, pseed = seeds & SEED_PATTERN;
      int wlo, whi;
      widened(wlo, whi);
      // the begin-free cells (endsfree_init / endsfree_null_m,
      // src/aligner.rs:3388-3465), after the recurrences: the widened cells
      // are NULL in every component, and a begin-free cell replaces the
      // cell on its diagonal that reaches no further (endsfree_init_offset,
      // src/aligner.rs:3421-3437; a medium/low-memory chain starts there)
      __syncthreads();  // the computed cells are written
      for (int k = wlo + tid; k <= whi; k += NT) {
        const int cm = rowM * R.wcap + (k & R.mask);
        const bool in = k >= lo && k <= hi;
        int m = in ? R.at(cm) : WFA_NULL;
        if (!in) {
          if (GAPS) {
            R.set((rI1 + rowE) * R.wcap + (k & R.mask), WFA_NULL);
            R.set((rD1 + rowE) * R.wcap + (k & R.mask), WFA_NULL);
          }
          if (two) {
            R.set((rI2 + rowE2) * R.wcap + (k & R.mask), WFA_NULL);
            R.set((rD2 + rowE2) * R.wcap + (k & R.mask), WFA_NULL);
          }
          if (MED) H.bytes[fbase + (k - lo)] = FL_M_NULLED;
          if (MED && zc) H.signs[fbase + (k - lo)] = 0;
        }
        const bool seed = (tseed && k == seed_k && m <= seed_k) || (pseed && k == -seed_k && m <= 0);
        if (seed) {
          m = k > 0 ? k : 0;
          if (MED) H.bytes[fbase + (k - lo)] = (H.bytes[fbase + (k - lo)] & ~7) | FL_M_SEED;
          // a seed beyond its sequence (a free begin gap longer than it)
          // is out of bounds, the end cell trim_ends drops
          // (src/wavefront.rs); a seed-only M keeps it (endsfree_null_m)
          if (!seed_only && !in_bounds(m, k, plen, tlen)) m = WFA_NULL;
        }
        if (!in || seed) R.set(cm, m);
      }
      __syncthreads();
      // a begin-free cell widens M's range
      auto widen = [&](int k) {
        if (uniform(R.at(rowM * R.wcap + (k & R.mask))) >= 0) {
          m_lo = imin(m_lo, k);
          m_hi = imax(m_hi, k);
        }
      };
      if (pseed) widen(-seed_k);
      if (tseed) widen(seed_k);
    }
    // trimmed-empty rows: lo > hi (but not META_ABSENT)
    auto fix = [](int& l, int& h) { if (l == INT32_MAX) { l = 1; h = 0; } };
    fix(m_lo, m_hi); fix(i_lo, i_hi); fix(d_lo, d_hi); fix(i2_lo, i2_hi); fix(d2_lo, d2_hi);
    __syncthreads();  // ring rows written before the extension reads them
    if (dom_on && (seeds < 0 || !(seeds & SEED_ONLY))) {
      // Corrected: drop from each component's ends the cells that reach no
      // further than that component did on their diagonal at an earlier
      // score; I/D record what they keep (endsfree_dominance_prune,
      // src/aligner.rs:1247-1312, order as src/aligner.rs:2183-2190)
      auto dom_prune = [&](int row, int c, bool record, int& l, int& h) {
        if (l > h) return;
        int32_t* best = dom + c * dlen - hist_lo;
        const int rb = row * R.wcap;
        int first = INT32_MAX, last = INT32_MIN;
        for (int kb = l; kb <= h; kb += 64) {
          const int k = kb + lane;
          const uint64_t b = ballot(k <= h && R.at(rb + (k & R.mask)) > best[k]);
          if (b) {
            if (first == INT32_MAX) first = kb + __builtin_ctzll(b);
            last = kb + 63 - __builtin_clzll(b);
          }
        }
        for (int k = l + lane; k <= h; k += 64) {
          const int cell = rb + (k & R.mask);
          if (k < first || k > last) {
            R.set(cell, WFA_NULL);
          } else if (record) {
            const int o = R.at(cell);
            if (o > best[k]) best[k] = o;
          }
        }
        if (first == INT32_MAX) {
          l = 1;
          h = 0;
        } else {
          l = first;
          h = last;
        }
      };
      dom_prune(rowM, 0, false, m_lo, m_hi);
      if (GAPS) {
        dom_prune(rI1 + rowE, 1, true, i_lo, i_hi);
        dom_prune(rD1 + rowE, 2, true, d_lo, d_hi);
      }
      if (two) {
        dom_prune(rI2 + rowE2, 3, true, i2_lo, i2_hi);
        dom_prune(rD2 + rowE2, 4, true, d2_lo, d2_hi);
      }
      __syncthreads();
    }
    if ((MED || (MODE == 1 && FUSE && fuse)) && tid == 0) H.hdr[s] = hrow;
    int zo = 0;
    if (zc && !(seeds >= 0 && (seeds & SEED_ONLY))) {
      // this step's occupancies (a begin-free-only M is a fresh wavefront)
      if (KIND == KIND_LINEAR) {
        zo = imax(zocc(s - p.x, 0), zocc(s - p.o1, 0)) + 1;
      } else if (GAPS) {
        // occ_gap: one more than the fuller of a gap's open and extend inputs
        auto gap = [&](int so, int se, int c) {
          return imax(zlive(so, 0) ? zocc(so, 0) + 1 : 0, zlive(se, c) ? zocc(se, c) + 1 : 0);
        };
        const int oi1 = gap(s - p.oe1, s - p.e1, 1), od1 = gap(s - p.oe1, s - p.e1, 2);
        const int oi2 = two && use2 ? gap(s - p.oe2, s - p.e2, 3) : 0;
        const int od2 = two && use2 ? gap(s - p.oe2, s - p.e2, 4) : 0;
        const int om = imax(imax(zocc(s - p.x, 0), oi1), imax(od1, imax(oi2, od2))) + 1;
        zo = om | oi1 << 5 | od1 << 10 | oi2 << 15 | od2 << 20;
      }
    }
    ztr_put(s, m_lo, m_hi, i_lo, i_hi, d_lo, d_hi, i2_lo, i2_hi, d2_lo, d2_hi, zo);
    if (tid == 0) {
      R.meta[rowM] = make_int4(m_lo, m_hi, m_lo, m_hi);
      if (GAPS) {
        R.meta[rI1 + rowE] = make_int4(i_lo, i_hi, i_lo, i_hi);
        R.meta[rD1 + rowE] = make_int4(d_lo, d_hi, d_lo, d_hi);
      }
      if (two) {
        R.meta[rI2 + rowE2] = make_int4(i2_lo, i2_hi, i2_lo, i2_hi);
        R.meta[rD2 + rowE2] = make_int4(d2_lo, d2_hi, d2_lo, d2_hi);
      }
    }
    if (s >= p.max_steps) { status = WFA_STATUS_MAX_STEPS_REACHED; result = s; __syncthreads(); break; }
    if (FUSE && fuse) {
      // M[s] is extended already
      __syncthreads();
      if (freach) { status = WFA_STATUS_END_REACHED; result = s; break; }
      continue;
    }
    // ---- extend M[s] (the reference extends at the top of the next
    // iteration; nothing reads M[s] in between) ----
    // In high memory the stored row keeps its computed range: M cells
    // outside the trimmed range are OFFSET_NULL there anyway.
    int elo = m_lo, ehi = m_hi;
    if (MODE == 1) widened(elo, ehi);
    if (MODE == 1 && !halloc(ehi - elo + 1, &hrow, elo, ehi)) { overflow = true; break; }
    if (corrected) {
      if (corrected_step(s, rowM, rowE, rowE2, elo, ehi, hrow)) break;
      continue;
    }
    const bool reached =
        extend_row(s, rowM, elo, ehi, hrow) && (ef || (m_lo <= k_end && k_end <= m_hi));
    __syncthreads();
    if (reached) { status = WFA_STATUS_END_REACHED; result = s; break; }
    if (heur != WFA_HEURISTIC_NONE && cutoff(s, rowM, rowE, rowE2)) {
      status = WFA_STATUS_END_UNREACHABLE;
      result = s;
      zdropped = true;
      break;
    }
  }

  if (overflow) {
    if (tid == 0) {
      const int slot = atomicAdd(&p.counters[1], 1);
      p.retry_list[slot] = pair;
      p.status[pair] = ST_RETRY;
    }
    return;
  }
  int out_status = status;
  int ncig = 0;
  // the end position (alignment_end_pos): the end reached, the Z-drop best
  // cell (found at an earlier score), or unset
  int pos_s = -1, pos_k = INT32_MAX, pos_off = WFA_NULL;
  if (status == WFA_STATUS_END_REACHED) {
    // (a Corrected candidate can be from an earlier score)
    pos_s = end_s >= 0 ? end_s : score; pos_k = end_k; pos_off = end_off;
  } else if (zdropped) {
    pos_s = hs.max_wf; pos_k = hs.best_k; pos_off = hs.best_off;
  }
  if (status == WFA_STATUS_END_REACHED || zdropped) {
    // a Z-dropped alignment is partial: backtraced from its best cell at
    // the score that cell was found at (high memory, Compat::C,
    // src/aligner.rs:1393-1406) and max-trimmed (src/aligner.rs:683-694,
    // 784-806); its score is the score it stopped at
    out_status = zdropped ? WFA_STATUS_PARTIAL : WFA_STATUS_COMPLETED;
    if (CIG && wave == 0) {
      __threadfence_block();
      uint8_t* out = p.cig + p.cig_off[pair];
      const int cap = plen + tlen;
      if (MED && zc && zdropped) {
        // Compat::C: the best cell's diagonal at the stop score, as it is
        // there (src/aligner.rs:1443-1460; HPCBio/wfa2lib-rs#22)
        ncig = pos_k < zalo || pos_k > zahi
                   ? 0
                   : piggyback_backtrace<KIND, BITS>(p, H, score + 1, sP, sT, plen, tlen, score,
                                                     pos_k, pos_off, out, cap, ztr);
      } else if (MED)
        ncig = piggyback_backtrace<KIND, BITS>(p, H, score + 1, sP, sT, plen, tlen, pos_s, pos_k,
                                               pos_off, out, cap);
      else if (!GAPS)
        ncig = backtrace_linear<BITS>(p, H, score + 1, plen, tlen, pos_s, pos_k, pos_off, out,
                                      cap);
      else
        ncig = singletrack<BITS>(p, PoolView{H, score + 1, nwf}, sP, sT, plen, tlen, pos_s,
                                 pos_k, pos_off, out, cap);
      if (ncig < 0) {
        // the stored flags do not lead back to the start: never expected;
        // report the pair as unattainable rather than a wrong CIGAR
        out_status = WFA_STATUS_END_UNREACHABLE;
        ncig = 0;
      }
      // move the ops from the tail of the region to its head
      const int gap = MED ? 0 : cap - ncig;
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
      if ((zdropped || (ef && ext)) && KIND != KIND_EDIT && ncig > 0) {
        __threadfence_block();
        bool trimmed = false;
        ncig = maxtrim_cigar(p, out, ncig, pb_free, tb_free, &trimmed);
        if (trimmed) out_status = WFA_STATUS_PARTIAL;
      } else if (!zdropped && !corrected && heur != WFA_HEURISTIC_NONE && GAPS) {
        // (a Corrected candidate's total is its score)
        __threadfence_block();
        const int trailing = ef ? (plen - (end_off - end_k)) + (tlen - end_off) : 0;
        result = rescore_cigar(p, out, ncig, ef, pb_free, tb_free, trailing);
      }
    }
  } else if (status == WFA_STATUS_END_UNREACHABLE && ef) {
    // an ends-free alignment that ran out of wavefront is partial
    // (src/aligner.rs:819-825)
    out_status = WFA_STATUS_PARTIAL;
  }
  if (tid == 0) {
    p.score[pair] = result;
    p.status[pair] = out_status;
    if (CIG) p.cig_len[pair] = ncig;
    if (p.end_pos) {
      p.end_pos[3 * pair] = pos_s;
      p.end_pos[3 * pair + 1] = pos_k;
      p.end_pos[3 * pair + 2] = pos_off;
    }
  }
}

template <int KIND, int MODE, bool LDS_RING, bool I16, bool PACKED, bool HEUR, int NW>
__global__ void __launch_bounds__(64 * NW, LDS_RING ? 5 : 4) wfa_align_kernel(KParams p) {
  using RT = typename std::conditional<I16, int16_t, int32_t>::type;
  constexpr bool CIG = MODE != 0;
  static_assert(NW == 1 || !LDS_RING, "several waves per pair only with rings in HBM");
  extern __shared__ uint64_t smem[];
  const int tid = threadIdx.x;
  uint64_t* sP = smem;
  uint64_t* sT = smem + p.seq_words;
  int4* meta = (int4*)(sT + p.seq_words);
  Rows<RT> R;
  R.meta = meta;
  R.wcap = p.wcap;
  R.mask = p.wcap - 1;
  R.cap = p.ring_rows * p.wcap;
  R.chk = p.counters;
  if constexpr (LDS_RING) R.ring = (RT*)(meta + p.nrows_meta);
  else R.ring = (RT*)(p.gring + (int64_t)blockIdx.x * p.ring_rows * p.wcap);
  // cross-wave reductions (NW > 1): 16 ints per wave after the metadata
  int* red = LDS_RING ? nullptr : (int*)(meta + p.nrows_meta);
  Hist H;
  const int64_t pool = (int64_t)p.pool_chunks * p.chunk;
  H.hdr = HBuf<int4>{CIG ? p.hdr + (int64_t)blockIdx.x * p.hdr_cap : nullptr, p.hdr_cap,
                     p.counters};
  H.cells = HBuf<int32_t>{MODE == 1 ? p.hist : nullptr, pool, p.counters};
  H.bytes = HBuf<uint8_t>{MODE == 2 ? (uint8_t*)p.hist : nullptr, pool, p.counters};
  H.chain = MODE == 2 ? p.chain + (int64_t)blockIdx.x * p.chain_cap : nullptr;
  H.chain_cap = p.chain_cap;
  H.signs = HBuf<uint8_t>{MODE == 2 ? p.zsign : nullptr, pool, p.counters};
  H.hdr_cap = p.hdr_cap;
  H.tab = CIG ? p.chunk_tab + (int64_t)blockIdx.x * p.tab_cap : nullptr;
  H.tab_cap = p.tab_cap;
  H.chunk = p.chunk;
  H.pool_chunks = p.pool_chunks;
  H.next = p.counters + 3;
  int owned = 0;  // pool chunks this slot holds
  for (;;) {
    int item = 0;
    if (NW == 1) {
      if (tid == 0) item = atomicAdd(&p.counters[0], 1);
      item = uniform(__shfl(item, 0));
    } else {
      if (tid == 0) red[15 * NW] = atomicAdd(&p.counters[0], 1);
      __syncthreads();
      item = uniform(red[15 * NW]);
      __syncthreads();
    }
    if (item >= p.num_pairs) break;
    // pair-level values are uniform: keep them in scalar registers
    const int pair = uniform(p.pair_list ? p.pair_list[item] : p.pair_base + item);
    const int plen = uniform(p.plen[pair]), tlen = uniform(p.tlen[pair]);
    const int bits = stage_sequences<NW>(p, pair, plen, tlen, sP, sT, red);
    if (bits == 0) {
      if (tid == 0) {
        const int slot = atomicAdd(&p.counters[1], 1);
        p.retry_list[slot] = pair;
        p.status[pair] = ST_RETRY;
      }
      continue;
    }
    // 2-bit packed input never needs the byte-wise path
    // HBM-resident rings are wide: keep several chunks of loads in flight
    constexpr int U = LDS_RING || NW > 1 ? 1 : MODE == 0 ? 4 : 2;
    constexpr bool BFREE = LDS_RING || MODE == 0;
    if (PACKED || bits == 2)
      align_pair<KIND, MODE, 2, RT, HEUR, U, NW, BFREE>(p, pair, plen, tlen, sP, sT, R, H, red,
                                                        owned);
    else
      align_pair<KIND, MODE, 8, RT, HEUR, U, NW, BFREE>(p, pair, plen, tlen, sP, sT, R, H, red,
                                                        owned);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------
// Instantiation: every storage/format/heuristic variant of one kind
// ---------------------------------------------------------------------
template <int K, int M, int NW>
static KernelFn pick_wide16(bool packed) {
  return packed ? wfa_align_kernel<K, M, false, true, true, false, NW>
                : wfa_align_kernel<K, M, false, true, false, false, NW>;
}

template <int K, int M, bool HEUR>
static KernelFn pick_kmh(bool lds, bool i16, bool packed, int nw) {
  if (!lds) {
    if constexpr (!HEUR) {
      // the launch shapes the autotuner tries for long reads (16-bit rows,
      // score only and medium/low memory): 4, 8 or 16 waves per pair
      if constexpr (M != 1) {
        if (nw == 4 && i16) return pick_wide16<K, M, 4>(packed);
        if (nw == 16 && i16) return pick_wide16<K, M, 16>(packed);
      }
      if (nw > 1 && i16) return pick_wide16<K, M, kWideWaves>(packed);
      if (nw > 1)
        return packed ? wfa_align_kernel<K, M, false, false, true, false, kWideWaves>
                      : wfa_align_kernel<K, M, false, false, false, false, kWideWaves>;
    }
    return packed ? wfa_align_kernel<K, M, false, false, true, HEUR, 1>
                  : wfa_align_kernel<K, M, false, false, false, HEUR, 1>;
  }
  if (i16) return packed ? wfa_align_kernel<K, M, true, true, true, HEUR, 1>
                         : wfa_align_kernel<K, M, true, true, false, HEUR, 1>;
  return packed ? wfa_align_kernel<K, M, true, false, true, HEUR, 1>
                : wfa_align_kernel<K, M, true, false, false, HEUR, 1>;
}

template <int K, int M>
static KernelFn pick_km(bool lds, bool i16, bool packed, bool heur, int nw) {
  return heur ? pick_kmh<K, M, true>(lds, i16, packed, nw)
              : pick_kmh<K, M, false>(lds, i16, packed, nw);
}

template <int K>
static KernelFn pick_kind(int mode, bool lds, bool i16, bool packed, bool heur, int nw) {
  if (mode == 1) return pick_km<K, 1>(lds, i16, packed, heur, nw);
  if (mode == 2) return pick_km<K, 2>(lds, i16, packed, heur, nw);
  return pick_km<K, 0>(lds, i16, packed, heur, nw);
}

}  // namespace wfa
