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
// * Scores in steps of g, the largest power of two that divides the
//   mismatch, gap-open + gap-extend and gap-extend penalties: every
//   wavefront of a score that g does not divide is absent (each one derives
//   from M0 by those penalties), so the reference's steps there are null
//   steps. The loop runs over reduced scores t (score t * g) with the
//   penalties divided by g, and accounts for the real scores it passes over
//   exactly: each counts towards the run of null steps (src/aligner.rs:
//   2031-2043) and meets the per-score limits in order.
// * Each pair gets a fixed window of 448 diagonals [base, base + 448),
//   centred on the diagonals its alignment must cross. Every ring row is an
//   LDS row over the whole window plus pad cells on each side, all NULL when
//   the pair starts. A step writes its three output rows over the
//   64-diagonal chunks it computes, OFFSET_NULL outside their trimmed
//   ranges, so a row's cells outside its range read as NULL with no range
//   test: the reference NULL-fills exactly those cells before it reads them
//   (src/wavefront.rs:422-457). An absent wavefront is a row of NULL cells
//   as well: a null step writes its rows NULL, and the rows of scores below
//   zero are the initial fill (a ring holds as many rows as the longest
//   lookback: a step reads all its inputs before it writes an output, so
//   its output may take the slot of its oldest input), so a step reads its
//   inputs from their ring slots with no presence test.
// * The chunks a step computes cover [lo - 1, hi + 1] of the union of every
//   row computed so far. That holds the reference's computed range (the
//   union of the inputs' ranges, widened by one, src/aligner.rs:2045-2062);
//   a cell outside the latter has only NULL inputs, so it comes out below
//   zero, outside the trimmed ranges, and is stored as NULL — as it would
//   be read in the reference. The union only grows, so no cell outside it
//   has been written since the pair started, and every write of a slot
//   covers what the slot's previous score wrote.
// * Which ring rows hold a wavefront (not absent, not trimmed empty) is kept
//   by age in two scalar masks, only to detect null steps.
// * In CIGAR scope the extended M row of every reduced score goes to the
//   slot's history in HBM, its computed chunks packed one after another
//   (16-bit cells, whole 128-byte lines), with one int32 header per reduced
//   score: the row's first diagonal, its chunk count and its first chunk.
//   The Singletrack backtrace reads it through WinView, which finds no
//   wavefront at the scores in between.
//
// Offsets are held in 16 bits as in the general kernel's short-read rings
// (sequences of at most 8000 bases, scores up to 16000; NULL-derived values
// saturate at -16384 and only their sign matters). A pair whose union of
// diagonals would leave its window, whose score passes the history's rows,
// or whose sequences hold other bytes than A/C/G/T is handed back through
// the retry list to the general kernel, which aligns it from the start.
#pragma once

#include <type_traits>

#include "wfa_align.h"

namespace wfa {

constexpr int SH_WIN = kShortWindow;     // diagonals of a pair's window
constexpr int SH_CHUNKS = SH_WIN / 64;   // most 64-diagonal chunks a step computes
constexpr int SH_ROW = SH_WIN + 8;       // LDS cells per row: pad, window, pad (16-byte rows)
constexpr int SH_NULL = -16384;          // OFFSET_NULL as held in 16 bits
constexpr int SH_GRAB = 4;               // score-only work items a wave takes at once (8: 1% slower)
constexpr int SH_GRAB_TAIL = 2;          // ... until this many runs per workgroup are left
constexpr int SH_HDR_EMPTY = 0;          // header of an absent M row: no chunks

// Ring classes: M rows and I1/D1 rows per ring (powers of two, so a slot is
// a mask of the reduced score), for reduced lookbacks x, o + e < M and
// e < E; class 2 is sized for the built-in penalties alone (x 2, o + e 4,
// e 1 reduced: lookbacks of exactly M and E rows).
template <int RC> struct ShRing;
template <> struct ShRing<0> { static constexpr int M = 8, E = 2; };
template <> struct ShRing<1> { static constexpr int M = 16, E = 8; };
template <> struct ShRing<2> { static constexpr int M = 4, E = 1; };

static __device__ __forceinline__ int16_t sh16(int v) { return (int16_t)imax(v, SH_NULL); }

// A history row's header: its first diagonal as an offset into the window
// (9 bits), its chunk count (4 bits; none for a score without an M row) and
// its first chunk in the slot's history (19 bits).
static __device__ __forceinline__ int32_t sh_hdr(int a_rel, int nch, int chunk) {
  return (int32_t)((uint32_t)a_rel | ((uint32_t)nch << 9) | ((uint32_t)chunk << 13));
}

// The history cell of diagonal k in the row of header h, or -1 outside the
// row's chunks.
static __device__ __forceinline__ int sh_cell(int32_t h, int base, int k) {
  const uint32_t u = (uint32_t)h;
  const int w = k - base - (int)(u & 511u);
  return (uint32_t)w < 64u * ((u >> 9) & 15u) ? (int)(u >> 13) * 64 + w : -1;
}

// One slot's high-memory history as the Singletrack backtrace reads it
// (st_m_offset, src/backtrace.rs:645-663): M[s] at diagonal k, if score s
// has a row covering k and the cell is not NULL. Rows exist for multiples
// of 2^gsh only.
struct WinView {
  const int16_t* cells;  // the slot's history: rows of packed 64-cell chunks
  const int32_t* hdr;    // per reduced score t (score t << gsh): the row's header (sh_hdr)
  int base, nscores, nwf, gsh;
  int64_t cap;           // cells in the slot's history (tested in WFA_DEVICE_CHECKS builds)
  int32_t* chk;          // the launch's counters
  __device__ __forceinline__ int cell(int i) const {
    return cells[checked_index<int16_t>(i, cap, chk)];
  }
  __device__ __forceinline__ bool m(int s, int k, int* out) const {
    if (s < 0 || s >= nscores || s >= nwf || (s & ((1 << gsh) - 1))) return false;
    const int i = sh_cell(hdr[s >> gsh], base, k);
    if (i < 0) return false;
    const int v = cell(i);
    if (v < 0) return false;
    *out = v;
    return true;
  }
};

// M[s] at diagonals k1 and k2 of one row (WinView::m twice), or -1 where
// absent, with the row headers read from `lhdr` (a copy in LDS).
static __device__ __forceinline__ void win_cells(const WinView& hv, const int32_t* lhdr, int s,
                                                 int k1, int k2, int& v1, int& v2) {
  v1 = v2 = -1;
  if (s < 0 || s >= hv.nscores || s >= hv.nwf || (s & ((1 << hv.gsh) - 1))) return;
  const int32_t h = lhdr[s >> hv.gsh];
  const int i1 = sh_cell(h, hv.base, k1), i2 = sh_cell(h, hv.base, k2);
  if (i1 >= 0) v1 = hv.cell(i1);
  if (i2 >= 0) v2 = hv.cell(i2);
}

// ops[pos], ops[pos - 1], ...: n1 copies of op1, then n2 of op2 (counts
// below zero write none), in one masked store when they fit the wave;
// pos moves past them
static __device__ __forceinline__ void put_ops2(uint8_t* out, int& pos, uint8_t op1, int n1,
                                                uint8_t op2, int n2) {
  const int lane = threadIdx.x;
  n1 = imax(n1, 0);
  n2 = imax(n2, 0);
  const int n = n1 + n2;
  if (n <= 64) {
    if (lane < n && pos - lane >= 0) out[pos - lane] = lane < n1 ? op1 : op2;
    pos -= n;
  } else {
    write_ops(out, pos, op1, n1);
    write_ops(out, pos, op2, n2);
  }
}

// Singletrack backtrace (src/backtrace.rs:751-858) of the short kernel's
// end-to-end gap-affine alignments: the steps of `singletrack` for one gap
// piece, arranged for latency. At each M cell the wave loads the mismatch
// source and the sources of gap lengths 1..64 (one lane per length,
// insertion and deletion) at once, and counts the matches before the cell
// in the staged sequences while those loads are in flight; a longer gap
// loads 64 more lengths at a time. The row headers come from `lhdr` (LDS).
// The ops are written backwards into `ops` (LDS), ending at ops[cap - 1];
// returns their number.
static __device__ __forceinline__ int short_backtrace(const KParams& p, const WinView& hv,
                                                      const int32_t* lhdr,
                                                      const uint64_t* sP, const uint64_t* sT,
                                                      int plen, int tlen, int score, int k,
                                                      int offset, uint8_t* ops, int cap) {
  const int lane = threadIdx.x;
  const int x = p.x, o = p.o1, e = p.e1;
  int pos = cap - 1;
  put_ops2(ops, pos, 'D', plen - (offset - k), 'I', tlen - offset);
  while (score != 0) {
    score = uniform(score);
    k = uniform(k);
    offset = uniform(offset);
    int mo, unused;
    win_cells(hv, lhdr, score - x, k, k, mo, unused);
    int L = 1 + lane;
    int sg = score - o - L * e;
    int mi, md;
    win_cells(hv, lhdr, sg, k - L, k + L, mi, md);
    const int nm = backward_matches<2>(sP, sT, offset - k, offset);
    const int orig = offset - nm;
    if (mo >= 0 && mo == orig - 1) {
      put_ops2(ops, pos, 'M', nm, 'X', 1);
      offset = orig - 1;
      score -= x;
      continue;
    }
    // the lowest gap length with a source (or past the score) decides,
    // insertion before deletion
    bool found = false;
    for (int l = 0;; l += 64) {
      if (l > 0) {
        L = l + 1 + lane;
        sg = score - o - L * e;
        win_cells(hv, lhdr, sg, k - L, k + L, mi, md);
      }
      const bool ri = mi >= 0 && mi + L >= orig && mi + L <= offset;
      const bool rd = md >= 0 && md >= orig && md <= offset;
      const uint64_t mask = ballot(ri || rd || sg < 0);
      if (!mask) continue;
      const int j = __builtin_ctzll(mask);
      const int Lj = l + 1 + j;
      const int fj = uniform(__shfl((ri ? 1 : 0) | (rd ? 2 : 0), j));
      const int m = uniform(__shfl(ri ? mi : md, j));
      if (fj & 1) {
        put_ops2(ops, pos, 'M', offset - (m + Lj), 'I', Lj);
        k -= Lj;
      } else if (fj & 2) {
        put_ops2(ops, pos, 'M', offset - m, 'D', Lj);
        k += Lj;
      } else {
        break;  // no source
      }
      offset = m;
      score -= o + Lj * e;
      found = true;
      break;
    }
    if (!found) break;
  }
  int v = offset - k, h = offset;
  const int nm = imin(v, h);
  v -= nm;
  h -= nm;
  put_ops2(ops, pos, 'M', nm, 'D', v);
  write_ops(ops, pos, 'I', h);
  return cap - 1 - pos;
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

// Bases [pos, pos + 32) of a 2-bit sequence staged as 32-bit words: two
// funnel shifts over three consecutive words (the staged sequences carry
// two zero words of padding past their end, so the third is in bounds).
static __device__ __forceinline__ uint64_t sh_w32(const uint32_t* s, uint32_t pos) {
  const uint32_t q = pos >> 4, sh = (pos & 15u) * 2u;
  const uint32_t w0 = s[q], w1 = s[q + 1], w2 = s[q + 2];
  return (uint64_t)__builtin_amdgcn_alignbit(w1, w0, sh) |
         ((uint64_t)__builtin_amdgcn_alignbit(w2, w1, sh) << 32);
}

// Matching bases before the first difference of two 16-base words (16 when
// they are equal).
static __device__ __forceinline__ int sh_match16(uint32_t x) {
  return x ? (int)(__builtin_ctz(x) >> 1) : 16;
}

// Matching bases before the first difference of two 32-base words (32 when
// they are equal).
static __device__ __forceinline__ int sh_match32(uint64_t x) {
  return x ? (int)(__builtin_ctzll(x) >> 1) : 32;
}

// Extend every lane's M cell `off` along its diagonal k (lanes with a NULL
// cell keep it): the matching prefix of pattern[off - k..] and text[off..]
// bounded by the sequence ends (extend_matches_packed_end2end's per-diagonal
// LCP, src/extend.rs:15-38). The first 16 bases are compared for every lane
// at once; only lanes still matching after them loop, LOOP (16 or 32) bases
// at a time. The score-only kernel loops 32 at a time (2% faster on 1M
// 100-300 bp pairs); the CIGAR kernel, held to 80 VGPRs, was 3% slower so.
template <int LOOP>
static __device__ __forceinline__ int sh_extend(const uint32_t* P, const uint32_t* T, int off,
                                                int k, int plen, int tlen) {
  const bool act = off >= 0;
  const uint32_t h = act ? (uint32_t)off : 0u, v = act ? (uint32_t)(off - k) : 0u;
  const int rem = act ? imin(plen - (int)v, tlen - (int)h) : 0;
  const int c = sh_match16(sh_w16(P, v) ^ sh_w16(T, h));
  int n = imin(c, rem);
  bool more = c == 16 && rem > 16;
  if (ballot(more)) {
    while (more) {
      const int left = rem - n;
      const int d = LOOP == 32 ? sh_match32(sh_w32(P, v + n) ^ sh_w32(T, h + n))
                               : sh_match16(sh_w16(P, v + n) ^ sh_w16(T, h + n));
      n += imin(d, left);
      more = d == LOOP && left > LOOP;
    }
  }
  return off + n;
}

// Penalty sets built in at compile time (lookbacks and masks become
// immediates): 0 none (the kernel reads them from KParams), 1 the default
// gap-affine penalties (mismatch 4, gap open 6, gap extend 2, the driver's
// default, src/bin/align_benchmark.rs:177-183; ring class 0, score step 2).
template <int PEN> struct ShPen;
template <> struct ShPen<1> { static constexpr int gsh = 1, x = 2, oe = 4, e = 1; };

template <int MODE, int RC, int PEN>
static __device__ void short_pair(const KParams& p, int pair, int plen, int tlen,
                                  const uint64_t* sP, const uint64_t* sT, int16_t* rows,
                                  int16_t* hcells, int32_t* hhdr, uint64_t& ext_cycles,
                                  uint64_t& bt_cycles) {
  constexpr bool CIG = MODE == 1;
  constexpr int RM = ShRing<RC>::M, RE = ShRing<RC>::E;
  // ring rows as LDS byte offsets from `rows`: reduced score t's M row at
  // (t mod RM) * RB, its I1 / D1 rows at IB0 / DB0 + (t mod RE) * RB
  constexpr int RB = SH_ROW * 2;
  constexpr int IB0 = RM * RB, DB0 = (RM + RE) * RB;
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
  // reduced penalties (score = t << gsh)
  int gsh, xr, oer, er;
  if constexpr (PEN != 0) {
    gsh = ShPen<PEN>::gsh;
    xr = ShPen<PEN>::x;
    oer = ShPen<PEN>::oe;
    er = ShPen<PEN>::e;
  } else {
    gsh = p.sh_gsh;
    xr = p.x >> gsh;
    oer = p.oe1 >> gsh;
    er = p.e1 >> gsh;
  }
  // every row NULL over the window and its pads
  {
    constexpr int n16 = (RM + 2 * RE) * SH_ROW / 8;
    const int nn = (int)(((uint32_t)(uint16_t)SH_NULL << 16) | (uint16_t)SH_NULL);
    int4* r4 = reinterpret_cast<int4*>(rows);
    for (int i = lane; i < n16; i += 64) r4[i] = make_int4(nn, nn, nn, nn);
  }
  const uint32_t* P32 = reinterpret_cast<const uint32_t*>(sP);
  const uint32_t* T32 = reinterpret_cast<const uint32_t*>(sT);
  auto cell = [&](int byte_off) {
    return reinterpret_cast<int16_t*>(reinterpret_cast<char*>(rows) + byte_off) + 1 - base;
  };
  const int nwf = uniform(high_mem_nwf(p, plen, tlen));
  // cells in the slot's history (index checks of WFA_DEVICE_CHECKS builds)
  const int64_t hist_cells = CIG ? (int64_t)p.hdr_cap * SH_WIN : 0;
  int status = WFA_STATUS_OK, result = 0, t = 0;
  bool overflow = false;
  __syncthreads();

  // score 0: M0 = {k 0: offset 0} (init_alignment, src/aligner.rs:846-1000),
  // extended
  {
    // (the same in every lane: keep it scalar, so the loop state stays
    // scalar)
    const int off = uniform(lcp<2>(sP, sT, 0, 0, plen, tlen));
    if (lane == 0) cell(0)[0] = sh16(off);
    // the history's first chunk: M0, then NULL cells
    if (CIG) hcells[lane] = lane == 0 ? sh16(off) : (int16_t)SH_NULL;
    if (k_end == 0 && off >= tlen) status = WFA_STATUS_END_REACHED;
  }
  __syncthreads();

  // Which earlier rows hold a wavefront, by age: bit d of agM is M[t - 1 - d]
  // at the top of step t, bit d of agE is I1 or D1 of t - 1 - d.
  uint32_t agM = 1;  // M0
  uint32_t agE = 0;
  const uint32_t needM = (1u << (xr - 1)) | (1u << (oer - 1));
  const uint32_t needE = 1u << (er - 1);
  int ulo = 0, uhi = 0;  // union of the diagonals of every row so far
  // CIGAR headers of 64 consecutive reduced scores, lane i holding
  // (t & ~63) + i, stored once the block is full (and at the end)
  int hbuf = lane == 0 ? sh_hdr(-base, 1, 0) : 0;
  int hnext = 1;  // the history's next free chunk
  // the first score at which a per-score limit applies
  int limit = imin(16001, p.max_steps);
  if (!p.modular) limit = imin(limit, nwf);
  if (CIG) limit = imin(limit, p.hdr_cap << gsh);
  limit = uniform(limit);
  // last non-null score, and the first score at which a check can fire: a
  // limit, or a run of null steps longer than the scope
  int lastnn = 0;
  int ulim = imin(limit, p.scope + 1);
  // Every exit of the loop is a uniform break that sets the outcome; the
  // loop-carried state is read back as scalar at the top of each step (the
  // same in every lane; the compiler cannot tell across the per-lane
  // extension, and would otherwise run the whole loop under exec masks).
  for (int step = status == WFA_STATUS_OK ? 1 : 0; step;) {
    t = uniform(t) + 1;
    ulo = uniform(ulo);
    uhi = uniform(uhi);
    agM = (uint32_t)uniform((int)agM);
    agE = (uint32_t)uniform((int)agE);
    lastnn = uniform(lastnn);
    ulim = uniform(ulim);
    if (CIG) hnext = uniform(hnext);
    const int r = t << gsh;
    // inputs M[t-x], M[t-o-e], I1/D1[t-e]; all absent: a null step
    const bool nul = !(agM & needM) && !(agE & needE);
    // the chunks: [ulo - 1, uhi + 1], inside the window
    const int clo = ulo - 1, chi = uhi + 1;
    if (clo < base || chi > base + SH_WIN - 1) { overflow = true; break; }
    const int nch = ((chi - clo) >> 6) + 1;
    const int a = imin(clo, base + SH_WIN - 64 * nch);
    const int tm = (t & (RM - 1)) * RB, te = (t & (RE - 1)) * RB;
    int16_t* Wm = cell(tm);
    int16_t* Wi = cell(IB0 + te);
    int16_t* Wd = cell(DB0 + te);
    bool last = false;  // max steps: this step is computed, then the loop ends
    if (r >= ulim || nul) {
      // The real scores of this step one at a time, as the reference meets
      // them: those since the previous step's are null steps, and so is r
      // when its inputs are all absent.
      bool stop = false;
      for (int q = r - (1 << gsh) + 1; q <= r; ++q) {
        if (q >= limit) {
          if (!p.modular && q >= nwf) { status = WFA_STATUS_MAX_STEPS_REACHED; result = q; stop = true; break; }
          if (CIG && q >= (p.hdr_cap << gsh)) { overflow = true; stop = true; break; }
          if (q > 16000) { overflow = true; stop = true; break; }
          last = q >= p.max_steps;
        }
        if (q < r || nul) {
          if (last) { status = WFA_STATUS_MAX_STEPS_REACHED; result = q; stop = true; break; }
          if (q - lastnn > p.scope) { status = WFA_STATUS_END_UNREACHABLE; result = q; stop = true; break; }
        }
      }
      if (stop) break;
      if (nul) {
        // null step (src/aligner.rs:2031-2043): the outputs are absent
#pragma unroll
        for (int j = 0; j < SH_CHUNKS; ++j) {
          if (j < nch) {
            const int k = a + 64 * j + lane;
            Wm[k] = (int16_t)SH_NULL;
            Wi[k] = (int16_t)SH_NULL;
            Wd[k] = (int16_t)SH_NULL;
          }
        }
        agM <<= 1;
        agE <<= 1;
        if (CIG) {
          hbuf = lane == (t & 63) ? SH_HDR_EMPTY : hbuf;
          if ((t & 63) == 63) hhdr[(t & ~63) + lane] = hbuf;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        continue;
      }
    }
    const int16_t* Px = cell(((t - xr) & (RM - 1)) * RB);
    const int16_t* Po = cell(((t - oer) & (RM - 1)) * RB);
    const int eb = ((t - er) & (RE - 1)) * RB;
    const int16_t* Pi = cell(IB0 + eb);
    const int16_t* Pd = cell(DB0 + eb);

    // The step over NCH chunks (one instance per chunk count, so the chunk
    // loops carry no guards); true when the alignment ends here.
    auto compute = [&](auto nch_c) -> bool {
      constexpr int NCH = decltype(nch_c)::value;
      // ---- compute_affine_idm (src/compute/affine.rs:318-351) over the
      // chunks, trims by ballot (trim_ends, src/compute/mod.rs:14-49) ----
      // M's trimmed range needs no tracking: a cell of M in bounds takes its
      // offset from an in-bounds I1 or D1 cell of this score or from a cell
      // of M[t - x], so it lies in their ranges or the union already; only
      // whether M is empty matters (null steps)
      int mv[NCH], iv[NCH], dv[NCH];
      uint64_t m_any = 0;
      int i_lo = INT32_MAX, i_hi = INT32_MIN, d_lo = INT32_MAX, d_hi = INT32_MIN;
#pragma unroll
      for (int j = 0; j < NCH; ++j) {
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
        m_any |= ballot(vm);
        sh_track(i_lo, i_hi, ballot(vcol(ins) <= (uint32_t)plen), a + 64 * j);
        sh_track(d_lo, d_hi, ballot(vcol(del) <= (uint32_t)plen), a + 64 * j);
      }
      const bool mne = m_any != 0, ine = i_lo != INT32_MAX, dne = d_lo != INT32_MAX;
      // I1 and D1 keep their cells inside their trimmed ranges (out-of-bounds
      // ones included), NULL elsewhere; M is NULL wherever it is out of bounds
      const Guard gi = ine ? Guard{i_lo, i_hi - i_lo} : Guard{INT32_MIN / 2, 0};
      const Guard gd = dne ? Guard{d_lo, d_hi - d_lo} : Guard{INT32_MIN / 2, 0};
#pragma unroll
      for (int j = 0; j < NCH; ++j) {
        const int k = a + 64 * j + lane;
        Wi[k] = (uint32_t)(k - gi.first) <= (uint32_t)gi.span ? sh16(iv[j]) : (int16_t)SH_NULL;
        Wd[k] = (uint32_t)(k - gd.first) <= (uint32_t)gd.span ? sh16(dv[j]) : (int16_t)SH_NULL;
      }
      agM = (agM << 1) | (uint32_t)mne;
      agE = (agE << 1) | (uint32_t)(ine | dne);
      // (an empty range is [INT32_MAX, INT32_MIN]: no effect)
      ulo = imin(ulo, imin(i_lo, d_lo));
      uhi = imax(uhi, imax(i_hi, d_hi));
      lastnn = r;
      ulim = imin(limit, r + p.scope + 1);
      if (last) {
        // the row is computed but not extended (the loop stops first)
        status = WFA_STATUS_MAX_STEPS_REACHED;
        result = r;
        return true;
      }
      // ---- extend M[t] (extend_matches_packed_end2end, src/extend.rs:47-69;
      // the reference extends at the top of the next iteration and nothing
      // reads M[t] in between), keep it, test the end cell
      // (termination_end2end, src/termination.rs:14-23) ----
      bool reached = false;
      const int64_t hrow = (int64_t)hnext * 64;
#ifdef WFA_PHASE_CLOCKS
      const uint64_t c0 = __builtin_readcyclecounter();
#endif
#pragma unroll
      for (int j = 0; j < NCH; ++j) {
        const int k = a + 64 * j + lane;
        const int off = sh_extend<CIG ? 16 : 32>(P32, T32, mv[j], k, plen, tlen);
        const int16_t o16 = sh16(off);
        Wm[k] = o16;
        if (CIG) hcells[checked_index<int16_t>(hrow + 64 * j + lane, hist_cells, p.counters)] = o16;
        reached |= k == k_end && off >= tlen;
      }
#ifdef WFA_PHASE_CLOCKS
      ext_cycles += __builtin_readcyclecounter() - c0;
#endif
      if (CIG) {
        hbuf = lane == (t & 63) ? sh_hdr(a - base, NCH, hnext) : hbuf;
        if ((t & 63) == 63) hhdr[(t & ~63) + lane] = hbuf;
        hnext += NCH;
      }
      // the step's rows are written before the next step reads them: one
      // wave's LDS accesses complete in order, so only the compiler must
      // keep them in program order
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      if (ballot(reached)) {
        status = WFA_STATUS_END_REACHED;
        result = r;
        return true;
      }
      return false;
    };
    bool done;
    switch (nch) {
      case 1: done = compute(std::integral_constant<int, 1>{}); break;
      case 2: done = compute(std::integral_constant<int, 2>{}); break;
      case 3: done = compute(std::integral_constant<int, 3>{}); break;
      case 4: done = compute(std::integral_constant<int, 4>{}); break;
      case 5: done = compute(std::integral_constant<int, 5>{}); break;
      case 6: done = compute(std::integral_constant<int, 6>{}); break;
      case 7: done = compute(std::integral_constant<int, 7>{}); break;
      default: done = compute(std::integral_constant<int, 8>{}); break;
    }
    if (done) break;
  }

  if (overflow) {
    retry();
    return;
  }
  int out_status = status;
  int ncig = 0;
  if (status == WFA_STATUS_END_REACHED) {
    out_status = WFA_STATUS_COMPLETED;
    if (CIG) {
#ifdef WFA_PHASE_CLOCKS
      const uint64_t b0 = __builtin_readcyclecounter();
#endif
      // (this wave's history writes land before it reads them back)
      __threadfence_block();
      uint8_t* out = p.cig + p.cig_off[pair];
      const int cap = plen + tlen;
      const WinView hv{hcells, hhdr, base, result + 1, nwf, gsh, hist_cells, p.counters};
      // The ring rows are done with: the row headers (rows 0..t) and the ops
      // go there when they fit, and the ops from there to the pair's region
      // in order.
      constexpr int LDS_ROWS = (RM + 2 * RE) * SH_ROW * 2;
      const int hbytes = (t + 1) * 4;
      const bool in_lds = hbytes + cap <= LDS_ROWS;
      if (in_lds) {
        int32_t* lhdr = reinterpret_cast<int32_t*>(rows);
        for (int b = 0; b < (t & ~63); b += 64) lhdr[b + lane] = hhdr[b + lane];
        if (lane <= (t & 63)) lhdr[(t & ~63) + lane] = hbuf;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        uint8_t* ops = reinterpret_cast<uint8_t*>(rows) + hbytes;
        ncig = short_backtrace(p, hv, lhdr, sP, sT, plen, tlen, result, k_end, tlen, ops, cap);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const uint8_t* src = ops + (cap - ncig);
        for (int i = lane; i < ncig; i += 64) out[i] = src[i];
      } else {
        // the headers of the last, partial block
        if (lane <= (t & 63)) hhdr[(t & ~63) + lane] = hbuf;
        __threadfence_block();
        ncig = singletrack<2>(p, hv, sP, sT, plen, tlen, result, k_end, tlen, out, cap);
      }
      // move the ops from the tail of the region to its head
      const int gap = cap - ncig;
      if (!in_lds && gap > 0) {
        for (int c = 0; c < ncig; c += 64) {
          const int i = c + lane;
          uint8_t v = 0;
          if (i < ncig) v = out[gap + i];
          __builtin_amdgcn_wave_barrier();
          if (i < ncig) out[i] = v;
          __builtin_amdgcn_wave_barrier();
        }
      }
#ifdef WFA_PHASE_CLOCKS
      bt_cycles += __builtin_readcyclecounter() - b0;
#endif
    }
  }
  if (lane == 0) {
    p.score[pair] = result;
    p.status[pair] = out_status;
    if (CIG) p.cig_len[pair] = ncig;
    if (p.end_pos) {
      // the end cell, or unset (src/aligner.rs:903-909)
      const bool end = status == WFA_STATUS_END_REACHED;
      p.end_pos[3 * pair] = end ? result : -1;
      p.end_pos[3 * pair + 1] = end ? k_end : INT32_MAX;
      p.end_pos[3 * pair + 2] = end ? tlen : WFA_NULL;
    }
  }
}

// One wave per pair, persistent over the launch's work queue. The slot's
// history: room for p.hdr_cap rows of SH_WIN 16-bit cells in p.hist (the
// rows' chunks packed from its start) and p.hdr_cap row headers in p.hdr,
// both per workgroup.
// The CIGAR kernel is held to 80 VGPRs (6 waves per SIMD: 24 per CU, about
// what its LDS allows); unbounded it takes 89 and runs 20 waves per CU.
template <int MODE, bool PACKED, int RC, int PEN>
__global__ void __launch_bounds__(64, MODE == 1 ? 6 : 1) wfa_short_kernel(KParams p) {
  extern __shared__ uint64_t smem[];
  uint64_t* sP = smem;
  uint64_t* sT = smem + p.seq_words;
  int16_t* rows = reinterpret_cast<int16_t*>(sT + p.seq_words);
  int16_t* hcells =
      MODE == 1 ? reinterpret_cast<int16_t*>(p.hist) + (int64_t)blockIdx.x * p.hdr_cap * SH_WIN
                : nullptr;
  int32_t* hhdr =
      MODE == 1 ? reinterpret_cast<int32_t*>(p.hdr) + (int64_t)blockIdx.x * p.hdr_cap : nullptr;
  // phase clocks (WFA_PHASE_CLOCKS builds): pairs, staging, extension and
  // whole-pair cycles of this wave
  uint64_t npairs = 0, stage_cycles = 0, ext_cycles = 0, pair_cycles = 0, bt_cycles = 0;
  // Score-only pairs take a few microseconds each, and one atomic per pair
  // on the launch's single queue counter limits them: they are taken
  // SH_GRAB at a time, one at a time again once the queue is nearly empty
  // (the last SH_GRAB_TAIL runs per workgroup), so the waves still finish
  // together. CIGAR pairs take one at a time.
  const int tail = p.num_pairs - SH_GRAB_TAIL * SH_GRAB * (int)gridDim.x;
  int next = 0, end = 0;
  for (;;) {
    int item = 0;
    if constexpr (MODE == 0) {
      if (uniform(next) == uniform(end)) {
        const int g = uniform(end) < tail ? SH_GRAB : 1;
        int got = 0;
        if (threadIdx.x == 0) got = atomicAdd(&p.counters[0], g);
        next = uniform(__shfl(got, 0));
        end = next + g;
      }
      item = uniform(next);
      next = item + 1;
    } else {
      if (threadIdx.x == 0) item = atomicAdd(&p.counters[0], 1);
      item = uniform(__shfl(item, 0));
    }
    if (item >= p.num_pairs) break;
#ifdef WFA_PHASE_CLOCKS
    const uint64_t c0 = __builtin_readcyclecounter();
#endif
    // pair-level values are uniform: keep them in scalar registers
    const int pair = uniform(p.pair_list ? p.pair_list[item] : p.pair_base + item);
    const int plen = uniform(p.plen[pair]), tlen = uniform(p.tlen[pair]);
    const int bits = stage_sequences<1>(p, pair, plen, tlen, sP, sT, nullptr);
#ifdef WFA_PHASE_CLOCKS
    const uint64_t c1 = __builtin_readcyclecounter();
    stage_cycles += c1 - c0;
#endif
    if (bits != 2) {
      // does not fit, or not A/C/G/T: the general kernel takes it
      if (threadIdx.x == 0) {
        const int slot = atomicAdd(&p.counters[1], 1);
        p.retry_list[slot] = pair;
        p.status[pair] = ST_RETRY;
      }
      continue;
    }
    short_pair<MODE, RC, PEN>(p, pair, plen, tlen, sP, sT, rows, hcells, hhdr, ext_cycles,
                              bt_cycles);
    __syncthreads();
#ifdef WFA_PHASE_CLOCKS
    pair_cycles += __builtin_readcyclecounter() - c0;
    npairs += 1;
#endif
  }
#ifdef WFA_PHASE_CLOCKS
  if (p.phase && threadIdx.x == 0) {
    atomicAdd(&p.phase[0], (unsigned long long)npairs);
    atomicAdd(&p.phase[1], (unsigned long long)stage_cycles);
    atomicAdd(&p.phase[2], (unsigned long long)ext_cycles);
    atomicAdd(&p.phase[3], (unsigned long long)pair_cycles);
    atomicAdd(&p.phase[4], (unsigned long long)bt_cycles);
  }
#else
  (void)bt_cycles;
  (void)npairs;
  (void)stage_cycles;
  (void)ext_cycles;
  (void)pair_cycles;
#endif
}

// LDS bytes of the short kernel's ring rows for ring class `rc`
static inline size_t short_lds_row_bytes(int rc) {
  const int rows = rc == 0   ? ShRing<0>::M + 2 * ShRing<0>::E
                   : rc == 1 ? ShRing<1>::M + 2 * ShRing<1>::E
                             : ShRing<2>::M + 2 * ShRing<2>::E;
  return (size_t)rows * SH_ROW * 2;
}

}  // namespace wfa
