// wfa_device.h — parameters shared by the host engine and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/wfa_mi355x.h"

namespace wfa {

// Kernel kinds: edit / indel (M from M[s-1]), gap-affine, 2-piece
// gap-affine, gap-linear (M from M[s-x] and M[s-indel], no gap components)
enum { KIND_EDIT = 0, KIND_AFFINE = 1, KIND_AFFINE2P = 2, KIND_LINEAR = 3 };

// wavefront components per score of a kind: M, then I1/D1, then I2/D2
constexpr int kind_components(int kind) {
  return kind == KIND_AFFINE ? 3 : kind == KIND_AFFINE2P ? 5 : 1;
}

// Everything one persistent launch needs, passed by value.
struct KParams {
  // penalties (internal, after the Eizenga transform) and derived sizes
  int metric;
  int x, o1, e1, o2, e2, oe1, oe2;
  int scope;                 // max_score_scope (src/components.rs:435-504)
  int ringM, ringE1, ringE2; // ring lengths of M, I1/D1, I2/D2 rows
  int modular;               // score scope: no history, no slot limit
  int max_steps;
  // span (ends-free with match == 0; extension)
  int endsfree, extension;
  int pb_free, pe_free, tb_free, te_free;
  // optional per-pair span and heuristic (wfa_pair_config_t, three int4
  // per pair) within the envelope of the launch's span and heuristic
  const int4* pair_cfg;
  // user penalties (max-trim of extension alignments and the heuristic
  // rescoring score with them)
  int u_match, u_x, u_o1, u_e1, u_o2, u_e2;
  // heuristic (WFA_HEURISTIC_*) and its three parameters
  int heuristic, hp0, hp1, hp2;
  // pairs
  int num_pairs;             // work items in this launch
  int pair_base;             // first pair when pair_list is null
  const int32_t* pair_list;  // optional indirection (retry passes)
  int fmt;
  const uint8_t* seq;
  const int64_t* poff;
  const int32_t* plen;
  const int64_t* toff;
  const int32_t* tlen;
  int32_t* score;
  int32_t* status;
  uint8_t* cig;
  const int64_t* cig_off;
  int32_t* cig_len;
  int32_t* end_pos;          // optional {score, k, offset} per pair
  int compat;                // WFA_COMPAT_*
  // Compat::Corrected begin-free dominance record: dom_stride ints per
  // workgroup (components x diagonals), or null
  int32_t* dom;
  int dom_stride;
  // scheduling
  int32_t* counters;         // [0] work-queue head, [1] retry count
  int32_t* retry_list;
  // storage
  int wcap;                  // ring width, power of two
  int seq_words;             // 64-bit words per staged sequence (incl. padding)
  int nrows_meta;            // metadata entries (= ring rows)
  int ring_rows;
  int32_t* gring;            // global ring when the ring is not in LDS
  int32_t* hist;             // CIGAR history cells, hist_cap per slot
  int64_t hist_cap;
  int4* hdr;                 // CIGAR history headers, hdr_cap per slot
  int hdr_cap;
  uint8_t* chain;            // medium/low memory op-chain scratch, chain_cap per slot
  int chain_cap;
  // Z-drop ends of medium/low-memory CIGARs in Compat::C: per slot and
  // score, the trimmed ranges {M, I1}, {D1, I2}, {D2, occupancies} (3 x
  // hdr_cap int4), and per cell of the history pool the signs of its five
  // components' offsets (zsign); else null
  int4* ztr;
  uint8_t* zsign;
  int low;                   // low memory (else medium) in CIGAR mode 2
  int sh_gsh;                // short-read kernel: scores in steps of 2^sh_gsh
  // CIGAR history pool (`hist`, pool_chunks chunks of `chunk` elements):
  // every slot takes chunks as its pairs need them and keeps them for its
  // next pairs; chunk_tab holds each slot's chunk ids (tab_cap per slot)
  int32_t* chunk_tab;
  int tab_cap;
  int chunk;
  int pool_chunks;
  // phase-clock builds (WFA_PHASE_CLOCKS) of the short-read kernel: cycles
  // summed over the launch, {pairs, sequence staging, extension, whole
  // pairs, backtrace}; null otherwise
  unsigned long long* phase;
};

// ---------------------------------------------------------------------
// BiWFA (ultralow memory)
// ---------------------------------------------------------------------
// wavefront components (ComponentType, src/bialign.rs:10-17)
enum { C_M = 0, C_I1 = 1, C_D1 = 2, C_I2 = 3, C_D2 = 4 };

// Outcome of a breakpoint search or a base case (BiStatus, src/bialign.rs)
enum { BI_OK = 0, BI_END_REACHED = 1, BI_UNREACHABLE = 2, BI_MAX_STEPS = 3,
       BI_UNATTAINABLE = 4, BI_RETRY = 5 };

// One sub-problem: pattern [pb, pe) x text [tb, te) of pair `pair`,
// starting in component cbeg and ending in cend (0 M, 1 I1, 2 D1, 3 I2, 4 D2)
struct BiJob {
  int pair, pb, pe, tb, te, cbeg, cend, pad;
};

struct BiResult {
  int status;      // BI_*
  int score;       // breakpoint score (BI_OK)
  int score_fwd, score_rev;
  int k_fwd, off_fwd;  // breakpoint cell in forward coordinates
  int comp;        // breakpoint component
  int end_score;   // BI_END_REACHED: the end score
};

struct BiParams {
  KParams k;                 // penalties, heuristic, max steps, sequences
  const BiJob* jobs;
  int njobs;
  int32_t* counters;         // [0] work-queue head
  int seq_words;             // LDS words per staged sequence
  int wcap;                  // ring width, power of two
  int32_t* ring;             // rings: per slot, (2 directions or 1) x NC x scope rows
  // breakpoint search
  BiResult* results;
  // base cases: full storage and fragments
  int4* hdr;                 // per slot, hdr_cap scores x NC components
  int hdr_cap;
  int32_t* hist;             // per slot, hist_cap cells
  int64_t hist_cap;
  uint8_t* frag;             // CIGAR fragment of job i at frag + frag_off[i]
  const int64_t* frag_off;
  int32_t* frag_len;
  int32_t* frag_status;      // BI_OK, BI_UNATTAINABLE or BI_RETRY
  int32_t* frag_score;
};

// A piece of an assembled BiWFA CIGAR: `len` bytes from frag + src, or
// (src < 0) `len` copies of `op`, to cig + dst
struct BiPiece {
  int64_t dst, src;
  int32_t len, op;
};

using KernelFn = void (*)(KParams);
using KernelFn2 = void (*)(BiParams);
// which: 0 breakpoint search, 1 base case
KernelFn2 pick_biwfa_edit(int which, bool heur);
KernelFn2 pick_biwfa_affine(int which, bool heur);
KernelFn2 pick_biwfa_affine2p(int which, bool heur);
KernelFn2 pick_biwfa_linear(int which, bool heur);
hipError_t launch_biwfa(int kind, int which, bool heur, const BiParams& bp, int grid,
                        size_t lds_bytes, hipStream_t stream);
int biwfa_blocks_per_cu(int kind, int which, bool heur, size_t lds_bytes);
hipError_t launch_biwfa_gather(const BiPiece* pieces, int n, const uint8_t* frag, uint8_t* cig,
                               hipStream_t stream);
// BiWFA CIGAR score (cigar_wfa_score, src/aligner.rs:2700-2719) of pairs
// pairs[0..n): into p.score
hipError_t launch_biwfa_rescore(const KParams& p, const int32_t* pairs, int n,
                                hipStream_t stream);
// the alignment kernel for one configuration, from the kind's kernel set
// (mode: 0 score only, 1 high memory CIGAR, 2 medium/low memory CIGAR)
KernelFn pick_edit(int mode, bool lds, bool i16, bool packed, bool heur, int nw);
KernelFn pick_affine(int mode, bool lds, bool i16, bool packed, bool heur, int nw);
KernelFn pick_affine2p(int mode, bool lds, bool i16, bool packed, bool heur, int nw);
KernelFn pick_linear(int mode, bool lds, bool i16, bool packed, bool heur, int nw);
// the short-read kernel (gap-affine, end-to-end, no heuristic; mode 0 or
// 1): its history is p.hdr_cap rows of kShortWindow 16-bit cells and
// p.hdr_cap int32 headers per slot
KernelFn pick_affine_short(int mode, bool packed, int rc);
size_t short_lds_row_bytes_host(int rc);
// the short kernel's ring class for penalties x, o + e, e (the score step
// 2^gsh in *gsh): 2 for the built-in default penalties, else 0 or 1, or -1
// when their reduced lookbacks exceed every class
int short_ring_class(int x, int oe, int e, int* gsh);
// diagonals of the short-read kernel's window: 7 chunks of 64 (reads100_300
// CIGAR launch 12.45 ms at 512, 11.86 ms at 448; at 384 one pair of 1M leaves
// the window and its retry pass takes the step to 12.74 ms)
constexpr int kShortWindow = 448;

hipError_t launch_reset(int32_t* counters, hipStream_t stream);
hipError_t launch_max_length(const int32_t* plen, const int32_t* tlen, int n, int32_t* out,
                             hipStream_t stream);
// mode: 0 score only, 1 high memory CIGAR, 2 medium/low memory CIGAR
// i16: ring offsets held in 16 bits (pairs of at most 8000 bases)
// Waves per workgroup (one pair each) of the wide kernels, whose rings live
// in HBM, without heuristics
constexpr int kWideWaves = 8;

// The engine's counters: [0] work queue head, [1] retries, [2] scratch,
// [3] next history pool chunk (all reset before each launch), and
// [kCheckCounter] the out-of-range indexes a WFA_DEVICE_CHECKS build found
// during one call (reset at the call's start)
constexpr int kCheckCounter = 4;

// Compat::Corrected ends-free alignment with a match reward: candidate
// ends, band and dominance pruning (src/aligner.rs:1104-1360)
__host__ __device__ inline bool corrected_endsfree(const KParams& p) {
  return p.compat == WFA_COMPAT_CORRECTED && p.u_match < 0 && p.endsfree;
}
// Launches that need the single-wave kernels with the per-score
// cut-off logic: heuristics, and ends-free alignment with a match reward
// (its begin-free cells and, in Compat::Corrected, its candidate ends).
// The other kernels carry no per-score code for them: they are bound by
// scalar instructions per score.
__host__ __device__ inline bool heur_kernel(const KParams& p) {
  return p.heuristic != WFA_HEURISTIC_NONE || (p.endsfree && p.u_match < 0);
}

// Waves per workgroup of a BiWFA launch (`which` 0 = breakpoint search,
// 1 = base cases): breakpoint searches without a heuristic take kWideWaves
// (heuristic cut-offs are single-wave); base cases one.
inline int biwfa_waves(int which, bool heur) { return which == 0 && !heur ? kWideWaves : 1; }

// Status WFA_STATUS_OOM (score 0, no CIGAR) for the `n` pairs of `list`
hipError_t launch_mark_oom(const int32_t* list, int n, int32_t* score, int32_t* status,
                           int32_t* cig_len, int32_t* end_pos, hipStream_t stream);

// Longest-first order of pairs first..first+n-1 into `order` (a device
// radix sort; keys/keys_sorted/vals hold n entries, temp lpt_temp_bytes(n))
size_t lpt_temp_bytes(int n);
hipError_t launch_lpt_order(const int32_t* plen, const int32_t* tlen, int first, int n,
                            uint32_t* keys, uint32_t* keys_sorted, int32_t* vals,
                            int32_t* order, void* temp, size_t temp_bytes, hipStream_t stream);

// `nw` waves per pair (1, or kWideWaves for HBM rings without heuristics)
hipError_t launch_align(const KParams& p, int kind, int mode, bool lds_ring, bool i16, int nw,
                        int grid, size_t lds_bytes, hipStream_t stream);
int max_blocks_per_cu(int kind, int mode, bool lds_ring, bool i16, bool packed, bool heur, int nw,
                      size_t lds_bytes);
hipError_t launch_short(const KParams& p, int mode, int rc, int grid, size_t lds_bytes,
                        hipStream_t stream);
int short_blocks_per_cu(int mode, bool packed, int rc, size_t lds_bytes);

}  // namespace wfa
