// wfa_launch.hip — small utility kernels and the launch dispatch over the
// per-kind kernel sets.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <hipcub/hipcub.hpp>

#include "wfa_align.h"
#include "wfa_device.h"

namespace wfa {

__global__ void wfa_reset_counters(int32_t* c) {
  if (threadIdx.x < 4) c[threadIdx.x] = 0;
}

// out[0] = max over pairs of max(plen, tlen); out must be zeroed first.
__global__ void wfa_max_length(const int32_t* plen, const int32_t* tlen, int n, int32_t* out) {
  int m = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    m = max(m, max(plen[i], tlen[i]));
  for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// Longest-first work order: key = ~(plen + tlen) so an ascending sort puts
// the longest pairs first; value = the pair index.
__global__ void wfa_lpt_keys(const int32_t* plen, const int32_t* tlen, int first, int n,
                             uint32_t* keys, int32_t* vals) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    keys[i] = ~(uint32_t)(plen[first + i] + tlen[first + i]);
    vals[i] = first + i;
  }
}

// Pairs that outgrew every pass: out of memory, no CIGAR, no end position
__global__ void wfa_mark_oom(const int32_t* list, int n, int32_t* score, int32_t* status,
                             int32_t* cig_len, int32_t* end_pos) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int pair = list[i];
    score[pair] = 0;
    status[pair] = WFA_STATUS_OOM;
    if (cig_len) cig_len[pair] = 0;
    if (end_pos) {
      end_pos[3 * pair] = -1;
      end_pos[3 * pair + 1] = INT32_MAX;
      end_pos[3 * pair + 2] = INT32_MIN / 2;
    }
  }
}

// BiWFA: copy the CIGAR pieces (fragments or gap runs) into place
__global__ void biwfa_gather(const BiPiece* pieces, int n, const uint8_t* frag, uint8_t* cig) {
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const BiPiece pc = pieces[i];
    for (int j = threadIdx.x; j < pc.len; j += blockDim.x)
      cig[pc.dst + j] = pc.src >= 0 ? frag[pc.src + j] : (uint8_t)pc.op;
  }
}

// BiWFA: cigar_wfa_score (src/aligner.rs:2700-2719) of whole CIGARs
__global__ void __launch_bounds__(64) biwfa_rescore(KParams p, const int32_t* pairs, int n) {
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const int pair = pairs[i];
    const uint8_t* ops = p.cig + p.cig_off[pair];
    const int len = p.cig_len[pair];
    int score;
    if (p.metric == WFA_METRIC_EDIT || p.metric == WFA_METRIC_INDEL) {
      int c = 0;
      for (int j = threadIdx.x; j < len; j += 64) c += ops[j] != 'M';
      score = wave_sum(c);
    } else {
      score = rescore_cigar(p, ops, len, false, 0, 0, 0);
    }
    if (threadIdx.x == 0) p.score[pair] = score;
  }
}

// ---------------------------------------------------------------------
// Host-side launch helpers
// ---------------------------------------------------------------------
static KernelFn pick(int kind, int mode, bool lds, bool i16, bool packed, bool heur, int nw) {
  switch (kind) {
    case KIND_EDIT: return pick_edit(mode, lds, i16, packed, heur, nw);
    case KIND_AFFINE: return pick_affine(mode, lds, i16, packed, heur, nw);
    case KIND_AFFINE2P: return pick_affine2p(mode, lds, i16, packed, heur, nw);
    case KIND_LINEAR: return pick_linear(mode, lds, i16, packed, heur, nw);
    default: return nullptr;
  }
}

size_t lpt_temp_bytes(int n) {
  size_t bytes = 0;
  hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                     (const int32_t*)nullptr, (int32_t*)nullptr, n);
  return bytes;
}

hipError_t launch_lpt_order(const int32_t* plen, const int32_t* tlen, int first, int n,
                            uint32_t* keys, uint32_t* keys_sorted, int32_t* vals,
                            int32_t* order, void* temp, size_t temp_bytes, hipStream_t stream) {
  int grid = (n + 255) / 256;
  if (grid > 2048) grid = 2048;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(wfa_lpt_keys, dim3(grid), dim3(256), 0, stream, plen, tlen, first, n, keys,
                     vals);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys, keys_sorted, vals, order, n,
                                            0, 32, stream);
}

hipError_t launch_mark_oom(const int32_t* list, int n, int32_t* score, int32_t* status,
                           int32_t* cig_len, int32_t* end_pos, hipStream_t stream) {
  const int grid = (n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024;
  hipLaunchKernelGGL(wfa_mark_oom, dim3(grid > 0 ? grid : 1), dim3(256), 0, stream, list, n,
                     score, status, cig_len, end_pos);
  return hipGetLastError();
}

hipError_t launch_reset(int32_t* counters, hipStream_t stream) {
  hipLaunchKernelGGL(wfa_reset_counters, dim3(1), dim3(64), 0, stream, counters);
  return hipGetLastError();
}

hipError_t launch_max_length(const int32_t* plen, const int32_t* tlen, int n, int32_t* out,
                             hipStream_t stream) {
  hipLaunchKernelGGL(wfa_reset_counters, dim3(1), dim3(64), 0, stream, out);
  int grid = (n + 255) / 256;
  if (grid > 1024) grid = 1024;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(wfa_max_length, dim3(grid), dim3(256), 0, stream, plen, tlen, n, out);
  return hipGetLastError();
}

hipError_t launch_align(const KParams& p, int kind, int mode, bool lds_ring, bool i16, int nw,
                        int grid, size_t lds_bytes, hipStream_t stream) {
  KernelFn fn = pick(kind, mode, lds_ring, i16, p.fmt == WFA_SEQ_PACKED2BITS, heur_kernel(p), nw);
  if (!fn) return hipErrorInvalidValue;
  hipLaunchKernelGGL(fn, dim3(grid), dim3(64 * nw), lds_bytes, stream, p);
  return hipGetLastError();
}

static KernelFn2 pick_biwfa(int kind, int which, bool heur) {
  switch (kind) {
    case KIND_EDIT: return pick_biwfa_edit(which, heur);
    case KIND_AFFINE: return pick_biwfa_affine(which, heur);
    case KIND_AFFINE2P: return pick_biwfa_affine2p(which, heur);
    case KIND_LINEAR: return pick_biwfa_linear(which, heur);
    default: return nullptr;
  }
}

hipError_t launch_biwfa(int kind, int which, bool heur, const BiParams& bp, int grid,
                        size_t lds_bytes, hipStream_t stream) {
  KernelFn2 fn = pick_biwfa(kind, which, heur);
  if (!fn) return hipErrorInvalidValue;
  hipLaunchKernelGGL(fn, dim3(grid), dim3(64 * biwfa_waves(which, heur)), lds_bytes, stream, bp);
  return hipGetLastError();
}

int biwfa_blocks_per_cu(int kind, int which, bool heur, size_t lds_bytes) {
  KernelFn2 fn = pick_biwfa(kind, which, heur);
  int n = 0;
  if (!fn || hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, 64 * biwfa_waves(which, heur),
                                                          lds_bytes) != hipSuccess)
    return 1;
  return n > 0 ? n : 1;
}

hipError_t launch_biwfa_gather(const BiPiece* pieces, int n, const uint8_t* frag, uint8_t* cig,
                               hipStream_t stream) {
  const int grid = std::min(n, 4096);
  hipLaunchKernelGGL(biwfa_gather, dim3(grid), dim3(256), 0, stream, pieces, n, frag, cig);
  return hipGetLastError();
}

hipError_t launch_biwfa_rescore(const KParams& p, const int32_t* pairs, int n,
                                hipStream_t stream) {
  const int grid = std::min(n, 4096);
  hipLaunchKernelGGL(biwfa_rescore, dim3(grid), dim3(64), 0, stream, p, pairs, n);
  return hipGetLastError();
}

hipError_t launch_short(const KParams& p, int mode, int rc, int grid, size_t lds_bytes,
                        hipStream_t stream) {
  KernelFn fn = pick_affine_short(mode, p.fmt == WFA_SEQ_PACKED2BITS, rc);
  if (!fn) return hipErrorInvalidValue;
  hipLaunchKernelGGL(fn, dim3(grid), dim3(64), lds_bytes, stream, p);
  return hipGetLastError();
}

int short_blocks_per_cu(int mode, bool packed, int rc, size_t lds_bytes) {
  KernelFn fn = pick_affine_short(mode, packed, rc);
  int n = 0;
  if (!fn || hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, 64, lds_bytes) != hipSuccess)
    return 1;
  return n > 0 ? n : 1;
}

int max_blocks_per_cu(int kind, int mode, bool lds_ring, bool i16, bool packed, bool heur, int nw,
                      size_t lds_bytes) {
  KernelFn fn = pick(kind, mode, lds_ring, i16, packed, heur, nw);
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, 64 * nw, lds_bytes) != hipSuccess)
    return 1;
  return n > 0 ? n : 1;
}

}  // namespace wfa
