// wfa_launch.hip — small utility kernels and the launch dispatch over the
// per-kind kernel sets.

#include <hip/hip_runtime.h>
#include <stdint.h>

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

// ---------------------------------------------------------------------
// Host-side launch helpers
// ---------------------------------------------------------------------
static KernelFn pick(int kind, int mode, bool lds, bool i16, bool packed, bool heur) {
  switch (kind) {
    case KIND_EDIT: return pick_edit(mode, lds, i16, packed, heur);
    case KIND_AFFINE: return pick_affine(mode, lds, i16, packed, heur);
    case KIND_AFFINE2P: return pick_affine2p(mode, lds, i16, packed, heur);
    default: return nullptr;
  }
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

hipError_t launch_align(const KParams& p, int kind, int mode, bool lds_ring, bool i16, int grid,
                        size_t lds_bytes, hipStream_t stream) {
  KernelFn fn = pick(kind, mode, lds_ring, i16, p.fmt == WFA_SEQ_PACKED2BITS,
                     p.heuristic != WFA_HEURISTIC_NONE);
  if (!fn) return hipErrorInvalidValue;
  hipLaunchKernelGGL(fn, dim3(grid), dim3(64), lds_bytes, stream, p);
  return hipGetLastError();
}

int max_blocks_per_cu(int kind, int mode, bool lds_ring, bool i16, bool packed, bool heur,
                      size_t lds_bytes) {
  KernelFn fn = pick(kind, mode, lds_ring, i16, packed, heur);
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, 64, lds_bytes) != hipSuccess) return 1;
  return n > 0 ? n : 1;
}

}  // namespace wfa
