// wfa_kernels_affine.hip — the affine kernel set (one translation unit per
// kind so the sets compile in parallel).
#include "wfa_align.h"
#include "wfa_biwfa.h"
#include "wfa_short.h"

namespace wfa {

KernelFn pick_affine(int mode, bool lds, bool i16, bool packed, bool heur, int nw) {
  return pick_kind<KIND_AFFINE>(mode, lds, i16, packed, heur, nw);
}

// The short-read kernel (gap-affine, end-to-end, no heuristic): score only
// or high-memory CIGAR
KernelFn pick_affine_short(int mode, bool packed) {
  if (mode == 0) return packed ? wfa_short_kernel<0, true> : wfa_short_kernel<0, false>;
  if (mode == 1) return packed ? wfa_short_kernel<1, true> : wfa_short_kernel<1, false>;
  return nullptr;
}

size_t short_lds_row_bytes_host(int ring_rows) { return short_lds_row_bytes(ring_rows); }

KernelFn2 pick_biwfa_affine(int which, bool heur) {
  return pick_biwfa_kind<KIND_AFFINE>(which, heur);
}

}  // namespace wfa
