// wfa_kernels_affine.hip — the affine kernel set (one translation unit per
// kind so the sets compile in parallel).
#include "wfa_align.h"
#include "wfa_biwfa.h"
#include "wfa_window.h"

namespace wfa {

KernelFn pick_affine(int mode, bool lds, bool i16, bool packed, bool heur, int nw) {
  return pick_kind<KIND_AFFINE>(mode, lds, i16, packed, heur, nw);
}

// The window kernel runs score-only launches (its high-memory CIGAR variant
// keeps a per-slot history and measures slower than the general kernel)
KernelFn pick_affine_window(int mode, bool packed) {
  if (mode != 0) return nullptr;
  return packed ? wfa_window_kernel<0, true> : wfa_window_kernel<0, false>;
}

size_t window_lds_bytes_host(int ring_rows) { return window_lds_bytes(ring_rows); }

KernelFn2 pick_biwfa_affine(int which, bool heur) {
  return pick_biwfa_kind<KIND_AFFINE>(which, heur);
}

}  // namespace wfa
