// wfa_kernels_affine.hip — the affine kernel set (one translation unit per
// kind so the sets compile in parallel).
#include "wfa_align.h"
#include "wfa_biwfa.h"

namespace wfa {

KernelFn pick_affine(int mode, bool lds, bool i16, bool packed, bool heur) {
  return pick_kind<KIND_AFFINE>(mode, lds, i16, packed, heur);
}

KernelFn2 pick_biwfa_affine(int which, bool heur) {
  return pick_biwfa_kind<KIND_AFFINE>(which, heur);
}

}  // namespace wfa
