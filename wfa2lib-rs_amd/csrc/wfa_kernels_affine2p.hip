// wfa_kernels_affine2p.hip — the affine2p kernel set (one translation unit per
// kind so the sets compile in parallel).
#include "wfa_align.h"
#include "wfa_biwfa.h"

namespace wfa {

KernelFn pick_affine2p(int mode, bool lds, bool i16, bool packed, bool heur, int nw) {
  return pick_kind<KIND_AFFINE2P>(mode, lds, i16, packed, heur, nw);
}

KernelFn2 pick_biwfa_affine2p(int which, bool heur) {
  return pick_biwfa_kind<KIND_AFFINE2P>(which, heur);
}

}  // namespace wfa
