// wfa_kernels_linear.hip — the gap-linear kernel set (one translation unit
// per kind so the sets compile in parallel).
#include "wfa_align.h"
#include "wfa_biwfa.h"

namespace wfa {

KernelFn pick_linear(int mode, bool lds, bool i16, bool packed, bool heur, int nw) {
  return pick_kind<KIND_LINEAR>(mode, lds, i16, packed, heur, nw);
}

KernelFn2 pick_biwfa_linear(int which, bool heur) {
  return pick_biwfa_kind<KIND_LINEAR>(which, heur);
}

}  // namespace wfa
