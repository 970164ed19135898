// wfa_kernels_edit.hip — the edit kernel set (one translation unit per
// kind so the sets compile in parallel).
#include "wfa_align.h"
#include "wfa_biwfa.h"

namespace wfa {

KernelFn pick_edit(int mode, bool lds, bool i16, bool packed, bool heur, int nw) {
  return pick_kind<KIND_EDIT>(mode, lds, i16, packed, heur, nw);
}

KernelFn2 pick_biwfa_edit(int which, bool heur) {
  return pick_biwfa_kind<KIND_EDIT>(which, heur);
}

}  // namespace wfa
