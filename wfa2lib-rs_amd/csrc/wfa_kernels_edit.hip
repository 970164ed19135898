// wfa_kernels_edit.hip — the edit kernel set (one translation unit per
// kind so the sets compile in parallel).
#include "wfa_align.h"

namespace wfa {

KernelFn pick_edit(int mode, bool lds, bool i16, bool packed, bool heur) {
  return pick_kind<KIND_EDIT>(mode, lds, i16, packed, heur);
}

}  // namespace wfa
