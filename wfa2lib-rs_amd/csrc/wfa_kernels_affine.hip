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
// or high-memory CIGAR, for ring class `rc` (2: the default penalties built
// in, with the smallest ring they allow)
KernelFn pick_affine_short(int mode, bool packed, int rc) {
  if (rc == 2) {
    if (mode == 0) return packed ? wfa_short_kernel<0, true, 2, 1> : wfa_short_kernel<0, false, 2, 1>;
    if (mode == 1) return packed ? wfa_short_kernel<1, true, 2, 1> : wfa_short_kernel<1, false, 2, 1>;
  } else if (rc == 0) {
    if (mode == 0) return packed ? wfa_short_kernel<0, true, 0, 0> : wfa_short_kernel<0, false, 0, 0>;
    if (mode == 1) return packed ? wfa_short_kernel<1, true, 0, 0> : wfa_short_kernel<1, false, 0, 0>;
  } else if (rc == 1) {
    if (mode == 0) return packed ? wfa_short_kernel<0, true, 1, 0> : wfa_short_kernel<0, false, 1, 0>;
    if (mode == 1) return packed ? wfa_short_kernel<1, true, 1, 0> : wfa_short_kernel<1, false, 1, 0>;
  }
  return nullptr;
}

size_t short_lds_row_bytes_host(int rc) { return short_lds_row_bytes(rc); }

int short_ring_class(int x, int oe, int e, int* gsh) {
  // g: the largest power of two dividing all three penalties
  int g = 0;
  while (g < 8 && ((x | oe | e) & ((2 << g) - 1)) == 0) ++g;
  *gsh = g;
  const int xr = x >> g, oer = oe >> g, er = e >> g;
  if (g == ShPen<1>::gsh && xr == ShPen<1>::x && oer == ShPen<1>::oe && er == ShPen<1>::e)
    return 2;
  if (xr < ShRing<0>::M && oer < ShRing<0>::M && er < ShRing<0>::E) return 0;
  if (xr < ShRing<1>::M && oer < ShRing<1>::M && er < ShRing<1>::E) return 1;
  return -1;
}

KernelFn2 pick_biwfa_affine(int which, bool heur) {
  return pick_biwfa_kind<KIND_AFFINE>(which, heur);
}

}  // namespace wfa
