// fvelm_kernels.hip -- the FV element kernels (fvelm_impl.h) with
// dnfn_extrap1's fixed weights, and the launchers that pick the variant with
// the fv_divf_extrap s1/s2/opt limit (fvelm_ex_kernels.hip) when it is on.
#include "fvelm_impl.h"

namespace nmx {

hipError_t launch_fvelm_stage(int np, int rec, int geom, const StageArgs &a, const double *wc, int grid,
                              hipStream_t s) {
  if (a.ex_on || a.recend != 0 || a.rec_first) return launch_fvelm_stage_ex(np, rec, geom, a, wc, grid, s);
  return launch_fvelm_stage_v<false>(np, rec, geom, a, wc, grid, s);
}

hipError_t launch_fvelm_amr(int np, int rec, int geom, const AmrArgs &a, const double *wc, int grid,
                            hipStream_t s) {
  if (a.ex_on || a.recend != 0 || a.rec_first) return launch_fvelm_amr_ex(np, rec, geom, a, wc, grid, s);
  return launch_fvelm_amr_v<false>(np, rec, geom, a, wc, grid, s);
}

}  // namespace nmx
