// fvelm_ex_kernels.hip -- the FV element kernels (fvelm_impl.h) with the
// fv_divf_extrap_s1 / _s2 / _opt limit of dnfn_extrap1 (rec1d.c:542-609);
// a separate translation unit so both variants compile in parallel and the
// common one keeps its registers.
#include "fvelm_impl.h"

namespace nmx {

hipError_t launch_fvelm_stage_ex(int np, int rec, int geom, const StageArgs &a, const double *wc, int grid,
                                 hipStream_t s) {
  return launch_fvelm_stage_v<true>(np, rec, geom, a, wc, grid, s);
}

hipError_t launch_fvelm_amr_ex(int np, int rec, int geom, const AmrArgs &a, const double *wc, int grid,
                               hipStream_t s) {
  return launch_fvelm_amr_v<true>(np, rec, geom, a, wc, grid, s);
}

}  // namespace nmx
