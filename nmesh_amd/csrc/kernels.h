// kernels.h -- launch interface between the libnmx host code and its kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace nmx {

// Arguments of one fused DG RHS + RK stage launch over elements [e0, e1).
// Output combinations (each optional):
//   R    = r                                      (RHS only)
//   Aout = (Ain ? Ain : win) + cA * r              (RK accumulator "u")
//   Wout = U + cW * r                   (wmode 1)  (next stage input "w")
//   Wout = 0.75 U + 0.25 (win + cW r)   (wmode 2)  (sspRK3 third stage input)
struct StageArgs {
  const double *win = nullptr;   // stage input state
  const double *U = nullptr;     // state at the start of the step (u_p)
  const double *Ain = nullptr;
  double *Aout = nullptr;
  double *Wout = nullptr;
  double *R = nullptr;
  const double *dXdx = nullptr;  // [e][9][np]
  const double *dXbdX = nullptr; // [e][3]
  const double *xyz = nullptr;   // [e][3][np]
  const int4 *face = nullptr;    // [e][6]: {nb elm (-1 none), local base, s1, s2}
  const double *D = nullptr;     // D[i*np + l]
  const double *tdev = nullptr;  // stage times on the device; t = tdev[tidx]
  int tidx = 0;
  double cA = 0, cW = 0, wend_inv = 0;
  double k[3] = {0, 0, 0};
  double om = 0;
  int e0 = 0, e1 = 0;
  int wmode = 0, numflux = 0, profile = 0;
};

hipError_t launch_dg_stage(int np, bool curv, const StageArgs &a, int grid, hipStream_t s);

}  // namespace nmx
