// fv.h -- launch interface of the FV WENO5 + LLF kernels (special-relativistic
// Euler, ideal gas) on a periodic grid of cubic cell blocks.
#pragma once
#include <hip/hip_runtime.h>

namespace nmx {

// Grid: nbx^3 blocks of B^3 cells (cell size 1, periodic).  A state register
// holds [block][var][B^3] doubles, block b = bx + nbx (by + nbx bz), cell
// i + B (j + B k) inside; conserved vars (D, Sx, Sy, Sz, tau), primitive vars
// (rho, vx, vy, vz, p).
struct FvGrid {
  int B = 16;        // cells per block edge
  int nbx = 16;      // blocks per grid edge
  double gam = 2.0;  // ideal-gas Gamma
};

// c2p of every cell: P = primitives of U, P's pressure as the Newton guess;
// blocks whose solve fails get bad[b] = 1 (never cleared here)
struct FvC2pArgs {
  FvGrid g;
  const double *U = nullptr;
  double *P = nullptr;
  int *bad = nullptr;
  double tol = 1e-12;
};

// fused RHS + RK stage on 8^3-cell tiles:
//   r = -sum_d (F_{i+1/2} - F_{i-1/2}) from the primitives P (WENO5 + LLF)
//   R = r; Aout = (Ain ? Ain : win) + cA r;
//   Wout = U + cW r (wmode 1) or 0.75 U + 0.25 (win + cW r) (wmode 2)
struct FvStageArgs {
  FvGrid g;
  const double *P = nullptr;
  const double *win = nullptr, *U = nullptr, *Ain = nullptr;
  double *Aout = nullptr, *Wout = nullptr, *R = nullptr;
  double cA = 0, cW = 0;
  int wmode = 0;
  int xcd_map = 1;
};

hipError_t launch_fv_c2p(const FvC2pArgs &a, hipStream_t s);
// the largest characteristic speed |lambda+-| over all cells and directions
// of primitives P: per-workgroup maxima in part[grid], the grid's in out[0]
int fv_speed_grid(const FvGrid &g);
hipError_t launch_fv_max_speed(const FvGrid &g, const double *P, double *part, double *out, hipStream_t s);
hipError_t launch_fv_p2c(const FvGrid &g, const double *P, double *U, hipStream_t s);
hipError_t launch_fv_stage(const FvStageArgs &a, hipStream_t s);
// flat [v][cell] (cell = i + N (j + N k), N = B nbx) <-> block layout
hipError_t launch_fv_layout(const FvGrid &g, const double *src, double *dst, int to_blocks, hipStream_t s);

}  // namespace nmx
