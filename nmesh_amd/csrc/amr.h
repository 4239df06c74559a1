// amr.h -- launch interface of the batched (order-bucketed) DG stage kernel.
#pragma once
#include <hip/hip_runtime.h>

namespace nmx {

constexpr int kAmrStage = 1600;   // doubles of LDS for staged neighbour planes per round

// face descriptor, 8 ints per element face:
//   [0] kind: 0 outer boundary, 1 conforming copy, 2 interpolation
//   [1] nb, [2] base node in nb, [3] s1, [4] s2   (kind 1: nb node = base + a*s1 + b*s2)
//   [5] nslot, [6] first slot, [7] staging round  (kind 2)
// slot, 12 ints:
//   [0] nb, [1] m (nb points per direction), [2] base, [3] t1, [4] t2 (nb plane node =
//   base + q1*t1 + q2*t2), [5] swap, [6..9] a0,a1,b0,b1 covered point ranges,
//   [10] offset of the rows for my a, [11] for my b, in Btab (rows of m doubles)
struct AmrArgs {
  const double *win = nullptr, *U = nullptr, *Ain = nullptr;
  double *Aout = nullptr, *Wout = nullptr, *R = nullptr;
  const int *elist = nullptr;
  int nlist = 0;
  const long *soff = nullptr;     // [e] first state double of element e (5 * node offset)
  const int *face = nullptr;      // [e][6][8]
  const int *erounds = nullptr;   // [e] staging rounds
  const int *slot = nullptr;      // [s][12]
  const double *Btab = nullptr;
  const double *dXbdX = nullptr;  // [e][3]
  const double *egeo = nullptr;   // [e][8]
  const int *edom = nullptr;      // [e]
  const double *dXdx = nullptr;   // element blocks of 9*np
  const double *xyz = nullptr;    // element blocks of 3*np
  const double *D = nullptr;      // this bucket's D[i*n + l]
  const double *xb = nullptr;     // this bucket's LGL nodes
  const double *tdev = nullptr;
  int tidx = 0;
  double cA = 0, cW = 0, wend_inv = 0;
  double k[3] = {0, 0, 0};
  double om = 0;
  int wmode = 0, numflux = 0, profile = 0, xcd_map = 0;
};

hipError_t launch_dg_amr(int np, int geom, const AmrArgs &a, int grid, hipStream_t s);

}  // namespace nmx
