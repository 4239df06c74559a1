// filter_kernels.hip -- the exponential filter of the FILTER bin, applied in
// place to the state after each full step.
//
// Per element and filtered variable (expfilter_var, filter.c:75-92):
//   u <- u * J^jpow, with J = 1 / det(dXb/dx)  (var_to_var_times_JtoPower,
//        utilities.c:885-909; det_dXbdx as coordinates.c:173-203 sets it);
//   u <- F_z F_y F_x u, F = S diag(f) A per direction: the normalized
//        Legendre analysis A and synthesis S on the LGL nodes
//        (LGL_AT_ST_matrices, Legendre.c:257-307) around the coefficient
//        scaling f(i) = exp(-alp (i/N)^s) (expfilter_coeff_array,
//        filter.c:20-55).  The reference applies A, the scaling and S as six
//        directional passes; the product S diag(f) A is the same operator in
//        three, one LDS round trip each;
//   u <- u * J^-jpow.
// One workgroup per element, one node per thread, the filtered variables
// staged in LDS.
#include <hip/hip_runtime.h>

#include "amr.h"
#include "dg_common.h"

namespace nmx {

#define NMX_CONST __attribute__((address_space(4)))

template <int NP, int GEOM, int NT>
__global__ __launch_bounds__(NT) void filter_kernel(FilterArgs a) {
  constexpr int NP2 = NP * NP;
  constexpr int NP3 = NP * NP * NP;
  constexpr int NC = (NP3 + NT - 1) / NT;
  __shared__ double s_u[5][NP3];
  __shared__ double s_F[NP2];
  __shared__ double s_xb[NP];
  const int tid = threadIdx.x;
  for (int q = tid; q < NP2; q += NT) s_F[q] = a.F[q];
  for (int q = tid; q < NP; q += NT) s_xb[q] = a.xb[q];
  const int mask = a.varmask;

  for (int it = blockIdx.x; it < a.nlist; it += gridDim.x) {
    const NMX_CONST SlotRec &R = *(const NMX_CONST SlotRec *)(a.rec + it);
    const long noff = R.noff;
    double *u = a.u + 5 * noff;
    __syncthreads();
    // J^jpow weighting into LDS
    double ooJ[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int node = tid + c * NT;
      ooJ[c] = 1.0;
      if (node >= NP3) continue;
      const double detb = R.dXbdX[0] * R.dXbdX[1] * R.dXbdX[2];
      if (GEOM == 0) {
        ooJ[c] = detb;
      } else {
        double M[3][3];
        if (GEOM == 1) {
          const double *g = a.dXdx + 9 * noff;
#pragma unroll
          for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int q = 0; q < 3; ++q) M[m][q] = g[(3 * m + q) * NP3 + node];
        } else {
          const int I[3] = {node % NP, (node / NP) % NP, node / NP2};
          double X[3];
#pragma unroll
          for (int d = 0; d < 3; ++d)
            X[d] = 0.5 * ((R.bb[2 * d + 1] - R.bb[2 * d]) * s_xb[I[d]] + (R.bb[2 * d + 1] + R.bb[2 * d]));
          shell_dXdx(R.dom, R.s0, R.s1, X, M);
        }
        ooJ[c] = detb * det3(M);
      }
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        if (!(mask >> v & 1)) continue;
        double val = u[v * NP3 + node];
        if (a.jpow == 1) val = val / ooJ[c];
        else if (a.jpow == -1) val = val * ooJ[c];
        s_u[v][node] = val;
      }
    }
    __syncthreads();
    // F along x, y, z
#pragma unroll
    for (int dir = 0; dir < 3; ++dir) {
      double t[NC][5];
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int node = tid + c * NT;
        if (node >= NP3) continue;
        const int i = node % NP, j = (node / NP) % NP, k = node / NP2;
        const int im = dir == 0 ? i : dir == 1 ? j : k;
        const int stride = dir == 0 ? 1 : dir == 1 ? NP : NP2;
        const int base = node - im * stride;
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          if (!(mask >> v & 1)) continue;
          double sum = 0.0;
#pragma unroll
          for (int l = 0; l < NP; ++l) sum += s_F[im * NP + l] * s_u[v][base + l * stride];
          t[c][v] = sum;
        }
      }
      __syncthreads();
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int node = tid + c * NT;
        if (node >= NP3) continue;
#pragma unroll
        for (int v = 0; v < 5; ++v)
          if (mask >> v & 1) s_u[v][node] = t[c][v];
      }
      __syncthreads();
    }
    // undo the weighting, back to HBM
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int node = tid + c * NT;
      if (node >= NP3) continue;
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        if (!(mask >> v & 1)) continue;
        double val = s_u[v][node];
        if (a.jpow == 1) val = val * ooJ[c];
        else if (a.jpow == -1) val = val / ooJ[c];
        u[v * NP3 + node] = val;
      }
    }
  }
}

static constexpr int filter_threads(int np3) { return np3 >= 1024 ? 1024 : (np3 + 63) / 64 * 64; }

template <int NP, int GEOM>
static hipError_t launch_filter_t(const FilterArgs &a, hipStream_t s) {
  constexpr int NT = filter_threads(NP * NP * NP);
  const int grid = a.nlist < 65536 ? a.nlist : 65536;
  hipLaunchKernelGGL((filter_kernel<NP, GEOM, NT>), dim3(grid), dim3(NT), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_filter(int np, int geom, const FilterArgs &a, hipStream_t s) {
  if (a.nlist == 0 || a.varmask == 0) return hipSuccess;
#define NMX_CASE(N)                                              \
  case N:                                                        \
    if (geom == 0) return launch_filter_t<N, 0>(a, s);           \
    if (geom == 1) return launch_filter_t<N, 1>(a, s);           \
    return launch_filter_t<N, 2>(a, s);
  switch (np) {
    NMX_CASE(2)
    NMX_CASE(3)
    NMX_CASE(4)
    NMX_CASE(5)
    NMX_CASE(6)
    NMX_CASE(7)
    NMX_CASE(8)
    NMX_CASE(9)
    NMX_CASE(10)
    default:
      return hipErrorInvalidValue;
  }
#undef NMX_CASE
}

}  // namespace nmx
