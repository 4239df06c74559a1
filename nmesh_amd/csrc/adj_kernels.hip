// adj_kernels.hip -- discrete adjoint of the scalarwave1 DG right-hand side,
// fused with the adjoint RK stage combination (gfx950).
//
// The forward RHS (dg_stage_kernel) is linear in u:
//   r_pi(p) = sum_m sum_q M_p[m][q] g_m (D_m c_q)(p) - sum_{f at p} s_f(p) [A(n_f(p)) (u(p) - u_adj(p))]_pi
//   r_cq(p) = sum_m M_p[m][q] g_m (D_m pi)(p)        - sum_{f at p} s_f(p) [A(n_f(p)) (u(p) - u_adj(p))]_cq
//   r_phi(p) = pi(p)
// with g_m = dXb/dX of the element, M = dX/dx, D_m the LGL derivative along
// direction m, s_f the lift scale sqrtgdiag / w_end and A(n) the flux
// Jacobian of the upwind or LLF flux (scalarwave1.c:139-209,324-499;
// numflux1d.c:85-112; dg.c:90-146).  A(n) is symmetric for both fluxes
// (upwind: 1/2 [[1, n^T], [n, n n^T]]; LLF: 1/2 [[1, n^T], [n, I]]), so its
// transpose is applied with the forward face_flux.  The transpose y = L^T lam:
//   y_cq(s) = sum_m (D_m^T a_mq)(s),   a_mq = g_m M[m][q] lam_pi
//   y_pi(s) = sum_m (D_m^T b_m)(s) + lam_phi(s),  b_m = g_m sum_q M[m][q] lam_cq
//   face nodes: y(s) -= s_f(s) A(n_f(s)) lam(s)
//               y(s') += s_f'(p) A(n_f'(p)) lam_nb(p) for the neighbour node p
//               whose face f' reads s (conforming faces: the relation is
//               mutual, so it is gathered through this element's own face
//               descriptor and the neighbour's face index)
//   y_phi = 0.
// The neighbour's normal and lift scale are taken from the neighbour's own
// geometry at p, so the transpose is exact to rounding.
#include <hip/hip_runtime.h>

#include "dg_common.h"
#include "kernels.h"

namespace nmx {

// dX/dx of element e at its node `node` (GEOM 1 stored, 2 recomputed shell)
template <int NP, int GEOM>
__device__ __forceinline__ void elm_metric(const AdjArgs &a, int e, int node, double M[3][3]) {
  constexpr int NP3 = NP * NP * NP;
  if (GEOM == 1) {
    const double *p = a.dXdx + 9L * NP3 * e + node;
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int q = 0; q < 3; ++q) M[m][q] = p[(3 * m + q) * NP3];
  } else if (GEOM == 2) {
    const int i = node % NP, j = (node / NP) % NP, k = node / (NP * NP);
    const double *bb = a.egeo + 8 * e;
    double X[3];
    X[0] = 0.5 * ((bb[1] - bb[0]) * a.xb[i] + (bb[1] + bb[0]));
    X[1] = 0.5 * ((bb[3] - bb[2]) * a.xb[j] + (bb[3] + bb[2]));
    X[2] = 0.5 * ((bb[5] - bb[4]) * a.xb[k] + (bb[5] + bb[4]));
    shell_dXdx(a.edom[e], bb[6], bb[7], X, M);
  }
}

template <int NP, int GEOM, int NT>
__global__ __launch_bounds__(NT, 2) void dg_adjoint_kernel(AdjArgs a) {
  constexpr int NP2 = NP * NP;
  constexpr int NP3 = NP * NP * NP;
  constexpr int NC = (NP3 + NT - 1) / NT;
  constexpr bool CURV = (GEOM != 0);
  __shared__ double s_t[4][NP3];       // staged fields of one direction
  __shared__ double s_f[4][6][NP2];    // neighbours' outgoing adjoint fluxes
  __shared__ double s_D[NP2];

  const int tid = threadIdx.x;
  for (int q = tid; q < NP2; q += NT) s_D[q] = a.D[q];

  int e_hi, e_step, e_first;
  if (a.xcd_map) {
    const int xcd = blockIdx.x & 7, q = blockIdx.x >> 3, g8 = gridDim.x >> 3;
    const int chunk = (a.e1 - a.e0 + 7) >> 3;
    const int e_lo = a.e0 + xcd * chunk;
    e_hi = min(e_lo + chunk, a.e1);
    e_first = e_lo + q;
    e_step = g8;
  } else {
    e_hi = a.e1;
    e_first = a.e0 + blockIdx.x;
    e_step = gridDim.x;
  }

  for (int e = e_first; e < e_hi; e += e_step) {
    const long ebase = 5L * NP3 * e;
    const double g[3] = {a.dXbdX[3 * e], a.dXbdX[3 * e + 1], a.dXbdX[3 * e + 2]};
    __syncthreads();

    // ---- phase 1: own lam to registers, neighbour outgoing fluxes to LDS ----
    double lam[NC][4];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int node = tid + c * NT;
      if (node < NP3) {
#pragma unroll
        for (int v = 0; v < 4; ++v) lam[c][v] = a.lam[ebase + node + v * NP3];
      }
    }
    for (int idx = tid; idx < 6 * NP2; idx += NT) {
      const int f = idx / NP2, q = idx % NP2;
      const int4 fd = a.face[6 * e + f];
      double F[4] = {0.0, 0.0, 0.0, 0.0};
      if (fd.x >= 0) {
        const int nb = fd.x, nf = a.nbface[6 * e + f];
        const int p = fd.y + (q % NP) * fd.z + (q / NP) * fd.w;
        const double *lp = a.lam + 5L * NP3 * nb + p;
        double M[3][3];
        elm_metric<NP, GEOM>(a, nb, p, M);
        double nrm[3], scale;
        face_normal<GEOM>(M, a.dXbdX + 3 * nb, nf, a.wend_inv, nrm, scale);
        const double lc[3] = {lp[NP3], lp[2 * NP3], lp[3 * NP3]};
        face_flux(a.numflux, nrm, lp[0], lc, F);
#pragma unroll
        for (int v = 0; v < 4; ++v) F[v] *= scale;
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) s_f[v][f][q] = F[v];
    }

    // own metric per node (registers)
    double M[NC][3][3];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int node = tid + c * NT;
      if (CURV && node < NP3) elm_metric<NP, GEOM>(a, e, node, M[c]);
    }

    // ---- phase 2: transposed volume term, direction by direction ----
    double y[NC][4];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int v = 0; v < 4; ++v) y[c][v] = 0.0;

#pragma unroll
    for (int m = 0; m < 3; ++m) {
      // Cartesian elements need one staging only: a_mm = g_m lam_pi and
      // b_m = g_m lam_cm differ from lam by the factor g_m
      if (!CURV && m > 0) break;
      if (m > 0) __syncthreads();   // previous direction's s_t is consumed
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int node = tid + c * NT;
        if (node < NP3) {
          if (CURV) {
            s_t[0][node] = g[m] * M[c][m][0] * lam[c][0];
            s_t[1][node] = g[m] * M[c][m][1] * lam[c][0];
            s_t[2][node] = g[m] * M[c][m][2] * lam[c][0];
            s_t[3][node] = g[m] * (M[c][m][0] * lam[c][1] + M[c][m][1] * lam[c][2] + M[c][m][2] * lam[c][3]);
          } else {
#pragma unroll
            for (int v = 0; v < 4; ++v) s_t[v][node] = lam[c][v];
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int node = tid + c * NT;
        if (node >= NP3) continue;
        const int I[3] = {node % NP, (node / NP) % NP, node / NP2};
        if (CURV) {
          const int im = I[m];
          const int stride = (m == 0) ? 1 : (m == 1) ? NP : NP2;
          const int base = node - im * stride;
          double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
#pragma unroll
          for (int l = 0; l < NP; ++l) {
            const double Dl = s_D[l * NP + im];
            const int src = base + l * stride;
            t0 += Dl * s_t[0][src];
            t1 += Dl * s_t[1][src];
            t2 += Dl * s_t[2][src];
            t3 += Dl * s_t[3][src];
          }
          y[c][1] += t0; y[c][2] += t1; y[c][3] += t2; y[c][0] += t3;
        } else {
#pragma unroll
          for (int mm = 0; mm < 3; ++mm) {
            const int im = I[mm];
            const int stride = (mm == 0) ? 1 : (mm == 1) ? NP : NP2;
            const int base = node - im * stride;
            double tp = 0.0, tc = 0.0;
#pragma unroll
            for (int l = 0; l < NP; ++l) {
              const double Dl = s_D[l * NP + im];
              const int src = base + l * stride;
              tp += Dl * s_t[0][src];
              tc += Dl * s_t[1 + mm][src];
            }
            y[c][1 + mm] += g[mm] * tp;
            y[c][0] += g[mm] * tc;
          }
        }
      }
    }

    // ---- face terms and the stage combination ----
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int node = tid + c * NT;
      if (node >= NP3) continue;
      const int I[3] = {node % NP, (node / NP) % NP, node / NP2};
      const long gi = ebase + node;
#pragma unroll
      for (int dir = 0; dir < 3; ++dir) {
#pragma unroll
        for (int side = 0; side < 2; ++side) {
          if (I[dir] != side * (NP - 1)) continue;
          const int f = 2 * dir + side;
          double nrm[3], scale;
          double Mz[3][3];
          if (CURV) {
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
              for (int q = 0; q < 3; ++q) Mz[r][q] = M[c][r][q];
          }
          face_normal<GEOM>(Mz, g, f, a.wend_inv, nrm, scale);
          const int ia = (dir == 0) ? I[1] : I[0];
          const int ib = (dir == 2) ? I[1] : I[2];
          const int q = ia + NP * ib;
          const double lc[3] = {lam[c][1], lam[c][2], lam[c][3]};
          double F[4];
          face_flux(a.numflux, nrm, lam[c][0], lc, F);
#pragma unroll
          for (int v = 0; v < 4; ++v) y[c][v] += s_f[v][f][q] - scale * F[v];
        }
      }
      const double yv[5] = {y[c][0] + a.lam[gi + 4 * NP3], y[c][1], y[c][2], y[c][3], 0.0};
      if (a.R) {
#pragma unroll
        for (int v = 0; v < 5; ++v) a.R[gi + v * NP3] = yv[v];
      }
      if (a.Aout) {
        double Ab[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) Ab[v] = a.Abase[gi + v * NP3];
#pragma unroll
        for (int v = 0; v < 5; ++v) a.Aout[gi + v * NP3] = Ab[v] + a.cA * yv[v];
      }
      if (a.Wout) {
        double Wb[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) Wb[v] = a.Wbase[gi + v * NP3];
#pragma unroll
        for (int v = 0; v < 5; ++v) a.Wout[gi + v * NP3] = a.cU * Wb[v] + a.cW * yv[v];
      }
    }
  }
}

template <int NP, int GEOM>
static hipError_t launch_adj_np(const AdjArgs &a, int grid, hipStream_t s) {
  constexpr int NP3 = NP * NP * NP;
  constexpr int NT = ((NP3 + 1) / 2 + 63) / 64 * 64;   // two nodes per thread
  hipLaunchKernelGGL((dg_adjoint_kernel<NP, GEOM, NT>), dim3(grid), dim3(NT), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_dg_adjoint(int np, int geom, const AdjArgs &a, int grid, hipStream_t s) {
#define NMX_CASE(N)                                           \
  case N:                                                     \
    if (geom == 0) return launch_adj_np<N, 0>(a, grid, s);    \
    if (geom == 1) return launch_adj_np<N, 1>(a, grid, s);    \
    return launch_adj_np<N, 2>(a, grid, s);
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
