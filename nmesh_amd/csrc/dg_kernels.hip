// dg_kernels.hip -- gfx950 kernels for the scalarwave1 DG right-hand side
// fused with the Runge-Kutta stage update.
//
// One workgroup owns one element.  The element's four flux-carrying fields
// (pi, cx, cy, cz) are staged in LDS; each thread owns one node (in chunks of
// NT nodes) and computes:
//   - the tensor-product derivatives D_m q along the three directions
//     (src/main/amr/array.c:36-300 via basis_array_deriv1, basis.c:61-95),
//   - the non-conservative Cartesian divergence sum_m dXdx[m][i] dXbdX[m] D_m
//     (src/main/coordinates/derivs.c:102-181,756-815) of the scalarwave1 fluxes
//     f_pi^i = -c_i, f_cj^i = -pi delta_ij (scalarwave1.c:39-48,62-134),
//   - on face nodes the upwind or LLF numerical flux against the neighbour's
//     face value and the strong-form lift with sqrtgdiag / w_end
//     (scalarwave1.c:139-209,287-321,324-499; dg.c:30-151; numflux1d.c:85-112),
//   - the RHS r = -(div f + lift), r_phi = pi (scalarwave1.c:242-284),
//   - and the RK stage combination of evolve_RK4_mesh / evolve_sspRK3_mesh /
//     evolve_Euler_mesh (RungeKutta.c:19-188) in the epilogue, so the RHS is
//     never written to HBM during a time step.
// State layout: element blocks, element e at 5*off[e] doubles, variable v at
// +v*np, node ijk = i + n*(j + n*k) inside.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace nmx {

// upwind: F = fnum - f_inner for the 4 flux-carrying fields.  With A = f.n
// Jacobian [[0,-n^T],[-n,0]], A+/- = (A +- |A|)/2 gives
// F_pi = (dpi + n.dc)/2 and F_c = n F_pi with d = inner - adjacent.
// LLF (amax = 1): F_pi = (dpi + n.dc)/2, F_c = (dc + n dpi)/2.
__device__ __forceinline__ void face_flux(int numflux, const double n[3], double dpi,
                                          const double dc[3], double F[4]) {
  double ndc = n[0] * dc[0] + n[1] * dc[1] + n[2] * dc[2];
  double Fp = 0.5 * (dpi + ndc);
  F[0] = Fp;
  if (numflux == 0) {
    F[1] = n[0] * Fp; F[2] = n[1] * Fp; F[3] = n[2] * Fp;
  } else {
    F[1] = 0.5 * (dc[0] + n[0] * dpi);
    F[2] = 0.5 * (dc[1] + n[1] * dpi);
    F[3] = 0.5 * (dc[2] + n[2] * dpi);
  }
}

// plane wave of scalarwave1_set_profile_pt (scalarwave1.c:503-524)
__device__ __forceinline__ void profile(const StageArgs &a, double x, double y, double z,
                                        double t, double u[4]) {
  double arg = a.k[0] * x + a.k[1] * y + a.k[2] * z - a.om * t;
  double s, c;
  sincos(arg, &s, &c);
  (void)s;
  u[0] = -a.om * c; u[1] = a.k[0] * c; u[2] = a.k[1] * c; u[3] = a.k[2] * c;
}

template <int NP, bool CURV, int NT>
__global__ __launch_bounds__(NT) void dg_stage_kernel(StageArgs a) {
  constexpr int NP2 = NP * NP;
  constexpr int NP3 = NP * NP * NP;
  __shared__ double s_u[4][NP3];
  __shared__ double s_D[NP2];

  const int tid = threadIdx.x;
  for (int i = tid; i < NP2; i += NT) s_D[i] = a.D[i];

  for (int e = a.e0 + blockIdx.x; e < a.e1; e += gridDim.x) {
    const long ebase = 5L * NP3 * e;
    const double *win = a.win + ebase;
    __syncthreads();
#pragma unroll
    for (int v = 0; v < 4; ++v)
      for (int i = tid; i < NP3; i += NT) s_u[v][i] = win[v * NP3 + i];
    __syncthreads();

    double dXbdX[3] = {a.dXbdX[3 * e], a.dXbdX[3 * e + 1], a.dXbdX[3 * e + 2]};

    for (int node = tid; node < NP3; node += NT) {
      const int k = node / NP2;
      const int j = (node / NP) % NP;
      const int i = node % NP;

      // tensor-product derivatives (computational coordinates)
      double dpi[3] = {0, 0, 0};
      double dc[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};   // dc[c][m] = D_m c_c
#pragma unroll
      for (int l = 0; l < NP; ++l) {
        const double D0 = s_D[i * NP + l], D1 = s_D[j * NP + l], D2 = s_D[k * NP + l];
        const int i0 = l + NP * (j + NP * k), i1 = i + NP * (l + NP * k), i2 = i + NP * (j + NP * l);
        dpi[0] += D0 * s_u[0][i0];
        dpi[1] += D1 * s_u[0][i1];
        dpi[2] += D2 * s_u[0][i2];
        if (CURV) {
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            dc[c][0] += D0 * s_u[1 + c][i0];
            dc[c][1] += D1 * s_u[1 + c][i1];
            dc[c][2] += D2 * s_u[1 + c][i2];
          }
        } else {
          dc[0][0] += D0 * s_u[1][i0];
          dc[1][1] += D1 * s_u[2][i1];
          dc[2][2] += D2 * s_u[3][i2];
        }
      }

      // Cartesian partials and RHS volume terms: r_pi = d_i c_i, r_cj = d_j pi
      double M[3][3];
      double rpi, rc[3];
      if (CURV) {
        const double *g = a.dXdx + 9L * NP3 * e + node;
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
          for (int q = 0; q < 3; ++q) M[m][q] = g[(3 * m + q) * NP3];
        double sp[3], sc[3][3];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          sp[m] = dXbdX[m] * dpi[m];
#pragma unroll
          for (int c = 0; c < 3; ++c) sc[c][m] = dXbdX[m] * dc[c][m];
        }
        rpi = 0.0;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          double dq_cq = M[0][q] * sc[q][0] + M[1][q] * sc[q][1] + M[2][q] * sc[q][2];
          rpi += dq_cq;
          rc[q] = M[0][q] * sp[0] + M[1][q] * sp[1] + M[2][q] * sp[2];
        }
      } else {
        rpi = dXbdX[0] * dc[0][0] + dXbdX[1] * dc[1][1] + dXbdX[2] * dc[2][2];
#pragma unroll
        for (int q = 0; q < 3; ++q) rc[q] = dXbdX[q] * dpi[q];
      }

      // faces this node lies on
      const int I[3] = {i, j, k};
#pragma unroll
      for (int dir = 0; dir < 3; ++dir) {
#pragma unroll
        for (int side = 0; side < 2; ++side) {
          if (I[dir] != side * (NP - 1)) continue;
          const int f = 2 * dir + side;
          const double sig = side ? 1.0 : -1.0;
          double nrm[3], scale;
          if (CURV) {
            double s2 = M[dir][0] * M[dir][0] + M[dir][1] * M[dir][1] + M[dir][2] * M[dir][2];
            double s = sqrt(s2);
            double smag = sig * s;
            nrm[0] = M[dir][0] / smag; nrm[1] = M[dir][1] / smag; nrm[2] = M[dir][2] / smag;
            scale = dXbdX[dir] * s * a.wend_inv;
          } else {
            nrm[0] = nrm[1] = nrm[2] = 0.0;
            nrm[dir] = sig;
            scale = dXbdX[dir] * a.wend_inv;
          }
          const int ia = (dir == 0) ? j : i;          // index along Dir1_norm(dir)
          const int ib = (dir == 2) ? j : k;          // index along Dir2_norm(dir)
          const int4 fd = a.face[6 * e + f];
          double ua[4];
          if (fd.x >= 0) {
            const double *nw = a.win + 5L * NP3 * fd.x + fd.y + ia * fd.z + ib * fd.w;
            ua[0] = nw[0]; ua[1] = nw[NP3]; ua[2] = nw[2 * NP3]; ua[3] = nw[3 * NP3];
          } else if (a.profile) {
            const double *xp = a.xyz + 3L * NP3 * e + node;
            profile(a, xp[0], xp[NP3], xp[2 * NP3], a.tdev[a.tidx], ua);
          } else {
            ua[0] = ua[1] = ua[2] = ua[3] = 0.0;
          }
          const double ui0 = s_u[0][node];
          const double dcv[3] = {s_u[1][node] - ua[1], s_u[2][node] - ua[2], s_u[3][node] - ua[3]};
          double F[4];
          face_flux(a.numflux, nrm, ui0 - ua[0], dcv, F);
          rpi -= F[0] * scale;
          rc[0] -= F[1] * scale;
          rc[1] -= F[2] * scale;
          rc[2] -= F[3] * scale;
        }
      }

      const double r[5] = {rpi, rc[0], rc[1], rc[2], s_u[0][node]};
      const long gi = ebase + node;
      if (a.R) {
#pragma unroll
        for (int v = 0; v < 5; ++v) a.R[gi + v * NP3] = r[v];
      }
      if (a.Aout) {
        const double *base = a.Ain ? a.Ain : a.win;
#pragma unroll
        for (int v = 0; v < 5; ++v) a.Aout[gi + v * NP3] = base[gi + v * NP3] + a.cA * r[v];
      }
      if (a.wmode == 1) {
#pragma unroll
        for (int v = 0; v < 5; ++v) a.Wout[gi + v * NP3] = a.U[gi + v * NP3] + a.cW * r[v];
      } else if (a.wmode == 2) {
#pragma unroll
        for (int v = 0; v < 5; ++v)
          a.Wout[gi + v * NP3] = 0.75 * a.U[gi + v * NP3] + 0.25 * (a.win[gi + v * NP3] + a.cW * r[v]);
      }
    }
  }
}

template <int NP, bool CURV>
static hipError_t launch_np(const StageArgs &a, int grid, hipStream_t s) {
  constexpr int NP3 = NP * NP * NP;
  constexpr int NT = NP3 <= 64 ? 64 : (NP3 <= 128 ? 128 : (NP3 <= 256 ? 256 : (NP3 <= 512 ? 512 : 1024)));
  hipLaunchKernelGGL((dg_stage_kernel<NP, CURV, NT>), dim3(grid), dim3(NT), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_dg_stage(int np, bool curv, const StageArgs &a, int grid, hipStream_t s) {
#define NMX_CASE(N)                                                    \
  case N:                                                              \
    return curv ? launch_np<N, true>(a, grid, s) : launch_np<N, false>(a, grid, s);
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
