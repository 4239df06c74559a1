// dg_kernels.hip -- gfx950 kernels for the scalarwave1 DG right-hand side
// fused with the Runge-Kutta stage update.
//
// One workgroup owns one element at a time.  The element's four
// flux-carrying fields (pi, cx, cy, cz) are staged in LDS; each thread owns
// one node per chunk of NT nodes and computes:
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
// Geometry (template GEOM): 0 Cartesian box (dXdx = 1), 1 dXdx streamed from
// HBM, 2 CubedShell dXdx recomputed from the element box and the cubed-sphere
// map (coordtrans_CubedSphere.c:259-509 with constant radii) instead of read.
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

// dX/dx of a CubedShell (constant radii s0, s1, centre 0) at patch point X
// (lam, A, B); the same expressions as the host geometry.  Written with
// compile-time component indices only (runtime-indexed private arrays would
// live in scratch memory).
__device__ __forceinline__ void shell_dXdx(int dom, double s0, double s1, const double X[3],
                                           double M[3][3]) {
  const double lam = X[0], A = X[1], B = X[2];
  const double pm = (dom & 1) ? 1.0 : -1.0;
  const double s = sqrt(1.0 + A * A + B * B);
  const double a0 = pm * s0 / s, a1 = pm * s1 / s;
  const double r = (a1 - a0) * lam + a0;
  const int dir = dom >> 1;
  const int ia = (dir == 1) ? 0 : 1;      // axis scaled by A
  const double ra = A * r, rb = B * r;
  double rv[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) rv[d] = (d == dir) ? r : ((d == ia) ? ra : rb);
  const double rc2 = rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
  const double ooda = 1.0 / (a1 - a0);
  const double oor = 1.0 / r, rr2 = r * r;
  const double mA = -ra / rr2, mB = -rb / rr2;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const bool isr = (d == dir), isa = (d == ia);
    const double md = isr ? (oor - r / rc2) : (-rv[d] / rc2);
    const double da0 = a0 * md, da1 = a1 * md;
    M[0][d] = ((isr ? 1.0 : 0.0) - da0 - (da1 - da0) * lam) * ooda;
    M[1][d] = isr ? mA : (isa ? oor : 0.0);
    M[2][d] = isr ? mB : ((!isa) ? oor : 0.0);
  }
}

template <int NP, int GEOM, int NT>
__global__ __launch_bounds__(NT, NMX_DG_WAVES) void dg_stage_kernel(StageArgs a) {
  constexpr int NP2 = NP * NP;
  constexpr int NP3 = NP * NP * NP;
  constexpr int NC = (NP3 + NT - 1) / NT;     // nodes per thread
  constexpr bool CURV = (GEOM != 0);
  __shared__ double s_u[4][NP3];        // own pi, cx, cy, cz
  __shared__ double s_f[4][6][NP2];     // adjacent values (ajsurf) on the 6 faces
  __shared__ double s_D[NP2];
  __shared__ double s_xb[NP];

  const int tid = threadIdx.x;
  for (int q = tid; q < NP2; q += NT) s_D[q] = a.D[q];
  for (int q = tid; q < NP; q += NT) s_xb[q] = a.xb[q];

  // elements of this block: contiguous element ranges per XCD when xcd_map
  // (workgroups are dispatched round-robin over the 8 XCDs), so neighbour
  // faces are mostly read from the same XCD's L2
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
    __syncthreads();   // LDS of the previous element is no longer read

    // ---- phase 1: every global load this element needs before its
    // derivatives is issued here, then parked in LDS ----
    double uown[NC][4];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int node = tid + c * NT;
      if (node < NP3) {
#pragma unroll
        for (int v = 0; v < 4; ++v) uown[c][v] = a.win[ebase + node + v * NP3];
      }
    }
    // neighbour face planes: s_f[v][f][a + NP*b] for my face point (a,b)
    for (int idx = tid; idx < 6 * NP2; idx += NT) {
      const int f = idx / NP2, q = idx % NP2;
      const int fa = q % NP, fb = q / NP;
      const int4 fd = a.face[6 * e + f];
      if (fd.x >= 0) {
        const double *nw = a.win + 5L * NP3 * fd.x + fd.y + fa * fd.z + fb * fd.w;
#pragma unroll
        for (int v = 0; v < 4; ++v) s_f[v][f][q] = nw[v * NP3];
      } else if (a.profile) {
        const int dir = f >> 1, pl = (f & 1) * (NP - 1);
        const int ii = (dir == 0) ? pl : fa;
        const int jj = (dir == 1) ? pl : ((dir == 0) ? fa : fb);
        const int kk = (dir == 2) ? pl : fb;
        const int mynode = ii + NP * (jj + NP * kk);
        const double *xp = a.xyz + 3L * NP3 * e + mynode;
        double ua[4];
        profile(a, xp[0], xp[NP3], xp[2 * NP3], a.tdev[a.tidx], ua);
#pragma unroll
        for (int v = 0; v < 4; ++v) s_f[v][f][q] = ua[v];
      } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) s_f[v][f][q] = 0.0;
      }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int node = tid + c * NT;
      if (node < NP3) {
#pragma unroll
        for (int v = 0; v < 4; ++v) s_u[v][node] = uown[c][v];
      }
    }
    const double dXbdX[3] = {a.dXbdX[3 * e], a.dXbdX[3 * e + 1], a.dXbdX[3 * e + 2]};
    __syncthreads();

    // ---- phase 2: compute from LDS and registers, node by node ----
#pragma unroll 1
    for (int c = 0; c < NC; ++c) {
      const int node = tid + c * NT;
      if (node >= NP3) continue;
      const int k = node / NP2;
      const int j = (node / NP) % NP;
      const int i = node % NP;
      const long gi = ebase + node;

      double M[3][3];
      if (GEOM == 1) {
        const double *g = a.dXdx + 9L * NP3 * e + node;
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
          for (int q = 0; q < 3; ++q) M[m][q] = g[(3 * m + q) * NP3];
      } else if (GEOM == 2) {
        const int I[3] = {i, j, k};
        double X[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const double lo = a.egeo[8 * e + 2 * d], hi = a.egeo[8 * e + 2 * d + 1];
          X[d] = 0.5 * ((hi - lo) * s_xb[I[d]] + (hi + lo));
        }
        shell_dXdx(a.edom[e], a.egeo[8 * e + 6], a.egeo[8 * e + 7], X, M);
      }

      // tensor-product derivative along one direction m at a time, folded at
      // once into the Cartesian partials: d_q f = sum_m dXdx[m][q] dXbdX[m] D_m f
      // (r_pi = d_q c_q, r_cq = d_q pi)
      double rpi = 0.0, rc[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        const int im = (m == 0) ? i : (m == 1) ? j : k;
        const int stride = (m == 0) ? 1 : (m == 1) ? NP : NP2;
        const int base = node - im * stride;
        double dp = 0.0, d0 = 0.0, d1 = 0.0, d2 = 0.0;
#pragma unroll
        for (int l = 0; l < NP; ++l) {
          const double Dl = s_D[im * NP + l];
          const int src = base + l * stride;
          dp += Dl * s_u[0][src];
          if (CURV) {
            d0 += Dl * s_u[1][src];
            d1 += Dl * s_u[2][src];
            d2 += Dl * s_u[3][src];
          } else {
            d0 += Dl * s_u[1 + m][src];
          }
        }
        if (CURV) {
          const double sp = dXbdX[m] * dp;
          rpi += M[m][0] * (dXbdX[m] * d0) + M[m][1] * (dXbdX[m] * d1) + M[m][2] * (dXbdX[m] * d2);
          rc[0] += M[m][0] * sp;
          rc[1] += M[m][1] * sp;
          rc[2] += M[m][2] * sp;
        } else {
          rpi += dXbdX[m] * d0;
          rc[m] = dXbdX[m] * dp;
        }
      }

      // numerical flux and lift on the faces this node lies on
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
          const int q = ia + NP * ib;
          const double dcv[3] = {s_u[1][node] - s_f[1][f][q], s_u[2][node] - s_f[2][f][q],
                                 s_u[3][node] - s_f[3][f][q]};
          double F[4];
          face_flux(a.numflux, nrm, s_u[0][node] - s_f[0][f][q], dcv, F);
          rpi -= F[0] * scale;
          rc[0] -= F[1] * scale;
          rc[1] -= F[2] * scale;
          rc[2] -= F[3] * scale;
        }
      }

      const double r[5] = {rpi, rc[0], rc[1], rc[2], s_u[0][node]};
      double Ubuf[5], Abuf[5];
      if (a.wmode) {
#pragma unroll
        for (int v = 0; v < 5; ++v) Ubuf[v] = a.U[gi + v * NP3];
      }
      if (a.Aout) {
        const double *bas = a.Ain ? a.Ain : a.win;
#pragma unroll
        for (int v = 0; v < 5; ++v) Abuf[v] = bas[gi + v * NP3];
      }
      if (a.R) {
#pragma unroll
        for (int v = 0; v < 5; ++v) a.R[gi + v * NP3] = r[v];
      }
      if (a.Aout) {
#pragma unroll
        for (int v = 0; v < 5; ++v) a.Aout[gi + v * NP3] = Abuf[v] + a.cA * r[v];
      }
      if (a.wmode == 1) {
#pragma unroll
        for (int v = 0; v < 5; ++v) a.Wout[gi + v * NP3] = Ubuf[v] + a.cW * r[v];
      } else if (a.wmode == 2) {
        // sspRK3: win's own values; phi is not staged in LDS
        const double wv[5] = {s_u[0][node], s_u[1][node], s_u[2][node], s_u[3][node],
                              a.win[gi + 4 * NP3]};
#pragma unroll
        for (int v = 0; v < 5; ++v) a.Wout[gi + v * NP3] = 0.75 * Ubuf[v] + 0.25 * (wv[v] + a.cW * r[v]);
      }
    }
  }
}

template <int NP, int GEOM, int NT>
static hipError_t launch_t(const StageArgs &a, int grid, hipStream_t s) {
  hipLaunchKernelGGL((dg_stage_kernel<NP, GEOM, NT>), dim3(grid), dim3(NT), 0, s, a);
  return hipGetLastError();
}

// threads per block: "full" one node per thread, "half" and "quarter" two
// and four nodes per thread (more, smaller blocks per CU)
static constexpr int tile_threads(int np3, int div) {
  return ((np3 + div - 1) / div + 63) / 64 * 64;
}

template <int NP, int GEOM>
static hipError_t launch_np(const StageArgs &a, int grid, int nt, hipStream_t s) {
  constexpr int NP3 = NP * NP * NP;
  constexpr int T1 = tile_threads(NP3, 1) > 1024 ? 1024 : tile_threads(NP3, 1);
  constexpr int T2 = tile_threads(NP3, 2);
  constexpr int T4 = tile_threads(NP3, 4);
  if (nt == T4) return launch_t<NP, GEOM, T4>(a, grid, s);
  if (nt == T2) return launch_t<NP, GEOM, T2>(a, grid, s);
  return launch_t<NP, GEOM, T1>(a, grid, s);
}

int dg_stage_threads(int np, int div) {
  int np3 = np * np * np;
  if (div != 2 && div != 4) div = 1;
  int t = tile_threads(np3, div);
  return t > 1024 ? 1024 : t;
}

hipError_t launch_dg_stage(int np, int geom, const StageArgs &a, int grid, int nt, hipStream_t s) {
#define NMX_CASE(N)                                                  \
  case N:                                                            \
    if (geom == 0) return launch_np<N, 0>(a, grid, nt, s);           \
    if (geom == 1) return launch_np<N, 1>(a, grid, nt, s);           \
    return launch_np<N, 2>(a, grid, nt, s);
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
