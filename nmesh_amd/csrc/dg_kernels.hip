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

#include "dg_common.h"
#include "kernels.h"

namespace nmx {

template <int NP, int GEOM, int NT>
__global__ __launch_bounds__(NT, NMX_DG_WAVES) void dg_stage_kernel(StageArgs a) {
  constexpr int NP2 = NP * NP;
  constexpr int NP3 = NP * NP * NP;
  constexpr int NC = (NP3 + NT - 1) / NT;     // nodes per thread
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
        plane_wave(a.k, a.om, xp[0], xp[NP3], xp[2 * NP3], a.tdev[a.tidx], ua);
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
    NodeOut o;
    o.win = a.win; o.U = a.U; o.Ain = a.Ain; o.Aout = a.Aout; o.Wout = a.Wout; o.R = a.R;
    o.cA = a.cA; o.cW = a.cW; o.wend_inv = a.wend_inv; o.wmode = a.wmode; o.numflux = a.numflux; o.nt = a.nt;
    ElmGeo g;
    g.dXbdX[0] = dXbdX[0]; g.dXbdX[1] = dXbdX[1]; g.dXbdX[2] = dXbdX[2];
    g.dXdx = a.dXdx + 9L * NP3 * e;
    if (GEOM == 2) {
#pragma unroll
      for (int q = 0; q < 6; ++q) g.bb[q] = a.egeo[8 * e + q];
      g.s0 = a.egeo[8 * e + 6];
      g.s1 = a.egeo[8 * e + 7];
      g.dom = a.edom[e];
    }
#pragma unroll 1
    for (int c = 0; c < NC; ++c) {
      const int node = tid + c * NT;
      if (node < NP3) node_update<NP, GEOM>(o, g, s_u, s_f, s_D, s_xb, ebase, node);
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
