// amr_kernels.hip -- batched DG RHS + RK stage for meshes with mixed
// refinement levels and mixed element orders.
//
// Elements are grouped into order buckets; one launch per bucket runs the
// kernel instantiated for that order over the bucket's element list (all
// buckets of a stage are captured in one hipGraph per RK step).  Per element:
//   phase 1: own fields and conforming neighbour faces (aliased or oriented
//            copies, surface.c:825-1072) go to LDS; for non-conforming or
//            mixed-order faces the neighbours' face planes are staged in LDS;
//   phase 1b: those faces get ajsurf by 2D tensor Lagrange interpolation of
//            each neighbour's plane at this element's face points, averaged
//            where several neighbours hold a point (surface.c:1081-1337,
//            interpolate.c:296-363); the 1D Lagrange rows come from a small
//            deduplicated table;
//   phase 2: the per-node RHS and stage update shared with the plain kernel.
#include <hip/hip_runtime.h>

#include "amr.h"
#include "dg_common.h"

namespace nmx {

template <int NP, int GEOM, int NT>
__global__ __launch_bounds__(NT, NMX_DG_WAVES) void dg_amr_kernel(AmrArgs a) {
  constexpr int NP2 = NP * NP;
  constexpr int NP3 = NP * NP * NP;
  constexpr int NC = (NP3 + NT - 1) / NT;
  __shared__ double s_u[4][NP3];
  __shared__ double s_f[4][6][NP2];
  __shared__ double s_stage[kAmrStage];
  __shared__ double s_D[NP2];
  __shared__ double s_xb[NP];

  const int tid = threadIdx.x;
  for (int q = tid; q < NP2; q += NT) s_D[q] = a.D[q];
  for (int q = tid; q < NP; q += NT) s_xb[q] = a.xb[q];

  int i_hi, i_step, i_first;
  if (a.xcd_map) {
    const int xcd = blockIdx.x & 7, q = blockIdx.x >> 3, g8 = gridDim.x >> 3;
    const int chunk = (a.nlist + 7) >> 3;
    const int lo = xcd * chunk;
    i_hi = min(lo + chunk, a.nlist);
    i_first = lo + q;
    i_step = g8;
  } else {
    i_hi = a.nlist;
    i_first = blockIdx.x;
    i_step = gridDim.x;
  }

  for (int it = i_first; it < i_hi; it += i_step) {
    const int e = a.elist[it];
    const long gbase = a.soff[e];
    const int *fdesc = a.face + 48 * e;
    const int nrounds = a.erounds[e];
    __syncthreads();

    // ---- phase 1 ----
    double uown[NC][4];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int node = tid + c * NT;
      if (node < NP3) {
#pragma unroll
        for (int v = 0; v < 4; ++v) uown[c][v] = a.win[gbase + node + v * NP3];
      }
    }
    for (int f = 0; f < 6; ++f) {
      const int kind = fdesc[8 * f];
      if (kind == 1) {          // conforming copy from a same-order neighbour
        const int nb = fdesc[8 * f + 1], base = fdesc[8 * f + 2];
        const int s1 = fdesc[8 * f + 3], s2 = fdesc[8 * f + 4];
        const double *nw = a.win + a.soff[nb] + base;
        for (int q = tid; q < NP2; q += NT) {
          const int fa = q % NP, fb = q / NP;
          const double *p = nw + fa * s1 + fb * s2;
#pragma unroll
          for (int v = 0; v < 4; ++v) s_f[v][f][q] = p[v * NP3];
        }
      } else if (kind == 0) {   // outer boundary: scalarwave1's exterior profile
        for (int q = tid; q < NP2; q += NT) {
          double ua[4] = {0.0, 0.0, 0.0, 0.0};
          if (a.profile) {
            const int fa = q % NP, fb = q / NP;
            const int dir = f >> 1, pl = (f & 1) * (NP - 1);
            const int ii = (dir == 0) ? pl : fa;
            const int jj = (dir == 1) ? pl : ((dir == 0) ? fa : fb);
            const int kk = (dir == 2) ? pl : fb;
            const int mynode = ii + NP * (jj + NP * kk);
            const double *xp = a.xyz + 3 * (gbase / 5) + mynode;
            plane_wave(a.k, a.om, xp[0], xp[NP3], xp[2 * NP3], a.tdev[a.tidx], ua);
          }
#pragma unroll
          for (int v = 0; v < 4; ++v) s_f[v][f][q] = ua[v];
        }
      }
    }
    // stage the neighbour planes of round 0
    auto stage_round = [&](int r) {
      int off = 0;
      for (int f = 0; f < 6; ++f) {
        if (fdesc[8 * f] != 2 || fdesc[8 * f + 7] != r) continue;
        const int ns = fdesc[8 * f + 5], s0 = fdesc[8 * f + 6];
        for (int si = 0; si < ns; ++si) {
          const int *sl = a.slot + 12 * (s0 + si);
          const int nb = sl[0], m = sl[1], base = sl[2], t1 = sl[3], t2 = sl[4];
          const int m2 = m * m, m3 = m2 * m;
          const double *nw = a.win + a.soff[nb] + base;
          for (int idx = tid; idx < 4 * m2; idx += NT) {
            const int v = idx / m2, q = idx % m2;
            s_stage[off + idx] = nw[v * m3 + (q % m) * t1 + (q / m) * t2];
          }
          off += 4 * m2;
        }
      }
    };
    if (nrounds > 0) stage_round(0);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int node = tid + c * NT;
      if (node < NP3) {
#pragma unroll
        for (int v = 0; v < 4; ++v) s_u[v][node] = uown[c][v];
      }
    }
    __syncthreads();

    // ---- phase 1b: interpolated faces, round by round ----
    for (int r = 0; r < nrounds; ++r) {
      if (r > 0) {
        __syncthreads();
        stage_round(r);
        __syncthreads();
      }
      int off = 0;
      for (int f = 0; f < 6; ++f) {
        if (fdesc[8 * f] != 2 || fdesc[8 * f + 7] != r) continue;
        const int ns = fdesc[8 * f + 5], s0 = fdesc[8 * f + 6];
        for (int idx = tid; idx < 4 * NP2; idx += NT) {
          const int v = idx / NP2, q = idx % NP2;
          const int fa = q % NP, fb = q / NP;
          double sum = 0.0;
          int cnt = 0, soff = off;
          for (int si = 0; si < ns; ++si) {
            const int *sl = a.slot + 12 * (s0 + si);
            const int m = sl[1], m2 = m * m;
            if (fa >= sl[6] && fa <= sl[7] && fb >= sl[8] && fb <= sl[9]) {
              const double *P = s_stage + soff + v * m2;
              const double *ra = a.Btab + sl[10] + fa * m;
              const double *rb = a.Btab + sl[11] + fb * m;
              const double *B1 = sl[5] ? rb : ra;
              const double *B2 = sl[5] ? ra : rb;
              double sv = 0.0;
              for (int q2 = 0; q2 < m; ++q2) {
                const double b2 = B2[q2];
                for (int q1 = 0; q1 < m; ++q1) sv += P[q1 + m * q2] * B1[q1] * b2;
              }
              sum += sv;
              ++cnt;
            }
            soff += 4 * m2;
          }
          if (cnt > 1) sum /= cnt;
          s_f[v][f][q] = sum;
        }
        for (int si = 0; si < ns; ++si) off += 4 * a.slot[12 * (s0 + si) + 1] * a.slot[12 * (s0 + si) + 1];
      }
    }
    if (nrounds > 0) __syncthreads();

    // ---- phase 2 ----
    NodeOut o;
    o.win = a.win; o.U = a.U; o.Ain = a.Ain; o.Aout = a.Aout; o.Wout = a.Wout; o.R = a.R;
    o.cA = a.cA; o.cW = a.cW; o.wend_inv = a.wend_inv; o.wmode = a.wmode; o.numflux = a.numflux;
    ElmGeo g;
    g.dXbdX[0] = a.dXbdX[3 * e]; g.dXbdX[1] = a.dXbdX[3 * e + 1]; g.dXbdX[2] = a.dXbdX[3 * e + 2];
    g.dXdx = a.dXdx + 9 * (gbase / 5);
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
      if (node < NP3) node_update<NP, GEOM>(o, g, s_u, s_f, s_D, s_xb, gbase, node);
    }
  }
}

template <int NP, int GEOM>
static hipError_t launch_amr_np(const AmrArgs &a, int grid, hipStream_t s) {
  constexpr int NP3 = NP * NP * NP;
  constexpr int NT = NP3 <= 64 ? 64 : (NP3 <= 128 ? 128 : 256);
  hipLaunchKernelGGL((dg_amr_kernel<NP, GEOM, NT>), dim3(grid), dim3(NT), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_dg_amr(int np, int geom, const AmrArgs &a, int grid, hipStream_t s) {
#define NMX_CASE(N)                                              \
  case N:                                                        \
    if (geom == 0) return launch_amr_np<N, 0>(a, grid, s);       \
    if (geom == 1) return launch_amr_np<N, 1>(a, grid, s);       \
    return launch_amr_np<N, 2>(a, grid, s);
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
