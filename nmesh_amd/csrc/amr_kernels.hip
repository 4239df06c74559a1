// amr_kernels.hip -- batched DG RHS + RK stage for meshes with mixed
// refinement levels and mixed element orders.
//
// Each RK stage, per order bucket (the buckets run concurrently on stream
// lanes):
//   1. face_interp_kernel: one wave per (interpolated face, neighbour)
//      computes ajsurf (the neighbour's values at this face's points) by 2D
//      tensor Lagrange interpolation of its face plane (surface.c:1081-1337,
//      interpolate.c:296-363) into a compact face buffer;
//   2. dg_amr_kernel: own fields to LDS, conforming faces copied from the
//      neighbour (surface.c:825-1072), interpolated faces read from the face
//      buffer (averaged where several neighbours hold a point), then the
//      per-node RHS and stage update shared with the plain kernel.
#include <hip/hip_runtime.h>

#include "amr.h"
#include "dg_common.h"

namespace nmx {

// ---------------------------------------------------------------------------
// face interpolation, one wave per (face, neighbour) job:
//   pass 1: T[v][x1][q2] = sum_q1 P[v](q1, q2) R1[x1][q1]; a lane owns one
//           (v, q2) line of the neighbour plane and keeps the x1 sums in
//           registers;
//   pass 2: V[v][a][b] = sum_q2 T[v][x1][q2] R2[x2][q2], (x1, x2) = (a, b),
//           or (b, a) when the neighbour's first in-plane axis runs along my
//           b, for the points this neighbour covers (0 elsewhere).
// The plane lines and both passes' rows are requested together, so a job
// waits on one round trip to memory.  A face with several neighbours has
// one region per neighbour, averaged by the stage kernel.
// ---------------------------------------------------------------------------
constexpr int kFaceThreads = 64;

struct JobView {
  int n, m, ajoff, nslot, base, t1, t2, sw, a0, a1, b0, b1, offA, offB;
  long soff;
  __device__ JobView(const NMX_CONST int *j) {
    n = j[0]; m = j[1]; ajoff = j[2]; nslot = j[3]; base = j[4]; t1 = j[5]; t2 = j[6];
    sw = j[7]; a0 = j[8]; a1 = j[9]; b0 = j[10]; b1 = j[11]; offA = j[12]; offB = j[13];
    soff = (long)(unsigned)j[14] | ((long)j[15] << 32);
  }
  __device__ int x1lo() const { return sw ? b0 : a0; }
  __device__ int nx1() const { return (sw ? b1 - b0 : a1 - a0) + 1; }
  __device__ int x2lo() const { return sw ? a0 : b0; }
  __device__ int nx2() const { return (sw ? a1 - a0 : b1 - b0) + 1; }
};

// One job with the neighbour's order M known at compile time: the loops
// over its points are unrolled without branches, so the LDS reads of a row
// are issued together instead of one round trip per term.
template <int M>
__device__ __forceinline__ void face_job(const FaceArgs &a, const JobView &J, int tid, double *s_T, double *s_R1,
                                         double *s_R2) {
  const int n = J.n, nx1 = J.nx1(), nx2 = J.nx2();
  const int n1 = nx1 * M, n2 = nx2 * M;
  // every load of this face: rows (at most 100 doubles each, two per lane)
  // and my (v, q2) line of the plane
  const double *R1 = a.Btab + (J.sw ? J.offB : J.offA) + J.x1lo() * M;
  const double *R2 = a.Btab + (J.sw ? J.offA : J.offB) + J.x2lo() * M;
  const double r1a = tid < n1 ? R1[tid] : 0.0, r1b = tid + 64 < n1 ? R1[tid + 64] : 0.0;
  const double r2a = tid < n2 ? R2[tid] : 0.0, r2b = tid + 64 < n2 ? R2[tid + 64] : 0.0;
  const bool act = tid < 4 * M;
  const int v = act ? tid / M : 0, q2 = act ? tid % M : 0;
  const double *P = a.win + J.soff + J.base + (long)v * M * M * M + q2 * J.t2;
  double p[M];
#pragma unroll
  for (int q1 = 0; q1 < M; ++q1) p[q1] = act ? P[q1 * J.t1] : 0.0;
  __syncthreads();   // the previous face is done with LDS
  if (tid < n1) s_R1[tid] = r1a;
  if (tid + 64 < n1) s_R1[tid + 64] = r1b;
  if (tid < n2) s_R2[tid] = r2a;
  if (tid + 64 < n2) s_R2[tid + 64] = r2b;
  __syncthreads();
  if (act) {
    double *T = s_T + v * nx1 * M + q2;
    for (int x = 0; x < nx1; ++x) {
      const double *row = s_R1 + x * M;
      double t = 0.0;
#pragma unroll
      for (int q1 = 0; q1 < M; ++q1) t += p[q1] * row[q1];
      T[x * M] = t;
    }
  }
  __syncthreads();
  const int nn = n * n;
  for (int idx = tid; idx < 4 * nn; idx += kFaceThreads) {
    const int vv = idx / nn, q = idx % nn;
    const int fa = q % n, fb = q / n;
    double sum = 0.0;
    if (fa >= J.a0 && fa <= J.a1 && fb >= J.b0 && fb <= J.b1) {
      const int x1 = (J.sw ? fb : fa) - J.x1lo(), x2 = (J.sw ? fa : fb) - J.x2lo();
      const double *T = s_T + (vv * nx1 + x1) * M;
      const double *row = s_R2 + x2 * M;
      double sv = 0.0;
#pragma unroll
      for (int k = 0; k < M; ++k) sv += T[k] * row[k];
      sum += sv;
    }
    a.aj[J.ajoff + idx] = sum;
  }
}

__global__ __launch_bounds__(kFaceThreads) void face_interp_kernel(FaceArgs a) {
  __shared__ double s_T[kFaceT1];    // [v][x1][m]
  __shared__ double s_R1[kFaceR1];   // [x1][m]
  __shared__ double s_R2[kFaceR1];   // [x2][m]
  const int tid = threadIdx.x;
  const int xcd = blockIdx.x & 7, g8 = gridDim.x >> 3;
  const int chunk = (a.nfaces + 7) >> 3;
  const int f_hi = min((xcd + 1) * chunk, a.nfaces);
  for (int fi = xcd * chunk + (blockIdx.x >> 3); fi < f_hi; fi += g8) {
    const JobView J(as_const(a.jobs) + kJobInts * fi);
    switch (J.m) {
#define NMX_CASE(M_)                                   \
  case M_:                                             \
    face_job<M_>(a, J, tid, s_T, s_R1, s_R2);          \
    break;
      NMX_CASE(2)
      NMX_CASE(3)
      NMX_CASE(4)
      NMX_CASE(5)
      NMX_CASE(6)
      NMX_CASE(7)
      NMX_CASE(8)
      NMX_CASE(9)
      NMX_CASE(10)
#undef NMX_CASE
      default:
        break;   // orders outside 2..10 are refused on the host
    }
  }
}

hipError_t launch_face_interp(const FaceArgs &a, hipStream_t s) {
  if (a.nfaces == 0) return hipSuccess;
  const int grid = 8 * ((a.nfaces + 7) / 8);   // a multiple of 8 (XCD mapping)
  hipLaunchKernelGGL(face_interp_kernel, dim3(grid), dim3(kFaceThreads), 0, s, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// stage kernel per order bucket
// ---------------------------------------------------------------------------
template <int NP, int GEOM, int NT>
__global__ __launch_bounds__(NT, NMX_AMR_WAVES) void dg_amr_kernel(AmrArgs a) {
  constexpr int NP2 = NP * NP;
  constexpr int NP3 = NP * NP * NP;
  constexpr int NC = (NP3 + NT - 1) / NT;
  __shared__ double s_u[4][NP3];
  __shared__ double s_f[4][6][NP2];
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
    const NMX_CONST SlotRec &R = *as_const(a.rec + it);
    const long noff = R.noff;
    const long gbase = 5 * noff;
    __syncthreads();

    // ---- phase 1: own fields and all six adjacent face planes ----
    double uown[NC][4];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int node = tid + c * NT;
      if (node < NP3) {
#pragma unroll
        for (int v = 0; v < 4; ++v) uown[c][v] = a.win[gbase + node + v * NP3];
      }
    }
    for (int idx = tid; idx < 6 * NP2; idx += NT) {
      const int f = idx / NP2, q = idx % NP2;
      const int kind = R.fkind[f];
      if (kind == 1) {          // conforming copy from a same-order neighbour
        const double *p = a.win + R.foff[f] + (q % NP) * R.fs1[f] + (q / NP) * R.fs2[f];
#pragma unroll
        for (int v = 0; v < 4; ++v) s_f[v][f][q] = p[v * NP3];
      } else if (kind == 2) {   // interpolated this stage by face_interp_kernel
        const double *p = a.aj + R.foff[f] + q;
#pragma unroll
        for (int v = 0; v < 4; ++v) s_f[v][f][q] = p[v * NP2];
      } else if (kind == 3) {   // several neighbours: averaged where several hold
                                // the point (surface.c:1322-1337)
        const int *md = a.mdesc + R.foff[f];
        const int ns = md[0];
        const int fa = q % NP, fb = q / NP;
        double sum[4] = {0.0, 0.0, 0.0, 0.0};
        int cnt = 0;
        for (int si = 0; si < ns; ++si) {
          const int *d = md + 1 + 5 * si;
          if (fa >= d[1] && fa <= d[2] && fb >= d[3] && fb <= d[4]) {
            const double *p = a.aj + d[0] + q;
#pragma unroll
            for (int v = 0; v < 4; ++v) sum[v] += p[v * NP2];
            ++cnt;
          }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) s_f[v][f][q] = cnt > 1 ? sum[v] / cnt : sum[v];
      } else {                  // outer boundary: scalarwave1's exterior profile
        double ua[4] = {0.0, 0.0, 0.0, 0.0};
        if (a.profile) {
          const int fa = q % NP, fb = q / NP;
          const int dir = f >> 1, pl = (f & 1) * (NP - 1);
          const int ii = (dir == 0) ? pl : fa;
          const int jj = (dir == 1) ? pl : ((dir == 0) ? fa : fb);
          const int kk = (dir == 2) ? pl : fb;
          const int mynode = ii + NP * (jj + NP * kk);
          const double *xp = a.xyz + 3 * noff + mynode;
          plane_wave(a.k, a.om, xp[0], xp[NP3], xp[2 * NP3], a.tdev[a.tidx], ua);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) s_f[v][f][q] = ua[v];
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
    __syncthreads();

    // ---- phase 2 ----
    NodeOut o;
    o.win = a.win; o.U = a.U; o.Ain = a.Ain; o.Aout = a.Aout; o.Wout = a.Wout; o.R = a.R;
    o.cA = a.cA; o.cW = a.cW; o.wend_inv = a.wend_inv; o.wmode = a.wmode; o.numflux = a.numflux; o.nt = a.nt;
    ElmGeo g;
    g.dXbdX[0] = R.dXbdX[0]; g.dXbdX[1] = R.dXbdX[1]; g.dXbdX[2] = R.dXbdX[2];
    g.dXdx = a.dXdx + 9 * noff;
    if (GEOM == 2) {
#pragma unroll
      for (int q = 0; q < 6; ++q) g.bb[q] = R.bb[q];
      g.s0 = R.s0;
      g.s1 = R.s1;
      g.dom = R.dom;
    }
#pragma unroll 1
    for (int c = 0; c < NC; ++c) {
      const int node = tid + c * NT;
      if (node < NP3) node_update<NP, GEOM>(o, g, s_u, s_f, s_D, s_xb, gbase, node);
    }
  }
}

template <int NP, int GEOM, int NT>
static hipError_t launch_amr_t(const AmrArgs &a, int grid, hipStream_t s) {
  hipLaunchKernelGGL((dg_amr_kernel<NP, GEOM, NT>), dim3(grid), dim3(NT), 0, s, a);
  return hipGetLastError();
}

// nodes per thread: div = 1, 2 or 4 (threads rounded to whole waves)
static constexpr int amr_threads(int np3, int div) {
  return (((np3 + div - 1) / div + 63) / 64 * 64) > 1024 ? 1024 : (((np3 + div - 1) / div + 63) / 64 * 64);
}

template <int NP, int GEOM>
static hipError_t launch_amr_np(const AmrArgs &a, int grid, int div, hipStream_t s) {
  constexpr int NP3 = NP * NP * NP;
  if (div == 4) return launch_amr_t<NP, GEOM, amr_threads(NP3, 4)>(a, grid, s);
  if (div == 2) return launch_amr_t<NP, GEOM, amr_threads(NP3, 2)>(a, grid, s);
  return launch_amr_t<NP, GEOM, amr_threads(NP3, 1)>(a, grid, s);
}

hipError_t launch_dg_amr(int np, int geom, const AmrArgs &a, int grid, int div, hipStream_t s) {
#define NMX_CASE(N)                                                   \
  case N:                                                             \
    if (geom == 0) return launch_amr_np<N, 0>(a, grid, div, s);       \
    if (geom == 1) return launch_amr_np<N, 1>(a, grid, div, s);       \
    return launch_amr_np<N, 2>(a, grid, div, s);
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
