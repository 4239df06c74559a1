// adj_amr_kernels.hip -- discrete adjoint of the batched (mixed-order,
// h-refined) DG right-hand side, fused with the adjoint RK stage
// combination (gfx950).
//
// The forward batched RHS (face_interp_kernel + dg_amr_kernel) reads, at each
// point p of face f of element E, the adjacent value u_adj(p): a copy of the
// partner's node on a conforming face, or the 2D tensor Lagrange
// interpolation R1 P R2^T of a neighbour's face plane P, averaged where
// several neighbours hold the point (surface.c:1081-1337,
// interpolate.c:296-363).  With G(p) = s_f(p) A(n_f(p)) lam_E(p) (A the
// symmetric flux Jacobian), the transpose y = L^T lam is
//   y_E(p) = (volume transpose)(p) - G(p) + sum over the points p' that read
//            p of the weight with which p' reads p, times G(p'),
// and the last sum is, per neighbour face plane:
//   copy faces: the partner's G at the same point;
//   interpolated faces: T = R1^T (w G) R2 of every (receiving face,
//   neighbour) job that reads the plane, w = 1 / count where several
//   neighbours hold a point, summed in a fixed order.
// Three launches per stage: adj_gflux_kernel (G at every face point),
// adj_jobs_kernel (T per job) and one dg_adjoint_amr_kernel per order.
// Every sum runs in a fixed order, so the result is deterministic.
#include <hip/hip_runtime.h>

#include "amr.h"
#include "dg_common.h"

namespace nmx {

// dX/dx of element e (first node noff, n points per direction) at its node
template <int GEOM>
__device__ __forceinline__ void amr_metric(const AdjAmrArgs &a, int e, long noff, int n, int node,
                                           const double *xb, double M[3][3]) {
  const int np3 = n * n * n;
  if (GEOM == 1) {
    const double *p = a.dXdx + 9 * noff + node;
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int q = 0; q < 3; ++q) M[m][q] = p[(3 * m + q) * np3];
  } else if (GEOM == 2) {
    const int i = node % n, j = (node / n) % n, k = node / (n * n);
    const double *bb = a.egeo + 8 * e;
    double X[3];
    X[0] = 0.5 * ((bb[1] - bb[0]) * xb[i] + (bb[1] + bb[0]));
    X[1] = 0.5 * ((bb[3] - bb[2]) * xb[j] + (bb[3] + bb[2]));
    X[2] = 0.5 * ((bb[5] - bb[4]) * xb[k] + (bb[5] + bb[4]));
    shell_dXdx(a.edom[e], bb[6], bb[7], X, M);
  }
}

// node of face f at face point (fa, fb): fa along Dir1_norm, fb along Dir2_norm
__device__ __forceinline__ int face_node(int n, int f, int fa, int fb) {
  const int dir = f >> 1, pl = (f & 1) * (n - 1);
  const int i = (dir == 0) ? pl : fa;
  const int j = (dir == 1) ? pl : ((dir == 0) ? fa : fb);
  const int k = (dir == 2) ? pl : fb;
  return i + n * (j + n * k);
}

// ---------------------------------------------------------------------------
// G = s_f A(n_f) lam at every point of the six faces of the listed elements
// ---------------------------------------------------------------------------
constexpr int kAdjGThreads = 128;

template <int GEOM>
__global__ __launch_bounds__(kAdjGThreads) void adj_gflux_kernel(AdjAmrArgs a) {
  for (int it = blockIdx.x; it < a.nlist; it += gridDim.x) {
    const int e = a.list[it];
    const int n = a.en[e], n2 = n * n, np3 = n2 * n;
    const long noff = a.eoff[e];
    const double *lam = a.lam + 5 * noff;
    const double *xb = a.xball + batch_xoff(n);
    const double wi = a.wend_inv[n];
    double *Gp = a.G + a.goff[e];
    const double g[3] = {a.dXbdX[3 * e], a.dXbdX[3 * e + 1], a.dXbdX[3 * e + 2]};
    for (int idx = threadIdx.x; idx < 6 * n2; idx += kAdjGThreads) {
      const int f = idx / n2, q = idx % n2;
      const int node = face_node(n, f, q % n, q / n);
      double M[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      if (GEOM != 0) amr_metric<GEOM>(a, e, noff, n, node, xb, M);
      double nrm[3], scale;
      face_normal<GEOM>(M, g, f, wi, nrm, scale);
      const double lc[3] = {lam[node + np3], lam[node + 2 * np3], lam[node + 3 * np3]};
      double F[4];
      face_flux(a.numflux, nrm, lam[node], lc, F);
#pragma unroll
      for (int v = 0; v < 4; ++v) Gp[(4 * f + v) * n2 + q] = scale * F[v];
    }
  }
}

// ---------------------------------------------------------------------------
// transposed interpolation, one wave per job:
//   W(x1, x2) = w(a, b) G(a, b) at the job's covered points,
//   S[x1][q2] = sum_x2 W(x1, x2) R2[x2][q2],
//   T(q1, q2) = sum_x1 R1[x1][q1] S[x1][q2]
// with (x1, x2) = (a, b), or (b, a) when the neighbour's first in-plane axis
// runs along the receiver's b: the transpose of face_interp_kernel's passes.
// ---------------------------------------------------------------------------
constexpr int kAdjJThreads = 64;

// One job with the receiver's order N known at compile time.  The covered
// ranges x1 < nx1, x2 < nx2 are padded to N with zero rows of W, R1, R2 and
// S, so the sums over x1 and x2 unroll without branches (their LDS reads are
// issued together) and add only exact zeros beyond the ranges.
template <int N>
__device__ __forceinline__ void adj_job(const AdjAmrArgs &a, const AdjJob &J, int tid, double *s_W, double *s_S,
                                        double *s_R1, double *s_R2) {
  const int m = J.m, n2 = N * N;
  const int x1lo = J.sw ? J.b0 : J.a0, x2lo = J.sw ? J.a0 : J.b0;
  const int nx1 = (J.sw ? J.b1 - J.b0 : J.a1 - J.a0) + 1;
  const int nx2 = (J.sw ? J.a1 - J.a0 : J.b1 - J.b0) + 1;
  const double *R1 = a.Btab + (J.sw ? J.offB : J.offA) + x1lo * m;
  const double *R2 = a.Btab + (J.sw ? J.offA : J.offB) + x2lo * m;
  const double *Gp = a.G + a.goff[J.e] + 4 * J.g * n2;
  __syncthreads();   // the previous job is done with LDS
  for (int idx = tid; idx < N * m; idx += kAdjJThreads) {
    s_R1[idx] = idx < nx1 * m ? R1[idx] : 0.0;
    s_R2[idx] = idx < nx2 * m ? R2[idx] : 0.0;
  }
  for (int idx = tid; idx < 4 * N * N; idx += kAdjJThreads) {
    const int v = idx / n2, r = idx % n2;
    const int x1 = r / N, x2 = r % N;
    double wg = 0.0;
    if (x1 < nx1 && x2 < nx2) {
      const int fa = J.sw ? x2 + x2lo : x1 + x1lo;
      const int fb = J.sw ? x1 + x1lo : x2 + x2lo;
      double w = 1.0;
      if (J.md >= 0) {   // several neighbours: the averaging weight of this point
        const int *md = a.mdesc + J.md;
        int cnt = 0;
        for (int si = 0; si < md[0]; ++si) {
          const int *d = md + 1 + 5 * si;
          if (fa >= d[1] && fa <= d[2] && fb >= d[3] && fb <= d[4]) ++cnt;
        }
        if (cnt > 1) w = 1.0 / cnt;
      }
      wg = w * Gp[v * n2 + fa + N * fb];
    }
    s_W[idx] = wg;
  }
  __syncthreads();
  for (int idx = tid; idx < 4 * N * m; idx += kAdjJThreads) {
    const int v = idx / (N * m), r = idx % (N * m);
    const int x1 = r / m, q2 = r % m;
    const double *W = s_W + (v * N + x1) * N;
    double t = 0.0;
#pragma unroll
    for (int x2 = 0; x2 < N; ++x2) t += W[x2] * s_R2[x2 * m + q2];
    s_S[idx] = t;
  }
  __syncthreads();
  double *T = a.T + J.toff;
  for (int idx = tid; idx < 4 * m * m; idx += kAdjJThreads) {
    const int v = idx / (m * m), r = idx % (m * m);
    const int q1 = r % m, q2 = r / m;
    double t = 0.0;
#pragma unroll
    for (int x1 = 0; x1 < N; ++x1) t += s_R1[x1 * m + q1] * s_S[(v * N + x1) * m + q2];
    T[idx] = t;
  }
}

__global__ __launch_bounds__(kAdjJThreads) void adj_jobs_kernel(AdjAmrArgs a) {
  __shared__ double s_W[4 * kMaxFaceN * kMaxFaceN];
  __shared__ double s_S[4 * kMaxFaceN * kMaxFaceN];
  __shared__ double s_R1[kMaxFaceN * kMaxFaceN];
  __shared__ double s_R2[kMaxFaceN * kMaxFaceN];
  const int tid = threadIdx.x;
  for (int ji = blockIdx.x; ji < a.njobs; ji += gridDim.x) {
    const AdjJob J = a.jobs[ji];
    switch (J.n) {
#define NMX_CASE(N_)                                    \
  case N_:                                              \
    adj_job<N_>(a, J, tid, s_W, s_S, s_R1, s_R2);       \
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

// ---------------------------------------------------------------------------
// the transposed RHS of one order bucket, fused with the adjoint stage
// combination (dg_adjoint_kernel's phases on the bucket's elements)
// ---------------------------------------------------------------------------
template <int NP, int GEOM, int NT>
__global__ __launch_bounds__(NT, 2) void dg_adjoint_amr_kernel(AdjAmrArgs a) {
  constexpr int NP2 = NP * NP;
  constexpr int NP3 = NP * NP * NP;
  constexpr int NC = (NP3 + NT - 1) / NT;
  constexpr bool CURV = (GEOM != 0);
  __shared__ double s_t[4][NP3];       // staged fields of one direction
  __shared__ double s_f[4][6][NP2];    // incoming adjoint face terms
  __shared__ double s_D[NP2];

  const int tid = threadIdx.x;
  const double *Dn = a.Dall + batch_doff(NP);
  const double *xb = a.xball + batch_xoff(NP);
  const double wi = a.wend_inv[NP];
  for (int q = tid; q < NP2; q += NT) s_D[q] = Dn[q];

  for (int it = blockIdx.x; it < a.nlist; it += gridDim.x) {
    const int e = a.list[it];
    const long noff = a.eoff[e];
    const long ebase = 5 * noff;
    const double g[3] = {a.dXbdX[3 * e], a.dXbdX[3 * e + 1], a.dXbdX[3 * e + 2]};
    __syncthreads();

    // ---- phase 1: own lam to registers, incoming face terms to LDS ----
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
      const int fa = q % NP, fb = q / NP;
      const AdjFace F = a.af[6 * e + f];
      double in[4] = {0.0, 0.0, 0.0, 0.0};
      if (F.kind == 1) {          // the partner's G at the same point
        const int p = F.base + fa * F.s1 + fb * F.s2;
        const int pi = p % NP, pj = (p / NP) % NP, pk = p / NP2;
        const int gd = F.g >> 1;
        const int pa = (gd == 0) ? pj : pi, pb = (gd == 2) ? pj : pk;
        const double *Gp = a.G + a.goff[F.nb] + 4 * F.g * NP2 + pa + NP * pb;
#pragma unroll
        for (int v = 0; v < 4; ++v) in[v] = Gp[v * NP2];
      } else if (F.kind == 2) {   // the transposed jobs that read this plane
        for (int jj = 0; jj < F.jcount; ++jj) {
          const AdjJob &J = a.jobs[a.sjobs[F.jfirst + jj]];
          const int q1 = J.A1 * (fa - J.A0) + J.B1 * (fb - J.B0);
          const int q2 = J.A2 * (fa - J.A0) + J.B2 * (fb - J.B0);
          const double *Tp = a.T + J.toff + q1 + NP * q2;
#pragma unroll
          for (int v = 0; v < 4; ++v) in[v] += Tp[v * NP2];
        }
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) s_f[v][f][q] = in[v];
    }

    // own metric per node (registers)
    double M[NC][3][3];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int node = tid + c * NT;
      if (CURV && node < NP3) amr_metric<GEOM>(a, e, noff, NP, node, xb, M[c]);
    }

    // ---- phase 2: transposed volume term, direction by direction ----
    double y[NC][4];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int v = 0; v < 4; ++v) y[c][v] = 0.0;

#pragma unroll
    for (int m = 0; m < 3; ++m) {
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

    // ---- own and incoming face terms, the stage combination ----
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
          double Mz[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
          if (CURV) {
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
              for (int q = 0; q < 3; ++q) Mz[r][q] = M[c][r][q];
          }
          face_normal<GEOM>(Mz, g, f, wi, nrm, scale);
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

hipError_t launch_adj_gflux(int geom, const AdjAmrArgs &a, hipStream_t s) {
  if (a.nlist <= 0) return hipSuccess;
  const dim3 grid(a.nlist < 16384 ? a.nlist : 16384);
  if (geom == 0) hipLaunchKernelGGL(adj_gflux_kernel<0>, grid, dim3(kAdjGThreads), 0, s, a);
  else if (geom == 1) hipLaunchKernelGGL(adj_gflux_kernel<1>, grid, dim3(kAdjGThreads), 0, s, a);
  else hipLaunchKernelGGL(adj_gflux_kernel<2>, grid, dim3(kAdjGThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_adj_jobs(const AdjAmrArgs &a, hipStream_t s) {
  if (a.njobs <= 0) return hipSuccess;
  const dim3 grid(a.njobs < 65536 ? a.njobs : 65536);
  hipLaunchKernelGGL(adj_jobs_kernel, grid, dim3(kAdjJThreads), 0, s, a);
  return hipGetLastError();
}

template <int NP, int GEOM>
static hipError_t launch_adj_amr_np(const AdjAmrArgs &a, int grid, hipStream_t s) {
  constexpr int NP3 = NP * NP * NP;
  constexpr int NT = ((NP3 + 1) / 2 + 63) / 64 * 64;   // two nodes per thread
  hipLaunchKernelGGL((dg_adjoint_amr_kernel<NP, GEOM, NT>), dim3(grid), dim3(NT), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_dg_adjoint_amr(int np, int geom, const AdjAmrArgs &a, int grid, hipStream_t s) {
  if (a.nlist <= 0) return hipSuccess;
#define NMX_CASE(N)                                              \
  case N:                                                        \
    if (geom == 0) return launch_adj_amr_np<N, 0>(a, grid, s);   \
    if (geom == 1) return launch_adj_amr_np<N, 1>(a, grid, s);   \
    return launch_adj_amr_np<N, 2>(a, grid, s);
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
