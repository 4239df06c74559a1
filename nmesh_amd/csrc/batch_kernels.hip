// batch_kernels.hip -- the batched DG RHS + RK stage for meshes with mixed
// refinement levels and mixed element orders, with every face gathered by
// the receiving element.
//
// This replaces nmesh's per-element VOLRHS/SURFRHS loop over the elements
// (evolve_setrhs_mesh, src/main/evolve/evolve.c:182-232) on the batched path.
// Two launch shapes share the per-element code:
//   - dg_batch_kernel: one persistent launch per RK stage for every DG
//     element.  Workgroup q of XCD x walks its XCD's element list
//     (contiguous in element order, by decreasing n) with stride g8 and runs
//     each order's code, a function of its own so each order keeps its own
//     register allocation;
//   - dg_bucket_kernel<NP>: one order per launch with LDS and threads sized for
//     that order, the buckets running side by side on stream lanes.
// Per element:
//   1. own pi, cx, cy, cz into registers (for the next element of the same
//      order this is issued before the current one's RHS);
//   2. conforming faces copied from the neighbour's plane (surface.c:825-1072),
//      outer faces from scalarwave1's exterior profile;
//   3. interpolated faces (set_ajsurf_vl's interpolation and averaging,
//      surface.c:1081-1337; basis_array_interp2d, interpolate.c:296-363), in
//      rounds of whole faces: every load of a round (the neighbour plane lines
//      and both passes' 1D rows) is issued before its first barrier; a thread
//      per (job, v, q2) line contracts along q1, then per face point the
//      second direction is summed over the face's neighbours in order and
//      divided where more than one holds the point.  The rows and first-pass
//      results live in the own-field buffer, free until step 4;
//   4. own fields to LDS and the per-node RHS + RK stage shared with the plain
//      kernel (node_update).
#include <hip/hip_runtime.h>

#include "amr.h"
#include "dg_common.h"

namespace nmx {

constexpr int kBT = 256;   // threads per workgroup of the persistent launch

struct RJobView {
  long nbase;
  int t1, t2, m, f, sw, a0, a1, b0, b1, fstart, offA, offB, line0, toff, r1off, r2off, nround;
  __device__ explicit RJobView(const NMX_CONST RJob &j) {
    nbase = j.nbase; t1 = j.t1; t2 = j.t2;
    const int pk = j.pk;
    m = pk & 15; f = (pk >> 4) & 7; sw = (pk >> 7) & 1;
    a0 = (pk >> 8) & 15; a1 = (pk >> 12) & 15; b0 = (pk >> 16) & 15; b1 = (pk >> 20) & 15;
    fstart = (pk >> 24) & 1;
    offA = j.offA; offB = j.offB; line0 = j.line0; toff = j.toff; r1off = j.r1off; r2off = j.r2off;
    nround = j.nround;
  }
  __device__ int x1lo() const { return sw ? b0 : a0; }
  __device__ int nx1() const { return (sw ? b1 - b0 : a1 - a0) + 1; }
  __device__ int x2lo() const { return sw ? a0 : b0; }
  __device__ int nx2() const { return (sw ? a1 - a0 : b1 - b0) + 1; }
};

// one round of interpolation jobs [j0, j0 + nr) of the current element:
// results into s_f of the round's faces.  Each of the NT threads takes up to
// LPT lines and RPT row values; the host keeps a round within kBatchLines
// lines and kBatchRows row values.
template <int NP, int NT>
__device__ __forceinline__ void batch_round(const NMX_CONST BatchArgs &a, int j0, int nr, double *s_scr,
                                            double (*s_f)[6][NP * NP]) {
  constexpr int NP2 = NP * NP;
  constexpr int LPT = (kBatchLines + NT - 1) / NT;
  constexpr int RPT = (kBatchRows + NT - 1) / NT;
  const int tid = threadIdx.x;
  const NMX_CONST RJob *J = as_const(a.jobs) + j0;
  const RJobView Vl(J[nr - 1]);
  const int lines = Vl.line0 + 4 * Vl.m;
  // every load of the round: my lines of the neighbour planes ...
  long nbase[LPT];
  int t1[LPT], t2[LPT], m[LPT], nx1[LPT], line0[LPT], toff[LPT], r1off[LPT];
#pragma unroll
  for (int u = 0; u < LPT; ++u) {
    nbase[u] = 0; t1[u] = t2[u] = m[u] = nx1[u] = line0[u] = toff[u] = r1off[u] = 0;
    const int l = tid + u * NT;
    for (int jj = 0; jj < nr; ++jj) {
      const RJobView V(J[jj]);
      if (l >= V.line0) {
        nbase[u] = V.nbase; t1[u] = V.t1; t2[u] = V.t2; m[u] = V.m; nx1[u] = V.nx1();
        line0[u] = V.line0; toff[u] = V.toff; r1off[u] = V.r1off;
      }
    }
  }
  double p[LPT][kMaxFaceN];
#pragma unroll
  for (int u = 0; u < LPT; ++u) {
    const int l = tid + u * NT;
    const bool act = l < lines;
    const int ll = act ? l - line0[u] : 0;
    const int mm = act ? m[u] : 1;
    const int v = ll / mm, q2 = ll % mm;
    const double *P = a.win + nbase[u] + (long)v * m[u] * m[u] * m[u] + (long)q2 * t2[u];
#pragma unroll
    for (int q1 = 0; q1 < kMaxFaceN; ++q1) p[u][q1] = (act && q1 < m[u]) ? P[(long)q1 * t1[u]] : 0.0;
  }
  // ... and the rows of both passes: value k of the round's row list (the
  // jobs in order, each its pass-1 rows then its pass-2 rows)
  double rv[RPT];
  int rdst[RPT];
#pragma unroll
  for (int u = 0; u < RPT; ++u) {
    const int k = tid + u * NT;
    rdst[u] = -1;
    rv[u] = 0.0;
    int kb = 0;
    for (int jj = 0; jj < nr; ++jj) {
      const RJobView V(J[jj]);
      const int n1 = V.nx1() * V.m, n2 = V.nx2() * V.m;
      if (k >= kb && k < kb + n1) {
        rv[u] = a.Btab[(V.sw ? V.offB : V.offA) + V.x1lo() * V.m + (k - kb)];
        rdst[u] = V.r1off + (k - kb);
      } else if (k >= kb + n1 && k < kb + n1 + n2) {
        rv[u] = a.Btab[(V.sw ? V.offA : V.offB) + V.x2lo() * V.m + (k - kb - n1)];
        rdst[u] = V.r2off + (k - kb - n1);
      }
      kb += n1 + n2;
    }
  }
  __syncthreads();   // the buffer's previous readers are done
#pragma unroll
  for (int u = 0; u < RPT; ++u)
    if (rdst[u] >= 0) s_scr[rdst[u]] = rv[u];
  __syncthreads();
  // pass 1: T[v][x1][q2] = sum_q1 P[v](q1, q2) R1[x1][q1]
#pragma unroll
  for (int u = 0; u < LPT; ++u) {
    const int l = tid + u * NT;
    if (l < lines) {
      const int ll = l - line0[u];
      const int v = ll / m[u], q2 = ll % m[u];
      double *T = s_scr + toff[u] + v * nx1[u] * m[u] + q2;
      const double *R1 = s_scr + r1off[u];
#pragma unroll
      for (int x = 0; x < kMaxFaceN; ++x) {
        if (x < nx1[u]) {
          double t = 0.0;
#pragma unroll
          for (int q1 = 0; q1 < kMaxFaceN; ++q1)
            if (q1 < m[u]) t += p[u][q1] * R1[x * m[u] + q1];
          T[x * m[u]] = t;
        }
      }
    }
  }
  __syncthreads();
  // pass 2, face by face: the neighbours' values summed in order, divided
  // where several hold the point (surface.c:1322-1337)
  for (int js = 0; js < nr;) {
    const int f = RJobView(J[js]).f;
    int je = js + 1;
    while (je < nr && !RJobView(J[je]).fstart) ++je;
    for (int idx = tid; idx < 4 * NP2; idx += NT) {
      const int vv = idx / NP2, q = idx % NP2;
      const int fa = q % NP, fb = q / NP;
      double sum = 0.0;
      int cnt = 0;
      for (int jj = js; jj < je; ++jj) {
        const RJobView V(J[jj]);
        if (fa >= V.a0 && fa <= V.a1 && fb >= V.b0 && fb <= V.b1) {
          const int x1 = (V.sw ? fb : fa) - V.x1lo(), x2 = (V.sw ? fa : fb) - V.x2lo();
          const double *T = s_scr + V.toff + (vv * V.nx1() + x1) * V.m;
          const double *row = s_scr + V.r2off + x2 * V.m;
          double sv = 0.0;
          for (int k = 0; k < V.m; ++k) sv += T[k] * row[k];
          sum += sv;
          ++cnt;
        }
      }
      s_f[vv][f][q] = cnt > 1 ? sum / cnt : sum;
    }
    js = je;
  }
}

template <int NP, int NT>
__device__ __forceinline__ void load_own(const double *win, long gbase, double (*uown)[4]) {
  constexpr int NP3 = NP * NP * NP;
  constexpr int NC = (NP3 + NT - 1) / NT;
  const int tid = threadIdx.x;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int node = tid + c * NT;
    if (node < NP3) {
#pragma unroll
      for (int v = 0; v < 4; ++v) uown[c][v] = win[gbase + node + v * NP3];
    }
  }
}

// Elements it, it + step, ... below hi (while their order is NP when
// check_n), with NT threads; s_buf holds max(4 NP^3, batch_scratch(NP))
// doubles.  Returns the first item not done.
template <int NP, int GEOM, int NT, bool PRE = false>
__device__ __forceinline__ int elem_loop(const NMX_CONST BatchArgs &a, int it, int hi, int step, bool check_n,
                                         double *s_buf, double *s_fbuf, double *s_D, double *s_xb) {
  constexpr int NP2 = NP * NP;
  constexpr int NP3 = NP * NP * NP;
  constexpr int NC = (NP3 + NT - 1) / NT;
  double(*s_u)[NP3] = reinterpret_cast<double(*)[NP3]>(s_buf);
  double(*s_f)[6][NP2] = reinterpret_cast<double(*)[6][NP2]>(s_fbuf);
  const int tid = threadIdx.x;
  if (it >= hi) return it;
  double uown[NC][4];
  load_own<NP, NT>(a.win, 5 * as_const(a.rec + it)->noff, uown);
  __syncthreads();   // earlier elements are done with LDS
  for (int q = tid; q < NP2; q += NT) s_D[q] = a.Dall[batch_doff(NP) + q];
  for (int q = tid; q < NP; q += NT) s_xb[q] = a.xball[batch_xoff(NP) + q];
  const double wend_inv = as_const(a.wend_inv)[NP];

  for (; it < hi; it += step) {
    const NMX_CONST SlotRec &R = *as_const(a.rec + it);
    if (check_n && R.n != NP) break;
    const long noff = R.noff;
    const long gbase = 5 * noff;
    // PRE: the epilogue's operands (u_p, the accumulator base) requested
    // now, in flight while the faces are gathered
    double Upre[PRE ? NC : 1][5], Apre[PRE ? NC : 1][5];
    if (PRE) {
      const double *bas = a.Ain ? a.Ain : a.win;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int node = tid + c * NT;
        const long gi = gbase + (node < NP3 ? node : 0);
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          Upre[c][v] = a.wmode ? a.U[gi + v * NP3] : 0.0;
          Apre[c][v] = a.Aout ? bas[gi + v * NP3] : 0.0;
        }
      }
    }
    __syncthreads();   // the previous element is done with LDS

    // conforming copies and outer faces
    for (int idx = tid; idx < 6 * NP2; idx += NT) {
      const int f = idx / NP2, q = idx % NP2;
      const int kind = R.fkind[f];
      if (kind == 1) {
        const double *p = a.win + R.foff[f] + (q % NP) * R.fs1[f] + (q / NP) * R.fs2[f];
#pragma unroll
        for (int v = 0; v < 4; ++v) s_f[v][f][q] = p[v * NP3];
      } else if (kind == 0) {   // outer boundary: scalarwave1's exterior profile
        double ua[4] = {0.0, 0.0, 0.0, 0.0};
        if (a.profile) {
          const int fa = q % NP, fb = q / NP;
          const int dir = f >> 1, pl = (f & 1) * (NP - 1);
          const int ii = (dir == 0) ? pl : fa;
          const int jj = (dir == 1) ? pl : ((dir == 0) ? fa : fb);
          const int kk = (dir == 2) ? pl : fb;
          const int mynode = ii + NP * (jj + NP * kk);
          const double *xp = a.xyz + 3 * noff + mynode;
          const double kv[3] = {a.k[0], a.k[1], a.k[2]};
          plane_wave(kv, a.om, xp[0], xp[NP3], xp[2 * NP3], a.tdev[a.tidx], ua);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) s_f[v][f][q] = ua[v];
      }
    }
    // interpolated faces, in rounds
    for (int j = R.jfirst; j < R.jfirst + R.jcount;) {
      const int nr = as_const(a.jobs)[j].nround;
      batch_round<NP, NT>(a, j, nr, s_buf, s_f);
      j += nr;
    }
    __syncthreads();   // the rounds' buffer reads are done
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int node = tid + c * NT;
      if (node < NP3) {
#pragma unroll
        for (int v = 0; v < 4; ++v) s_u[v][node] = uown[c][v];
      }
    }
    // the next element: its own fields in flight during this RHS
    const int nx = it + step;
    if (nx < hi && (!check_n || as_const(a.rec + nx)->n == NP))
      load_own<NP, NT>(a.win, 5 * as_const(a.rec + nx)->noff, uown);
    __syncthreads();

    NodeOut o;
    o.win = a.win; o.U = a.U; o.Ain = a.Ain; o.Aout = a.Aout; o.Wout = a.Wout; o.R = a.R;
    o.cA = a.cA; o.cW = a.cW; o.wend_inv = wend_inv; o.wmode = a.wmode; o.numflux = a.numflux; o.nt = a.nt;
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
    if (PRE) {
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int node = tid + c * NT;
        if (node < NP3) node_update<NP, GEOM>(o, g, s_u, s_f, s_D, s_xb, gbase, node, Upre[c], Apre[c]);
      }
    } else {
#pragma unroll 1
      for (int c = 0; c < NC; ++c) {
        const int node = tid + c * NT;
        if (node < NP3) node_update<NP, GEOM>(o, g, s_u, s_f, s_D, s_xb, gbase, node);
      }
    }
  }
  return it;
}

// ---------------------------------------------------------------------------
// persistent launch: all orders
// ---------------------------------------------------------------------------
typedef __attribute__((address_space(3))) double lds_double;

// the workgroup's consecutive items of order NP.  Not inlined: each order
// gets a register allocation of its own instead of the union of all nine.
template <int NP, int GEOM>
__device__ __attribute__((noinline)) int batch_run(const NMX_CONST BatchArgs *ap, int it, int hi, int step,
                                                   lds_double *l_buf, lds_double *l_fbuf, lds_double *l_D,
                                                   lds_double *l_xb) {
  return elem_loop<NP, GEOM, kBT>(*ap, it, hi, step, true, (double *)l_buf, (double *)l_fbuf, (double *)l_D,
                                  (double *)l_xb);
}

template <int GEOM>
__global__ __launch_bounds__(kBT, 3) void dg_batch_kernel(BatchArgs a) {
  // the arguments where they lie, in the kernarg segment (the first explicit
  // argument sits at its offset 0): the address of the parameter itself
  // would be that of a private copy
  const NMX_CONST BatchArgs *ap = (const NMX_CONST BatchArgs *)__builtin_amdgcn_kernarg_segment_ptr();
  __shared__ double s_buf[kBatchScratch];                 // own fields, or the rounds' buffer
  __shared__ double s_fbuf[4 * 6 * kBatchMaxN * kBatchMaxN];
  __shared__ double s_D[kBatchMaxN * kBatchMaxN];
  __shared__ double s_xb[kBatchMaxN];
  const int xcd = blockIdx.x & 7, step = gridDim.x >> 3;
  const int hi = as_const(a.xlo)[xcd + 1];
  int it = as_const(a.xlo)[xcd] + (blockIdx.x >> 3);
  while (it < hi) {
    switch (as_const(a.rec + it)->n) {
#define NMX_CASE(N)                                                                              \
  case N:                                                                                        \
    it = batch_run<N, GEOM>(ap, it, hi, step, (lds_double *)s_buf, (lds_double *)s_fbuf,         \
                            (lds_double *)s_D, (lds_double *)s_xb);                            \
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
        it = hi;   // orders outside 2..10 are refused on the host
    }
  }
}

hipError_t launch_dg_batch(int geom, const BatchArgs &a, int grid, hipStream_t s) {
  if (grid % 8) return hipErrorInvalidValue;
  if (geom == 0)
    hipLaunchKernelGGL(dg_batch_kernel<0>, dim3(grid), dim3(kBT), 0, s, a);
  else if (geom == 1)
    hipLaunchKernelGGL(dg_batch_kernel<1>, dim3(grid), dim3(kBT), 0, s, a);
  else
    hipLaunchKernelGGL(dg_batch_kernel<2>, dim3(grid), dim3(kBT), 0, s, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// one order per launch: items [a.xlo[0], a.xlo[1]) in 8 contiguous XCD
// chunks, threads and LDS sized for NP
// ---------------------------------------------------------------------------
template <int NP, int GEOM, int NT, bool PRE>
__global__ __launch_bounds__(NT, PRE ? 3 : NMX_AMR_WAVES) void dg_bucket_kernel(BatchArgs a) {
  constexpr int NBUF = (4 * NP * NP * NP > batch_scratch(NP)) ? 4 * NP * NP * NP : batch_scratch(NP);
  __shared__ double s_buf[NBUF];
  __shared__ double s_fbuf[4 * 6 * NP * NP];
  __shared__ double s_D[NP * NP];
  __shared__ double s_xb[NP];
  const NMX_CONST BatchArgs *ap = (const NMX_CONST BatchArgs *)__builtin_amdgcn_kernarg_segment_ptr();
  const int lo = as_const(a.xlo)[0], n = as_const(a.xlo)[1] - lo;
  const int xcd = blockIdx.x & 7, step = gridDim.x >> 3;
  const int chunk = (n + 7) >> 3;
  const int it0 = lo + xcd * chunk;
  const int hi = min(it0 + chunk, lo + n);
  elem_loop<NP, GEOM, NT, PRE>(*ap, it0 + (blockIdx.x >> 3), hi, step, false, s_buf, s_fbuf, s_D, s_xb);
}

// threads: nodes per thread div (1, 2 or 4), whole waves, 128 to 1024
static constexpr int bucket_threads(int np3, int div) {
  return (((np3 + div - 1) / div + 63) / 64 * 64) > 1024  ? 1024
         : (((np3 + div - 1) / div + 63) / 64 * 64) < 128 ? 128
                                                          : (((np3 + div - 1) / div + 63) / 64 * 64);
}

// tile: 1 or 2 nodes per thread, and 3 or 4 with the epilogue's operands
// requested at the element's start
template <int NP, int GEOM>
static hipError_t launch_bucket_np(const BatchArgs &a, int grid, int tile, hipStream_t s) {
  constexpr int NP3 = NP * NP * NP;
  if (tile == 2)
    hipLaunchKernelGGL((dg_bucket_kernel<NP, GEOM, bucket_threads(NP3, 2), false>), dim3(grid),
                       dim3(bucket_threads(NP3, 2)), 0, s, a);
  else if (tile == 3)
    hipLaunchKernelGGL((dg_bucket_kernel<NP, GEOM, bucket_threads(NP3, 1), true>), dim3(grid),
                       dim3(bucket_threads(NP3, 1)), 0, s, a);
  else if (tile == 4)
    hipLaunchKernelGGL((dg_bucket_kernel<NP, GEOM, bucket_threads(NP3, 2), true>), dim3(grid),
                       dim3(bucket_threads(NP3, 2)), 0, s, a);
  else
    hipLaunchKernelGGL((dg_bucket_kernel<NP, GEOM, bucket_threads(NP3, 1), false>), dim3(grid),
                       dim3(bucket_threads(NP3, 1)), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_dg_bucket(int np, int geom, const BatchArgs &a, int grid, int tile, hipStream_t s) {
  if (grid % 8) return hipErrorInvalidValue;
#define NMX_CASE(N)                                                  \
  case N:                                                            \
    if (geom == 0) return launch_bucket_np<N, 0>(a, grid, tile, s);   \
    if (geom == 1) return launch_bucket_np<N, 1>(a, grid, tile, s);   \
    return launch_bucket_np<N, 2>(a, grid, tile, s);
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
