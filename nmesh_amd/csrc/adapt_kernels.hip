// adapt_kernels.hip -- the evolved state moved onto an adapted mesh, gfx950.
//
// nmesh interpolates the variables when elements change (update_node_n_pt_typ
// storage.c:858-931 for a new n; make_child_elm storage.c:493-585 for the
// children of a refined element; make_parent_elm storage.c:627-760 for the
// parent of 8 coarsened siblings).  Here every new element has one or more
// jobs, each a source element of the old register with three 1D Lagrange
// matrices (new points x old points, one per direction) and the box of new
// points the source holds:
//   u_new[k][j][i] = sum_{c,b,a} A2[k][c] A1[j][b] A0[i][a] u_old[c][b][a]
// summed over the jobs that hold the point and divided by their number when
// there are several (the parent points on the children's common faces).  A
// job with no matrices copies the element unchanged (same place, same n).
// One workgroup per new element; one variable at a time through LDS.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace nmx {

constexpr int kAdaptThreads = 256;
constexpr int kAdaptMaxN = 10;

__global__ __launch_bounds__(kAdaptThreads) void adapt_kernel(AdaptArgs a) {
  constexpr int M3 = kAdaptMaxN * kAdaptMaxN * kAdaptMaxN;
  __shared__ double s_in[M3], s_t1[M3], s_t2[M3], s_acc[M3];
  __shared__ int s_cnt[M3];
  const int tid = threadIdx.x;
  for (int e = blockIdx.x; e < a.nelem; e += gridDim.x) {
    const int j0 = a.jfirst[e], j1 = a.jfirst[e + 1];
    const AdaptJob J0 = a.jobs[j0];
    const int nd = J0.nd, nd3 = nd * nd * nd;
    double *dst = a.unew + 5 * J0.dst;
    if (j1 - j0 == 1 && J0.a0 < 0) {   // same place, same n: copy
      const double *src = a.uold + 5 * J0.src;
      for (int q = tid; q < 5 * nd3; q += kAdaptThreads) dst[q] = src[q];
      continue;
    }
    for (int v = 0; v < 5; ++v) {
      __syncthreads();
      for (int q = tid; q < nd3; q += kAdaptThreads) {
        s_acc[q] = 0.0;
        s_cnt[q] = 0;
      }
      for (int jb = j0; jb < j1; ++jb) {
        const AdaptJob J = a.jobs[jb];
        const int ns = J.ns, ns3 = ns * ns * ns;
        const double *A0 = a.atab + J.a0, *A1 = a.atab + J.a1, *A2 = a.atab + J.a2;
        const double *src = a.uold + 5 * J.src + (long)v * ns3;
        __syncthreads();
        for (int q = tid; q < ns3; q += kAdaptThreads) s_in[q] = src[q];
        __syncthreads();
        // along i: t1[c][b][i]
        for (int q = tid; q < ns * ns * nd; q += kAdaptThreads) {
          const int i = q % nd, bc = q / nd;
          double s = 0.0;
          for (int l = 0; l < ns; ++l) s += A0[i * ns + l] * s_in[bc * ns + l];
          s_t1[q] = s;
        }
        __syncthreads();
        // along j: t2[c][j][i]
        for (int q = tid; q < ns * nd * nd; q += kAdaptThreads) {
          const int i = q % nd, j = (q / nd) % nd, c = q / (nd * nd);
          double s = 0.0;
          for (int l = 0; l < ns; ++l) s += A1[j * ns + l] * s_t1[i + nd * (l + ns * c)];
          s_t2[q] = s;
        }
        __syncthreads();
        // along k, into the points this source holds
        for (int q = tid; q < nd3; q += kAdaptThreads) {
          const int i = q % nd, j = (q / nd) % nd, k = q / (nd * nd);
          if (i < J.lo[0] || i > J.hi[0] || j < J.lo[1] || j > J.hi[1] || k < J.lo[2] || k > J.hi[2])
            continue;
          double s = 0.0;
          for (int l = 0; l < ns; ++l) s += A2[k * ns + l] * s_t2[i + nd * (j + nd * l)];
          s_acc[q] += s;
          s_cnt[q] += 1;
        }
      }
      __syncthreads();
      for (int q = tid; q < nd3; q += kAdaptThreads)
        dst[(long)v * nd3 + q] = s_cnt[q] > 1 ? s_acc[q] / s_cnt[q] : s_acc[q];
    }
  }
}

hipError_t launch_adapt(const AdaptArgs &a, hipStream_t s) {
  if (a.nelem <= 0) return hipSuccess;
  const int grid = a.nelem < 65536 ? a.nelem : 65536;
  hipLaunchKernelGGL(adapt_kernel, dim3(grid), dim3(kAdaptThreads), 0, s, a);
  return hipGetLastError();
}

}  // namespace nmx
