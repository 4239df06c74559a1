// dt_kernels.hip -- the time step's reductions on the device, gfx950.
//
// DG meshes (update_elm_eid_dt / adapt_node_dt_and_mesh_dt, storage.c:
// 1305-1386, with find_hmin, utilities.c:1028-1128): per element the
// smallest distance between a corner node and its neighbour along each
// direction, starting from the space diagonal / sqrt(3); the element's dt is
// uniform_dtfac times it on an FV element whose minimum is such a pair, else
// dtfac times it; the mesh dt is the minimum.  One wave per element: 24
// lanes take the 8 corners x 3 directions, the minimum comes from lane
// shuffles, then the workgroup's and the grid's minima.  The distances are
// evaluated without fused multiply-adds, as the host's are.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace nmx {

constexpr int kDtThreads = 256;
constexpr double kDtHuge = 1.7976931348623157e308;

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fmin(v, __shfl_xor(v, m, 64));
  return v;
}

__device__ __forceinline__ double node_dist(const double *xe, int np3, int a, int b) {
#pragma clang fp contract(off)
  double s = 0.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double t = xe[d * np3 + a] - xe[d * np3 + b];
    s += t * t;
  }
  return sqrt(s);
}

__global__ __launch_bounds__(kDtThreads) void dt_elm_kernel(DtArgs a) {
#pragma clang fp contract(off)
  __shared__ double s_min[kDtThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double best = kDtHuge;
  for (int e = blockIdx.x * (kDtThreads / 64) + wave; e < a.nelem; e += gridDim.x * (kDtThreads / 64)) {
    const int n = a.en[e], np3 = n * n * n;
    const double *xe = a.xyz + 3 * a.eoff[e];
    double dd = kDtHuge;
    if (lane < 24) {
      const int c = lane / 3, d = lane % 3;
      int I[3] = {(c & 1) ? n - 1 : 0, (c & 2) ? n - 1 : 0, (c & 4) ? n - 1 : 0};
      int J[3] = {I[0], I[1], I[2]};
      J[d] += I[d] == 0 ? 1 : -1;   // the one neighbour inside the element
      dd = node_dist(xe, np3, I[0] + n * (I[1] + n * I[2]), J[0] + n * (J[1] + n * J[2]));
    }
    const double mind = wave_min(dd);
    const double diag = node_dist(xe, np3, 0, np3 - 1) / sqrt(3.);
    const double h = mind < diag ? mind : diag;
    const bool pair = mind < diag;
    const double edt = (a.efv[e] && pair ? a.uniform_dtfac : a.dtfac) * h;
    best = fmin(best, edt);
  }
  if (lane == 0) s_min[wave] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = s_min[0];
    for (int w = 1; w < kDtThreads / 64; ++w) m = fmin(m, s_min[w]);
    a.part[blockIdx.x] = m;
  }
}

__global__ __launch_bounds__(kDtThreads) void dt_min_kernel(const double *part, int n, double *out) {
  __shared__ double s_min[kDtThreads / 64];
  double m = kDtHuge;
  for (int i = threadIdx.x; i < n; i += kDtThreads) m = fmin(m, part[i]);
  m = wave_min(m);
  if ((threadIdx.x & 63) == 0) s_min[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = s_min[0];
    for (int w = 1; w < kDtThreads / 64; ++w) r = fmin(r, s_min[w]);
    out[0] = r;
  }
}

int dt_grid(int nelem) {
  const int g = (nelem + kDtThreads / 64 - 1) / (kDtThreads / 64);
  return g < 2048 ? (g > 0 ? g : 1) : 2048;
}

hipError_t launch_dt(const DtArgs &a, int grid, hipStream_t s) {
  if (a.nelem <= 0 || grid <= 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(dt_elm_kernel, dim3(grid), dim3(kDtThreads), 0, s, a);
  hipLaunchKernelGGL(dt_min_kernel, dim3(1), dim3(kDtThreads), 0, s, a.part, grid, a.out);
  return hipGetLastError();
}

}  // namespace nmx
