// trouble_kernels.hip -- nmesh's troubled-element machinery on the device,
// gfx950.
//
// evolve_set_trouble_score (src/main/evolve/trouble.c:27-113) per element:
// the TROUBLE value tr in {1, 0, -1} (1: switch to the robust method FV,
// 0: keep, -1: all good, switch back to DG) moves the element's trouble
// score: tr = 1 sets it to 1 or raises it by 1, tr = 0 sets it to 0, tr = -1
// sets it to -1 or lowers it by 1, clamped to +-2^30.  The mesh's score
// (evolve_read_trouble_score_mesh, trouble.c:129-161) is the maximum when
// it is positive, else the minimum.
//
// TROUBLE values come from a project's function; scalarwave1 registers
// none.  The library's own indicator (rdmp_kernel) is a relaxed discrete
// maximum principle on the data evolve_collect_u_p_data_mesh gathers
// (trouble.c:355-400: limdata_MRS of u_p, the element minima and maxima): an
// element is out of bounds when a value of u leaves [m - d, M + d], with m
// and M the extremes of u_p over the element and its face neighbours and
// d = max(delta0, eps (M - m)) per variable.  DG elements out of bounds are
// troubled (1), FV elements in bounds are good (-1), the rest 0.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace nmx {

constexpr int kTrbThreads = 128;
constexpr int kMaxTrouble = 1073741824;

__global__ __launch_bounds__(kTrbThreads) void rdmp_kernel(TroubleArgs a) {
  __shared__ double lo[5], hi[5];
  const int tid = threadIdx.x;
  for (int e = blockIdx.x; e < a.nelem; e += gridDim.x) {
    const int n = a.en[e], np = n * n * n;
    __syncthreads();   // the previous element is done with lo, hi
    if (tid < 5) {
      const int v = tid;
      double m = a.ldat[20L * e + 4 * v + 1], M = a.ldat[20L * e + 4 * v + 2];
      for (int k = 0; k < 24; ++k) {
        const int nb = a.nbr[24 * e + k];
        if (nb < 0) continue;
        m = fmin(m, a.ldat[20L * nb + 4 * v + 1]);
        M = fmax(M, a.ldat[20L * nb + 4 * v + 2]);
      }
      const double d = fmax(a.delta0, a.eps * (M - m));
      lo[v] = m - d;
      hi[v] = M + d;
    }
    __syncthreads();
    const double *eu = a.u + 5 * a.eoff[e];
    int out = 0;
    for (int t = tid; t < 5 * np; t += kTrbThreads) {
      const int v = t / np;
      const double x = eu[t];
      if (!(x >= lo[v] && x <= hi[v])) out = 1;   // NaN is out of bounds too
    }
    out = __syncthreads_or(out);
    if (tid == 0) a.tr[e] = a.efv[e] ? (out ? 0 : -1) : (out ? 1 : 0);
  }
}

// the scores from the TROUBLE values (tr == nullptr: read them only), then
// the mesh's score in out[0]
__global__ __launch_bounds__(1024) void trouble_score_kernel(const int *tr, int *score, int n, int *out) {
  __shared__ int smax[16], smin[16];
  int mx = -2147483647 - 1, mn = 2147483647;
  for (int e = threadIdx.x; e < n; e += 1024) {
    int s = score[e];
    if (tr) {
      const int t = tr[e];
      if (t > 0) s = s <= 0 ? 1 : s + 1;
      else if (t == 0) s = 0;
      else s = s >= 0 ? -1 : s - 1;
      if (s < -kMaxTrouble) s = -kMaxTrouble;
      if (s > kMaxTrouble) s = kMaxTrouble;
      score[e] = s;
    }
    mx = max(mx, s);
    mn = min(mn, s);
  }
  for (int m = 32; m >= 1; m >>= 1) {
    mx = max(mx, __shfl_xor(mx, m, 64));
    mn = min(mn, __shfl_xor(mn, m, 64));
  }
  if ((threadIdx.x & 63) == 0) { smax[threadIdx.x >> 6] = mx; smin[threadIdx.x >> 6] = mn; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < 16; ++i) { mx = max(mx, smax[i]); mn = min(mn, smin[i]); }
    mx = max(mx, smax[0]);
    mn = min(mn, smin[0]);
    out[0] = mx > 0 ? mx : mn;
  }
}

// y = x + s r  (mode 0), y = x  (mode 1)  or  y = 0.75 x + 0.25 (x2 + s r)
// (mode 2) on the elements whose score (or mask) is positive, in the stage
// epilogue's arithmetic
__global__ __launch_bounds__(256) void trouble_fix_kernel(TroubleFix f) {
  for (int e = blockIdx.x; e < f.nelem; e += gridDim.x) {
    if (f.score[e] <= 0) continue;
    const int n = f.en[e];
    const long np5 = 5L * n * n * n, o = 5 * f.eoff[e];
    for (long t = threadIdx.x; t < np5; t += blockDim.x) {
      const long g = o + t;
      if (f.mode == 2) f.y[g] = 0.75 * f.x[g] + 0.25 * (f.x2[g] + f.s * f.r[g]);
      else if (f.mode == 1) f.y[g] = f.x[g];
      else f.y[g] = f.x[g] + f.s * f.r[g];
    }
  }
}

hipError_t launch_rdmp(const TroubleArgs &a, hipStream_t s) {
  if (a.nelem <= 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(rdmp_kernel, dim3(a.nelem < 8192 ? a.nelem : 8192), dim3(kTrbThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_trouble_score(const int *tr, int *score, int n, int *out, hipStream_t s) {
  hipLaunchKernelGGL(trouble_score_kernel, dim3(1), dim3(1024), 0, s, tr, score, n, out);
  return hipGetLastError();
}

hipError_t launch_trouble_fix(const TroubleFix &f, hipStream_t s) {
  if (f.nelem <= 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(trouble_fix_kernel, dim3(f.nelem < 8192 ? f.nelem : 8192), dim3(256), 0, s, f);
  return hipGetLastError();
}

}  // namespace nmx
