// limiter_kernels.hip -- the MRS limiter of scalarwave1_limiter = MRS on the
// device, gfx950.
//
// evolve_limiter_mesh (evolve.c:430-541) with limdata_MRS and limiter_MRS
// (limiter.c:56-196):
//   lim_data_kernel: per element and variable the node average (the LGL
//     quadrature along x, then y, then z, over the Xb volume 8;
//     var_nodeaverage, basis.c:462-510), the minimum and the maximum;
//   lim_apply_kernel: per element, over its variables, theta = min(1,
//     phi((M - wbar)/(wmax - wbar)), phi((m - wbar)/(wmin - wbar))) with
//     M, m the neighbours' extremes widened to wbar +- alpha h^1.5 and
//     phi(y) = min(y / 1.1, 1) (0 for y < 0; a zero denominator gives 1);
//     then q <- qbar + theta (q - qbar) for every variable.
// The averages, theta and the update are evaluated in the reference's
// operation order without fused multiply-adds, so the limiter decides
// exactly as the reference does on the same state (a constant that the
// quadrature does not reproduce to the last bit flips theta between 0 and
// 1 there as well).
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace nmx {

constexpr int kLimThreads = 128;

__global__ __launch_bounds__(kLimThreads) void lim_data_kernel(LimArgs a) {
#pragma clang fp contract(off)
  __shared__ double s1[100], s2[10];
  __shared__ double smn[kLimThreads / 64], smx[kLimThreads / 64];
  const int tid = threadIdx.x;
  for (int e = blockIdx.x; e < a.nelem; e += gridDim.x) {
    const int n = a.en[e], n2 = n * n, np = n2 * n;
    const double *w = a.wq + (n - 2) * 10;
    const double *eu = a.u + 5 * a.eoff[e];
    for (int v = 0; v < 5; ++v) {
      const double *q = eu + (long)v * np;
      __syncthreads();
      if (tid < n2) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += w[i] * q[i + n * tid];
        s1[tid] = s;
      }
      double mn = q[0], mx = q[0];
      for (int t = tid; t < np; t += kLimThreads) {
        mn = fmin(mn, q[t]);
        mx = fmax(mx, q[t]);
      }
      for (int m = 32; m >= 1; m >>= 1) {
        mn = fmin(mn, __shfl_xor(mn, m, 64));
        mx = fmax(mx, __shfl_xor(mx, m, 64));
      }
      if ((tid & 63) == 0) { smn[tid >> 6] = mn; smx[tid >> 6] = mx; }
      __syncthreads();
      if (tid < n) {
        double s = 0.0;
        for (int j = 0; j < n; ++j) s += w[j] * s1[j + n * tid];
        s2[tid] = s;
      }
      __syncthreads();
      if (tid == 0) {
        double s = 0.0;
        for (int k = 0; k < n; ++k) s += w[k] * s2[k];
        double mn2 = smn[0], mx2 = smx[0];
        for (int i = 1; i < kLimThreads / 64; ++i) { mn2 = fmin(mn2, smn[i]); mx2 = fmax(mx2, smx[i]); }
        double *o = a.ldat + 15L * e + 3 * v;
        o[0] = s * 0.125;
        o[1] = mn2;
        o[2] = mx2;
      }
    }
  }
}

__device__ __forceinline__ double mrs_theta(double Mi, double wbar, double wMi) {
#pragma clang fp contract(off)
  const double den = wMi - wbar;
  if (den == 0.0) return 1.0;
  const double y = (Mi - wbar) / den;
  if (y < 0.0) return 0.0;
  return fmin(y / 1.1, 1.0);
}

__global__ __launch_bounds__(kLimThreads) void lim_apply_kernel(LimArgs a) {
#pragma clang fp contract(off)
  const int tid = threadIdx.x;
  for (int e = blockIdx.x; e < a.nelem; e += gridDim.x) {
    const int n = a.en[e], np = n * n * n;
    const double *ld = a.ldat + 15L * e;
    const int *nb = a.nbr + 24L * e;
    const double ah = a.alpha * a.h15[e];
    double thM = 1.7976931348623157e308, thm = 1.7976931348623157e308;
    for (int v = 0; v < 5; ++v) {
      const double wbar = ld[3 * v], wmi = ld[3 * v + 1], wMi = ld[3 * v + 2];
      double Mi = -1.7976931348623157e308, mi = 1.7976931348623157e308;
      for (int i = 0; i < 24; ++i) {
        const int j = nb[i];
        if (j < 0) continue;
        const double ma = a.ldat[15L * j + 3 * v + 1], Ma = a.ldat[15L * j + 3 * v + 2];
        if (ma < mi) mi = ma;
        if (Ma > Mi) Mi = Ma;
      }
      mi = fmin(wbar - ah, mi);
      Mi = fmax(wbar + ah, Mi);
      const double t1 = mrs_theta(Mi, wbar, wMi), t2 = mrs_theta(mi, wbar, wmi);
      if (t1 < thM) thM = t1;
      if (t2 < thm) thm = t2;
    }
    const double th = fmin(1.0, fmin(thm, thM));
    double *eu = a.u + 5 * a.eoff[e];
    for (int v = 0; v < 5; ++v) {
      const double qbar = ld[3 * v];
      double *q = eu + (long)v * np;
      for (int t = tid; t < np; t += kLimThreads) q[t] = qbar + th * (q[t] - qbar);
    }
  }
}

hipError_t launch_limiter(const LimArgs &a, hipStream_t s) {
  if (a.nelem <= 0) return hipSuccess;
  const int grid = a.nelem < 65536 ? a.nelem : 65536;
  hipLaunchKernelGGL(lim_data_kernel, dim3(grid), dim3(kLimThreads), 0, s, a);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(lim_apply_kernel, dim3(grid), dim3(kLimThreads), 0, s, a);
  return hipGetLastError();
}

}  // namespace nmx
