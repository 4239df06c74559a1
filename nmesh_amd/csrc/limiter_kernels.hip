// limiter_kernels.hip -- the limiters of scalarwave1_limiter = MRS and
// minmodB on the device, gfx950.
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
// minmodB (limdata_c000_100_010_001 and limiter_minmodB, limiter.c:
// 200-380):
//   mm_data_kernel: per element and variable the normalized Legendre
//     coefficients c000, c100, c010, c001 (the analysis along x, y, z);
//   mm_apply_kernel: per variable the three slopes through the bounded
//     minmod against beta / sqrt(3) times the differences to the face
//     neighbours' mean c000 (0 where a face has none); when a slope changes
//     by more than 1e-10 the variable becomes the synthesis of c000 and the
//     limited slopes.
// The averages, coefficients, theta and the updates are evaluated in the
// reference's operation order without fused multiply-adds, so a limiter
// decides exactly as the reference does on the same state (for MRS, a
// constant that the quadrature does not reproduce to the last bit flips
// theta between 0 and 1 there as well).
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace nmx {

constexpr int kLimThreads = 128;

// One wave per element: lane l sums the x lines l, l + 64, ... of the five
// variables side by side straight from the element block (a line is
// contiguous, so the block is read once through the caches) and takes their
// minima and maxima on the way; the y sums of the 5 n rows and the z sums follow from
// LDS.  Every sum runs in the order of the one-variable quadrature.
constexpr int kLdThreads = 64;

__global__ __launch_bounds__(kLdThreads) void lim_data_kernel(LimArgs a) {
#pragma clang fp contract(off)
  __shared__ double s1[5 * 100], s2[5 * 10];
  const int lane = threadIdx.x;
  for (int e = blockIdx.x; e < a.nelem; e += gridDim.x) {
    const int n = a.en[e], n2 = n * n, np = n2 * n;
    const double *w = a.wq + (n - 2) * 10;
    const double *eu = a.u + 5 * a.eoff[e];
    double mn[5], mx[5], s[5];
    __syncthreads();   // the previous element is done with s1, s2
#pragma unroll
    for (int v = 0; v < 5; ++v) mn[v] = mx[v] = eu[(long)v * np];
    for (int l = lane; l < n2; l += kLdThreads) {
      // line l of the five variables, point by point: five loads in flight
#pragma unroll
      for (int v = 0; v < 5; ++v) s[v] = 0.0;
#pragma unroll 4
      for (int i = 0; i < n; ++i) {
        double x[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) x[v] = eu[(long)v * np + n * l + i];
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          s[v] += w[i] * x[v];
          mn[v] = fmin(mn[v], x[v]);
          mx[v] = fmax(mx[v], x[v]);
        }
      }
#pragma unroll
      for (int v = 0; v < 5; ++v) s1[v * n2 + l] = s[v];
    }
#pragma unroll
    for (int v = 0; v < 5; ++v)
      for (int m = 32; m >= 1; m >>= 1) {
        mn[v] = fmin(mn[v], __shfl_xor(mn[v], m, 64));
        mx[v] = fmax(mx[v], __shfl_xor(mx[v], m, 64));
      }
    __syncthreads();
    if (lane < 5 * n) {
      const int v = lane / n, k = lane - v * n;
      const double *q = s1 + v * n2 + n * k;
      double y = 0.0;
      for (int j = 0; j < n; ++j) y += w[j] * q[j];
      s2[v * n + k] = y;
    }
    __syncthreads();
    if (lane < 5) {
      const int v = lane;
      double z = 0.0;
      for (int k = 0; k < n; ++k) z += w[k] * s2[v * n + k];
      double *o = a.ldat + 20L * e + 4 * v;
      o[0] = z * 0.125;
      // every lane holds the reduced extremes; lane v writes variable v's
      o[1] = v == 0 ? mn[0] : v == 1 ? mn[1] : v == 2 ? mn[2] : v == 3 ? mn[3] : mn[4];
      o[2] = v == 0 ? mx[0] : v == 1 ? mx[1] : v == 2 ? mx[2] : v == 3 ? mx[3] : mx[4];
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
    const double *ld = a.ldat + 20L * e;
    const int *nb = a.nbr + 24L * e;
    const double ah = a.alpha * a.h15[e];
    double thM = 1.7976931348623157e308, thm = 1.7976931348623157e308;
    for (int v = 0; v < 5; ++v) {
      const double wbar = ld[4 * v], wmi = ld[4 * v + 1], wMi = ld[4 * v + 2];
      double Mi = -1.7976931348623157e308, mi = 1.7976931348623157e308;
      for (int i = 0; i < 24; ++i) {
        const int j = nb[i];
        if (j < 0) continue;
        const double ma = a.ldat[20L * j + 4 * v + 1], Ma = a.ldat[20L * j + 4 * v + 2];
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
      const double qbar = ld[4 * v];
      double *q = eu + (long)v * np;
      for (int t = tid; t < np; t += kLimThreads) q[t] = qbar + th * (q[t] - qbar);
    }
  }
}

// minmodB: c000, c100, c010, c001 of every variable.  The analysis along x
// (rows i = 0, 1), then y, then z, each sum in ascending order, as the full
// contraction computes these four coefficients.
__global__ __launch_bounds__(kLimThreads) void mm_data_kernel(LimArgs a) {
#pragma clang fp contract(off)
  __shared__ double t1[2 * 100], t2[3 * 10];
  const int tid = threadIdx.x;
  for (int e = blockIdx.x; e < a.nelem; e += gridDim.x) {
    const int n = a.en[e], n2 = n * n, np = n2 * n;
    const double *A = a.amat + (n - 2) * 100;
    const double *eu = a.u + 5 * a.eoff[e];
    for (int v = 0; v < 5; ++v) {
      const double *q = eu + (long)v * np;
      __syncthreads();
      for (int t = tid; t < 2 * n2; t += kLimThreads) {   // t1[i][b + n c]
        const int i = t / n2, bc = t % n2;
        double s = 0.0;
        for (int l = 0; l < n; ++l) s += A[i * n + l] * q[l + n * bc];
        t1[t] = s;
      }
      __syncthreads();
      if (tid < 3 * n) {   // t2[(i,j)][c] for (0,0), (1,0), (0,1)
        const int m = tid / n, c = tid % n;
        const int i = (m == 1) ? 1 : 0, j = (m == 2) ? 1 : 0;
        double s = 0.0;
        for (int l = 0; l < n; ++l) s += A[j * n + l] * t1[i * n2 + l + n * c];
        t2[tid] = s;
      }
      __syncthreads();
      if (tid < 4) {   // c000, c100, c010 from (i,j) rows with k = 0; c001 with k = 1
        const int m = (tid == 3) ? 0 : tid, k = (tid == 3) ? 1 : 0;
        double s = 0.0;
        for (int l = 0; l < n; ++l) s += A[k * n + l] * t2[m * n + l];
        a.ldat[20L * e + 4 * v + tid] = s;
      }
    }
  }
}

__device__ __forceinline__ double minmod3(double x, double y, double z) {
#pragma clang fp contract(off)
  if (x * y >= 0 && y * z >= 0) return (x > 0. ? 1. : -1.) * fmin(fabs(x), fmin(fabs(y), fabs(z)));
  return 0.;
}

__device__ __forceinline__ double minmod3B(double x, double y, double z, double bound) {
  if (fabs(x) <= bound) return x;
  return minmod3(x, y, z);
}

__device__ __forceinline__ bool dequal(double x, double y) {
  return !(x < y - 1e-10 || y < x - 1e-10);
}

__global__ __launch_bounds__(kLimThreads) void mm_apply_kernel(LimArgs a) {
#pragma clang fp contract(off)
  const int tid = threadIdx.x;
  const double bos3 = a.beta / sqrt(3.);
  for (int e = blockIdx.x; e < a.nelem; e += gridDim.x) {
    const int n = a.en[e], n2 = n * n, np = n2 * n;
    const double *S = a.smat + (n - 2) * 100;
    const int *nb = a.nbr + 24L * e;
    const double Mt_h = a.alpha * a.hw[e];
    double *eu = a.u + 5 * a.eoff[e];
    for (int v = 0; v < 5; ++v) {
      const double *my = a.ldat + 20L * e + 4 * v;
      double w0[6], w0L1 = 0.;
      for (int f = 0; f < 6; ++f) {
        double av = 0.;
        int cnt = 0;
        for (int ni = 0; ni < 4; ++ni) {
          const int j = nb[4 * f + ni];
          if (j < 0) break;
          av += a.ldat[20L * j + 4 * v];
          ++cnt;
        }
        if (cnt > 0) av = av / cnt;
        w0[f] = av;
        w0L1 += fabs(av) * 0.166666666666667;
      }
      const double bound = a.scale_bound ? Mt_h * w0L1 : Mt_h;
      const double w0c = my[0];
      const double l1 = minmod3B(my[1], bos3 * (w0c - w0[0]), bos3 * (w0[1] - w0c), bound);
      const double l2 = minmod3B(my[2], bos3 * (w0c - w0[2]), bos3 * (w0[3] - w0c), bound);
      const double l3 = minmod3B(my[3], bos3 * (w0c - w0[4]), bos3 * (w0[5] - w0c), bound);
      if (dequal(l1, my[1]) && dequal(l2, my[2]) && dequal(l3, my[3])) continue;
      // the synthesis of c000 e000 + l1 e100 + l2 e010 + l3 e001 along x,
      // then y, then z (the zero coefficients add exact zeros)
      double *q = eu + (long)v * np;
      for (int t = tid; t < np; t += kLimThreads) {
        const int ia = t % n, ib = (t / n) % n, ic = t / n2;
        const double u00 = S[ia * n] * w0c + S[ia * n + 1] * l1;   // (j, k) = (0, 0)
        const double u10 = S[ia * n] * l2;                         // (1, 0)
        const double u01 = S[ia * n] * l3;                         // (0, 1)
        const double v0 = S[ib * n] * u00 + S[ib * n + 1] * u10;   // k = 0
        const double v1 = S[ib * n] * u01;                         // k = 1
        q[t] = S[ic * n] * v0 + S[ic * n + 1] * v1;
      }
    }
  }
}

// the MRS data pass alone (limdata_MRS): average, minimum, maximum per
// element and variable into a.ldat
hipError_t launch_lim_data(const LimArgs &a, hipStream_t s) {
  if (a.nelem <= 0) return hipSuccess;
  const int grid = a.nelem < 65536 ? a.nelem : 65536;
  hipLaunchKernelGGL(lim_data_kernel, dim3(grid), dim3(kLdThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_limiter(const LimArgs &a, hipStream_t s) {
  if (a.nelem <= 0) return hipSuccess;
  const int grid = a.nelem < 65536 ? a.nelem : 65536;
  if (a.kind == 1)
    hipLaunchKernelGGL(mm_data_kernel, dim3(grid), dim3(kLimThreads), 0, s, a);
  else
    hipLaunchKernelGGL(lim_data_kernel, dim3(grid), dim3(kLdThreads), 0, s, a);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  if (a.kind == 1)
    hipLaunchKernelGGL(mm_apply_kernel, dim3(grid), dim3(kLimThreads), 0, s, a);
  else
    hipLaunchKernelGGL(lim_apply_kernel, dim3(grid), dim3(kLimThreads), 0, s, a);
  return hipGetLastError();
}

}  // namespace nmx
