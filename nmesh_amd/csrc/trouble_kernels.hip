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

constexpr int kTrbThreads = 64;   // rdmp_kernel: one wave per element
constexpr int kMaxTrouble = 1073741824;

// One wave per element, no barriers.  The lanes fetch the 5 x 24
// neighbour extremes of u_p (two pairs each) while they stream u for the
// per-variable minimum, maximum and a NaN flag; wave reductions follow.
// The element is out of bounds when a variable's minimum is below m - d or
// its maximum above M + d (or a value is NaN), which is when some value
// leaves [m - d, M + d].
__device__ __forceinline__ double wave_min(double x) {
  for (int m = 32; m >= 1; m >>= 1) x = fmin(x, __shfl_xor(x, m, 64));
  return x;
}
__device__ __forceinline__ double wave_max(double x) {
  for (int m = 32; m >= 1; m >>= 1) x = fmax(x, __shfl_xor(x, m, 64));
  return x;
}

__global__ __launch_bounds__(kTrbThreads) void rdmp_kernel(TroubleArgs a) {
  const int lane = threadIdx.x;
  for (int e = blockIdx.x; e < a.nelem; e += gridDim.x) {
    const int n = a.en[e], np = n * n * n;
    const double *own = a.ldat + 20L * e;
    // neighbour pairs p = lane and lane + 64 (of 120: p = 24 v + k)
    double pm[2], pM[2];
    int pv[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int p = lane + 64 * h;
      pv[h] = p < 120 ? p / 24 : -1;
      const int nb = p < 120 ? a.nbr[24 * e + (p - 24 * (p / 24))] : -1;
      const double *src = nb < 0 ? own : a.ldat + 20L * nb;   // a missing neighbour: the element's own
      const int v = p < 120 ? pv[h] : 0;
      pm[h] = src[4 * v + 1];
      pM[h] = src[4 * v + 2];
    }
    const double *eu = a.u + 5 * a.eoff[e];
    int nan = 0;
    int out = 0;
    double umn[5], umx[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) umn[v] = umx[v] = eu[(long)v * np];
    // the five variables' values at t requested together
#pragma unroll 2
    for (int t = lane; t < np; t += kTrbThreads) {
      double x[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) x[v] = eu[(long)v * np + t];
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        nan |= x[v] != x[v];
        umn[v] = fmin(umn[v], x[v]);
        umx[v] = fmax(umx[v], x[v]);
      }
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const double mn = wave_min(umn[v]);
      const double mx = wave_max(umx[v]);
      double m = own[4 * v + 1], M = own[4 * v + 2];
#pragma unroll
      for (int h = 0; h < 2; ++h)
        if (pv[h] == v) { m = fmin(m, pm[h]); M = fmax(M, pM[h]); }
      m = wave_min(m);
      M = wave_max(M);
      const double d = fmax(a.delta0, a.eps * (M - m));
      if (!(mn >= m - d && mx <= M + d)) out = 1;
    }
    for (int m = 32; m >= 1; m >>= 1) nan |= __shfl_xor(nan, m, 64);
    if (lane == 0) a.tr[e] = a.efv[e] ? (out || nan ? 0 : -1) : (out || nan ? 1 : 0);
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
  hipLaunchKernelGGL(rdmp_kernel, dim3(a.nelem < 65536 ? a.nelem : 65536), dim3(kTrbThreads), 0, s, a);
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
