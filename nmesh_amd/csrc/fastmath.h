// fastmath.h -- reciprocal and reciprocal square root in FP64 for the
// VALU-bound FV kernels, gfx950.
//
// 1 / b from the hardware's reciprocal estimate refined by two Newton steps
// (each doubles the correct bits; within about an ulp for normal b), and
// 1 / sqrt(x) likewise.  Fewer instructions than an IEEE division (scaling,
// two Newton steps, a residual correction, fix-up).  Callers keep their
// operands normal and nonzero.
#pragma once
#include <hip/hip_runtime.h>

namespace nmx {

__device__ __forceinline__ double rcp_nr(double b) {
  double r = __builtin_amdgcn_rcp(b);
  r = fma(fma(-b, r, 1.0), r, r);
  return fma(fma(-b, r, 1.0), r, r);
}

__device__ __forceinline__ double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  y = fma(0.5 * y, fma(-x * y, y, 1.0), y);
  return fma(0.5 * y, fma(-x * y, y, 1.0), y);
}

}  // namespace nmx
