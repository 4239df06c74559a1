// state_kernels.hip -- state layout conversion on the device.
//
// The C ABI's transfer layout is flat, u[v*nnodes + node] with node = element
// offset + ijk (nmesh_amr_loops.h:279); the device registers hold element
// blocks, element e at 5*off[e] with variable v at +v*np_e.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace nmx {

__global__ __launch_bounds__(256) void layout_kernel(const double *src, double *dst, const long *eoff,
                                                     const int *en, int nelem, long nn, int to_blocks) {
  for (int e = blockIdx.x; e < nelem; e += gridDim.x) {
    const long off = eoff[e];
    const int n = en[e];
    const int np = n * n * n;
    for (int idx = threadIdx.x; idx < 5 * np; idx += blockDim.x) {
      const int v = idx / np, q = idx % np;
      const long b = 5 * off + idx, f = v * nn + off + q;
      if (to_blocks) dst[b] = src[f];
      else dst[f] = src[b];
    }
  }
}

hipError_t launch_layout(const double *src, double *dst, const long *eoff, const int *en, int nelem,
                         long nnodes, int to_blocks, hipStream_t s) {
  if (nelem <= 0) return hipSuccess;
  const int grid = nelem < 65535 ? nelem : 65535;
  hipLaunchKernelGGL(layout_kernel, dim3(grid), dim3(256), 0, s, src, dst, eoff, en, nelem, nnodes, to_blocks);
  return hipGetLastError();
}

}  // namespace nmx
