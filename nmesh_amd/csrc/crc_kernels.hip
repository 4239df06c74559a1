// crc_kernels.hip -- the checkpoint's "vars" CRC-64 of a state register on
// the device, gfx950.
//
// nmesh's CRC (crc64.c:153-256; norms.c:440-458 crc_VarList_local) is the
// reflected CRC-64 with the ECMA-182 polynomial (0xc96c5795d7870f42
// reversed), pre- and post-inverted, over the byte stream "every element in
// order, each variable of the list in order, its doubles as stored".  For
// scalarwave1 the list is phi, pi, cx, cy, cz (variable-index order); the
// register holds element blocks [e][pi, cx, cy, cz, phi][np_e].
//
// The CRC is affine in the data and the CRC of a concatenation is the XOR
// of each piece's CRC (from 0) multiplied by x^(8 * bytes after the piece)
// mod P (crc64_0start_combine), so the stream is cut into independent
// pieces:
//   - one wave per element, its lanes interleaved over the element's words
//     so that every load is coalesced (crc_elm_kernel below);
//   - the element's CRC times x^(320 * nodes after the element) (the
//     product of four table factors, one per base-256 digit of the node
//     count), written per element;
//   - a second launch XOR-reduces the elements' contributions.
// Only loads, vector stores and wave shuffles: no atomics.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"

namespace nmx {

constexpr uint64_t kCrcPoly = 0xc96c5795d7870f42ULL;

// a * b mod P, reflected (bit 63 is x^0): b runs through b, b x, b x^2, ...
// while the bits of a are read from x^0 up
__device__ __forceinline__ uint64_t gf2_mul(uint64_t a, uint64_t b) {
  uint64_t r = 0;
#pragma unroll 8
  for (int i = 63; i >= 0; --i) {
    if ((a >> i) & 1) r ^= b;
    b = (b >> 1) ^ ((b & 1) ? kCrcPoly : 0);
  }
  return r;
}

__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
  const uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
  const uint32_t l2 = __shfl_xor(lo, m, 64), h2 = __shfl_xor(hi, m, 64);
  return ((uint64_t)h2 << 32) | l2;
}

// block: 4 waves, one element each at a time
constexpr int kCrcThreads = 256;

// One wave per element.  The element's stream is W = 5 np words w_t (phi,
// pi, cx, cy, cz); lane l reads words l, l + 64, l + 128, ... so that every
// load of the wave is 512 contiguous bytes.  With S(w) = w x^64 (eight
// slicing lookups) the register after the element is
//   (~0) x^(64 W)  +  sum_t S(w_t) x^(64 (W - 1 - t)),
// and lane l folds its words Horner-wise, acc <- acc x^4096 + S(w) (eight
// more lookups in tables of byte x^4096), then scales acc by
// x^(64 (W - 1 - its last word)).  The element's CRC (from 0) is the
// inverted sum; it is then moved past the nodes after the element.
__global__ __launch_bounds__(kCrcThreads) void crc_elm_kernel(CrcArgs a) {
  __shared__ uint64_t T[8][256];   // S(w): T[7 - i][byte i of w]
  __shared__ uint64_t V[8][256];   // V[j][b] = (b at byte j) x^4096
  for (int q = threadIdx.x; q < 8 * 256; q += kCrcThreads) {
    T[q >> 8][q & 255] = a.tab[q];
    V[q >> 8][q & 255] = a.vtab[q];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wid = blockIdx.x * (kCrcThreads / 64) + (threadIdx.x >> 6);
  const int nw = gridDim.x * (kCrcThreads / 64);
  for (int e = wid; e < a.nelem; e += nw) {
    const long noff = a.eoff[e];
    const int n = a.en[e];
    const int np = n * n * n;
    const int W = 5 * np;
    const double *eu = a.u + 5 * noff;
    // word t = lane + 64 i -> segment k (phi, pi, cx, cy, cz), index q
    int k = 0, q = lane;
    while (q >= np) { q -= np; ++k; }
    uint64_t acc = 0;
    for (int t = lane; t < W; t += 64) {
      const int var = (k == 0) ? 4 : k - 1;
      const uint64_t w = __double_as_longlong(eu[(long)var * np + q]);
      q += 64;
      while (q >= np) { q -= np; ++k; }
      const uint64_t sw = T[7][w & 0xff] ^ T[6][(w >> 8) & 0xff] ^ T[5][(w >> 16) & 0xff] ^
                          T[4][(w >> 24) & 0xff] ^ T[3][(w >> 32) & 0xff] ^ T[2][(w >> 40) & 0xff] ^
                          T[1][(w >> 48) & 0xff] ^ T[0][w >> 56];
      const uint64_t sh = V[0][acc & 0xff] ^ V[1][(acc >> 8) & 0xff] ^ V[2][(acc >> 16) & 0xff] ^
                          V[3][(acc >> 24) & 0xff] ^ V[4][(acc >> 32) & 0xff] ^ V[5][(acc >> 40) & 0xff] ^
                          V[6][(acc >> 48) & 0xff] ^ V[7][acc >> 56];
      acc = sh ^ sw;
    }
    uint64_t c = gf2_mul(acc, a.klane[(n - 2) * 64 + lane]);   // 0 for an idle lane
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) c ^= shfl_xor64(c, m);
    c = ~(c ^ a.aw[n - 2]);
    // x^(320 m), m = nodes after this element: one factor per base-256 digit
    const uint64_t after = (uint64_t)(a.nnodes - noff - np);
    uint64_t f = 1ULL << 63;   // x^0
    if (lane < 4) f = a.ypow[lane * 256 + ((after >> (8 * lane)) & 255)];
    f = gf2_mul(f, shfl_xor64(f, 1));
    f = gf2_mul(f, shfl_xor64(f, 2));
    const uint64_t contrib = gf2_mul(c, f);
    if (lane == 0) a.part[e] = contrib;
  }
}

__global__ __launch_bounds__(1024) void crc_xor_kernel(const uint64_t *part, int n, uint64_t *out) {
  __shared__ uint64_t s[16];
  uint64_t v = 0;
  for (int i = threadIdx.x; i < n; i += 1024) v ^= part[i];
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) v ^= shfl_xor64(v, m);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x < 64) {
    v = threadIdx.x < 16 ? s[threadIdx.x] : 0;
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) v ^= shfl_xor64(v, m);
    if (threadIdx.x == 0) out[0] = v;
  }
}

// ---------------------------------------------------------------------------
// The scratch of nmx_vars_crc_noalloc: the CRC tables, the per-element
// contributions and the result word in one static device array, so that
// entry allocates nothing.  Each device holds its own copy for the life of
// the process; its address is taken once per device with that device
// current and kept here.  In host code `scratch` names a host-side stand-in,
// never the device memory: only the recorded addresses are used.
// ---------------------------------------------------------------------------
static __device__ __align__(256) uint64_t scratch[kCrcScratchWords];
constexpr int kMaxDevices = 64;
static uint64_t *g_scratch_ptr[kMaxDevices] = {};

hipError_t crc_scratch_setup(int device, const uint64_t *tables, long nwords) {
  if (device < 0 || device >= kMaxDevices || nwords != kCrcScrPart) return hipErrorInvalidValue;
  if (g_scratch_ptr[device]) return hipSuccess;
  uint64_t *ptr = nullptr;
  hipError_t err = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
  if (err != hipSuccess) return err;
  err = hipMemcpy(ptr, tables, nwords * sizeof(uint64_t), hipMemcpyHostToDevice);
  if (err != hipSuccess) return err;
  g_scratch_ptr[device] = ptr;
  return hipSuccess;
}

uint64_t *crc_scratch(int device) {
  return (device >= 0 && device < kMaxDevices) ? g_scratch_ptr[device] : nullptr;
}

hipError_t launch_vars_crc(const CrcArgs &a, uint64_t *out, hipStream_t s) {
  if (a.nelem <= 0) return hipErrorInvalidValue;
  int grid = (a.nelem + kCrcThreads / 64 - 1) / (kCrcThreads / 64);
  if (grid > 8192) grid = 8192;
  hipLaunchKernelGGL(crc_elm_kernel, dim3(grid), dim3(kCrcThreads), 0, s, a);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(crc_xor_kernel, dim3(1), dim3(1024), 0, s, a.part, a.nelem, out);
  return hipGetLastError();
}

}  // namespace nmx
