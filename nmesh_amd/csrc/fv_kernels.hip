// fv_kernels.hip -- FV WENO5 + LLF for special-relativistic Euler (ideal
// gas) on a periodic grid of 16^3-cell blocks, gfx950, FP64.
//
// Per RK stage:
//   fv_c2p_kernel    primitives (rho, v, p) of every cell: Newton on the
//                    pressure, guessed from the previous stage's value;
//   fv_stage_kernel  one 64-lane workgroup per 8^3-cell tile: for each
//                    direction a lane walks one line of the tile with a
//                    six-cell register window of primitives, reconstructs
//                    both face states with WENO5 (rec1d_p_WENO5 /
//                    rec1d_m_WENO5 with uniform weights, rec1d.c:1190-1352,
//                    u_scale = 1 as in fv.c:141), takes the LLF flux
//                    (numflux1d.c:85-112) and accumulates the flux
//                    difference of each cell in LDS; the epilogue applies the
//                    RK stage combination (RungeKutta.c:19-188).
// The WENO5 formulas are the reference's; two rearrangements save work: the
// nonlinear weights are normalised with one division instead of four, and
// the right-side smoothness indicators of a face are kept for the left side
// of the next face, where the same stencils appear in the same roles.
// The SR-Euler flux, eigenvalues and c2p are textbook (Marti & Mueller,
// Living Rev. Relativ. 6 (2003) 7); the reference has no SR-Euler system.
#include <hip/hip_runtime.h>

#include "fastmath.h"
#include "fv.h"

namespace nmx {

constexpr int kFvB = 16;                     // cells per block edge
constexpr int kFvB3 = kFvB * kFvB * kFvB;
constexpr int kFvT = 8;                      // cells per tile edge
constexpr int kFvT3 = kFvT * kFvT * kFvT;
constexpr double kWeno5Eps = 1e-40;
constexpr int kC2pMaxIt = 100;

// WENO5 smoothness indicators of a 3-cell stencil (x, y, z) by the position
// of the midpoint it reconstructs: right end (role 0), centre (role 1), left
// end (role 2); rec1d_p_WENO5 uses roles 0, 1, 2 on cells im-2.., im-1..,
// im..; rec1d_m_WENO5 uses roles 0, 1, 2 on cells im-1.., im.., im+1..
// (rec1d.c:1195-1300)
__device__ __forceinline__ double beta_r0(double x, double y, double z) {
  const double s1 = x - 2. * y + z, s2 = x - 4. * y + 3. * z;
  return (13. / 12.) * s1 * s1 + 0.25 * s2 * s2;
}
__device__ __forceinline__ double beta_r1(double x, double y, double z) {
  const double s1 = x - 2. * y + z, s2 = x - z;
  return (13. / 12.) * s1 * s1 + 0.25 * s2 * s2;
}
__device__ __forceinline__ double beta_r2(double x, double y, double z) {
  const double s1 = x - 2. * y + z, s2 = 3. * x - 4. * y + z;
  return (13. / 12.) * s1 * s1 + 0.25 * s2 * s2;
}

// Nonlinear combination with one reciprocal: the weights g_k / (beta_k +
// eps)^2 normalised equal  g_k prod_{j != k} (beta_j + eps)^2  normalised
// (for the data here every product stays between 1e-240 and 1e+60, so the
// sum is a normal number for rcp_nr, fastmath.h).
__device__ __forceinline__ double weno5_combine(double g0, double g1, double g2, double b0, double b1,
                                                double b2, double q0, double q1, double q2) {
  const double e0 = b0 + kWeno5Eps, e1 = b1 + kWeno5Eps, e2 = b2 + kWeno5Eps;
  const double E0 = e0 * e0, E1 = e1 * e1, E2 = e2 * e2;
  const double A0 = g0 * (E1 * E2), A1 = g1 * (E0 * E2), A2 = g2 * (E0 * E1);
  return (A0 * q0 + A1 * q1 + A2 * q2) * rcp_nr(A0 + A1 + A2);
}

// midpoint from the left over (a, b, c, d, e) = u[im-2 .. im+2], ideal
// weights (1, 6, 3), stencils (2,-7,11)/6, (-1,5,2)/6, (2,5,-1)/6
__device__ __forceinline__ double weno5_p(double a, double b, double c, double d, double e, double b0,
                                          double b1, double b2) {
  const double o6 = 1. / 6.;
  const double q0 = (2. * o6) * a + (-7. * o6) * b + (11. * o6) * c;
  const double q1 = (-o6) * b + (5. * o6) * c + (2. * o6) * d;
  const double q2 = (2. * o6) * c + (5. * o6) * d + (-o6) * e;
  return weno5_combine(1., 6., 3., b0, b1, b2, q0, q1, q2);
}

// midpoint from the right over (a, b, c, d, e) = u[im-1 .. im+3], ideal
// weights (3, 6, 1) on the stencils starting at im-1, im, im+1
__device__ __forceinline__ double weno5_m(double a, double b, double c, double d, double e, double b0,
                                          double b1, double b2) {
  const double o6 = 1. / 6.;
  const double q2 = (11. * o6) * c + (-7. * o6) * d + (2. * o6) * e;
  const double q1 = (2. * o6) * b + (5. * o6) * c + (-o6) * d;
  const double q0 = (-o6) * a + (5. * o6) * b + (2. * o6) * c;
  return weno5_combine(3., 6., 1., b0, b1, b2, q0, q1, q2);
}

__device__ __forceinline__ void sr_p2c(double gam, const double P[5], double U[5]) {
  const double rho = P[0], vx = P[1], vy = P[2], vz = P[3], p = P[4];
  const double v2 = vx * vx + vy * vy + vz * vz;
  const double W = 1. / sqrt(1. - v2);
  const double h = 1. + gam / (gam - 1.) * p / rho;
  const double rhW2 = rho * h * W * W;
  U[0] = rho * W;
  U[1] = rhW2 * vx;
  U[2] = rhW2 * vy;
  U[3] = rhW2 * vz;
  U[4] = rhW2 - p - U[0];
}

// flux along DIR, conserved state, and max |acoustic eigenvalue|;
// g1 = gam / (gam - 1)
template <int DIR>
__device__ __forceinline__ double sr_flux(double gam, double g1, const double P[5], double F[5], double U[5]) {
  const double rho = P[0], vx = P[1], vy = P[2], vz = P[3], p = P[4];
  const double v2 = vx * vx + vy * vy + vz * vz;
  const double W = rsqrt_nr(1. - v2);
  const double ph = p * rcp_nr(rho);
  const double h = 1. + g1 * ph;
  const double rhW2 = rho * h * W * W;
  U[0] = rho * W;
  U[1] = rhW2 * vx;
  U[2] = rhW2 * vy;
  U[3] = rhW2 * vz;
  U[4] = rhW2 - p - U[0];
  const double vd = P[1 + DIR];
  F[0] = U[0] * vd;
  F[1] = U[1] * vd;
  F[2] = U[2] * vd;
  F[3] = U[3] * vd;
  F[1 + DIR] += p;
  F[4] = (U[4] + p) * vd;
  const double cs2 = gam * ph * rcp_nr(h);
  const double cs = sqrt(cs2);
  const double oden = rcp_nr(1. - v2 * cs2);
  const double root = cs * sqrt((1. - v2) * (1. - v2 * cs2 - vd * vd * (1. - cs2)));
  const double lm = (vd * (1. - cs2) - root) * oden;
  const double lp = (vd * (1. - cs2) + root) * oden;
  return fmax(fabs(lm), fabs(lp));
}

__device__ __forceinline__ bool physical(const double P[5]) {
  const double v2 = P[1] * P[1] + P[2] * P[2] + P[3] * P[3];
  return P[0] > 0. && P[4] > 0. && v2 < 1.;
}

// ---------------------------------------------------------------------------
// c2p
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void fv_c2p_kernel(FvC2pArgs a) {
  const long ncell = (long)a.g.nbx * a.g.nbx * a.g.nbx * kFvB3;
  const double gam = a.g.gam;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += (long)gridDim.x * blockDim.x) {
    const long b = c / kFvB3, q = c % kFvB3;
    const long base = b * 5 * kFvB3 + q;
    const double D = a.U[base], Sx = a.U[base + kFvB3], Sy = a.U[base + 2 * kFvB3],
                 Sz = a.U[base + 3 * kFvB3];
    const double E = a.U[base + 4 * kFvB3] + D;
    const double S2 = Sx * Sx + Sy * Sy + Sz * Sz;
    const double pmin = fmax(sqrt(S2) - E, 0.0);
    double p = a.P[base + 4 * kFvB3];
    if (!(p > pmin)) p = pmin + 1e-10 * E;
    int it;
    // (the reciprocals and 1/sqrt by rcp_nr / rsqrt_nr: 1/W = (1 - v^2) W)
    for (it = 1; it <= kC2pMaxIt; ++it) {
      const double Ep = E + p;
      const double iEp = rcp_nr(Ep);
      const double v2 = S2 * iEp * iEp;
      const double W = rsqrt_nr(1. - v2);
      const double iW = (1. - v2) * W;
      const double f = (gam - 1.) * (E - S2 * iEp - D * iW) - p;
      const double df = (gam - 1.) * v2 * (1. - D * W * iEp) - 1.;
      double pn = p - f * rcp_nr(df);
      if (pn <= pmin) pn = 0.5 * (p + pmin);
      const double dp = fabs(pn - p);
      p = pn;
      if (dp <= a.tol * p) break;
    }
    if (it > kC2pMaxIt) a.bad[b] = 1;
    const double Ep = E + p;
    const double iEp = rcp_nr(Ep);
    const double v2 = S2 * iEp * iEp;
    const double iW = (1. - v2) * rsqrt_nr(1. - v2);
    a.P[base] = D * iW;
    a.P[base + kFvB3] = Sx * iEp;
    a.P[base + 2 * kFvB3] = Sy * iEp;
    a.P[base + 3 * kFvB3] = Sz * iEp;
    a.P[base + 4 * kFvB3] = p;
  }
}

hipError_t launch_fv_c2p(const FvC2pArgs &a, hipStream_t s) {
  const long ncell = (long)a.g.nbx * a.g.nbx * a.g.nbx * kFvB3;
  long grid = (ncell + 255) / 256;
  if (grid > 1 << 20) grid = 1 << 20;
  hipLaunchKernelGGL(fv_c2p_kernel, dim3((unsigned)grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void fv_p2c_kernel(FvGrid g, const double *P, double *U) {
  const long ncell = (long)g.nbx * g.nbx * g.nbx * kFvB3;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += (long)gridDim.x * blockDim.x) {
    const long b = c / kFvB3, q = c % kFvB3;
    const long base = b * 5 * kFvB3 + q;
    double Q[5], V[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) Q[v] = P[base + v * kFvB3];
    sr_p2c(g.gam, Q, V);
#pragma unroll
    for (int v = 0; v < 5; ++v) U[base + v * kFvB3] = V[v];
  }
}

hipError_t launch_fv_p2c(const FvGrid &g, const double *P, double *U, hipStream_t s) {
  const long ncell = (long)g.nbx * g.nbx * g.nbx * kFvB3;
  long grid = (ncell + 255) / 256;
  if (grid > 1 << 20) grid = 1 << 20;
  hipLaunchKernelGGL(fv_p2c_kernel, dim3((unsigned)grid), dim3(256), 0, s, g, P, U);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void fv_layout_kernel(FvGrid g, const double *src, double *dst, int to_blocks) {
  const int N = g.nbx * kFvB;
  const long ncell = (long)N * N * N;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < 5 * ncell;
       idx += (long)gridDim.x * blockDim.x) {
    const int v = (int)(idx / ncell);
    const long c = idx % ncell;
    const int i = (int)(c % N), j = (int)((c / N) % N), k = (int)(c / ((long)N * N));
    const long b = (i / kFvB) + (long)g.nbx * ((j / kFvB) + (long)g.nbx * (k / kFvB));
    const long q = (i % kFvB) + kFvB * ((j % kFvB) + kFvB * (k % kFvB));
    const long bi = b * 5 * kFvB3 + (long)v * kFvB3 + q;
    if (to_blocks) dst[bi] = src[idx];
    else dst[idx] = src[bi];
  }
}

hipError_t launch_fv_layout(const FvGrid &g, const double *src, double *dst, int to_blocks, hipStream_t s) {
  const long n = 5L * g.nbx * g.nbx * g.nbx * kFvB3;
  long grid = (n + 255) / 256;
  if (grid > 1 << 20) grid = 1 << 20;
  hipLaunchKernelGGL(fv_layout_kernel, dim3((unsigned)grid), dim3(256), 0, s, g, src, dst, to_blocks);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// fused RHS + RK stage on 8^3 tiles
// ---------------------------------------------------------------------------

// address of primitive v at global cell (x, y, z), periodic
__device__ __forceinline__ long prim_addr(int nbx, int N, int x, int y, int z) {
  x = x < 0 ? x + N : (x >= N ? x - N : x);
  y = y < 0 ? y + N : (y >= N ? y - N : y);
  z = z < 0 ? z + N : (z >= N ? z - N : z);
  const long b = (x >> 4) + (long)nbx * ((y >> 4) + (long)nbx * (z >> 4));
  return b * 5 * kFvB3 + ((x & 15) + kFvB * ((y & 15) + kFvB * (z & 15)));
}

template <int DIR>
__device__ __forceinline__ void load_cell(const FvStageArgs &a, int N, int x0, int y0, int z0, int ta, int tb,
                                          int s, double out[5]) {
  const int x = x0 + (DIR == 0 ? s : ta);
  const int y = y0 + (DIR == 1 ? s : (DIR == 0 ? ta : tb));
  const int z = z0 + (DIR == 2 ? s : tb);
  const double *p = a.P + prim_addr(a.g.nbx, N, x, y, z);
#pragma unroll
  for (int v = 0; v < 5; ++v) out[v] = p[v * kFvB3];
}

// one direction: lane (ta, tb) walks the tile line along DIR; FIRST writes
// the accumulator, later directions add to it
template <int DIR, bool FIRST>
__device__ __forceinline__ void fv_sweep(const FvStageArgs &a, int N, int x0, int y0, int z0,
                                         double (*acc)[kFvT3]) {
  const int lane = threadIdx.x;
  const int ta = lane % kFvT, tb = lane / kFvT;
  const double gam = a.g.gam;
  double q[5][6], nx[5];
  {
    double t[5];
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      load_cell<DIR>(a, N, x0, y0, z0, ta, tb, s - 3, t);
#pragma unroll
      for (int v = 0; v < 5; ++v) q[v][s] = t[v];
    }
    load_cell<DIR>(a, N, x0, y0, z0, ta, tb, 3, nx);
  }
  // right-side smoothness indicators of the previous face = left-side ones
  // of this face (same stencils, same roles)
  double bm[5][3];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    bm[v][0] = beta_r0(q[v][0], q[v][1], q[v][2]);
    bm[v][1] = beta_r1(q[v][1], q[v][2], q[v][3]);
    bm[v][2] = beta_r2(q[v][2], q[v][3], q[v][4]);
  }
  const double g1 = gam / (gam - 1.);
  double Fprev[5] = {0, 0, 0, 0, 0};
#pragma unroll 1
  for (int f = 0; f <= kFvT; ++f) {   // face between tile cells f-1 and f
    if (f > 0) {
#pragma unroll
      for (int v = 0; v < 5; ++v) {
#pragma unroll
        for (int s = 0; s < 5; ++s) q[v][s] = q[v][s + 1];
        q[v][5] = nx[v];
      }
      if (f < kFvT) load_cell<DIR>(a, N, x0, y0, z0, ta, tb, f + 3, nx);
    }
    double PL[5], PR[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const double bp0 = bm[v][0], bp1 = bm[v][1], bp2 = bm[v][2];
      bm[v][0] = beta_r0(q[v][1], q[v][2], q[v][3]);
      bm[v][1] = beta_r1(q[v][2], q[v][3], q[v][4]);
      bm[v][2] = beta_r2(q[v][3], q[v][4], q[v][5]);
      PL[v] = weno5_p(q[v][0], q[v][1], q[v][2], q[v][3], q[v][4], bp0, bp1, bp2);
      PR[v] = weno5_m(q[v][1], q[v][2], q[v][3], q[v][4], q[v][5], bm[v][0], bm[v][1], bm[v][2]);
    }
    if (!physical(PL)) {
#pragma unroll
      for (int v = 0; v < 5; ++v) PL[v] = q[v][2];
    }
    if (!physical(PR)) {
#pragma unroll
      for (int v = 0; v < 5; ++v) PR[v] = q[v][3];
    }
    double FL[5], FR[5], UL[5], UR[5];
    const double aL = sr_flux<DIR>(gam, g1, PL, FL, UL);
    const double aR = sr_flux<DIR>(gam, g1, PR, FR, UR);
    const double am = fmax(aL, aR);
    double F[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) F[v] = 0.5 * (FL[v] + FR[v] - am * (UR[v] - UL[v]));
    if (f > 0) {
      const int s = f - 1;
      const int cx = DIR == 0 ? s : ta;
      const int cy = DIR == 1 ? s : (DIR == 0 ? ta : tb);
      const int cz = DIR == 2 ? s : tb;
      const int cell = cx + kFvT * (cy + kFvT * cz);
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        if (FIRST) acc[v][cell] = Fprev[v] - F[v];
        else acc[v][cell] += Fprev[v] - F[v];
      }
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) Fprev[v] = F[v];
  }
}

__global__ __launch_bounds__(64, 2) void fv_stage_kernel(FvStageArgs a) {
  __shared__ double acc[5][kFvT3];
  const int nbx = a.g.nbx, N = nbx * kFvB;
  const int ntiles = nbx * nbx * nbx * 8;
  int t_hi, t_step, t_first;
  if (a.xcd_map) {
    const int xcd = blockIdx.x & 7, g8 = gridDim.x >> 3;
    const int chunk = (ntiles + 7) >> 3;
    t_first = xcd * chunk + (blockIdx.x >> 3);
    t_hi = min((xcd + 1) * chunk, ntiles);
    t_step = g8;
  } else {
    t_first = blockIdx.x;
    t_hi = ntiles;
    t_step = gridDim.x;
  }
  for (int t = t_first; t < t_hi; t += t_step) {
    const int b = t >> 3, sub = t & 7;
    const int bx = b % nbx, by = (b / nbx) % nbx, bz = b / (nbx * nbx);
    const int x0 = bx * kFvB + (sub & 1) * kFvT;
    const int y0 = by * kFvB + ((sub >> 1) & 1) * kFvT;
    const int z0 = bz * kFvB + (sub >> 2) * kFvT;
    __syncthreads();
    fv_sweep<0, true>(a, N, x0, y0, z0, acc);
    __syncthreads();
    fv_sweep<1, false>(a, N, x0, y0, z0, acc);
    __syncthreads();
    fv_sweep<2, false>(a, N, x0, y0, z0, acc);
    __syncthreads();
    // stage combination, 8 cells per lane
    const int lane = threadIdx.x;
    const int cx = lane % kFvT, cy = lane / kFvT;
#pragma unroll 1
    for (int cz = 0; cz < kFvT; ++cz) {
      const int cell = cx + kFvT * (cy + kFvT * cz);
      const long gi = (long)b * 5 * kFvB3 + ((x0 & 15) + cx) + kFvB * (((y0 & 15) + cy) + kFvB * ((z0 & 15) + cz));
      double r[5], Ubuf[5], Abuf[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) r[v] = acc[v][cell];
      if (a.wmode) {
#pragma unroll
        for (int v = 0; v < 5; ++v) Ubuf[v] = a.U[gi + v * kFvB3];
      }
      if (a.Aout) {
        const double *bas = a.Ain ? a.Ain : a.win;
#pragma unroll
        for (int v = 0; v < 5; ++v) Abuf[v] = bas[gi + v * kFvB3];
      }
      if (a.R) {
#pragma unroll
        for (int v = 0; v < 5; ++v) a.R[gi + v * kFvB3] = r[v];
      }
      if (a.Aout) {
#pragma unroll
        for (int v = 0; v < 5; ++v) a.Aout[gi + v * kFvB3] = Abuf[v] + a.cA * r[v];
      }
      if (a.wmode == 1) {
#pragma unroll
        for (int v = 0; v < 5; ++v) a.Wout[gi + v * kFvB3] = Ubuf[v] + a.cW * r[v];
      } else if (a.wmode == 2) {
#pragma unroll
        for (int v = 0; v < 5; ++v)
          a.Wout[gi + v * kFvB3] = 0.75 * Ubuf[v] + 0.25 * (a.win[gi + v * kFvB3] + a.cW * r[v]);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// CFL: the largest |lambda+-| over cells and directions (the eigenvalues of
// sr_flux), reduced with lane shuffles, then per workgroup and over the grid
// ---------------------------------------------------------------------------
constexpr int kSpeedThreads = 256;

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, 64));
  return v;
}

__device__ __forceinline__ double sr_speed(double gam, double g1, const double P[5], int d) {
  const double rho = P[0], p = P[4];
  const double v2 = P[1] * P[1] + P[2] * P[2] + P[3] * P[3];
  const double ph = p / rho;
  const double h = 1. + g1 * ph;
  const double cs2 = gam * ph / h;
  const double cs = sqrt(cs2);
  const double vd = P[1 + d];
  const double oden = 1. / (1. - v2 * cs2);
  const double root = cs * sqrt((1. - v2) * (1. - v2 * cs2 - vd * vd * (1. - cs2)));
  const double lm = (vd * (1. - cs2) - root) * oden;
  const double lp = (vd * (1. - cs2) + root) * oden;
  return fmax(fabs(lm), fabs(lp));
}

__global__ __launch_bounds__(kSpeedThreads) void fv_speed_kernel(FvGrid g, const double *P, double *part) {
  __shared__ double s_max[kSpeedThreads / 64];
  const long ncell = (long)g.nbx * g.nbx * g.nbx * kFvB3;
  const double gam = g.gam, g1 = gam / (gam - 1.);
  double m = 0.0;
  for (long c = (long)blockIdx.x * kSpeedThreads + threadIdx.x; c < ncell; c += (long)gridDim.x * kSpeedThreads) {
    const long b = c / kFvB3, q = c % kFvB3;
    const double *pc = P + b * 5 * kFvB3 + q;
    const double Pc[5] = {pc[0], pc[kFvB3], pc[2 * kFvB3], pc[3 * kFvB3], pc[4 * kFvB3]};
#pragma unroll
    for (int d = 0; d < 3; ++d) m = fmax(m, sr_speed(gam, g1, Pc, d));
  }
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = s_max[0];
    for (int w = 1; w < kSpeedThreads / 64; ++w) r = fmax(r, s_max[w]);
    part[blockIdx.x] = r;
  }
}

__global__ __launch_bounds__(kSpeedThreads) void fv_speed_max_kernel(const double *part, int n, double *out) {
  __shared__ double s_max[kSpeedThreads / 64];
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += kSpeedThreads) m = fmax(m, part[i]);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = s_max[0];
    for (int w = 1; w < kSpeedThreads / 64; ++w) r = fmax(r, s_max[w]);
    out[0] = r;
  }
}

int fv_speed_grid(const FvGrid &g) {
  const long ncell = (long)g.nbx * g.nbx * g.nbx * kFvB3;
  const long gr = (ncell + kSpeedThreads - 1) / kSpeedThreads;
  return gr < 4096 ? (int)gr : 4096;
}

hipError_t launch_fv_max_speed(const FvGrid &g, const double *P, double *part, double *out, hipStream_t s) {
  const int grid = fv_speed_grid(g);
  hipLaunchKernelGGL(fv_speed_kernel, dim3(grid), dim3(kSpeedThreads), 0, s, g, P, part);
  hipLaunchKernelGGL(fv_speed_max_kernel, dim3(1), dim3(kSpeedThreads), 0, s, part, grid, out);
  return hipGetLastError();
}

hipError_t launch_fv_stage(const FvStageArgs &a, hipStream_t s) {
  const int ntiles = a.g.nbx * a.g.nbx * a.g.nbx * 8;
  const int grid = a.xcd_map ? 8 * ((ntiles + 7) / 8) : ntiles;
  hipLaunchKernelGGL(fv_stage_kernel, dim3(grid), dim3(64), 0, s, a);
  return hipGetLastError();
}

}  // namespace nmx
