// dg_common.h -- device code shared by the plain and the batched DG stage
// kernels: numerical flux, boundary profile, cubed-shell Jacobian, and the
// per-node RHS + RK stage update from LDS-staged element data.
#pragma once
#include <hip/hip_runtime.h>

namespace nmx {

// Tables that no kernel writes are read through the constant address space:
// wave-uniform reads from them become scalar loads issued together, instead
// of a chain of vector loads each waiting for the last.
#define NMX_CONST __attribute__((address_space(4)))
template <class T>
__device__ __forceinline__ const NMX_CONST T *as_const(const T *p) {
  return (const NMX_CONST T *)p;
}

// upwind: F = fnum - f_inner for the 4 flux-carrying fields.  With A = f.n
// Jacobian [[0,-n^T],[-n,0]], A+/- = (A +- |A|)/2 gives
// F_pi = (dpi + n.dc)/2 and F_c = n F_pi with d = inner - adjacent
// (scalarwave1_numflux1d_upwind, scalarwave1.c:324-499).
// LLF (amax = 1): F_pi = (dpi + n.dc)/2, F_c = (dc + n dpi)/2
// (numflux1d_LLF, numflux1d.c:85-112).
__device__ __forceinline__ void face_flux(int numflux, const double n[3], double dpi,
                                          const double dc[3], double F[4]) {
  double ndc = n[0] * dc[0] + n[1] * dc[1] + n[2] * dc[2];
  double Fp = 0.5 * (dpi + ndc);
  F[0] = Fp;
  if (numflux == 0) {
    F[1] = n[0] * Fp; F[2] = n[1] * Fp; F[3] = n[2] * Fp;
  } else {
    F[1] = 0.5 * (dc[0] + n[0] * dpi);
    F[2] = 0.5 * (dc[1] + n[1] * dpi);
    F[3] = 0.5 * (dc[2] + n[2] * dpi);
  }
}

// plane wave of scalarwave1_set_profile_pt (scalarwave1.c:503-524)
__device__ __forceinline__ void plane_wave(const double k[3], double om, double x, double y,
                                           double z, double t, double u[4]) {
  double arg = k[0] * x + k[1] * y + k[2] * z - om * t;
  double s, c;
  sincos(arg, &s, &c);
  (void)s;
  u[0] = -om * c; u[1] = k[0] * c; u[2] = k[1] * c; u[3] = k[2] * c;
}

// dX/dx of a CubedShell (constant radii s0, s1, centre 0) at patch point X
// (lam, A, B) (coordtrans_CubedSphere.c:259-509 with dsigma = 0).  With
// s = sqrt(1 + A^2 + B^2), the radius-like coordinate r = (pm/s)(s0 + (s1 -
// s0) lam) along the patch axis dir and x_dir = r, x_ia = A r, x_ib = B r,
// the reference's rows reduce to
//   dlam/dx = pm / ((s1 - s0) s) (1, A, B),
//   dA/dx   = (pm s / R) (-A, 1, 0),   dB/dx = (pm s / R) (-B, 0, 1),
// R = s0 + (s1 - s0) lam, components ordered (dir, ia, ib): two divisions
// and one square root per point instead of nine divisions.  Written with
// compile-time component indices only: runtime-indexed private arrays would
// live in scratch memory.
__device__ __forceinline__ void shell_dXdx(int dom, double s0, double s1, const double X[3],
                                           double M[3][3]) {
  const double lam = X[0], A = X[1], B = X[2];
  const double pm = (dom & 1) ? 1.0 : -1.0;
  const double s = sqrt(1.0 + A * A + B * B);
  const double c0 = pm / ((s1 - s0) * s);
  const double c1 = pm * s / (s0 + (s1 - s0) * lam);
  const int dir = dom >> 1;
  const int ia = (dir == 1) ? 0 : 1;      // axis scaled by A
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const bool isr = (d == dir), isa = (d == ia);
    M[0][d] = isr ? c0 : (isa ? A * c0 : B * c0);
    M[1][d] = isr ? -A * c1 : (isa ? c1 : 0.0);
    M[2][d] = isr ? -B * c1 : (isa ? 0.0 : c1);
  }
}

// det of a 3x3 matrix (cofactor expansion along the first row)
__device__ __forceinline__ double det3(const double M[3][3]) {
  return M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
         M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
         M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
}

// Everything the per-node update needs besides LDS.
struct NodeOut {
  const double *win, *U, *Ain;
  double *Aout, *Wout, *R;
  double cA, cW, wend_inv;
  int wmode, numflux;
  // streaming hints for the registers each node touches once per stage:
  // bit 0 loads of u_p and the accumulator base, bit 1 stores of the
  // accumulator and the RHS, bit 2 stores of the next stage input
  int nt = 0;
};

__device__ __forceinline__ double ld_stream(const double *p, bool nt) {
  return nt ? __builtin_nontemporal_load(p) : *p;
}
__device__ __forceinline__ void st_stream(double *p, double v, bool nt) {
  if (nt) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// Geometry of one element.
// outward unit normal and lift scale sqrtgdiag / w_end of face f at a node
// with metric M (dX/dx; unused on Cartesian elements, GEOM 0), as the
// stage kernels compute them (dg.c:90-146, utilities.c:48-81)
template <int GEOM>
__device__ __forceinline__ void face_normal(const double M[3][3], const double *dXbdX, int f,
                                            double wend_inv, double nrm[3], double &scale) {
  const int dir = f >> 1;
  const double sig = (f & 1) ? 1.0 : -1.0;
  if (GEOM == 0) {
    nrm[0] = (dir == 0) ? sig : 0.0;
    nrm[1] = (dir == 1) ? sig : 0.0;
    nrm[2] = (dir == 2) ? sig : 0.0;
    scale = dXbdX[dir] * wend_inv;
  } else {
    const double r0 = (dir == 0) ? M[0][0] : (dir == 1) ? M[1][0] : M[2][0];
    const double r1 = (dir == 0) ? M[0][1] : (dir == 1) ? M[1][1] : M[2][1];
    const double r2 = (dir == 0) ? M[0][2] : (dir == 1) ? M[1][2] : M[2][2];
    const double s = sqrt(r0 * r0 + r1 * r1 + r2 * r2);
    const double ism = 1.0 / (sig * s);
    nrm[0] = r0 * ism; nrm[1] = r1 * ism; nrm[2] = r2 * ism;
    scale = dXbdX[dir] * s * wend_inv;
  }
}

struct ElmGeo {
  double dXbdX[3];
  const double *dXdx;   // GEOM 1: element's 9*np block
  double bb[6];         // GEOM 2: element box in patch coordinates
  double s0, s1;
  int dom;
};

// RHS of scalarwave1 at one node and the RK stage combination.
// s_u[v][np3]: own pi,cx,cy,cz; s_f[v][f][np2]: adjacent values on face f;
// gbase: element's first state double; node: ijk.
// Upre/Apre: the epilogue's u_p and accumulator base at this node, loaded
// earlier by the caller (nullptr: loaded here)
template <int NP, int GEOM>
__device__ __forceinline__ void node_update(const NodeOut &o, const ElmGeo &g,
                                            const double (*s_u)[NP * NP * NP],
                                            const double (*s_f)[6][NP * NP],
                                            const double *s_D, const double *s_xb, long gbase,
                                            int node, const double *Upre = nullptr,
                                            const double *Apre = nullptr) {
  constexpr int NP2 = NP * NP;
  constexpr int NP3 = NP * NP * NP;
  constexpr bool CURV = (GEOM != 0);
  const int k = node / NP2;
  const int j = (node / NP) % NP;
  const int i = node % NP;
  const long gi = gbase + node;

  double M[3][3];
  if (GEOM == 1) {
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int q = 0; q < 3; ++q) M[m][q] = g.dXdx[(3 * m + q) * NP3 + node];
  } else if (GEOM == 2) {
    const int I[3] = {i, j, k};
    double X[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
      X[d] = 0.5 * ((g.bb[2 * d + 1] - g.bb[2 * d]) * s_xb[I[d]] + (g.bb[2 * d + 1] + g.bb[2 * d]));
    shell_dXdx(g.dom, g.s0, g.s1, X, M);
  }

  // tensor-product derivative along one direction m at a time
  // (mm_array_indir, array.c:100-300), folded at once into the Cartesian
  // partials d_q f = sum_m dXdx[m][q] dXbdX[m] D_m f (derivs.c:102-181):
  // r_pi = d_q c_q, r_cq = d_q pi
  double rpi = 0.0, rc[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const int im = (m == 0) ? i : (m == 1) ? j : k;
    const int stride = (m == 0) ? 1 : (m == 1) ? NP : NP2;
    const int base = node - im * stride;
    double dp = 0.0, d0 = 0.0, d1 = 0.0, d2 = 0.0;
#pragma unroll
    for (int l = 0; l < NP; ++l) {
      const double Dl = s_D[im * NP + l];
      const int src = base + l * stride;
      dp += Dl * s_u[0][src];
      if (CURV) {
        d0 += Dl * s_u[1][src];
        d1 += Dl * s_u[2][src];
        d2 += Dl * s_u[3][src];
      } else {
        d0 += Dl * s_u[1 + m][src];
      }
    }
    if (CURV) {
      const double sp = g.dXbdX[m] * dp;
      rpi += M[m][0] * (g.dXbdX[m] * d0) + M[m][1] * (g.dXbdX[m] * d1) + M[m][2] * (g.dXbdX[m] * d2);
      rc[0] += M[m][0] * sp;
      rc[1] += M[m][1] * sp;
      rc[2] += M[m][2] * sp;
    } else {
      rpi += g.dXbdX[m] * d0;
      rc[m] = g.dXbdX[m] * dp;
    }
  }

  // numerical flux and strong-form lift on the faces this node lies on
  // (dg_add_surface_fluxes_sign_fvflag, dg.c:90-146; normal dg.c:470-481 via
  // node_normal_from_idXdx_at_ijk, utilities.c:48-81)
  const int I[3] = {i, j, k};
#pragma unroll
  for (int dir = 0; dir < 3; ++dir) {
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      if (I[dir] != side * (NP - 1)) continue;
      const int f = 2 * dir + side;
      const double sig = side ? 1.0 : -1.0;
      double nrm[3], scale;
      if (CURV) {
        double s2 = M[dir][0] * M[dir][0] + M[dir][1] * M[dir][1] + M[dir][2] * M[dir][2];
        double s = sqrt(s2);
        const double ism = 1.0 / (sig * s);
        nrm[0] = M[dir][0] * ism; nrm[1] = M[dir][1] * ism; nrm[2] = M[dir][2] * ism;
        scale = g.dXbdX[dir] * s * o.wend_inv;
      } else {
        nrm[0] = nrm[1] = nrm[2] = 0.0;
        nrm[dir] = sig;
        scale = g.dXbdX[dir] * o.wend_inv;
      }
      const int ia = (dir == 0) ? j : i;          // index along Dir1_norm(dir)
      const int ib = (dir == 2) ? j : k;          // index along Dir2_norm(dir)
      const int q = ia + NP * ib;
      const double dcv[3] = {s_u[1][node] - s_f[1][f][q], s_u[2][node] - s_f[2][f][q],
                             s_u[3][node] - s_f[3][f][q]};
      double F[4];
      face_flux(o.numflux, nrm, s_u[0][node] - s_f[0][f][q], dcv, F);
      rpi -= F[0] * scale;
      rc[0] -= F[1] * scale;
      rc[1] -= F[2] * scale;
      rc[2] -= F[3] * scale;
    }
  }

  // r = -(div f + lift), r_phi = pi (scalarwave1.c:242-321) and the stage
  // combination (RungeKutta.c:19-188, variables.c:932-1041)
  const double r[5] = {rpi, rc[0], rc[1], rc[2], s_u[0][node]};
  const bool ntl = o.nt & 1, nta = o.nt & 2, ntw = o.nt & 4;
  double Ubuf[5], Abuf[5];
  if (o.wmode) {
#pragma unroll
    for (int v = 0; v < 5; ++v) Ubuf[v] = Upre ? Upre[v] : ld_stream(o.U + gi + v * NP3, ntl);
  }
  if (o.Aout) {
    const double *bas = o.Ain ? o.Ain : o.win;
#pragma unroll
    for (int v = 0; v < 5; ++v) Abuf[v] = Apre ? Apre[v] : ld_stream(bas + gi + v * NP3, ntl);
  }
  if (o.R) {
#pragma unroll
    for (int v = 0; v < 5; ++v) st_stream(o.R + gi + v * NP3, r[v], nta);
  }
  if (o.Aout) {
#pragma unroll
    for (int v = 0; v < 5; ++v) st_stream(o.Aout + gi + v * NP3, Abuf[v] + o.cA * r[v], nta);
  }
  if (o.wmode == 1) {
#pragma unroll
    for (int v = 0; v < 5; ++v) st_stream(o.Wout + gi + v * NP3, Ubuf[v] + o.cW * r[v], ntw);
  } else if (o.wmode == 2) {
    const double wv[5] = {s_u[0][node], s_u[1][node], s_u[2][node], s_u[3][node],
                          o.win[gi + 4 * NP3]};
#pragma unroll
    for (int v = 0; v < 5; ++v) st_stream(o.Wout + gi + v * NP3, 0.75 * Ubuf[v] + 0.25 * (wv[v] + o.cW * r[v]), ntw);
  }
}

}  // namespace nmx
