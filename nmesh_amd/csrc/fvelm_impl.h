// fvelm_impl.h -- scalarwave1 RHS + RK stage on FV elements (uniform
// points including both faces) of a Cartesian or CubedShell mesh, gfx950.
// Instantiated by fvelm_kernels.hip (dnfn_extrap1 without the s1/s2/opt
// limit) and fvelm_ex_kernels.hip (with it), which compile in parallel.
//
// nmesh's fv_divf (src/main/dg/fv.c:115-624) with scalarwave1's callbacks
// (src/tests/scalarwave1/scalarwave1.c:211-237,755-790) and its default
// parameters (src/main/dg/nmesh_dg.c:26-66, dg.c:285-400): fv_flux = fnum,
// fv_divf_extrap = dnfn_extrap1 (so the face fluxes are set here and
// dg_add_surface_fluxes adds nothing), fv_divf_use_only_right_flux = yes,
// WENOm3 optimal weights (1, 2).  Per element (one workgroup):
//   phase 1: own pi, c to LDS; the adjacent face values (ajsurf) to LDS;
//   phase 2: one task per (direction, line): reconstruct every midpoint
//            from both sides (WENOm5_2, WENOmZ_2 or WENOm3_2, rec1d.c:
//            115-232,875-1370,1627-1798), take upwind / LLF fnum with the
//            midpoint normal,
//            the face fluxes against ajsurf at both ends,
//            df_i = Jg_R fnumR_i - Jg_L fnumR_{i-1} with Jg = sqrtgdiag /
//            (ooJ dXb_i), dnfn_extrap1 at both ends (rec1d.c:542-626) and
//            divf_d = df ooJ, one LDS slab per direction;
//   epilogue: r_{pi,c} = -(divf_x + divf_y + divf_z), r_phi = pi, RK stage.
// On shell elements the metric factors and normals are recomputed at each
// node and midpoint of a line (no stored geometry is read).
#pragma once
#include <hip/hip_runtime.h>

#include "amr.h"
#include "dg_common.h"
#include "fastmath.h"
#include "kernels.h"

namespace nmx {

#define NMX_CONST __attribute__((address_space(4)))

constexpr double kWeno3Eps = 1e-6;    // WENO3_epsilon (rec1d.c:35), u_scale = 1
constexpr double kWeno5EpsE = 1e-40;  // WENO5_epsilon (rec1d.c:1193)

// The reconstructions below normalise their nonlinear weights with one
// reciprocal: a weight d_k / b_k^p over their sum equals (d_k prod_{j!=k}
// b_j^p) over the sum of those products, whose reciprocal is rcp_nr's
// (fastmath.h; the smoothness indicators are at least the epsilon, 1e-6 or
// 1e-40, so the products stay normal numbers).  Only rounding differs from
// the reference's per-weight divisions.

// WENOm3 (rec1d_p_WENO3 with the uniform WENOm3 weights)
__device__ __forceinline__ double wm3_p(double um1, double u0, double up1) {
  const double d1 = u0 - um1, d2 = up1 - u0;
  const double b1 = d1 * d1 + kWeno3Eps, b2 = d2 * d2 + kWeno3Eps;
  const double p1 = b2 * b2, p2 = 2. * (b1 * b1);   // weights 1/b1^2, 2/b2^2
  return (p1 * (-0.5 * um1 + 1.5 * u0) + p2 * (0.5 * u0 + 0.5 * up1)) * rcp_nr(p1 + p2);
}
__device__ __forceinline__ double wm3_m(double u0, double up1, double up2) {
  const double d1 = up1 - u0, d2 = up2 - up1;
  const double b1 = d1 * d1 + kWeno3Eps, b2 = d2 * d2 + kWeno3Eps;
  const double p1 = 2. * (b2 * b2), p2 = b1 * b1;   // weights 2/b1^2, 1/b2^2
  return (p1 * (0.5 * u0 + 0.5 * up1) + p2 * (1.5 * up1 - 0.5 * up2)) * rcp_nr(p1 + p2);
}

// WENO5 (rec1d_p/m_WENO5 with uniform weights): ideal weights (1, 6, 3)
// over the squared smoothness indicators plus 1e-40
__device__ __forceinline__ double w5_p(double a, double b, double c, double d, double e) {
  const double s01 = a - 2. * b + c, s02 = a - 4. * b + 3. * c;
  const double s11 = b - 2. * c + d, s12 = b - d;
  const double s21 = c - 2. * d + e, s22 = 3. * c - 4. * d + e;
  const double e0 = (13. / 12.) * s01 * s01 + 0.25 * s02 * s02 + kWeno5EpsE;
  const double e1 = (13. / 12.) * s11 * s11 + 0.25 * s12 * s12 + kWeno5EpsE;
  const double e2 = (13. / 12.) * s21 * s21 + 0.25 * s22 * s22 + kWeno5EpsE;
  const double E0 = e0 * e0, E1 = e1 * e1, E2 = e2 * e2;
  const double p0 = E1 * E2, p1 = 6. * (E0 * E2), p2 = 3. * (E0 * E1);
  const double o6 = 1. / 6.;
  const double q0 = (2. * o6) * a + (-7. * o6) * b + (11. * o6) * c;
  const double q1 = (-o6) * b + (5. * o6) * c + (2. * o6) * d;
  const double q2 = (2. * o6) * c + (5. * o6) * d + (-o6) * e;
  return (p0 * q0 + p1 * q1 + p2 * q2) * rcp_nr(p0 + p1 + p2);
}
__device__ __forceinline__ double w5_m(double a, double b, double c, double d, double e) {
  const double s21 = c - 2. * d + e, s22 = 3. * c - 4. * d + e;
  const double s11 = b - 2. * c + d, s12 = b - d;
  const double s01 = a - 2. * b + c, s02 = a - 4. * b + 3. * c;
  const double e2 = (13. / 12.) * s21 * s21 + 0.25 * s22 * s22 + kWeno5EpsE;
  const double e1 = (13. / 12.) * s11 * s11 + 0.25 * s12 * s12 + kWeno5EpsE;
  const double e0 = (13. / 12.) * s01 * s01 + 0.25 * s02 * s02 + kWeno5EpsE;
  const double E0 = e0 * e0, E1 = e1 * e1, E2 = e2 * e2;
  const double p2 = E0 * E1, p1 = 6. * (E0 * E2), p0 = 3. * (E1 * E2);
  const double o6 = 1. / 6.;
  const double q2 = (11. * o6) * c + (-7. * o6) * d + (2. * o6) * e;
  const double q1 = (2. * o6) * b + (5. * o6) * c + (-o6) * d;
  const double q0 = (-o6) * a + (5. * o6) * b + (2. * o6) * c;
  return (p0 * q0 + p1 * q1 + p2 * q2) * rcp_nr(p0 + p1 + p2);
}

// WENOT (rec1d_p/m_WENOT with uniform weights, rec1d.c:1412-1578): the
// weights d_k (1 + |b0 - b2|) / (b_k + 1e-40) normalize to d_k / g_k over
// the sum, i.e. WENO5 with the smoothness indicators to the first power
__device__ __forceinline__ double wt_p(double a, double b, double c, double d, double e) {
  const double s01 = a - 2. * b + c, s02 = a - 4. * b + 3. * c;
  const double s11 = b - 2. * c + d, s12 = b - d;
  const double s21 = c - 2. * d + e, s22 = 3. * c - 4. * d + e;
  const double g0 = (13. / 12.) * s01 * s01 + 0.25 * s02 * s02 + kWeno5EpsE;
  const double g1 = (13. / 12.) * s11 * s11 + 0.25 * s12 * s12 + kWeno5EpsE;
  const double g2 = (13. / 12.) * s21 * s21 + 0.25 * s22 * s22 + kWeno5EpsE;
  const double p0 = g1 * g2, p1 = 6. * (g0 * g2), p2 = 3. * (g0 * g1);
  const double o6 = 1. / 6.;
  const double q0 = (2. * o6) * a + (-7. * o6) * b + (11. * o6) * c;
  const double q1 = (-o6) * b + (5. * o6) * c + (2. * o6) * d;
  const double q2 = (2. * o6) * c + (5. * o6) * d + (-o6) * e;
  return (p0 * q0 + p1 * q1 + p2 * q2) * rcp_nr(p0 + p1 + p2);
}
__device__ __forceinline__ double wt_m(double a, double b, double c, double d, double e) {
  const double s21 = c - 2. * d + e, s22 = 3. * c - 4. * d + e;
  const double s11 = b - 2. * c + d, s12 = b - d;
  const double s01 = a - 2. * b + c, s02 = a - 4. * b + 3. * c;
  const double g2 = (13. / 12.) * s21 * s21 + 0.25 * s22 * s22 + kWeno5EpsE;
  const double g1 = (13. / 12.) * s11 * s11 + 0.25 * s12 * s12 + kWeno5EpsE;
  const double g0 = (13. / 12.) * s01 * s01 + 0.25 * s02 * s02 + kWeno5EpsE;
  const double p2 = g0 * g1, p1 = 6. * (g0 * g2), p0 = 3. * (g1 * g2);
  const double o6 = 1. / 6.;
  const double q2 = (11. * o6) * c + (-7. * o6) * d + (2. * o6) * e;
  const double q1 = (2. * o6) * b + (5. * o6) * c + (-o6) * d;
  const double q0 = (-o6) * a + (5. * o6) * b + (2. * o6) * c;
  return (p0 * q0 + p1 * q1 + p2 * q2) * rcp_nr(p0 + p1 + p2);
}

// WENO-Z (rec1d_p/m_WENOZ with uniform weights, rec1d.c:1627-1781): weights
// d_k (1 + tau5 / (beta_k + 1e-40)), tau5 = |beta0 - beta2|
__device__ __forceinline__ double wz_p(double a, double b, double c, double d, double e) {
  const double s01 = a - 2. * b + c, s02 = a - 4. * b + 3. * c;
  const double s11 = b - 2. * c + d, s12 = b - d;
  const double s21 = c - 2. * d + e, s22 = 3. * c - 4. * d + e;
  const double g0 = (13. / 12.) * s01 * s01 + 0.25 * s02 * s02 + kWeno5EpsE;
  const double g1 = (13. / 12.) * s11 * s11 + 0.25 * s12 * s12 + kWeno5EpsE;
  const double g2 = (13. / 12.) * s21 * s21 + 0.25 * s22 * s22 + kWeno5EpsE;
  const double tau5 = fabs(g0 - g2);
  // d_k (g_k + tau5) / g_k over the sum, times g0 g1 g2
  const double p0 = (g0 + tau5) * (g1 * g2), p1 = 6. * ((g1 + tau5) * (g0 * g2)),
               p2 = 3. * ((g2 + tau5) * (g0 * g1));
  const double o6 = 1. / 6.;
  const double q0 = (2. * o6) * a + (-7. * o6) * b + (11. * o6) * c;
  const double q1 = (-o6) * b + (5. * o6) * c + (2. * o6) * d;
  const double q2 = (2. * o6) * c + (5. * o6) * d + (-o6) * e;
  return (p0 * q0 + p1 * q1 + p2 * q2) * rcp_nr(p0 + p1 + p2);
}
__device__ __forceinline__ double wz_m(double a, double b, double c, double d, double e) {
  const double s21 = c - 2. * d + e, s22 = 3. * c - 4. * d + e;
  const double s11 = b - 2. * c + d, s12 = b - d;
  const double s01 = a - 2. * b + c, s02 = a - 4. * b + 3. * c;
  const double g2 = (13. / 12.) * s21 * s21 + 0.25 * s22 * s22 + kWeno5EpsE;
  const double g1 = (13. / 12.) * s11 * s11 + 0.25 * s12 * s12 + kWeno5EpsE;
  const double g0 = (13. / 12.) * s01 * s01 + 0.25 * s02 * s02 + kWeno5EpsE;
  const double tau5 = fabs(g0 - g2);
  const double p2 = (g2 + tau5) * (g0 * g1), p1 = 6. * ((g1 + tau5) * (g0 * g2)),
               p0 = 3. * ((g0 + tau5) * (g1 * g2));
  const double o6 = 1. / 6.;
  const double q2 = (11. * o6) * c + (-7. * o6) * d + (2. * o6) * e;
  const double q1 = (2. * o6) * b + (5. * o6) * c + (-o6) * d;
  const double q0 = (-o6) * a + (5. * o6) * b + (2. * o6) * c;
  return (p0 * q0 + p1 * q1 + p2 * q2) * rcp_nr(p0 + p1 + p2);
}

// both sides of midpoint im of a line q[0..NP-1]: WENOm5_2 (REC 5),
// WENOmZ_2 (REC 7) or WENOm3_2 (REC 3)
template <int NP, int REC>
__device__ __forceinline__ void rec_mid(const double (&q)[NP], int im, double &up, double &um) {
  if (REC == 5 && im >= 2 && im <= NP - 3) up = w5_p(q[im - 2], q[im - 1], q[im], q[im + 1], q[im + 2]);
  else if (REC == 7 && im >= 2 && im <= NP - 3) up = wz_p(q[im - 2], q[im - 1], q[im], q[im + 1], q[im + 2]);
  else if (im >= 1) up = wm3_p(q[im - 1], q[im], q[im + 1]);
  else up = 0.5 * q[0] + 0.5 * q[1];
  if (REC == 5 && im >= 1 && im <= NP - 4) um = w5_m(q[im - 1], q[im], q[im + 1], q[im + 2], q[im + 3]);
  else if (REC == 7 && im >= 1 && im <= NP - 4) um = wz_m(q[im - 1], q[im], q[im + 1], q[im + 2], q[im + 3]);
  else if (im <= NP - 3) um = wm3_m(q[im], q[im + 1], q[im + 2]);
  else um = 0.5 * q[NP - 2] + 0.5 * q[NP - 1];
}

// fnum = f(ui).n + F(ui - ua) for n = sig e_dir (scalarwave1 fluxes,
// scalarwave1.c:39-48; F as in face_flux)
template <int DIR>
__device__ __forceinline__ void fnum_dir(int numflux, double sig, const double ui[4], const double ua[4],
                                         double fn[4]) {
  double nrm[3] = {0.0, 0.0, 0.0};
  nrm[DIR] = sig;
  const double dc[3] = {ui[1] - ua[1], ui[2] - ua[2], ui[3] - ua[3]};
  double F[4];
  face_flux(numflux, nrm, ui[0] - ua[0], dc, F);
  fn[0] = -(ui[1 + DIR] * sig) + F[0];
  fn[1] = (DIR == 0 ? -ui[0] * sig : 0.0) + F[1];
  fn[2] = (DIR == 1 ? -ui[0] * sig : 0.0) + F[2];
  fn[3] = (DIR == 2 ? -ui[0] * sig : 0.0) + F[3];
}

// the neighbour's (or the exterior's) values at face f, point (fa, fb), of
// element e: the DG kernel's phase-1 gather for one point
template <int NP>
__device__ __forceinline__ void face_value(const StageArgs &a, int e, int f, int fa, int fb,
                                           double ua[4]) {
  constexpr int NP3 = NP * NP * NP;
  const int4 fd = a.face[6 * e + f];
  if (fd.x >= 0) {
    const double *nw = a.win + 5L * NP3 * fd.x + fd.y + fa * fd.z + fb * fd.w;
#pragma unroll
    for (int v = 0; v < 4; ++v) ua[v] = nw[v * NP3];
  } else if (a.profile) {
    const int dir = f >> 1, pl = (f & 1) * (NP - 1);
    const int ii = (dir == 0) ? pl : fa;
    const int jj = (dir == 1) ? pl : ((dir == 0) ? fa : fb);
    const int kk = (dir == 2) ? pl : fb;
    const double *xp = a.xyz + 3L * NP3 * e + ii + NP * (jj + NP * kk);
    plane_wave(a.k, a.om, xp[0], xp[NP3], xp[2 * NP3], a.tdev[a.tidx], ua);
  } else {
#pragma unroll
    for (int v = 0; v < 4; ++v) ua[v] = 0.0;
  }
}

// fnum = f(ui).n + F(ui - ua) for a unit normal n (curvilinear elements)
__device__ __forceinline__ void fnum_n(int numflux, const double n[3], const double ui[4],
                                       const double ua[4], double fn[4]) {
  const double dc[3] = {ui[1] - ua[1], ui[2] - ua[2], ui[3] - ua[3]};
  double F[4];
  face_flux(numflux, n, ui[0] - ua[0], dc, F);
  fn[0] = -(ui[1] * n[0] + ui[2] * n[1] + ui[3] * n[2]) + F[0];
  fn[1] = -ui[0] * n[0] + F[1];
  fn[2] = -ui[0] * n[1] + F[2];
  fn[3] = -ui[0] * n[2] + F[3];
}

// Geometry of one element for its FV lines.  GEOM 0 (Cartesian): dXbdX and
// ooJ = det(dXb/dx) are constant.  GEOM 2 (CubedShell): the metric is
// recomputed at each point from the element's box in patch coordinates.
// fv_divf_extrap_s1, _s2, _opt of dnfn_extrap1 (fv.c:147-150, 558-573)
// and the end-midpoint reconstruction of fv_rec (dg.c:295-326):
// recend 0 for the _2 modes (average of the two end points), 1 for the _1
// modes (the end point), 2 for WENOm3 / WENOm5 / WENOmZ (the average unless
// fv_rec_WENOm_s1 / _s2 / _opt take the end point, rec1d.c:1096-1186, which
// then also keeps that end's df from dnfn_extrap1, fv.c:45-79, 558-573)
struct FvEx {
  double s1, s2;
  int opt, on;   // on: s1 > 0 or s2 < DBL_MAX
  int recend;
  double rs1, rs2;
  int ropt, ron; // ron: fv_rec_WENOm_s1 > 0 or _s2 < DBL_MAX
  int first;     // fv_rec = 1: the point left / right of each midpoint (rec1d.c:16-27)
};

// rec at an end midpoint from the end point u0 and the next two inward
// (rec1d_p_WENOm3 / rec1d_m_WENOm3 with rec1d_u_in1_weightfac,
// rec1d.c:1096-1186; the _1 and _2 modes, rec1d.c:1032-1090, 909-1030);
// *kept: the end point itself was taken under the WENOm limit
template <bool EX>
__device__ __forceinline__ double rec_end(double u0, double u1, double u2, const FvEx &ex, bool *kept) {
  *kept = false;
  if (!EX || ex.recend == 0) return 0.5 * u0 + 0.5 * u1;
  if (ex.recend == 1) return u0;
  double w1fac = 1.;
  if (ex.ron) {
    const double d2 = u2 - u1, d1 = u1 - u0;
    const double s = fabs(d1) / (fabs(d2) + 1e-100);
    if (s < ex.rs1 || s >= ex.rs2) w1fac = 0.;
    if (ex.ropt == -1) w1fac = 0.;
    else if (ex.ropt == 1 && d1 * d2 <= 0.) w1fac = 0.;
    else if (ex.ropt == 2 && d1 * d2 >= 0.) w1fac = 0.;
  }
  const double w1 = 0.5 * w1fac, w0 = 1. - w1;
  const double r = w0 * u0 + w1 * u1;
  *kept = ex.ron && r == u0;   // fv_stat_WENOm_1or2: the rec equals the end point
  return r;
}

// The face value of df from the end value u0 and the next two inward, u1
// and u2 (rec1d_compute_1s1_u with forward = 0, rec1d.c:542-609): u0 sits a
// quarter spacing inside the face, so u_face = u0 / 0.75 + (1 - 1 / 0.75) u1,
// unless s = (4/3)|u1 - u0| / |u2 - u1| falls outside [s1, s2) or opt says
// otherwise, which keeps u0 (0th order).
template <bool EX>
__device__ __forceinline__ double fv_extrap1(double u0, double u1, double u2, const FvEx &ex) {
  double w0 = 1. / 0.75;
  double w1 = 1. - w0;
  if (EX && ex.on) {
    const double d2 = u2 - u1, d1 = u1 - u0;
    const double s = 1.3333333333 * fabs(d1) / (fabs(d2) + 1e-100);
    if (s < ex.s1 || s >= ex.s2) w1 = 0.;
    if (ex.opt == -1) w1 = 0.;
    else if (ex.opt == 1 && d1 * d2 <= 0.) w1 = 0.;
    else if (ex.opt == 2 && d1 * d2 >= 0.) w1 = 0.;
    w0 = 1. - w1;
  }
  return w0 * u0 + w1 * u1;
}

struct FvGeo {
  double dXbdX[3];
  double detb;          // dXbdX[0] dXbdX[1] dXbdX[2]
  double bb[6];         // GEOM 2: element box in patch coordinates
  double s0, s1;
  int dom;
};

// sqrtgdiag / det(dXb/dx) along DIR, det(dXb/dx) and the unit normal (row
// DIR of dX/dx over its length) at patch point X of a shell element, flat
// metric (coordinates.c:133-224,284-415; coordinates_set_sqrtgdiag_array
// coordinates.c:728-802; node_normal_from_idXdx_at_ijk utilities.c:48-81)
template <int DIR>
__device__ __forceinline__ void shell_metric(const FvGeo &g, const double X[3], double &soj,
                                             double &ooJ, double nrm[3]) {
  double M[3][3];
  shell_dXdx(g.dom, g.s0, g.s1, X, M);
  ooJ = g.detb * det3(M);
  double s = sqrt(M[DIR][0] * M[DIR][0] + M[DIR][1] * M[DIR][1] + M[DIR][2] * M[DIR][2]);
  soj = g.dXbdX[DIR] * s / ooJ;
  if (s == 0.0) s = 1.0;
  const double is = 1.0 / s;
  nrm[0] = M[DIR][0] * is;
  nrm[1] = M[DIR][1] * is;
  nrm[2] = M[DIR][2] * is;
}

// patch coordinate d of element point Xb (XYZ_of_XbYbZb, get_coords.c:384-394)
__device__ __forceinline__ double patch_X(const FvGeo &g, int d, double xb) {
  return 0.5 * ((g.bb[2 * d + 1] - g.bb[2 * d]) * xb + (g.bb[2 * d + 1] + g.bb[2 * d]));
}

// One line of direction DIR at (la, lb) (along Dir1_norm, Dir2_norm of DIR)
// of element e, streamed from HBM through a window of six points per
// variable (points i0-2 .. i0+3 at midpoint i0) with the next point
// requested one midpoint ahead.  Adds df * ooJ of the line to acc (FIRST:
// writes it), with dnfn_extrap1 at both ends.
//
// Metric factors (fv.c:347-401): Jgdow_R = sqrtgdiag / (ooJ wc) at the
// midpoint right of i0 (the node itself at the right end), Jgdow_L the same
// at the midpoint left of i0 (the node itself at the left end); the flux at
// a midpoint takes the midpoint's normal, the end faces the end node's.
// With coordinates_midpoint_data = no (a.fv_mid = 0) the metric and normal
// of midpoint i0+1/2 are node i0's (coordinates.c:467-483).  On Cartesian
// elements all of these are the constant dXbdX[DIR] / ooJ and +-e_DIR.
// eu: the element's state (5 variables of NP^3 points); face(f, fa, fb, ua)
// gives the adjacent values at point (fa, fb) of face f.
template <int NP, int REC, int DIR, bool FIRST, int GEOM, bool EX, class Face>
__device__ __forceinline__ void fv_line(const double *eu, const Face &face, int numflux, int fv_mid,
                                        const FvEx &ex, const double *wc, const double *xb, const FvGeo &G, int la,
                                        int lb, double (*acc)[NP * NP * NP]) {
  constexpr int NP2 = NP * NP;
  constexpr int NP3 = NP * NP * NP;
  constexpr int D1 = DIR == 0 ? 1 : 0;   // Dir1_norm: la runs along it
  constexpr int D2 = DIR == 2 ? 1 : 2;   // Dir2_norm: lb runs along it
  const int stride = DIR == 0 ? 1 : DIR == 1 ? NP : NP2;
  const int base = DIR == 0 ? (la * NP + lb * NP2) : DIR == 1 ? (la + lb * NP2) : (la + lb * NP);
  const double *g = eu + base;
  double lo[4], hi[4];
  face(2 * DIR, la, lb, lo);
  face(2 * DIR + 1, la, lb, hi);
  double q[4][6], nx[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      const int i = s - 2;
      q[v][s] = (i >= 0 && i < NP) ? g[v * NP3 + i * stride] : 0.0;
    }
    nx[v] = (4 < NP) ? g[v * NP3 + 4 * stride] : 0.0;
  }
  // patch coordinates of the line off its direction
  double X[3] = {0.0, 0.0, 0.0};
  if (GEOM == 2) {
    X[D1] = patch_X(G, D1, xb[la]);
    X[D2] = patch_X(G, D2, xb[lb]);
  }
  auto put = [&](int i, int l, double d, double oj) {
    if (FIRST) acc[l][base + i * stride] = d * oj;
    else acc[l][base + i * stride] += d * oj;
  };
  static_assert(NP >= 3, "FV lines have at least 3 points");
  // d0, d1: points 0 and 1, held until point 2 completes the left end;
  // dm3, dm2, dm1: points NP-3 .. NP-1 for the right end (NP - 3 is put as
  // well when it is not one of the held points)
  double fR[4], fL[4], d0[4], d1[4], dm3[4], dm2[4], dm1[4];
  bool keepL[4] = {false, false, false, false}, keepR[4] = {false, false, false, false};
  double oj0 = G.detb, oj1 = G.detb, ojm2 = G.detb, ojm1 = G.detb;
  double sojL = G.dXbdX[DIR] / G.detb;   // sqrtgdiag / ooJ left of i0
#pragma unroll 1
  for (int i0 = 0; i0 < NP; ++i0) {
    const int im = i0;
    // node metric; the right midpoint's metric and normal
    double ojn = G.detb, sojn = G.dXbdX[DIR] / G.detb, nn[3] = {0.0, 0.0, 0.0};
    double sojR, nR[3];
    nn[DIR] = 1.0;
    if (GEOM == 2) {
      X[DIR] = patch_X(G, DIR, xb[i0]);
      shell_metric<DIR>(G, X, sojn, ojn, nn);
    }
    sojR = sojn;
    nR[0] = nn[0]; nR[1] = nn[1]; nR[2] = nn[2];
    if (GEOM == 2 && fv_mid && i0 < NP - 1) {
      double ojm;
      X[DIR] = patch_X(G, DIR, 0.5 * (xb[i0] + xb[i0 + 1]));
      shell_metric<DIR>(G, X, sojR, ojm, nR);
    }
    if (i0 == 0) sojL = sojn;
    const double iwc = 1.0 / wc[i0];
    const double JR = sojR * iwc, JL = sojL * iwc;
    sojL = sojR;
    if (i0 > 0) {
#pragma unroll
      for (int l = 0; l < 4; ++l) fL[l] = -fR[l];
    } else {
      double ui[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) ui[v] = q[v][2];
      if (GEOM == 0) {
        fnum_dir<DIR>(numflux, -1.0, ui, lo, fL);
      } else {
        const double nL[3] = {-nn[0], -nn[1], -nn[2]};
        fnum_n(numflux, nL, ui, lo, fL);
      }
    }
    double ui[4], ua[4];
    if (i0 < NP - 1) {
      // both sides of midpoint im (WENOm5_2, WENOmZ_2 or WENOm3_2): the
      // window holds points im-2 .. im+3 at q[.][0..5]
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double *w = q[v];
        if (EX && ex.first) {
          ui[v] = w[2];
          ua[v] = w[3];
          continue;
        }
        if (REC == 5 && im >= 2 && im <= NP - 3) ui[v] = w5_p(w[0], w[1], w[2], w[3], w[4]);
        else if (REC == 7 && im >= 2 && im <= NP - 3) ui[v] = wz_p(w[0], w[1], w[2], w[3], w[4]);
        else if (REC == 9 && im >= 2 && im <= NP - 3) ui[v] = wt_p(w[0], w[1], w[2], w[3], w[4]);
        else if (im >= 1) ui[v] = wm3_p(w[1], w[2], w[3]);
        else ui[v] = rec_end<EX>(w[2], w[3], w[4], ex, &keepL[v]);
        if (REC == 5 && im >= 1 && im <= NP - 4) ua[v] = w5_m(w[1], w[2], w[3], w[4], w[5]);
        else if (REC == 7 && im >= 1 && im <= NP - 4) ua[v] = wz_m(w[1], w[2], w[3], w[4], w[5]);
        else if (REC == 9 && im >= 1 && im <= NP - 4) ua[v] = wt_m(w[1], w[2], w[3], w[4], w[5]);
        else if (im <= NP - 3) ua[v] = wm3_m(w[2], w[3], w[4]);
        else ua[v] = rec_end<EX>(w[3], w[2], w[1], ex, &keepR[v]);
      }
    } else {
#pragma unroll
      for (int v = 0; v < 4; ++v) { ui[v] = q[v][2]; ua[v] = hi[v]; }
    }
    if (GEOM == 0) fnum_dir<DIR>(numflux, 1.0, ui, ua, fR);
    else fnum_n(numflux, nR, ui, ua, fR);
    if (i0 == 0) oj0 = ojn;
    else if (i0 == 1) oj1 = ojn;
    if (i0 == NP - 2) ojm2 = ojn;
    else if (i0 == NP - 1) ojm1 = ojn;
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const double d = JR * fR[l] + JL * fL[l];
      if (i0 == 0) d0[l] = d;
      else if (i0 == 1) d1[l] = d;
      else if (i0 == NP - 2) dm2[l] = d;
      else if (i0 == NP - 1) dm1[l] = d;
      else put(i0, l, d, ojn);
      if (EX && NP >= 5 && i0 == NP - 3) dm3[l] = d;
      if (i0 == 2) {
        // the left end (rec1d_LR_uface_to_uin_1_Carray, left first)
        if (!(EX && keepL[l])) d0[l] = fv_extrap1<EX>(d0[l], d1[l], d, ex);
        put(0, l, d0[l], oj0);
        if (NP >= 4) put(1, l, d1[l], oj1);
      }
    }
    // slide the window by one point; request the one after next
#pragma unroll
    for (int v = 0; v < 4; ++v) {
#pragma unroll
      for (int s = 0; s < 5; ++s) q[v][s] = q[v][s + 1];
      q[v][5] = nx[v];
      nx[v] = (i0 + 5 < NP) ? g[v * NP3 + (i0 + 5) * stride] : 0.0;
    }
  }
  // the right end; with 3 points its inner neighbours are point 1 and the
  // extrapolated point 0, with 4 points point 1 itself
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    const double b1 = NP == 3 ? d1[l] : dm2[l];
    const double a2 = NP == 3 ? d0[l] : (NP == 4 ? d1[l] : (EX ? dm3[l] : 0.0));
    put(NP - 1, l, (EX && keepR[l]) ? dm1[l] : fv_extrap1<EX>(dm1[l], b1, a2, ex), ojm1);
    put(NP - 2, l, b1, ojm2);
  }
}

// Outputs of the RK stage (as StageArgs / AmrArgs name them).
struct FvOut {
  const double *win, *U, *Ain;
  double *Aout, *Wout, *R;
  double cA, cW;
  int wmode;
};

// One FV element whose state starts at eu = win + gbase: the three sweeps
// into acc, then r_{pi,c} = -divf, r_phi = pi and the RK stage.
template <int NP, int REC, int NT, int GEOM, bool EX, class Face>
__device__ __forceinline__ void fv_element(const FvOut &o, long gbase, const Face &face, int numflux,
                                           int fv_mid, const FvEx &ex, const double *s_wc,
                                           const double *s_xb,
                                           const FvGeo &G, double (*acc)[NP * NP * NP]) {
  constexpr int NP2 = NP * NP;
  constexpr int NP3 = NP * NP * NP;
  constexpr int NC = (NP3 + NT - 1) / NT;
  const int tid = threadIdx.x;
  const double *eu = o.win + gbase;
  const int la = tid % NP, lb = tid / NP;
  __syncthreads();   // the previous element's epilogue is done with acc
  if (tid < NP2) fv_line<NP, REC, 0, true, GEOM, EX>(eu, face, numflux, fv_mid, ex, s_wc, s_xb, G, la, lb, acc);
  __syncthreads();
  if (tid < NP2) fv_line<NP, REC, 1, false, GEOM, EX>(eu, face, numflux, fv_mid, ex, s_wc, s_xb, G, la, lb, acc);
  __syncthreads();
  if (tid < NP2) fv_line<NP, REC, 2, false, GEOM, EX>(eu, face, numflux, fv_mid, ex, s_wc, s_xb, G, la, lb, acc);
  __syncthreads();
  // ---- epilogue: r_{pi,c} = -divf, r_phi = pi, the RK stage ----
#pragma unroll 1
  for (int c = 0; c < NC; ++c) {
    const int node = tid + c * NT;
    if (node >= NP3) continue;
    const long gi = gbase + node;
    double r[5];
#pragma unroll
    for (int v = 0; v < 4; ++v) r[v] = -acc[v][node];
    r[4] = o.win[gi];
    double Ubuf[5], Abuf[5];
    if (o.wmode) {
#pragma unroll
      for (int v = 0; v < 5; ++v) Ubuf[v] = o.U[gi + v * NP3];
    }
    if (o.Aout) {
      const double *bas = o.Ain ? o.Ain : o.win;
#pragma unroll
      for (int v = 0; v < 5; ++v) Abuf[v] = bas[gi + v * NP3];
    }
    if (o.R) {
#pragma unroll
      for (int v = 0; v < 5; ++v) o.R[gi + v * NP3] = r[v];
    }
    if (o.Aout) {
#pragma unroll
      for (int v = 0; v < 5; ++v) o.Aout[gi + v * NP3] = Abuf[v] + o.cA * r[v];
    }
    if (o.wmode == 1) {
#pragma unroll
      for (int v = 0; v < 5; ++v) o.Wout[gi + v * NP3] = Ubuf[v] + o.cW * r[v];
    } else if (o.wmode == 2) {
#pragma unroll
      for (int v = 0; v < 5; ++v) o.Wout[gi + v * NP3] = 0.75 * Ubuf[v] + 0.25 * (o.win[gi + v * NP3] + o.cW * r[v]);
    }
  }
}

// element-range kernel of meshes that are FV throughout (one order)
template <int NP, int REC, int NT, int GEOM, bool EX>
// two waves per SIMD: the shell instance otherwise takes 256 VGPRs plus 12
// AGPRs and runs one wave per SIMD
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2))) void fvelm_stage_kernel(StageArgs a,
                                                                                                  const double *wc) {
  constexpr int NP3 = NP * NP * NP;
  __shared__ double acc[4][NP3];   // divf, summed over x, then y, then z
  __shared__ double s_wc[NP];
  __shared__ double s_xb[NP];
  const int tid = threadIdx.x;
  for (int q = tid; q < NP; q += NT) {
    s_wc[q] = wc[q];
    s_xb[q] = GEOM == 2 ? a.xb[q] : 0.0;
  }

  int e_hi, e_step, e_first;
  if (a.xcd_map) {
    const int xcd = blockIdx.x & 7, q = blockIdx.x >> 3, g8 = gridDim.x >> 3;
    const int chunk = (a.e1 - a.e0 + 7) >> 3;
    const int e_lo = a.e0 + xcd * chunk;
    e_hi = min(e_lo + chunk, a.e1);
    e_first = e_lo + q;
    e_step = g8;
  } else {
    e_hi = a.e1;
    e_first = a.e0 + blockIdx.x;
    e_step = gridDim.x;
  }
  const FvOut o = {a.win, a.U, a.Ain, a.Aout, a.Wout, a.R, a.cA, a.cW, a.wmode};

  for (int e = e_first; e < e_hi; e += e_step) {
    FvGeo G;
    G.dXbdX[0] = a.dXbdX[3 * e];
    G.dXbdX[1] = a.dXbdX[3 * e + 1];
    G.dXbdX[2] = a.dXbdX[3 * e + 2];
    G.detb = G.dXbdX[0] * G.dXbdX[1] * G.dXbdX[2];
    if (GEOM == 2) {
#pragma unroll
      for (int d = 0; d < 6; ++d) G.bb[d] = a.egeo[8 * e + d];
      G.s0 = a.egeo[8 * e + 6];
      G.s1 = a.egeo[8 * e + 7];
      G.dom = a.edom[e];
    }
    auto face = [&](int f, int fa, int fb, double ua[4]) { face_value<NP>(a, e, f, fa, fb, ua); };
    fv_element<NP, REC, NT, GEOM, EX>(o, 5L * NP3 * e, face, a.numflux, a.fv_mid,
                                  FvEx{a.ex_s1, a.ex_s2, a.ex_opt, a.ex_on, a.recend, a.rec_s1, a.rec_s2, a.rec_opt, a.rec_on,
                                       a.rec_first}, s_wc, s_xb, G, acc);
  }
}

// Order-bucket kernel of meshes with DG and FV elements (or AMR): the
// bucket's elements from their SlotRec; faces copied from a same-kind
// neighbour (kind 1), interpolated this stage by face_interp_kernel (kind
// 2), averaged over several neighbours (kind 3), or the exterior (kind 0),
// as the DG bucket kernel gathers them.
template <int NP, int REC, int NT, int GEOM, bool EX>
__global__ __launch_bounds__(NT) void fvelm_amr_kernel(AmrArgs a, const double *wc) {
  constexpr int NP2 = NP * NP;
  constexpr int NP3 = NP * NP * NP;
  __shared__ double acc[4][NP3];
  __shared__ double s_wc[NP];
  __shared__ double s_xb[NP];
  const int tid = threadIdx.x;
  for (int q = tid; q < NP; q += NT) {
    s_wc[q] = wc[q];
    s_xb[q] = GEOM == 2 ? a.xb[q] : 0.0;
  }
  int i_hi, i_step, i_first;
  if (a.xcd_map) {
    const int xcd = blockIdx.x & 7, q = blockIdx.x >> 3, g8 = gridDim.x >> 3;
    const int chunk = (a.nlist + 7) >> 3;
    const int lo = xcd * chunk;
    i_hi = min(lo + chunk, a.nlist);
    i_first = lo + q;
    i_step = g8;
  } else {
    i_hi = a.nlist;
    i_first = blockIdx.x;
    i_step = gridDim.x;
  }
  const FvOut o = {a.win, a.U, a.Ain, a.Aout, a.Wout, a.R, a.cA, a.cW, a.wmode};

  for (int it = i_first; it < i_hi; it += i_step) {
    const NMX_CONST SlotRec &R = *(const NMX_CONST SlotRec *)(a.rec + it);
    const long noff = R.noff;
    FvGeo G;
    G.dXbdX[0] = R.dXbdX[0];
    G.dXbdX[1] = R.dXbdX[1];
    G.dXbdX[2] = R.dXbdX[2];
    G.detb = G.dXbdX[0] * G.dXbdX[1] * G.dXbdX[2];
    if (GEOM == 2) {
#pragma unroll
      for (int d = 0; d < 6; ++d) G.bb[d] = R.bb[d];
      G.s0 = R.s0;
      G.s1 = R.s1;
      G.dom = R.dom;
    }
    auto face = [&](int f, int fa, int fb, double ua[4]) {
      const int kind = R.fkind[f];
      const int q = fa + NP * fb;
      if (kind == 1) {
        const double *p = a.win + R.foff[f] + fa * R.fs1[f] + fb * R.fs2[f];
#pragma unroll
        for (int v = 0; v < 4; ++v) ua[v] = p[v * NP3];
      } else if (kind == 2) {
        const double *p = a.aj + R.foff[f] + q;
#pragma unroll
        for (int v = 0; v < 4; ++v) ua[v] = p[v * NP2];
      } else if (kind == 3) {
        const int *md = a.mdesc + R.foff[f];
        const int ns = md[0];
        double sum[4] = {0.0, 0.0, 0.0, 0.0};
        int cnt = 0;
        for (int si = 0; si < ns; ++si) {
          const int *d = md + 1 + 5 * si;
          if (fa >= d[1] && fa <= d[2] && fb >= d[3] && fb <= d[4]) {
            const double *p = a.aj + d[0] + q;
#pragma unroll
            for (int v = 0; v < 4; ++v) sum[v] += p[v * NP2];
            ++cnt;
          }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) ua[v] = cnt > 1 ? sum[v] / cnt : sum[v];
      } else if (a.profile) {
        const int dir = f >> 1, pl = (f & 1) * (NP - 1);
        const int ii = (dir == 0) ? pl : fa;
        const int jj = (dir == 1) ? pl : ((dir == 0) ? fa : fb);
        const int kk = (dir == 2) ? pl : fb;
        const double *xp = a.xyz + 3 * noff + ii + NP * (jj + NP * kk);
        plane_wave(a.k, a.om, xp[0], xp[NP3], xp[2 * NP3], a.tdev[a.tidx], ua);
      } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) ua[v] = 0.0;
      }
    };
    fv_element<NP, REC, NT, GEOM, EX>(o, 5 * noff, face, a.numflux, a.fv_mid,
                                  FvEx{a.ex_s1, a.ex_s2, a.ex_opt, a.ex_on, a.recend, a.rec_s1, a.rec_s2, a.rec_opt, a.rec_on,
                                       a.rec_first}, s_wc, s_xb, G, acc);
  }
}

template <int NP, int REC, bool EX>
static hipError_t launch_fvelm_t(int geom, const StageArgs &a, const double *wc, int grid, hipStream_t s) {
  constexpr int NT = (NP * NP + 63) / 64 * 64;   // one thread per line
  if (geom == 0)
    hipLaunchKernelGGL((fvelm_stage_kernel<NP, REC, NT, 0, EX>), dim3(grid), dim3(NT), 0, s, a, wc);
  else
    hipLaunchKernelGGL((fvelm_stage_kernel<NP, REC, NT, 2, EX>), dim3(grid), dim3(NT), 0, s, a, wc);
  return hipGetLastError();
}

template <int NP, int REC, bool EX>
static hipError_t launch_fvelm_amr_t(int geom, const AmrArgs &a, const double *wc, int grid, hipStream_t s) {
  constexpr int NT = (NP * NP + 63) / 64 * 64;
  if (geom == 0)
    hipLaunchKernelGGL((fvelm_amr_kernel<NP, REC, NT, 0, EX>), dim3(grid), dim3(NT), 0, s, a, wc);
  else
    hipLaunchKernelGGL((fvelm_amr_kernel<NP, REC, NT, 2, EX>), dim3(grid), dim3(NT), 0, s, a, wc);
  return hipGetLastError();
}

// the FV launches of one variant (EX: with the fv_divf_extrap s1/s2/opt limit)
template <bool EX>
static hipError_t launch_fvelm_stage_v(int np, int rec, int geom, const StageArgs &a, const double *wc,
                                       int grid, hipStream_t s) {
#define NMX_CASE(N)                                                                 \
  case N:                                                                           \
    return rec == 5   ? launch_fvelm_t<N, 5, EX>(geom, a, wc, grid, s)              \
           : rec == 7 ? launch_fvelm_t<N, 7, EX>(geom, a, wc, grid, s)              \
           : rec == 9 ? launch_fvelm_t<N, 9, EX>(geom, a, wc, grid, s)              \
                      : launch_fvelm_t<N, 3, EX>(geom, a, wc, grid, s);
  switch (np) {
    NMX_CASE(3)
    NMX_CASE(4)
    NMX_CASE(5)
    NMX_CASE(6)
    NMX_CASE(7)
    NMX_CASE(8)
    NMX_CASE(9)
    NMX_CASE(10)
    default:
      return hipErrorInvalidValue;
  }
#undef NMX_CASE
}

template <bool EX>
static hipError_t launch_fvelm_amr_v(int np, int rec, int geom, const AmrArgs &a, const double *wc, int grid,
                                     hipStream_t s) {
  if (a.nlist == 0) return hipSuccess;
#define NMX_CASE(N)                                                                 \
  case N:                                                                           \
    return rec == 5   ? launch_fvelm_amr_t<N, 5, EX>(geom, a, wc, grid, s)          \
           : rec == 7 ? launch_fvelm_amr_t<N, 7, EX>(geom, a, wc, grid, s)          \
           : rec == 9 ? launch_fvelm_amr_t<N, 9, EX>(geom, a, wc, grid, s)          \
                      : launch_fvelm_amr_t<N, 3, EX>(geom, a, wc, grid, s);
  switch (np) {
    NMX_CASE(3)
    NMX_CASE(4)
    NMX_CASE(5)
    NMX_CASE(6)
    NMX_CASE(7)
    NMX_CASE(8)
    NMX_CASE(9)
    NMX_CASE(10)
    default:
      return hipErrorInvalidValue;
  }
#undef NMX_CASE
}

// defined in fvelm_ex_kernels.hip
hipError_t launch_fvelm_stage_ex(int np, int rec, int geom, const StageArgs &a, const double *wc, int grid,
                                 hipStream_t s);
hipError_t launch_fvelm_amr_ex(int np, int rec, int geom, const AmrArgs &a, const double *wc, int grid,
                               hipStream_t s);

}  // namespace nmx
