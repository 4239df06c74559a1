/* fv_oracle.c -- CPU restatement of the FV WENO5 + LLF evolution step, with
 * special-relativistic Euler physics (D, S_i, tau; ideal gas).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP FV path is checked
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links or calls it.
 *
 * What follows the reference (paths relative to the nmesh source tree):
 *   - WENO5-JS reconstruction to a midpoint from the left (p) and from the
 *     right (m): rec1d_p_WENO5 / rec1d_m_WENO5 with the uniform-grid weights
 *     of rec1d_p_WENO5_uniform / rec1d_m_WENO5_uniform (src/main/dg/rec1d.c:
 *     1190-1352): smoothness indicators with 13/12 and 1/4, epsilon
 *     1e-40 * u_scale^2, ideal weights (1, 6, 3) normalised after division by
 *     (beta + eps)^2, stencils (2,-7,11)/6, (-1,5,2)/6, (2,5,-1)/6;
 *     u_scale = 1 as fv_divf sets q_scale (src/main/dg/fv.c:141).
 *   - LLF numerical flux F = 1/2 (f_L + f_R - a (u_R - u_L)) with a the
 *     largest |eigenvalue| of both states (numflux1d_LLF,
 *     src/main/dg/numflux1d.c:85-112).
 *   - the sspRK3 stage sequence of evolve_sspRK3_mesh
 *     (src/main/evolve/RungeKutta.c:130-188).
 * What the reference does not have (its SR-Euler projects are absent,
 * doc/nmesh.txt:103-111): the SR-Euler fluxes, eigenvalues and the Newton
 * conserved-to-primitive solve below are textbook (Marti & Mueller, Living
 * Rev. Relativ. 6 (2003) 7, sections 2.1 and 4.3) and are pinned under
 * tests/ by physics: c2p round trips, constant-state preservation, the
 * eigenvalues of the flux Jacobian, and convergence to an exact advected
 * solution.
 *
 * Grid: N^3 periodic cells of size 1, cell (i, j, k) at flat index
 * c = i + N (j + N k); states are double u[var*N^3 + c], conserved
 * (D, Sx, Sy, Sz, tau) or primitive (rho, vx, vy, vz, p).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define FV_WENO5_EPS 1e-40
#define FV_C2P_MAXIT 100

/* ------------------------------------------------------------------ */
/* WENO5 (rec1d.c:1195-1352)                                           */
/* ------------------------------------------------------------------ */

/* midpoint im+1/2 from the left: u[im-2 .. im+2] */
double orc_rec1d_p_WENO5(const double *u, int im)
{
  const double a = u[im - 2], b = u[im - 1], c = u[im], d = u[im + 1], e = u[im + 2];
  const double s01 = a - 2. * b + c, s02 = a - 4. * b + 3. * c;
  const double s11 = b - 2. * c + d, s12 = b - d;
  const double s21 = c - 2. * d + e, s22 = 3. * c - 4. * d + e;
  const double beta0 = (13. / 12.) * s01 * s01 + 0.25 * s02 * s02;
  const double beta1 = (13. / 12.) * s11 * s11 + 0.25 * s12 * s12;
  const double beta2 = (13. / 12.) * s21 * s21 + 0.25 * s22 * s22;
  const double e0 = beta0 + FV_WENO5_EPS, e1 = beta1 + FV_WENO5_EPS, e2 = beta2 + FV_WENO5_EPS;
  const double w0 = 1. / (e0 * e0), w1 = 6. / (e1 * e1), w2 = 3. / (e2 * e2);
  const double inv = 1. / (w0 + w1 + w2);
  const double o6 = 1. / 6.;
  const double q0 = (2. * o6) * a + (-7. * o6) * b + (11. * o6) * c;
  const double q1 = (-o6) * b + (5. * o6) * c + (2. * o6) * d;
  const double q2 = (2. * o6) * c + (5. * o6) * d + (-o6) * e;
  return (w0 * inv) * q0 + (w1 * inv) * q1 + (w2 * inv) * q2;
}

/* midpoint im+1/2 from the right: u[im-1 .. im+3] */
double orc_rec1d_m_WENO5(const double *u, int im)
{
  const double a = u[im - 1], b = u[im], c = u[im + 1], d = u[im + 2], e = u[im + 3];
  const double s21 = c - 2. * d + e, s22 = 3. * c - 4. * d + e;
  const double s11 = b - 2. * c + d, s12 = b - d;
  const double s01 = a - 2. * b + c, s02 = a - 4. * b + 3. * c;
  const double beta2 = (13. / 12.) * s21 * s21 + 0.25 * s22 * s22;
  const double beta1 = (13. / 12.) * s11 * s11 + 0.25 * s12 * s12;
  const double beta0 = (13. / 12.) * s01 * s01 + 0.25 * s02 * s02;
  const double e2 = beta2 + FV_WENO5_EPS, e1 = beta1 + FV_WENO5_EPS, e0 = beta0 + FV_WENO5_EPS;
  /* mirrored ideal weights: gamma (3, 6, 1) on stencils 0, 1, 2 */
  const double w2 = 1. / (e2 * e2), w1 = 6. / (e1 * e1), w0 = 3. / (e0 * e0);
  const double inv = 1. / (w0 + w1 + w2);
  const double o6 = 1. / 6.;
  const double q2 = (11. * o6) * c + (-7. * o6) * d + (2. * o6) * e;
  const double q1 = (2. * o6) * b + (5. * o6) * c + (-o6) * d;
  const double q0 = (-o6) * a + (5. * o6) * b + (2. * o6) * c;
  return (w0 * inv) * q0 + (w1 * inv) * q1 + (w2 * inv) * q2;
}

/* WENO-Z (rec1d_p/m_WENOZ with uniform weights, rec1d.c:1627-1781):
   tau5 = |beta0 - beta2|, unnormalised weights g_k (1 + tau5 / (beta_k + eps)) */
double orc_rec1d_p_WENOZ(const double *u, int im)
{
  const double a = u[im - 2], b = u[im - 1], c = u[im], d = u[im + 1], e = u[im + 2];
  const double s01 = a - 2. * b + c, s02 = a - 4. * b + 3. * c;
  const double s11 = b - 2. * c + d, s12 = b - d;
  const double s21 = c - 2. * d + e, s22 = 3. * c - 4. * d + e;
  const double beta0 = (13. / 12.) * s01 * s01 + 0.25 * s02 * s02;
  const double beta1 = (13. / 12.) * s11 * s11 + 0.25 * s12 * s12;
  const double beta2 = (13. / 12.) * s21 * s21 + 0.25 * s22 * s22;
  const double e0 = beta0 + FV_WENO5_EPS, e1 = beta1 + FV_WENO5_EPS, e2 = beta2 + FV_WENO5_EPS;
  const double tau5 = fabs(beta0 - beta2);
  const double w0 = 1. * (1. + tau5 / e0), w1 = 6. * (1. + tau5 / e1), w2 = 3. * (1. + tau5 / e2);
  const double inv = 1. / (w0 + w1 + w2);
  const double o6 = 1. / 6.;
  const double q0 = (2. * o6) * a + (-7. * o6) * b + (11. * o6) * c;
  const double q1 = (-o6) * b + (5. * o6) * c + (2. * o6) * d;
  const double q2 = (2. * o6) * c + (5. * o6) * d + (-o6) * e;
  return (w0 * inv) * q0 + (w1 * inv) * q1 + (w2 * inv) * q2;
}

double orc_rec1d_m_WENOZ(const double *u, int im)
{
  const double a = u[im - 1], b = u[im], c = u[im + 1], d = u[im + 2], e = u[im + 3];
  const double s21 = c - 2. * d + e, s22 = 3. * c - 4. * d + e;
  const double s11 = b - 2. * c + d, s12 = b - d;
  const double s01 = a - 2. * b + c, s02 = a - 4. * b + 3. * c;
  const double beta2 = (13. / 12.) * s21 * s21 + 0.25 * s22 * s22;
  const double beta1 = (13. / 12.) * s11 * s11 + 0.25 * s12 * s12;
  const double beta0 = (13. / 12.) * s01 * s01 + 0.25 * s02 * s02;
  const double e2 = beta2 + FV_WENO5_EPS, e1 = beta1 + FV_WENO5_EPS, e0 = beta0 + FV_WENO5_EPS;
  const double tau5 = fabs(beta0 - beta2);
  const double w2 = 1. * (1. + tau5 / e2), w1 = 6. * (1. + tau5 / e1), w0 = 3. * (1. + tau5 / e0);
  const double inv = 1. / (w0 + w1 + w2);
  const double o6 = 1. / 6.;
  const double q2 = (11. * o6) * c + (-7. * o6) * d + (2. * o6) * e;
  const double q1 = (2. * o6) * b + (5. * o6) * c + (-o6) * d;
  const double q0 = (-o6) * a + (5. * o6) * b + (2. * o6) * c;
  return (w0 * inv) * q0 + (w1 * inv) * q1 + (w2 * inv) * q2;
}

/* ------------------------------------------------------------------ */
/* SR-Euler, ideal gas p = (gam - 1) rho eps                           */
/* ------------------------------------------------------------------ */

void orc_sr_p2c(double gam, const double *P, double *U)
{
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

/* Newton on the pressure: f(p) = (gam-1) [E - S^2/(E+p) - D sqrt(1 - v^2)] - p,
 * v^2 = S^2/(E+p)^2, E = tau + D; f'(p) = (gam-1) v^2 (1 - D W/(E+p)) - 1.
 * P[4] holds the initial guess on entry.  Stops when |dp| <= tol p.
 * Returns the iterations taken, or -1 if it did not converge. */
int orc_sr_c2p(double gam, const double *U, double *P, double tol)
{
  const double D = U[0], Sx = U[1], Sy = U[2], Sz = U[3];
  const double E = U[4] + D;
  const double S2 = Sx * Sx + Sy * Sy + Sz * Sz;
  const double pmin = fmax(sqrt(S2) - E, 0.0);
  double p = P[4];
  if (!(p > pmin)) p = pmin + 1e-10 * E;
  int it;
  for (it = 1; it <= FV_C2P_MAXIT; ++it) {
    const double Ep = E + p;
    const double v2 = S2 / (Ep * Ep);
    const double W = 1. / sqrt(1. - v2);
    const double f = (gam - 1.) * (E - S2 / Ep - D / W) - p;
    const double df = (gam - 1.) * v2 * (1. - D * W / Ep) - 1.;
    double pn = p - f / df;
    if (pn <= pmin) pn = 0.5 * (p + pmin);
    const double dp = fabs(pn - p);
    p = pn;
    if (dp <= tol * p) break;
  }
  {
    const double Ep = E + p;
    const double v2 = S2 / (Ep * Ep);
    const double W = 1. / sqrt(1. - v2);
    P[0] = D / W;
    P[1] = Sx / Ep;
    P[2] = Sy / Ep;
    P[3] = Sz / Ep;
    P[4] = p;
  }
  return it <= FV_C2P_MAXIT ? it : -1;
}

/* flux along direction d, conserved state and the two acoustic eigenvalues
 * lam = (v_d (1 - cs^2) -+ cs sqrt((1 - v^2)(1 - v^2 cs^2 - v_d^2 (1 - cs^2)))) / (1 - v^2 cs^2) */
void orc_sr_flux(double gam, int d, const double *P, double *F, double *U, double *lam)
{
  orc_sr_p2c(gam, P, U);
  const double rho = P[0], p = P[4];
  const double vd = P[1 + d];
  const double v2 = P[1] * P[1] + P[2] * P[2] + P[3] * P[3];
  F[0] = U[0] * vd;
  F[1] = U[1] * vd;
  F[2] = U[2] * vd;
  F[3] = U[3] * vd;
  F[1 + d] += p;
  F[4] = (U[4] + p) * vd;
  const double h = 1. + gam / (gam - 1.) * p / rho;
  const double cs2 = gam * p / (rho * h);
  const double cs = sqrt(cs2);
  const double den = 1. - v2 * cs2;
  const double root = cs * sqrt((1. - v2) * (1. - v2 * cs2 - vd * vd * (1. - cs2)));
  lam[0] = (vd * (1. - cs2) - root) / den;
  lam[1] = (vd * (1. - cs2) + root) / den;
}

/* LLF between the reconstructed primitive states PL | PR */
void orc_sr_llf(double gam, int d, const double *PL, const double *PR, double *F)
{
  double FL[5], FR[5], UL[5], UR[5], lL[2], lR[2];
  orc_sr_flux(gam, d, PL, FL, UL, lL);
  orc_sr_flux(gam, d, PR, FR, UR, lR);
  const double a = fmax(fmax(fabs(lL[0]), fabs(lL[1])), fmax(fabs(lR[0]), fabs(lR[1])));
  for (int v = 0; v < 5; ++v) F[v] = 0.5 * (FL[v] + FR[v] - a * (UR[v] - UL[v]));
}

static int physical(const double *P)
{
  const double v2 = P[1] * P[1] + P[2] * P[2] + P[3] * P[3];
  return P[0] > 0. && P[4] > 0. && v2 < 1.;
}

/* face states at the face between cells c(-1) and c(0) of one line: q[s]
 * holds cells -3 .. +2 (s = 0..5) of the 5 primitives, q[v][s] */
static void face_states(const double q[5][6], double *PL, double *PR)
{
  for (int v = 0; v < 5; ++v) {
    PL[v] = orc_rec1d_p_WENO5(q[v], 2);
    PR[v] = orc_rec1d_m_WENO5(q[v], 2);
  }
  /* a non-physical reconstruction falls back to the cell value */
  if (!physical(PL))
    for (int v = 0; v < 5; ++v) PL[v] = q[v][2];
  if (!physical(PR))
    for (int v = 0; v < 5; ++v) PR[v] = q[v][3];
}

static inline long cidx(int N, int i, int j, int k)
{
  i = (i % N + N) % N;
  j = (j % N + N) % N;
  k = (k % N + N) % N;
  return i + (long)N * (j + (long)N * k);
}

/* primitives of u (guess: P on entry); returns the number of cells whose
 * Newton solve did not converge */
int orc_fv_c2p(int N, double gam, const double *u, double *P, double tol)
{
  const long n3 = (long)N * N * N;
  int bad = 0;
#pragma omp parallel for reduction(+ : bad)
  for (long c = 0; c < n3; ++c) {
    double U[5], Q[5];
    for (int v = 0; v < 5; ++v) { U[v] = u[v * n3 + c]; Q[v] = P[v * n3 + c]; }
    if (orc_sr_c2p(gam, U, Q, tol) < 0) bad++;
    for (int v = 0; v < 5; ++v) P[v * n3 + c] = Q[v];
  }
  return bad;
}

void orc_fv_p2c(int N, double gam, const double *P, double *u)
{
  const long n3 = (long)N * N * N;
#pragma omp parallel for
  for (long c = 0; c < n3; ++c) {
    double Q[5], U[5];
    for (int v = 0; v < 5; ++v) Q[v] = P[v * n3 + c];
    orc_sr_p2c(gam, Q, U);
    for (int v = 0; v < 5; ++v) u[v * n3 + c] = U[v];
  }
}

/* r = -sum_d (F_{i+1/2} - F_{i-1/2}) from primitives P (cell size 1) */
void orc_fv_rhs_prim(int N, double gam, const double *P, double *r)
{
  const long n3 = (long)N * N * N;
  memset(r, 0, 5 * n3 * sizeof(double));
  for (int d = 0; d < 3; ++d) {
#pragma omp parallel for collapse(2)
    for (int b = 0; b < N; ++b)
      for (int a = 0; a < N; ++a) {
        /* the line along d through transverse position (a, b) */
        double Fprev[5];
        for (int f = 0; f <= N; ++f) {   /* face between cells f-1 and f */
          double q[5][6], PL[5], PR[5], F[5];
          for (int s = 0; s < 6; ++s) {
            int ii = f - 3 + s;
            int i = d == 0 ? ii : a, j = d == 1 ? ii : (d == 0 ? a : b), k = d == 2 ? ii : b;
            long c = cidx(N, i, j, k);
            for (int v = 0; v < 5; ++v) q[v][s] = P[v * n3 + c];
          }
          face_states(q, PL, PR);
          orc_sr_llf(gam, d, PL, PR, F);
          if (f > 0) {
            int ii = f - 1;
            int i = d == 0 ? ii : a, j = d == 1 ? ii : (d == 0 ? a : b), k = d == 2 ? ii : b;
            long c = cidx(N, i, j, k);
            for (int v = 0; v < 5; ++v) r[v * n3 + c] += Fprev[v] - F[v];
          }
          memcpy(Fprev, F, sizeof(Fprev));
        }
      }
  }
}

/* r = RHS(u); P holds the c2p guess on entry and the primitives on exit */
int orc_fv_rhs(int N, double gam, const double *u, double *P, double *r, double tol)
{
  int bad = orc_fv_c2p(N, gam, u, P, tol);
  orc_fv_rhs_prim(N, gam, P, r);
  return bad;
}

/* nsteps of evolve_sspRK3_mesh (RungeKutta.c:130-188):
 *   A = u + dt/6 r1; w = u + dt r1; A += dt/6 r2; w = 3/4 u + 1/4 (w + dt r2);
 *   A += 2/3 dt r3; u = A */
int orc_fv_evolve(int N, double gam, double dt, int nsteps, double *u, double *P, double tol)
{
  const long n = 5L * N * N * N;
  double *r = malloc(n * sizeof(double)), *A = malloc(n * sizeof(double)), *w = malloc(n * sizeof(double));
  int bad = 0;
  for (int s = 0; s < nsteps; ++s) {
    bad += orc_fv_rhs(N, gam, u, P, r, tol);
#pragma omp parallel for
    for (long q = 0; q < n; ++q) { A[q] = u[q] + dt / 6. * r[q]; w[q] = u[q] + dt * r[q]; }
    bad += orc_fv_rhs(N, gam, w, P, r, tol);
#pragma omp parallel for
    for (long q = 0; q < n; ++q) { A[q] += dt / 6. * r[q]; w[q] = 0.75 * u[q] + 0.25 * (w[q] + dt * r[q]); }
    bad += orc_fv_rhs(N, gam, w, P, r, tol);
#pragma omp parallel for
    for (long q = 0; q < n; ++q) u[q] = A[q] + 2. * dt / 3. * r[q];
  }
  free(r);
  free(A);
  free(w);
  return bad;
}
