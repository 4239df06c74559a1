/* nmx_oracle.c -- CPU restatement of nmesh's DG evolution hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle the HIP product path is
 * checked against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker or the
 * timed CPU baseline; the product library never links or calls it.
 *
 * It restates, in plain C99 + OpenMP, the algorithm of the reference files
 * cited at each function (paths relative to the nmesh source tree).  The
 * reference itself cannot be compiled in this image (every source includes
 * nmesh.h, which includes the build-generated nmesh_automatic_include.h), so
 * this restatement is pinned by closed forms, invariants and convergence
 * tests under tests/ instead of by a reference build.
 *
 * Layout used at this file's C-ABI: a state is double u[var*nnodes + node]
 * with node = elm_offset + ijk and ijk = i + n*(j + n*k) (nmesh_amr_loops.h:279).
 * Variable order is scalarwave1's vlu: (pi, cx, cy, cz, phi)
 * (src/tests/scalarwave1/scalarwave1.c:670-672).
 */
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAXN 24
#define ORC_MAXNB 8
#define DEQUALEPS 1e-10 /* src/main/main/nmesh_main.h:32 */

/* ------------------------------------------------------------------ */
/* basis: LGL points, weights, barycentric weights, transposed D       */
/* ------------------------------------------------------------------ */

typedef struct {
  int set;
  double Xb[ORC_MAXN], Wq[ORC_MAXN], WL[ORC_MAXN], DT[ORC_MAXN * ORC_MAXN];
  double Xu[ORC_MAXN], Wu[ORC_MAXN];        /* uniform points (FV) */
} OBasis;

static OBasis g_basis[ORC_MAXN + 1];

/* P_l, P_l', Q = P_{l+1}-P_{l-1}, Q'  (src/main/basis/Legendre.c:61-98) */
static void leg_P_dP_Q_dQ(int l, double x, double *P, double *dP, double *Q,
                          double *dQ)
{
  double P0 = 1., P1 = x, dP0 = 0., dP1 = 1.;
  double Pk = 0., dPk = 0.;
  int k;
  if (l == 0) { *P = 1.; *dP = 0.; return; }
  if (l == 1) { *P = x; *dP = 1.; return; }
  for (k = 2; k <= l + 1; k++) {
    Pk = (P1 * x * (2 * k - 1) - P0 * (k - 1)) / k;
    dPk = P1 * (2 * k - 1) + dP0;
    if (k >= l + 1) break;
    P0 = P1; P1 = Pk; dP0 = dP1; dP1 = dPk;
  }
  *P = P1; *dP = dP1; *Q = Pk - P0; *dQ = dPk - dP0;
}

/* LGL nodes and weights on [-1,1] (src/main/basis/Legendre.c:149-213) */
void orc_lgl(int npoints, double *x, double *w)
{
  int N = npoints - 1, i, j;
  double P, dP, Q, dQ, y, dy;
  if (N == 0) { x[0] = 0.; if (w) w[0] = 2.; return; }
  x[0] = -1.; x[N] = 1.;
  if (w) { w[0] = 2.0 / (N * (N + 1.0)); w[N] = w[0]; }
  if (N == 1) return;
  for (j = 1; j <= N / 2; j++) {
    if (j == N / 2.0) {
      y = 0.;
    } else {
      y = -cos(M_PI * (j + 0.25) / N - 3 / (8 * N * M_PI * (j + 0.25)));
      for (i = 0; i < 100; i++) {
        leg_P_dP_Q_dQ(N, y, &P, &dP, &Q, &dQ);
        dy = -Q / dQ;
        y += dy;
        if (fabs(dy) < fabs(1e-15 * y)) break;
      }
    }
    x[N - j] = -y;
    x[j] = y;
    leg_P_dP_Q_dQ(N, y, &P, &dP, &Q, &dQ);
    if (w) { w[j] = 2.0 / ((N * (N + 1)) * P * P); w[N - j] = w[j]; }
  }
}

/* barycentric weights (src/main/basis/Lagrange.c:15-28) */
static void lagrange_winterp(int n, const double *x, double *wi)
{
  for (int i = 0; i < n; i++) {
    double denom = 1.;
    for (int m = 0; m < n; m++) if (m != i) denom *= (x[i] - x[m]);
    wi[i] = 1. / denom;
  }
}

/* DT[j + i*n] = D_ij,  D_ii = -sum_{j!=i} D_ij (src/main/basis/Lagrange.c:126-147) */
static void lagrange_DT(int n, const double *x, const double *wi, double *DT)
{
  for (int i = 0; i < n; i++) {
    double Dii = 0;
    for (int j = 0; j < n; j++)
      if (i != j) {
        double Dij = (wi[j] / wi[i]) / (x[i] - x[j]);
        Dii -= Dij;
        DT[j + i * n] = Dij;
      }
    DT[i * n + i] = Dii;
  }
}

/* Lagrange basis l_k(x) (src/main/basis/Lagrange.c:151-158) */
static double lagrange_of_x(int k, double x, int np, const double *xp,
                            const double *wi)
{
  double prod = 1.;
  for (int m = 0; m < np; m++) if (m != k) prod *= (x - xp[m]);
  return prod * wi[k];
}

/* piecewise linear basis function k on points xp (basis_pw_linear,
   src/main/basis/interpolate.c:53-102): the two end intervals extend past
   the ends, x in [xp[k], xp[k+1]) belongs to the interval on the right */
static double pw_linear_of_x(int k, double x, int np, const double *xp)
{
  if (np <= 1) return 1.;
  if (x < xp[1]) {
    if (k == 0) return (x - xp[1]) / (xp[0] - xp[1]);
    if (k == 1) return (x - xp[0]) / (xp[1] - xp[0]);
    return 0.;
  }
  if (x >= xp[np - 2]) {
    int n0 = np - 2, n1 = np - 1;
    if (k == n0) return (x - xp[n1]) / (xp[n0] - xp[n1]);
    if (k == n1) return (x - xp[n0]) / (xp[n1] - xp[n0]);
    return 0.;
  }
  if (k < np - 1 && x >= xp[k] && x < xp[k + 1]) return (x - xp[k + 1]) / (xp[k] - xp[k + 1]);
  if (k > 0 && x >= xp[k - 1] && x < xp[k]) return (x - xp[k - 1]) / (xp[k] - xp[k - 1]);
  return 0.;
}

/* piecewise parabolic basis function k on points xp (basis_pw_parab,
   src/main/basis/interpolate.c:104-224): near node m, between the
   midpoints to its neighbours, the parabola through m-1, m, m+1; the first
   and last three points at the ends */
static double parab3(double x, const double *xp, int a, int b, int c)
{
  return (x - xp[b]) * (x - xp[c]) / ((xp[a] - xp[b]) * (xp[a] - xp[c]));
}

static double pw_parab_of_x(int k, double x, int np, const double *xp)
{
  if (np < 3) return pw_linear_of_x(k, x, np, xp);
  if (np == 3 || x < 0.5 * (xp[0] + xp[1])) {
    if (k == 0) return parab3(x, xp, 0, 1, 2);
    if (k == 1) return parab3(x, xp, 1, 0, 2);
    if (k == 2) return parab3(x, xp, 2, 0, 1);
    return 0.;
  }
  if (x >= 0.5 * (xp[np - 2] + xp[np - 1])) {
    int n0 = np - 3, n1 = np - 2, n2 = np - 1, l = k - n0;
    if (l == 0) return parab3(x, xp, n0, n1, n2);
    if (l == 1) return parab3(x, xp, n1, n0, n2);
    if (l == 2) return parab3(x, xp, n2, n0, n1);
    return 0.;
  }
  int m, m1, m2;
  double xml, xmr, xmll, xmrr;
  if (k <= 0) {
    m = 0;
    xmr = 0.5 * (xp[m] + xp[m + 1]);
    xml = xp[m] - (xmr - xp[m]);
    xmll = xml - (xmr - xml);
    xmrr = 0.5 * (xp[m + 1] + xp[m + 2]);
    m1 = 1; m2 = 2;
  } else if (k < np - 1) {
    m = k;
    xml = 0.5 * (xp[m] + xp[m - 1]);
    xmr = 0.5 * (xp[m] + xp[m + 1]);
    xmrr = (k < np - 2) ? 0.5 * (xp[m + 1] + xp[m + 2]) : xmr + (xmr - xml);
    xmll = (k >= 2) ? 0.5 * (xp[m - 1] + xp[m - 2]) : xml - (xmr - xml);
    m1 = k - 1; m2 = k + 1;
  } else {
    m = np - 1;
    xml = 0.5 * (xp[m - 1] + xp[m]);
    xmr = xp[m] + (xp[m] - xml);
    xmll = 0.5 * (xp[m - 2] + xp[m - 1]);
    xmrr = xmr + (xmr - xml);
    m1 = m - 2; m2 = m - 1;
  }
  if (x >= xml && x < xmr) return parab3(x, xp, m, m1, m2);
  if (x >= xmll && x < xml) return (m >= 2) ? parab3(x, xp, m, m - 2, m - 1) : 0.;
  if (x >= xmr && x < xmrr) return (m + 2 < np) ? parab3(x, xp, m, m + 1, m + 2) : 0.;
  return 0.;
}

double orc_pw_parab(int k, double x, int np, const double *xp) { return pw_parab_of_x(k, x, np, xp); }

/* fv_surface_interp for faces taken from FV neighbours: 0 linear (the
   default), 1 parabolic (surface.c:1182-1196) */
static int g_fv_surface_interp = 0, g_fv_lagrange_if_dg = 0;
void orc_set_fv_surface_interp(int mode) { g_fv_surface_interp = mode & 1; g_fv_lagrange_if_dg = (mode >> 1) & 1; }

/* Lagrange basis function q of the points xp at x, as a product */
static double lagrange_on_points(int q, double x, int n, const double *xp)
{
  double r = 1.;
  for (int m = 0; m < n; m++)
    if (m != q) r *= (x - xp[m]) / (xp[q] - xp[m]);
  return r;
}

static OBasis *basis(int n)
{
  OBasis *b = &g_basis[n];
  if (!b->set) {
#pragma omp critical(orc_basis)
    if (!b->set) {
      orc_lgl(n, b->Xb, b->Wq);
      lagrange_winterp(n, b->Xb, b->WL);
      lagrange_DT(n, b->Xb, b->WL, b->DT);
      /* uniform points incl. both faces (uniform_x, Legendre.c:626-663:
         ends at -+1, -1 + j h on the left half, mirrored, 0 exactly at the
         centre), trapezoid weights (gridpoints.c:49-53) */
      {
        int N = n - 1;
        if (N == 0) b->Xu[0] = 0.;
        else {
          b->Xu[0] = -1.;
          b->Xu[N] = 1.;
          if (N > 1) {
            double h = 2. / N;
            for (int j = 1; j <= N / 2; j++) {
              double y = (j == N / 2.0) ? 0. : -1. + j * h;
              b->Xu[N - j] = -y;
              b->Xu[j] = y;
            }
          }
        }
      }
      for (int i = 0; i < n; i++)
        b->Wu[i] = (n > 1) ? ((i == 0 || i == n - 1) ? 1. / (n - 1)
                                                      : 2. / (n - 1))
                           : 2.;
      b->set = 1;
    }
  }
  return b;
}

void orc_dt_matrix(int n, double *DT)
{
  OBasis *b = basis(n);
  memcpy(DT, b->DT, sizeof(double) * n * n);
}

/* ------------------------------------------------------------------ */
/* patches and coordinate maps                                         */
/* ------------------------------------------------------------------ */

enum { PAT_BOX = 0, PAT_CUBEDSHELL = 1 };

typedef struct {
  int type;
  int dom;          /* cubed-sphere domain 0..5 */
  double s0, s1;    /* CI->s[0], CI->s[1] */
  double xc[3];
  double bbox[6];
  int bound[6];     /* 1: patch face is an outer boundary */
} OPat;

/* x(X) for a CubedShell with constant sigma
   (src/main/coordinates/coordtrans_CubedSphere.c:38-121) */
static void cubsph_xyz(const OPat *P, const double X[3], double x[3])
{
  double lam = X[0], A = X[1], B = X[2];
  int dir = P->dom / 2;
  double pm = 2 * (P->dom % 2) - 1;
  double s = sqrt(1.0 + A * A + B * B);
  double a0 = pm * P->s0 / s, a1 = pm * P->s1 / s;
  double r = (a1 - a0) * lam + a0;
  if (dir == 0) { x[0] = P->xc[0] + r; x[1] = P->xc[1] + A * r; x[2] = P->xc[2] + B * r; }
  else if (dir == 1) { x[1] = P->xc[1] + r; x[0] = P->xc[0] + A * r; x[2] = P->xc[2] + B * r; }
  else { x[2] = P->xc[2] + r; x[1] = P->xc[1] + A * r; x[0] = P->xc[0] + B * r; }
}

/* inverse map X(x) for a CubedShell with constant sigma
   (src/main/coordinates/coordtrans_CubedSphere.c:125-256) */
static void cubsph_lamAB(const OPat *P, const double x[3], double X[3])
{
  int dir = P->dom / 2;
  double pm = 2 * (P->dom % 2) - 1;
  double rx = x[0] - P->xc[0], ry = x[1] - P->xc[1], rz = x[2] - P->xc[2];
  double r0, A, B;
  if (dir == 0) { r0 = rx; A = ry / rx; B = rz / rx; }
  else if (dir == 1) { r0 = ry; A = rx / ry; B = rz / ry; }
  else { r0 = rz; A = ry / rz; B = rx / rz; }
  double s = sqrt(1.0 + A * A + B * B);
  double a0 = pm * P->s0 / s, a1 = pm * P->s1 / s;
  X[0] = (r0 - a0) / (a1 - a0);
  X[1] = A;
  X[2] = B;
}

/* dX/dx for a CubedShell with constant sigma (dsigma = 0)
   (src/main/coordinates/coordtrans_CubedSphere.c:259-509) */
static void cubsph_dXdx(const OPat *P, const double X[3], double x[3],
                        double M[3][3])
{
  double lam = X[0], A = X[1], B = X[2];
  int dir = P->dom / 2;
  double pm = 2 * (P->dom % 2) - 1;
  double s = sqrt(1.0 + A * A + B * B);
  double a0 = pm * P->s0 / s, a1 = pm * P->s1 / s;
  double rx, ry, rz, rx2, ry2, rz2, rc2, mx, my, mz;
  double da0x = a0, da0y = a0, da0z = a0, da1x = a1, da1y = a1, da1z = a1;
  if (dir == 0) {
    rx = (a1 - a0) * lam + a0;
    x[0] = P->xc[0] + rx; x[1] = P->xc[1] + A * rx; x[2] = P->xc[2] + B * rx;
    ry = x[1] - P->xc[1]; rz = x[2] - P->xc[2];
    rx2 = rx * rx; ry2 = ry * ry; rz2 = rz * rz; rc2 = rx2 + ry2 + rz2;
    mx = (1.0 / rx - rx / rc2); my = (-ry / rc2); mz = (-rz / rc2);
    da0x *= (mx + 0.0); da0y *= (my + 0.0); da0z *= (mz + 0.0);
    da1x *= (mx + 0.0); da1y *= (my + 0.0); da1z *= (mz + 0.0);
    M[0][0] = (1.0 - da0x - (da1x - da0x) * lam) / (a1 - a0);
    M[0][1] = (-da0y - (da1y - da0y) * lam) / (a1 - a0);
    M[0][2] = (-da0z - (da1z - da0z) * lam) / (a1 - a0);
    M[1][0] = -ry / rx2; M[1][1] = 1.0 / rx; M[1][2] = 0.0;
    M[2][0] = -rz / rx2; M[2][1] = 0.0; M[2][2] = M[1][1];
  } else if (dir == 1) {
    ry = (a1 - a0) * lam + a0;
    x[1] = P->xc[1] + ry; x[0] = P->xc[0] + A * ry; x[2] = P->xc[2] + B * ry;
    rx = x[0] - P->xc[0]; rz = x[2] - P->xc[2];
    rx2 = rx * rx; ry2 = ry * ry; rz2 = rz * rz; rc2 = rx2 + ry2 + rz2;
    mx = (-rx / rc2); my = (1.0 / ry - ry / rc2); mz = (-rz / rc2);
    da0x *= (mx + 0.0); da0y *= (my + 0.0); da0z *= (mz + 0.0);
    da1x *= (mx + 0.0); da1y *= (my + 0.0); da1z *= (mz + 0.0);
    M[0][0] = (-da0x - (da1x - da0x) * lam) / (a1 - a0);
    M[0][1] = (1.0 - da0y - (da1y - da0y) * lam) / (a1 - a0);
    M[0][2] = (-da0z - (da1z - da0z) * lam) / (a1 - a0);
    M[1][0] = 1.0 / ry; M[1][1] = -rx / ry2; M[1][2] = 0.0;
    M[2][0] = 0.0; M[2][1] = -rz / ry2; M[2][2] = M[1][0];
  } else {
    rz = (a1 - a0) * lam + a0;
    x[2] = P->xc[2] + rz; x[1] = P->xc[1] + A * rz; x[0] = P->xc[0] + B * rz;
    rx = x[0] - P->xc[0]; ry = x[1] - P->xc[1];
    rx2 = rx * rx; ry2 = ry * ry; rz2 = rz * rz; rc2 = rx2 + ry2 + rz2;
    mx = (-rx / rc2); my = (-ry / rc2); mz = (1.0 / rz - rz / rc2);
    da0x *= (mx + 0.0); da0y *= (my + 0.0); da0z *= (mz + 0.0);
    da1x *= (mx + 0.0); da1y *= (my + 0.0); da1z *= (mz + 0.0);
    M[0][0] = (-da0x - (da1x - da0x) * lam) / (a1 - a0);
    M[0][1] = (-da0y - (da1y - da0y) * lam) / (a1 - a0);
    M[0][2] = (1.0 - da0z - (da1z - da0z) * lam) / (a1 - a0);
    M[1][0] = 0.0; M[1][1] = 1.0 / rz; M[1][2] = -ry / rz2;
    M[2][0] = M[1][1]; M[2][1] = 0.0; M[2][2] = -rx / rz2;
  }
}

static void pat_xyz(const OPat *P, const double X[3], double x[3])
{
  if (P->type == PAT_CUBEDSHELL) cubsph_xyz(P, X, x);
  else { x[0] = X[0]; x[1] = X[1]; x[2] = X[2]; }
}

static void pat_XYZ(const OPat *P, const double x[3], double X[3])
{
  if (P->type == PAT_CUBEDSHELL) cubsph_lamAB(P, x, X);
  else { X[0] = x[0]; X[1] = x[1]; X[2] = x[2]; }
}

/* ------------------------------------------------------------------ */
/* mesh: oct-tree leaves in Morton order                               */
/* ------------------------------------------------------------------ */

typedef struct {
  int p, l;
  char loc[32];     /* child octant digits per level (connections.c:1219-1233) */
  double bbox[6];
  int n;            /* points per direction */
  int use_fv;       /* 1: uniform points and fv_divf */
  long off;         /* first node of this element */
  int nfnb[6];
  int fnb[6][ORC_MAXNB];
  int nbf[6][ORC_MAXNB]; /* face of the neighbour that faces me */
  int ajmode[6];    /* 0 none, 1 alias, 2 oriented copy, 3 interpolation */
  int orient[6];    /* intrch*4 + rev1*2 + rev2 for ajmode 2 */
  double hmin;
  int hmin_pair;    /* hmin is a distance between two points (not the diagonal) */
} OElm;

typedef struct {
  int npat;
  OPat pat[8];
  int nelem;
  OElm *e;
  long nnodes;
  double *X, *x, *dXdx, *sqrtgdiag, *det;
  /* per-element face-plane offsets for the exchange buffers */
  long *foff;       /* foff[6*e+f]: start of face plane (n*n values) */
  long nface;       /* total face-plane values */
} OMesh;

static int elm_np(const OElm *E) { return E->n * E->n * E->n; }

/* element bbox from its location (src/main/amr/connections.c:1175-1206) */
static void elm_bbox_from_loc(const OPat *P, OElm *E)
{
  double LX[3];
  for (int f = 0; f < 6; f++) E->bbox[f] = P->bbox[f];
  for (int d = 0; d < 3; d++) LX[d] = E->bbox[2 * d + 1] - E->bbox[2 * d];
  for (int ll = 1; ll <= E->l; ll++) {
    int ijk = E->loc[ll - 1] - '0';
    for (int d = 0; d < 3; d++) {
      int bt = 1 << d;
      LX[d] /= 2.;
      if (ijk & bt) E->bbox[2 * d] = E->bbox[2 * d + 1] - LX[d];
      else E->bbox[2 * d + 1] = E->bbox[2 * d] + LX[d];
    }
  }
}

/* replace every flagged element by its 8 children in place, so the element
   list stays in depth-first (Morton) order
   (src/main/amr/storage.c:586-621, src/main/amr/refine.c:174-221) */
static void refine_flagged(OMesh *M, const int *flag)
{
  int cnt = 0;
  for (int e = 0; e < M->nelem; e++) cnt += flag[e] ? 8 : 1;
  OElm *ne = calloc(cnt, sizeof(OElm));
  int k = 0;
  for (int e = 0; e < M->nelem; e++) {
    if (!flag[e]) { ne[k++] = M->e[e]; continue; }
    for (int c = 0; c < 8; c++) {
      OElm C;
      memset(&C, 0, sizeof(C));
      C.p = M->e[e].p;
      C.l = M->e[e].l + 1;
      memcpy(C.loc, M->e[e].loc, sizeof(C.loc));
      C.loc[C.l - 1] = '0' + c;
      C.loc[C.l] = 0;
      C.n = M->e[e].n;
      C.use_fv = M->e[e].use_fv;
      elm_bbox_from_loc(&M->pat[C.p], &C);
      ne[k++] = C;
    }
  }
  free(M->e);
  M->e = ne;
  M->nelem = cnt;
}

static void mesh_add_roots(OMesh *M, int n)
{
  M->nelem = M->npat;
  M->e = calloc(M->nelem, sizeof(OElm));
  for (int p = 0; p < M->npat; p++) {
    OElm *E = &M->e[p];
    E->p = p; E->l = 0; E->loc[0] = 0; E->n = n;
    elm_bbox_from_loc(&M->pat[p], E);
  }
}

static void refine_uniform(OMesh *M, int luni)
{
  for (int l = 0; l < luni; l++) {
    int *flag = calloc(M->nelem, sizeof(int));
    for (int e = 0; e < M->nelem; e++) flag[e] = (M->e[e].l < luni);
    refine_flagged(M, flag);
    free(flag);
  }
}

/* ---- geometry helpers ---- */

static void elm_X_of_Xb(const OElm *E, const double Xb[3], double X[3])
{ /* src/main/coordinates/get_coords.c:384-394 */
  for (int d = 0; d < 3; d++) {
    int f = 2 * d;
    X[d] = 0.5 * ((E->bbox[f + 1] - E->bbox[f]) * Xb[d] + (E->bbox[f + 1] + E->bbox[f]));
  }
}

static void elm_dXbdX(const OElm *E, double dXbdX[3])
{ /* src/main/coordinates/get_coords.c:409-419 */
  for (int d = 0; d < 3; d++) dXbdX[d] = 2. / (E->bbox[2 * d + 1] - E->bbox[2 * d]);
}

static const double *elm_Xb1(const OElm *E)
{
  OBasis *b = basis(E->n);
  return E->use_fv ? b->Xu : b->Xb;
}

static void elm_point_X(const OElm *E, int i, int j, int k, double X[3])
{
  const double *xb = elm_Xb1(E);
  double Xb[3] = {xb[i], xb[j], xb[k]};
  elm_X_of_Xb(E, Xb, X);
}

/* does patch face f of patch p touch another patch? (bfaces) */
static int elm_on_patface(const OMesh *M, const OElm *E, int f)
{
  return E->bbox[f] == M->pat[E->p].bbox[f];
}

/* ---- neighbour search ---- */

static int rect_overlap(const double a[4], const double b[4], double tol)
{
  double lo0 = fmax(a[0], b[0]), hi0 = fmin(a[1], b[1]);
  double lo1 = fmax(a[2], b[2]), hi1 = fmin(a[3], b[3]);
  return (hi0 - lo0 > tol) && (hi1 - lo1 > tol);
}

/* face f of element E as an in-plane rectangle in patch coordinates */
static void face_rect(const OElm *E, int f, double r[4])
{
  int dir = f / 2, d1 = (dir == 0), d2 = (dir == 2) ? 1 : 2;
  r[0] = E->bbox[2 * d1]; r[1] = E->bbox[2 * d1 + 1];
  r[2] = E->bbox[2 * d2]; r[3] = E->bbox[2 * d2 + 1];
}

/* physical point at in-plane fractions (s,t) of face f */
static void face_point(const OMesh *M, const OElm *E, int f, double s,
                       double t, double x[3])
{
  int dir = f / 2, d1 = (dir == 0), d2 = (dir == 2) ? 1 : 2;
  double X[3];
  X[dir] = E->bbox[f];
  X[d1] = E->bbox[2 * d1] + s * (E->bbox[2 * d1 + 1] - E->bbox[2 * d1]);
  X[d2] = E->bbox[2 * d2] + t * (E->bbox[2 * d2 + 1] - E->bbox[2 * d2]);
  pat_xyz(&M->pat[E->p], X, x);
}

/* is physical point x on face f2 of element E2 (inside its rectangle)? */
static int point_on_elmface(const OMesh *M, const OElm *E2, int f2,
                            const double x[3], double tol)
{
  double X[3];
  int dir = f2 / 2, d1 = (dir == 0), d2 = (dir == 2) ? 1 : 2;
  pat_XYZ(&M->pat[E2->p], x, X);
  double L = fmax(fabs(E2->bbox[f2]), 1.0);
  if (fabs(X[dir] - E2->bbox[f2]) > tol * L) return 0;
  if (X[d1] < E2->bbox[2 * d1] - tol || X[d1] > E2->bbox[2 * d1 + 1] + tol) return 0;
  if (X[d2] < E2->bbox[2 * d2] - tol || X[d2] > E2->bbox[2 * d2 + 1] + tol) return 0;
  return 1;
}

/* face neighbours: same patch -> touching opposite faces with overlapping
   rectangles; other patch -> faces that share a 2D piece of surface.
   Neighbour lists are kept in element-id order. */
static void find_neighbours(OMesh *M)
{
  int ne = M->nelem;
#pragma omp parallel for schedule(dynamic, 16)
  for (int e = 0; e < ne; e++) {
    OElm *E = &M->e[e];
    for (int f = 0; f < 6; f++) {
      E->nfnb[f] = 0;
      double r[4];
      face_rect(E, f, r);
      int onpf = elm_on_patface(M, E, f);
      for (int e2 = 0; e2 < ne; e2++) {
        if (e2 == e) continue;
        OElm *E2 = &M->e[e2];
        int hit = -1;
        if (E2->p == E->p && !onpf) {
          int f2 = f ^ 1;
          if (E2->bbox[f2] == E->bbox[f]) {
            double r2[4];
            face_rect(E2, f2, r2);
            if (rect_overlap(r, r2, 1e-13)) hit = f2;
          }
        } else if (E2->p != E->p && onpf) {
          for (int f2 = 0; f2 < 6 && hit < 0; f2++) {
            if (!elm_on_patface(M, E2, f2)) continue;
            double xc[3], xc2[3];
            face_point(M, E, f, 0.5, 0.5, xc);
            face_point(M, E2, f2, 0.5, 0.5, xc2);
            if (point_on_elmface(M, E2, f2, xc, 1e-10) ||
                point_on_elmface(M, E, f, xc2, 1e-10)) {
              /* make sure it is not only an edge contact: an interior
                 point of the smaller face must be on the other face */
              double q[3];
              int ok = 0;
              face_point(M, E, f, 0.5, 0.5, q);
              if (point_on_elmface(M, E2, f2, q, 1e-10)) ok = 1;
              face_point(M, E2, f2, 0.5, 0.5, q);
              if (point_on_elmface(M, E, f, q, 1e-10)) ok = 1;
              if (ok) hit = f2;
            }
          }
        }
        if (hit >= 0) {
          if (E->nfnb[f] >= ORC_MAXNB) { fprintf(stderr, "too many nb\n"); abort(); }
          E->fnb[f][E->nfnb[f]] = e2;
          E->nbf[f][E->nfnb[f]] = hit;
          E->nfnb[f]++;
        }
      }
    }
  }
}

/* patch faces that touch no other patch are outer boundaries
   (src/main/amr/bfaces.c:894-910) */
static void mark_boundaries(OMesh *M)
{
  for (int p = 0; p < M->npat; p++)
    for (int f = 0; f < 6; f++) M->pat[p].bound[f] = 1;
  for (int e = 0; e < M->nelem; e++) {
    OElm *E = &M->e[e];
    for (int f = 0; f < 6; f++)
      if (elm_on_patface(M, E, f) && E->nfnb[f] > 0) M->pat[E->p].bound[f] = 0;
  }
}

/* nearest corner of the neighbour's face plane to physical point x0
   (src/main/amr/surface.c:885-930, nearest_corner_of_xyz_inplaneN) */
static double nearest_corner(const OMesh *M, const OElm *N, int nf,
                             const double x0[3], int inb[3])
{
  int dir = nf / 2, pl = (N->n - 1) * (nf % 2);
  double best = 1e300;
  for (int c2 = 0; c2 < 2; c2++)
    for (int c1 = 0; c1 < 2; c1++) {
      int I[3];
      int d1 = (dir == 0), d2 = (dir == 2) ? 1 : 2;
      I[dir] = pl; I[d1] = c1 * (N->n - 1); I[d2] = c2 * (N->n - 1);
      double X[3], x[3];
      elm_point_X(N, I[0], I[1], I[2], X);
      pat_xyz(&M->pat[N->p], X, x);
      double d = sqrt((x[0] - x0[0]) * (x[0] - x0[0]) + (x[1] - x0[1]) * (x[1] - x0[1]) +
                      (x[2] - x0[2]) * (x[2] - x0[2]));
      if (d < best) { best = d; inb[0] = I[0]; inb[1] = I[1]; inb[2] = I[2]; }
    }
  return best;
}

/* decide how ajsurf is built for every face (src/main/amr/surface.c:799-1080) */
static void set_ajmodes(OMesh *M)
{
#pragma omp parallel for schedule(dynamic, 16)
  for (int e = 0; e < M->nelem; e++) {
    OElm *E = &M->e[e];
    for (int f = 0; f < 6; f++) {
      E->ajmode[f] = 0; E->orient[f] = 0;
      if (E->nfnb[f] == 0) continue;
      E->ajmode[f] = 3;
      if (E->nfnb[f] != 1) continue;
      OElm *N = &M->e[E->fnb[f][0]];
      int nbf = E->nbf[f][0];
      if (N->p == E->p) {
        if (N->l == E->l && N->n == E->n && N->use_fv == E->use_fv) E->ajmode[f] = 1;
        continue;
      }
      if (N->nfnb[nbf] != 1 || N->n != E->n || N->use_fv != E->use_fv) continue;
      int dir = f / 2, pl = (E->n - 1) * (f % 2);
      int d1 = (dir == 0), d2 = (dir == 2) ? 1 : 2;
      int n1 = E->n, n2 = E->n;
      int odir = nbf / 2, od1 = (odir == 0), od2 = (odir == 2) ? 1 : 2;
      int on1 = N->n, on2 = N->n;
      int inb[3][3];
      double dist[3];
      int pts[3][2] = {{0, 0}, {n1 - 1, 0}, {0, n2 - 1}};
      for (int q = 0; q < 3; q++) {
        int I[3];
        I[dir] = pl; I[d1] = pts[q][0]; I[d2] = pts[q][1];
        double X[3], x0[3];
        elm_point_X(E, I[0], I[1], I[2], X);
        pat_xyz(&M->pat[E->p], X, x0);
        dist[q] = nearest_corner(M, N, nbf, x0, inb[q]);
        double mx0 = sqrt(x0[0] * x0[0] + x0[1] * x0[1] + x0[2] * x0[2]);
        if (mx0 > 0.) dist[q] /= mx0;
      }
      if (!(dist[0] <= DEQUALEPS && dist[1] <= DEQUALEPS && dist[2] <= DEQUALEPS)) continue;
      int a00 = inb[0][od1], a02 = inb[0][od2];
      int a10 = inb[1][od1], a12 = inb[1][od2];
      int a20 = inb[2][od1], a22 = inb[2][od2];
      int code = -1;
      if (a00 == 0 && a02 == 0) {
        if (a10 == on1 - 1 && a12 == 0 && a20 == 0 && a22 == on2 - 1) code = 0;
        else if (a10 == 0 && a12 == on2 - 1 && a20 == on1 - 1 && a22 == 0) code = 4;
      }
      if (code < 0 && a00 == on1 - 1 && a02 == 0) {
        if (a10 == 0 && a12 == 0 && a20 == on1 - 1 && a22 == on2 - 1) code = 2;
        else if (a10 == on1 - 1 && a12 == on2 - 1 && a20 == 0 && a22 == 0) code = 6;
      }
      if (code < 0 && a00 == 0 && a02 == on2 - 1) {
        if (a10 == on1 - 1 && a12 == on2 - 1 && a20 == 0 && a22 == 0) code = 1;
        else if (a10 == 0 && a12 == 0 && a20 == on1 - 1 && a22 == on2 - 1) code = 5;
      }
      if (code < 0 && a00 == on1 - 1 && a02 == on2 - 1) {
        if (a10 == 0 && a12 == on2 - 1 && a20 == on1 - 1 && a22 == 0) code = 3;
        else if (a10 == on1 - 1 && a12 == 0 && a20 == 0 && a22 == on2 - 1) code = 7;
      }
      if (code >= 0) { E->ajmode[f] = 2; E->orient[f] = code; }
    }
  }
}

/* smallest axial spacing at the 8 corners (src/main/coordinates/utilities.c:1028-1124) */
static double elm_find_hmin(const OMesh *M, const OElm *E, int *pair)
{
  const OPat *P = &M->pat[E->p];
  int n = E->n;
  double X0[3] = {E->bbox[0], E->bbox[2], E->bbox[4]};
  double X1[3] = {E->bbox[1], E->bbox[3], E->bbox[5]};
  double x0[3], x1[3];
  pat_xyz(P, X0, x0); pat_xyz(P, X1, x1);
  double d2 = 0.;
  for (int d = 0; d < 3; d++) d2 += (x1[d] - x0[d]) * (x1[d] - x0[d]);
  double hmin = sqrt(d2) / sqrt(3.);
  *pair = 0;
  int cs[8][3] = {{0, 0, 0}, {n - 1, 0, 0}, {0, n - 1, 0}, {n - 1, n - 1, 0},
                  {0, 0, n - 1}, {n - 1, 0, n - 1}, {0, n - 1, n - 1}, {n - 1, n - 1, n - 1}};
  for (int c = 0; c < 8; c++) {
    int i = cs[c][0], j = cs[c][1], k = cs[c][2];
    double Xa[3], xa[3];
    elm_point_X(E, i, j, k, Xa);
    pat_xyz(P, Xa, xa);
    for (int kk = -1; kk <= 1; kk++) {
      int kn = k + kk;
      if (kn < 0 || kn >= n) continue;
      for (int jj = -1; jj <= 1; jj++) {
        int jn = j + jj;
        if (jn < 0 || jn >= n) continue;
        for (int ii = -1; ii <= 1; ii++) {
          int in = i + ii;
          if (in < 0 || in >= n) continue;
          if (in == i && jn == j && kn == k) continue;
          if (in != i && jn != j) continue;
          if (in != i && kn != k) continue;
          if (jn != j && kn != k) continue;
          double Xb_[3], xb_[3];
          elm_point_X(E, in, jn, kn, Xb_);
          pat_xyz(P, Xb_, xb_);
          double dd = 0.;
          for (int d = 0; d < 3; d++) dd += (xb_[d] - xa[d]) * (xb_[d] - xa[d]);
          dd = sqrt(dd);
          if (dd < hmin) { hmin = dd; *pair = 1; }
        }
      }
    }
  }
  return hmin;
}

/* per-node geometry (src/main/coordinates/coordinates.c:133-224,728-805) */
static void mesh_geometry(OMesh *M)
{
  long off = 0, foff = 0;
  for (int e = 0; e < M->nelem; e++) {
    M->e[e].off = off;
    off += elm_np(&M->e[e]);
  }
  M->nnodes = off;
  free(M->foff);
  M->foff = malloc(sizeof(long) * 6 * M->nelem);
  for (int e = 0; e < M->nelem; e++)
    for (int f = 0; f < 6; f++) { M->foff[6 * e + f] = foff; foff += M->e[e].n * M->e[e].n; }
  M->nface = foff;
  long nn = M->nnodes;
  free(M->X); free(M->x); free(M->dXdx); free(M->sqrtgdiag); free(M->det);
  M->X = malloc(sizeof(double) * 3 * nn);
  M->x = malloc(sizeof(double) * 3 * nn);
  M->dXdx = malloc(sizeof(double) * 9 * nn);
  M->sqrtgdiag = malloc(sizeof(double) * 3 * nn);
  M->det = malloc(sizeof(double) * nn);
#pragma omp parallel for schedule(dynamic, 16)
  for (int e = 0; e < M->nelem; e++) {
    OElm *E = &M->e[e];
    const OPat *P = &M->pat[E->p];
    int n = E->n;
    double dXbdX[3];
    elm_dXbdX(E, dXbdX);
    double detb = dXbdX[0] * dXbdX[1] * dXbdX[2];
    for (int k = 0; k < n; k++)
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
          long g = E->off + i + n * (j + n * k);
          double X[3], x[3], D[3][3] = {{0}};
          elm_point_X(E, i, j, k, X);
          if (P->type == PAT_CUBEDSHELL) {
            cubsph_dXdx(P, X, x, D);
          } else {
            for (int d = 0; d < 3; d++) { x[d] = X[d]; D[d][d] = 1.; }
          }
          for (int d = 0; d < 3; d++) { M->X[d * nn + g] = X[d]; M->x[d * nn + g] = x[d]; }
          for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) M->dXdx[(3 * a + b) * nn + g] = D[a][b];
          double det3 = D[0][0] * (D[1][1] * D[2][2] - D[1][2] * D[2][1]) -
                        D[0][1] * (D[1][0] * D[2][2] - D[1][2] * D[2][0]) +
                        D[0][2] * (D[1][0] * D[2][1] - D[1][1] * D[2][0]);
          M->det[g] = (P->type == PAT_CUBEDSHELL) ? detb * det3 : detb;
          for (int a = 0; a < 3; a++) {
            double gam = 0.0;
            for (int l = 0; l < 3; l++) gam += D[a][l] * D[a][l];
            M->sqrtgdiag[a * nn + g] = dXbdX[a] * sqrt(gam);
          }
        }
    E->hmin = elm_find_hmin(M, E, &E->hmin_pair);
  }
}

static void mesh_finish(OMesh *M)
{
  find_neighbours(M);
  mark_boundaries(M);
  set_ajmodes(M);
  mesh_geometry(M);
}

/* BoxMesh with one patch xc +- dout (src/main/amr/mesh.c:354-445,
   src/main/coordinates/setup_Boxes.c:11-32) */
OMesh *orc_mesh_box(int luni, int n, double dout, const double xc[3])
{
  OMesh *M = calloc(1, sizeof(OMesh));
  M->npat = 1;
  OPat *P = &M->pat[0];
  P->type = PAT_BOX;
  for (int d = 0; d < 3; d++) {
    P->xc[d] = xc ? xc[d] : 0.;
    P->bbox[2 * d] = P->xc[d] - dout;
    P->bbox[2 * d + 1] = P->xc[d] + dout;
  }
  mesh_add_roots(M, n);
  refine_uniform(M, luni);
  mesh_finish(M);
  return M;
}

/* CubedSphere shell of 6 CubedShell patches with constant sigma
   (src/main/amr/mesh.c:625-669, src/main/coordinates/setup_CubedSpheres.c:428-443,653-684,935-944) */
static OMesh *shell_base(int n, double rin, double rout)
{
  OMesh *M = calloc(1, sizeof(OMesh));
  M->npat = 6;
  for (int p = 0; p < 6; p++) {
    OPat *P = &M->pat[p];
    P->type = PAT_CUBEDSHELL;
    P->dom = p;
    P->s0 = rin; P->s1 = rout;
    P->xc[0] = P->xc[1] = P->xc[2] = 0.;
    /* set_AB_min_max_from_Din with all Din equal gives [-1,1]x[-1,1] */
    double b[6] = {0., 1., -rin / rin, rin / rin, -rin / rin, rin / rin};
    memcpy(P->bbox, b, sizeof(b));
  }
  mesh_add_roots(M, n);
  return M;
}

OMesh *orc_mesh_shell(int luni, int n, double rin, double rout)
{
  OMesh *M = shell_base(n, rin, rout);
  refine_uniform(M, luni);
  mesh_finish(M);
  return M;
}

/* shell refined uniformly to luni, then in nested spheres about the origin:
   elements whose centre lies within radius*2^(levels-1-i) are refined in
   pass i (src/main/amr/refine.c:1025-1100) */
OMesh *orc_mesh_amr_shell(int luni, int n, double rin, double rout,
                          double radius, int levels)
{
  OMesh *M = shell_base(n, rin, rout);
  refine_uniform(M, luni);
  double r = radius * pow(2., levels - 1);
  for (int i = 0; i < levels; i++) {
    int *flag = calloc(M->nelem, sizeof(int));
    for (int e = 0; e < M->nelem; e++) {
      OElm *E = &M->e[e];
      double Xc[3], xcn[3];
      for (int d = 0; d < 3; d++) Xc[d] = 0.5 * (E->bbox[2 * d + 1] + E->bbox[2 * d]);
      pat_xyz(&M->pat[E->p], Xc, xcn);
      double rr = sqrt(xcn[0] * xcn[0] + xcn[1] * xcn[1] + xcn[2] * xcn[2]);
      flag[e] = (rr <= r);
    }
    refine_flagged(M, flag);
    free(flag);
    r *= 0.5;
  }
  mesh_finish(M);
  return M;
}

/* change the number of points per direction of every element (p-refinement,
   src/main/amr/storage.c update_node_n_pt_typ) and rebuild tables */
void orc_mesh_set_n(OMesh *M, const int *n)
{
  for (int e = 0; e < M->nelem; e++) M->e[e].n = n[e];
  set_ajmodes(M);
  mesh_geometry(M);
}

/* elements of every patch in use_fv mode on uniform points (amr_uniform_p
   and amr_fv_p naming all patches: mesh.c:40-80) */
void orc_mesh_set_fv(OMesh *M, int use_fv)
{
  for (int e = 0; e < M->nelem; e++) M->e[e].use_fv = use_fv;
  set_ajmodes(M);
  mesh_geometry(M);
}

/* FV on uniform points where use_fv[e] is set, DG elsewhere (the element
   switch of evolve_switch_troubled_nodes_mesh, trouble.c:186-261) */
void orc_mesh_set_fv_elems(OMesh *M, const int *use_fv)
{
  for (int e = 0; e < M->nelem; e++) M->e[e].use_fv = use_fv[e] != 0;
  set_ajmodes(M);
  mesh_geometry(M);
}

/* FV on uniform points in the patches of bit mask patmask, DG elsewhere
   (amr_uniform_p = amr_fv_p naming some patches, mesh.c:40-80) */
void orc_mesh_set_fv_patches(OMesh *M, int patmask)
{
  for (int e = 0; e < M->nelem; e++) M->e[e].use_fv = (patmask >> M->e[e].p) & 1;
  set_ajmodes(M);
  mesh_geometry(M);
}

void orc_mesh_free(OMesh *M)
{
  if (!M) return;
  free(M->e); free(M->X); free(M->x); free(M->dXdx); free(M->sqrtgdiag);
  free(M->det); free(M->foff); free(M);
}

/* ---- queries for the test harness ---- */
int orc_mesh_nelem(const OMesh *M) { return M->nelem; }
long orc_mesh_nnodes(const OMesh *M) { return M->nnodes; }
int orc_mesh_npat(const OMesh *M) { return M->npat; }

/* info[0..]: p, l, n, off(lo 31 bits), nfnb[6], ajmode[6], orient[6] */
void orc_mesh_elm_info(const OMesh *M, int e, long *info)
{
  const OElm *E = &M->e[e];
  info[0] = E->p; info[1] = E->l; info[2] = E->n; info[3] = E->off;
  for (int f = 0; f < 6; f++) {
    info[4 + f] = E->nfnb[f];
    info[10 + f] = E->ajmode[f];
    info[16 + f] = E->orient[f];
  }
}

void orc_mesh_elm_bbox(const OMesh *M, int e, double *bb)
{
  memcpy(bb, M->e[e].bbox, 6 * sizeof(double));
}

int orc_mesh_elm_fnb(const OMesh *M, int e, int f, int *nb, int *nbf)
{
  const OElm *E = &M->e[e];
  for (int i = 0; i < E->nfnb[f]; i++) { nb[i] = E->fnb[f][i]; if (nbf) nbf[i] = E->nbf[f][i]; }
  return E->nfnb[f];
}

int orc_mesh_pat_bound(const OMesh *M, int p, int f) { return M->pat[p].bound[f]; }

void orc_mesh_coords(const OMesh *M, double *X, double *x)
{
  if (X) memcpy(X, M->X, sizeof(double) * 3 * M->nnodes);
  if (x) memcpy(x, M->x, sizeof(double) * 3 * M->nnodes);
}

void orc_mesh_geom(const OMesh *M, double *dXdx, double *sqrtgdiag, double *det)
{
  if (dXdx) memcpy(dXdx, M->dXdx, sizeof(double) * 9 * M->nnodes);
  if (sqrtgdiag) memcpy(sqrtgdiag, M->sqrtgdiag, sizeof(double) * 3 * M->nnodes);
  if (det) memcpy(det, M->det, sizeof(double) * M->nnodes);
}

/* dt = min over elements of dtfac * hmin, or uniform_dtfac * hmin where
   hmin lies between two points along a uniform (FV) direction
   (src/main/amr/storage.c:1305-1366, src/main/coordinates/utilities.c:
   1218-1304; the FV elements here are uniform along all three directions) */
double orc_mesh_dt(const OMesh *M, double dtfac, double uniform_dtfac)
{
  double dt = 0.1 * 1.7976931348623157e308;
  for (int e = 0; e < M->nelem; e++) {
    const OElm *E = &M->e[e];
    double dtm = (E->use_fv && E->hmin_pair) ? uniform_dtfac * E->hmin : dtfac * E->hmin;
    if (dtm < dt) dt = dtm;
  }
  return dt;
}

double orc_mesh_elm_hmin(const OMesh *M, int e) { return M->e[e].hmin; }

/* ------------------------------------------------------------------ */
/* scalarwave1                                                         */
/* ------------------------------------------------------------------ */

typedef struct {
  int numflux;     /* 0 upwind (scalarwave1_numflux1d_upwind), 1 LLF */
  double k[3];     /* scalarwave1_k */
} OSW;

#define NV 5

/* plane wave (src/tests/scalarwave1/scalarwave1.c:503-524) */
static void sw_profile_pt(const double k[3], const double xyz[3], double t, double *u)
{
  double kx = k[0], ky = k[1], kz = k[2];
  double om = sqrt(kx * kx + ky * ky + kz * kz);
  double arg = kx * xyz[0] + ky * xyz[1] + kz * xyz[2] - om * t;
  double s = sin(arg), c = cos(arg);
  u[4] = s; u[0] = -om * c; u[1] = kx * c; u[2] = ky * c; u[3] = kz * c;
}

/* d/dt of the plane wave (src/tests/scalarwave1/scalarwave1.c:544-566) */
static void sw_dt_profile_pt(const double k[3], const double xyz[3], double t, double *du)
{
  double u[5];
  double om = sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
  sw_profile_pt(k, xyz, t, u);
  du[0] = -om * om * u[4]; du[1] = om * k[0] * u[4]; du[2] = om * k[1] * u[4];
  du[3] = om * k[2] * u[4]; du[4] = u[0];
}

void orc_set_profile(const OMesh *M, const double k[3], double t, double *u, double *dudt)
{
  long nn = M->nnodes;
#pragma omp parallel for
  for (long g = 0; g < nn; g++) {
    double xyz[3] = {M->x[g], M->x[nn + g], M->x[2 * nn + g]}, v[5];
    sw_profile_pt(k, xyz, t, v);
    for (int l = 0; l < NV; l++) u[l * nn + g] = v[l];
    if (dudt) {
      sw_dt_profile_pt(k, xyz, t, v);
      for (int l = 0; l < NV; l++) dudt[l * nn + g] = v[l];
    }
  }
}

/* f.n for u = (pi,cx,cy,cz,phi) (src/tests/scalarwave1/scalarwave1.c:39-48) */
static void sw_flux1d(double *f, const double nrm[3], const double *u)
{
  f[0] = -(u[1] * nrm[0] + u[2] * nrm[1] + u[3] * nrm[2]);
  f[1] = -u[0] * nrm[0];
  f[2] = -u[0] * nrm[1];
  f[3] = -u[0] * nrm[2];
  f[4] = 0.;
}

/* eigen-decomposition upwind flux (src/tests/scalarwave1/scalarwave1.c:324-499) */
static void sw_numflux_upwind(const double nrm[3], const double *lami, const double *ui,
                              const double *ua, double *fnum)
{
  const int nf = 4;
  double lp[4][4] = {{0}}, ln[4][4] = {{0}}, ip[4][4] = {{0}}, in_[4][4] = {{0}};
  double sp[4][4] = {{0}}, sn[4][4] = {{0}}, fi[4] = {0}, fa[4] = {0};
  double nx = nrm[0], ny = nrm[1], nz = nrm[2];
  double R[4][4], iR[4][4];
  for (int l = 0; l < nf; l++) {
    if (lami[l] > 0.) lp[l][l] = lami[l];
    else ln[l][l] = lami[l];
  }
  if (nz != 0) {
    double R_[4][4] = {{1.0, 1.0, 0.0, 0.0}, {-nx, nx, 0.0, nz}, {-ny, ny, nz, 0.0}, {-nz, nz, -ny, -nx}};
    double iR_[4][4] = {{1.0 / 2.0, -nx / 2.0, -ny / 2.0, -nz / 2.0},
                        {1.0 / 2.0, nx / 2.0, ny / 2.0, nz / 2.0},
                        {0.0, -(nx * ny) / nz, (nx * nx + nz * nz) / nz, -ny},
                        {0.0, (ny * ny + nz * nz) / nz, -(nx * ny) / nz, -nx}};
    memcpy(R, R_, sizeof(R)); memcpy(iR, iR_, sizeof(iR));
  } else if (nx != 0) {
    double R_[4][4] = {{1.0, 1.0, 0.0, 0.0}, {-nx, nx, -nz, -ny}, {-ny, ny, 0.0, nx}, {-nz, nz, nx, 0.0}};
    double iR_[4][4] = {{1.0 / 2.0, -nx / 2.0, -ny / 2.0, -nz / 2.0},
                        {1.0 / 2.0, nx / 2.0, ny / 2.0, nz / 2.0},
                        {0.0, -nz, -(ny * nz) / nx, (nx * nx + ny * ny) / nx},
                        {0.0, -ny, (nx * nx + nz * nz) / nx, -(ny * nz) / nx}};
    memcpy(R, R_, sizeof(R)); memcpy(iR, iR_, sizeof(iR));
  } else {
    double R_[4][4] = {{1.0, 1.0, 0.0, 0.0}, {-nx, nx, 0.0, ny}, {-ny, ny, -nz, -nx}, {-nz, nz, ny, 0.0}};
    double iR_[4][4] = {{1.0 / 2.0, -nx / 2.0, -ny / 2.0, -nz / 2.0},
                        {1.0 / 2.0, nx / 2.0, ny / 2.0, nz / 2.0},
                        {0.0, -(nx * nz) / ny, -nz, (nx * nx + ny * ny) / ny},
                        {0.0, (ny * ny + nz * nz) / ny, -nx, -(nx * nz) / ny}};
    memcpy(R, R_, sizeof(R)); memcpy(iR, iR_, sizeof(iR));
  }
  for (int l = 0; l < nf; l++)
    for (int m = 0; m < nf; m++)
      for (int s = 0; s < nf; s++) { ip[l][m] += lp[l][s] * iR[s][m]; in_[l][m] += ln[l][s] * iR[s][m]; }
  for (int l = 0; l < nf; l++)
    for (int m = 0; m < nf; m++)
      for (int s = 0; s < nf; s++) { sp[l][m] += R[l][s] * ip[s][m]; sn[l][m] += R[l][s] * in_[s][m]; }
  for (int l = 0; l < nf; l++)
    for (int m = 0; m < nf; m++) { fi[l] += sp[l][m] * ui[m]; fa[l] += sn[l][m] * ua[m]; }
  for (int l = 0; l < nf; l++) fnum[l] = fi[l] + fa[l];
  fnum[4] = 0.;
}

/* LLF flux (src/main/dg/numflux1d.c:85-112) */
static void numflux_LLF(int nf, const double *lami, const double *lama, const double *ui,
                        const double *ua, const double *fi, const double *fa, double *fnum)
{
  double apl = 0., ami = 0.;
  for (int i = 0; i < nf; i++) {
    apl = fmax(apl, fmax(lama[i], lami[i]));
    ami = fmin(ami, fmin(lama[i], lami[i]));
  }
  apl = fabs(apl); ami = fabs(ami);
  double amax = fmax(ami, apl);
  for (int i = 0; i < nf; i++) fnum[i] = 0.5 * (fa[i] + fi[i] - amax * (ua[i] - ui[i]));
}

/* out = D_dir u (src/main/amr/array.c:36-45,100-300 via src/main/basis/basis.c:61-95) */
static void deriv1(int n, const double *DT, int dir, const double *u, double *du)
{
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        double sum = 0.0;
        for (int l = 0; l < n; l++) {
          int src = (dir == 0) ? l + n * (j + n * k) : (dir == 1) ? i + n * (l + n * k)
                                                                  : i + n * (j + n * l);
          int ii = (dir == 0) ? i : (dir == 1) ? j : k;
          sum += DT[l + n * ii] * u[src];
        }
        du[i + n * (j + n * k)] = sum;
      }
}

/* Cartesian partials of one scalar: du[m] = sum_i dXdx[i][m] dXbdX[i] D_i u
   (src/main/coordinates/derivs.c:102-181) */
static void cart_partials(const OMesh *M, const OElm *E, const double *DT, const double *u,
                          double *du[3])
{
  int n = E->n, np = elm_np(E);
  double dXbdX[3];
  elm_dXbdX(E, dXbdX);
  for (int d = 0; d < 3; d++) deriv1(n, DT, d, u, du[d]);
  for (int i = 0; i < np; i++)
    for (int m = 0; m < 3; m++) du[m][i] *= dXbdX[m];
  if (M->pat[E->p].type == PAT_CUBEDSHELL) {
    long nn = M->nnodes;
    for (int i = 0; i < np; i++) {
      long g = E->off + i;
      double dw[3];
      for (int m = 0; m < 3; m++) {
        dw[m] = 0.;
        for (int a = 0; a < 3; a++) dw[m] += M->dXdx[(3 * a + m) * nn + g] * du[a][i];
      }
      for (int m = 0; m < 3; m++) du[m][i] = dw[m];
    }
  }
}

/* divU = d_x Ux + d_y Uy + d_z Uz (src/main/coordinates/derivs.c:756-815) */
static void cart_divergence(const OMesh *M, const OElm *E, const double *DT, const double *Ux,
                            const double *Uy, const double *Uz, double *divU, double *tmp)
{
  int np = elm_np(E);
  double *t1 = tmp, *t2 = tmp + np, *t3 = tmp + 2 * np;
  double *dU[3] = {divU, t2, t3};
  cart_partials(M, E, DT, Ux, dU);
  dU[0] = t1;
  cart_partials(M, E, DT, Uy, dU);
  for (int i = 0; i < np; i++) divU[i] += t2[i];
  cart_partials(M, E, DT, Uz, dU);
  for (int i = 0; i < np; i++) divU[i] += t3[i];
}

/* volume part (src/tests/scalarwave1/scalarwave1.c:62-134,242-284) */
static void sw_vol_rhs(const OMesh *M, const OElm *E, const double *u, double *r,
                       double *divf, double *work)
{
  int n = E->n, np = elm_np(E);
  long nn = M->nnodes, o = E->off;
  const double *DT = basis(n)->DT;
  double *fpi[3], *fc[3][3];
  for (int d = 0; d < 3; d++) fpi[d] = work + d * np;
  for (int c = 0; c < 3; c++)
    for (int d = 0; d < 3; d++) fc[c][d] = work + (3 + 3 * c + d) * np;
  double *tmp = work + 12 * np;
  for (int i = 0; i < np; i++) {
    double uu[5] = {u[0 * nn + o + i], u[1 * nn + o + i], u[2 * nn + o + i], u[3 * nn + o + i],
                    u[4 * nn + o + i]};
    for (int d = 0; d < 3; d++) {
      double no[3] = {d == 0, d == 1, d == 2}, f[5];
      sw_flux1d(f, no, uu);
      fpi[d][i] = f[0];
      for (int c = 0; c < 3; c++) fc[c][d][i] = f[1 + c];
    }
  }
  cart_divergence(M, E, DT, fpi[0], fpi[1], fpi[2], divf + 0 * np, tmp);
  for (int c = 0; c < 3; c++)
    cart_divergence(M, E, DT, fc[c][0], fc[c][1], fc[c][2], divf + (1 + c) * np, tmp);
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        int ijk = i + n * (j + n * k);
        int onface = (i == 0 || i == n - 1 || j == 0 || j == n - 1 || k == 0 || k == n - 1);
        double interior = !onface;
        r[4 * nn + o + ijk] = u[0 * nn + o + ijk];
        for (int l = 0; l < 4; l++) r[l * nn + o + ijk] = -divf[l * np + ijk] * interior;
      }
}

/* fill face plane f of element E from state u (set_mysurf, surface.c:231-256) */
static void set_mysurf(const OMesh *M, const OElm *E, int f, const double *u, int v,
                       double *plane)
{
  int n = E->n, dir = f / 2, p = (n - 1) * (f % 2);
  long nn = M->nnodes;
  for (int b = 0; b < n; b++)
    for (int a = 0; a < n; a++) {
      int I[3];
      int d1 = (dir == 0), d2 = (dir == 2) ? 1 : 2;
      I[dir] = p; I[d1] = a; I[d2] = b;
      /* Ind_n_norm layout: index a + n*b in both cases */
      plane[a + n * b] = u[v * nn + E->off + I[0] + n * (I[1] + n * I[2])];
    }
}

/* oriented copy (src/main/amr/surface.c:1361-1466); face planes use the
   in-plane index a + n1*b with a along Dir1_norm(dir), b along Dir2_norm(dir) */
static void copy_oriented(int n, int odir, int code, const double *nb, double *aj)
{
  int intrch = (code >> 2) & 1, rev1 = (code >> 1) & 1, rev2 = code & 1;
  for (int I2 = 0; I2 < n; I2++)
    for (int I1 = 0; I1 < n; I1++) {
      int oi, oj, ok;
      if (odir == 0) { oi = 0; oj = I1; ok = I2; }
      else if (odir == 1) { oi = I1; oj = 0; ok = I2; }
      else { oi = I1; oj = I2; ok = 0; }
      int i2 = oi, j2 = oj, k2 = ok;
      if (odir == 0) {
        if (!intrch) { if (rev1) j2 = n - 1 - oj; if (rev2) k2 = n - 1 - ok; }
        else { j2 = ok; k2 = oj; if (rev1) j2 = n - 1 - ok; if (rev2) k2 = n - 1 - oj; }
      } else if (odir == 1) {
        if (!intrch) { if (rev1) i2 = n - 1 - oi; if (rev2) k2 = n - 1 - ok; }
        else { i2 = ok; k2 = oi; if (rev1) i2 = n - 1 - ok; if (rev2) k2 = n - 1 - oi; }
      } else {
        if (!intrch) { if (rev1) i2 = n - 1 - oi; if (rev2) j2 = n - 1 - oj; }
        else { i2 = oj; j2 = oi; if (rev1) i2 = n - 1 - oj; if (rev2) j2 = n - 1 - oi; }
      }
      /* neighbour plane has n[odir]=1: Ind_n over (n0,n1) with the normal
         dimension collapsed */
      int ind2;
      if (odir == 0) ind2 = j2 + n * k2;
      else if (odir == 1) ind2 = i2 + n * k2;
      else ind2 = i2 + n * j2;
      aj[I1 + n * I2] = nb[ind2];
    }
}

/* build ajsurf of face f for variable plane data (src/main/amr/surface.c:799-1357) */
static void build_ajsurf(const OMesh *M, const OElm *E, int f, const double *planes, int v,
                         double *aj)
{
  int n = E->n;
  int mode = E->ajmode[f];
  long nf = M->nface;
  if (mode == 1) {
    int e2 = E->fnb[f][0], f2 = E->nbf[f][0];
    memcpy(aj, planes + v * nf + M->foff[6 * e2 + f2], sizeof(double) * n * n);
    return;
  }
  if (mode == 2) {
    int e2 = E->fnb[f][0], f2 = E->nbf[f][0];
    copy_oriented(n, f2 / 2, E->orient[f], planes + v * nf + M->foff[6 * e2 + f2], aj);
    return;
  }
  /* interpolation from every neighbour, averaged where several hold a point */
  int dir = f / 2, pl = (n - 1) * (f % 2), d1 = (dir == 0), d2 = (dir == 2) ? 1 : 2;
  const OPat *P = &M->pat[E->p];
  for (int b = 0; b < n; b++)
    for (int a = 0; a < n; a++) {
      int I[3];
      I[dir] = pl; I[d1] = a; I[d2] = b;
      double X[3], x[3];
      elm_point_X(E, I[0], I[1], I[2], X);
      pat_xyz(P, X, x);
      double sum = 0.;
      int cnt = 0;
      for (int ni = 0; ni < E->nfnb[f]; ni++) {
        const OElm *N = &M->e[E->fnb[f][ni]];
        int nbf = E->nbf[f][ni], nd = nbf / 2;
        int nd1 = (nd == 0), nd2 = (nd == 2) ? 1 : 2;
        double XN[3];
        if (N->p == E->p) { XN[0] = X[0]; XN[1] = X[1]; XN[2] = X[2]; }
        else pat_XYZ(&M->pat[N->p], x, XN);
        /* inside neighbour (array_find_Xplane_in_node): within its face
           rectangle, boundaries included */
        double tol = 1e-10;
        if (XN[nd1] < N->bbox[2 * nd1] - tol || XN[nd1] > N->bbox[2 * nd1 + 1] + tol) continue;
        if (XN[nd2] < N->bbox[2 * nd2] - tol || XN[nd2] > N->bbox[2 * nd2 + 1] + tol) continue;
        double Cb0 = (2. * XN[nd1] - (N->bbox[2 * nd1 + 1] + N->bbox[2 * nd1])) /
                     (N->bbox[2 * nd1 + 1] - N->bbox[2 * nd1]);
        double Cb1 = (2. * XN[nd2] - (N->bbox[2 * nd2 + 1] + N->bbox[2 * nd2])) /
                     (N->bbox[2 * nd2 + 1] - N->bbox[2 * nd2]);
        /* 2D tensor interpolation (src/main/basis/interpolate.c:296-363):
           Lagrange from a DG neighbour, piecewise linear from an FV one on
           uniform points (fv_surface_interp = linear, surface.c:1182-1245) */
        int nn_ = N->n;
        OBasis *bn = basis(nn_);
        double B1[ORC_MAXN], B2[ORC_MAXN];
        for (int q = 0; q < nn_; q++) {
          if (N->use_fv && g_fv_lagrange_if_dg && !E->use_fv) {
            /* Lagrange on the uniform points into a DG element
               (surface.c:1198-1238, LagrangeUNIorder 12 >= n) */
            B1[q] = lagrange_on_points(q, Cb0, nn_, bn->Xu);
            B2[q] = lagrange_on_points(q, Cb1, nn_, bn->Xu);
          } else if (N->use_fv && g_fv_surface_interp == 1) {
            B1[q] = pw_parab_of_x(q, Cb0, nn_, bn->Xu);
            B2[q] = pw_parab_of_x(q, Cb1, nn_, bn->Xu);
          } else if (N->use_fv) {
            B1[q] = pw_linear_of_x(q, Cb0, nn_, bn->Xu);
            B2[q] = pw_linear_of_x(q, Cb1, nn_, bn->Xu);
          } else {
            B1[q] = lagrange_of_x(q, Cb0, nn_, bn->Xb, bn->WL);
            B2[q] = lagrange_of_x(q, Cb1, nn_, bn->Xb, bn->WL);
          }
        }
        const double *np_ = planes + v * nf + M->foff[6 * E->fnb[f][ni] + nbf];
        double s = 0.;
        for (int q2 = 0; q2 < nn_; q2++)
          for (int q1 = 0; q1 < nn_; q1++) s += np_[q1 + nn_ * q2] * B1[q1] * B2[q2];
        sum += s;
        cnt++;
      }
      if (cnt > 1) sum /= cnt;
      aj[a + n * b] = sum;
    }
}

/* the adjacent values ajsurf of variable v on face f of element e for state
   u (set_ajsurf_vl, surface.c:799-1358): out[a + n b] */
void orc_face_values(const OMesh *M, const double *u, int e, int f, int v, double *out)
{
  double *planes = malloc(sizeof(double) * NV * M->nface);
  for (int e2 = 0; e2 < M->nelem; e2++)
    for (int f2 = 0; f2 < 6; f2++)
      for (int v2 = 0; v2 < NV; v2++)
        set_mysurf(M, &M->e[e2], f2, u, v2, planes + v2 * M->nface + M->foff[6 * e2 + f2]);
  build_ajsurf(M, &M->e[e], f, planes, v, out);
  free(planes);
}

/* surface part: numerical flux and lift on each face node, then r -= divf
   on all face points (src/tests/scalarwave1/scalarwave1.c:139-209,287-321,
   src/main/dg/dg.c:30-151, src/main/main/variables.c:1044-1071) */
static void sw_surf_rhs(const OMesh *M, const OSW *sw, const OElm *E, double t, const double *u,
                        const double *planes, double *r, double *divf)
{
  int n = E->n, np = elm_np(E);
  long nn = M->nnodes, o = E->off;
  OBasis *bb = basis(n);
  double aj[NV][ORC_MAXN * ORC_MAXN];
  for (int f = 0; f < 6; f++) {
    int dir = f / 2, p = (n - 1) * (f % 2), d1 = (dir == 0), d2 = (dir == 2) ? 1 : 2;
    double sig = 2 * (f % 2) - 1;
    double oow = 1. / bb->Wq[p];
    int have_aj = (E->nfnb[f] > 0);
    int onbound = elm_on_patface(M, E, f) && M->pat[E->p].bound[f];
    if (have_aj)
      for (int v = 0; v < NV; v++) build_ajsurf(M, E, f, planes, v, aj[v]);
    for (int b = 0; b < n; b++)
      for (int a = 0; a < n; a++) {
        int I[3];
        I[dir] = p; I[d1] = a; I[d2] = b;
        int ijk = I[0] + n * (I[1] + n * I[2]);
        int JK = a + n * b;
        long g = o + ijk;
        double nrm[3];
        for (int d = 0; d < 3; d++) nrm[d] = M->dXdx[(3 * dir + d) * nn + g];
        double smag = sig * sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
        if (smag == 0.0) smag = sig;
        for (int d = 0; d < 3; d++) nrm[d] /= smag;
        double lam[5] = {1.0, -1.0, 0.0, 0.0, 0.0};
        double ui[5], ua[5] = {0, 0, 0, 0, 0}, fi[5], fa[5], fnum[5];
        for (int l = 0; l < NV; l++) ui[l] = u[l * nn + g];
        if (have_aj) {
          for (int l = 0; l < NV; l++) ua[l] = aj[l][JK];
        } else if (onbound) {
          double xyz[3] = {M->x[g], M->x[nn + g], M->x[2 * nn + g]};
          sw_profile_pt(sw->k, xyz, t, ua);
        }
        sw_flux1d(fi, nrm, ui);
        sw_flux1d(fa, nrm, ua);
        if (sw->numflux == 1) numflux_LLF(NV, lam, lam, ui, ua, fi, fa, fnum);
        else sw_numflux_upwind(nrm, lam, ui, ua, fnum);
        double gd_ow = M->sqrtgdiag[dir * nn + g] * oow;
        for (int l = 0; l < 4; l++) {
          double F = (fnum[l] - fi[l] * 1.0) * 1.0;
          divf[l * np + ijk] += F * (0.0 * 0.0 + 1.0 * gd_ow);
        }
      }
  }
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++)
        if (i == 0 || i == n - 1 || j == 0 || j == n - 1 || k == 0 || k == n - 1) {
          int ijk = i + n * (j + n * k);
          for (int l = 0; l < 4; l++) r[l * nn + o + ijk] += -1. * divf[l * np + ijk];
        }
}

static OSW sw_from(int numflux, const double *k)
{
  OSW s;
  s.numflux = numflux;
  for (int d = 0; d < 3; d++) s.k[d] = k ? k[d] : 0.;
  return s;
}

/* ------------------------------------------------------------------ */
/* FV elements: fv_divf for scalarwave1 (src/main/dg/fv.c:115-624,     */
/* src/tests/scalarwave1/scalarwave1.c:211-237,755-790) with nmesh's   */
/* defaults fv_flux = fnum, fv_divf_extrap = dnfn_extrap1,             */
/* fv_divf_use_only_right_flux = yes, fv_WENOm3_opt_weightratio = 2,   */
/* fv_rec_WENOm_s1/s2 and fv_divf_extrap_s1/s2/opt off                 */
/* (src/main/dg/nmesh_dg.c:26-66, dg.c:285-400); Cartesian elements    */
/* (the midpoint metric equals the node metric)                        */
/* ------------------------------------------------------------------ */

double orc_rec1d_p_WENO5(const double *u, int im);
double orc_rec1d_m_WENO5(const double *u, int im);
double orc_rec1d_p_WENOZ(const double *u, int im);
double orc_rec1d_m_WENOZ(const double *u, int im);

#define ORC_WENO3_EPS 1e-6 /* WENO3_epsilon, rec1d.c:35, u_scale = 1 */

/* rec1d_p_WENO3 / rec1d_m_WENO3 with the WENOm3 uniform weights
   (rec1d.c:115-232,875-905): optimal weights (1, 2) */
double orc_rec1d_p_WENOm3(const double *u, int im)
{
  double u_im1 = u[im - 1], u_i = u[im], u_ip1 = u[im + 1];
  double d1 = u_i - u_im1, d2 = u_ip1 - u_i;
  double b1 = d1 * d1 + ORC_WENO3_EPS, b2 = d2 * d2 + ORC_WENO3_EPS;
  double o1 = 1. / (b1 * b1), o2 = 2. / (b2 * b2);
  double s = o1 + o2;
  double w1 = o1 / s, w2 = o2 / s;
  double v1 = -0.5 * u_im1 + 1.5 * u_i;
  double v2 = 0.5 * u_i + 0.5 * u_ip1;
  return w1 * v1 + w2 * v2;
}

double orc_rec1d_m_WENOm3(const double *u, int im)
{
  double u_i = u[im], u_ip1 = u[im + 1], u_ip2 = u[im + 2];
  double d1 = u_ip1 - u_i, d2 = u_ip2 - u_ip1;
  double b1 = d1 * d1 + ORC_WENO3_EPS, b2 = d2 * d2 + ORC_WENO3_EPS;
  double o1 = 2. / (b1 * b1), o2 = 1. / (b2 * b2);
  double s = o1 + o2;
  double w1 = o1 / s, w2 = o2 / s;
  double v1 = 0.5 * u_i + 0.5 * u_ip1;
  double v2 = 1.5 * u_ip1 - 0.5 * u_ip2;
  return w1 * v1 + w2 * v2;
}

/* WENOm3_2 at midpoint im = 0..n-2 (rec1d.c:909-1030): WENOm3 inside,
   the average of the two points next to a face */
/* the end midpoints by fv_rec (dg.c:295-326): 0 the _2 modes, 1 the _1
   modes (the end point, rec1d.c:1032-1090), 2 WENOm3 / WENOm5 / WENOmZ
   (rec1d_p_WENOm3 / rec1d_m_WENOm3 with rec1d_u_in1_weightfac,
   rec1d.c:1096-1186) with fv_rec_WENOm_s1 / _s2 / _opt */
static int g_rec_end = 0;
static double g_rec_s1 = 0., g_rec_s2 = DBL_MAX;
static int g_rec_opt = 0;
void orc_set_fv_rec_end(int mode, double s1, double s2, int opt)
{
  g_rec_end = mode; g_rec_s1 = s1; g_rec_s2 = s2; g_rec_opt = opt;
}
static int rec_limit_on(void) { return g_rec_end == 2 && (g_rec_s1 > 0. || g_rec_s2 < DBL_MAX); }

/* rec1d_u_in1_weightfac: 1 if u[i0in2] - u[i0in1] is like u[i0in1] - u[i0] */
static double rec_weightfac(int n, const double *u, int i0)
{
  int sign = (i0 > 0) ? 1 : -1;
  int i1 = i0 - sign;
  double w1fac;
  if (n <= 2) return 0.;
  w1fac = 1.;
  if (g_rec_s1 > 0. || g_rec_s2 < DBL_MAX) {
    int i2 = i1 - sign;
    double d2 = u[i2] - u[i1];
    double d1 = u[i1] - u[i0];
    double s = fabs(d1) / (fabs(d2) + 1e-100 * 1.0);
    if (s < g_rec_s1 || s >= g_rec_s2) w1fac = 0.;
    if (g_rec_opt == -1) w1fac = 0.;
    if (g_rec_opt == 1 && d1 * d2 <= 0.) w1fac = 0.;
    if (g_rec_opt == 2 && d1 * d2 >= 0.) w1fac = 0.;
  }
  return w1fac;
}

static double rec_p_WENOm3_2(int n, const double *u, int im)
{
  if (im >= 1 && im <= n - 2) return orc_rec1d_p_WENOm3(u, im);
  /* im == 0 */
  if (g_rec_end == 1) return u[0];
  if (g_rec_end == 2) {
    double w1 = 0.5 * rec_weightfac(n, u, 0), w0 = 1. - w1;
    return w0 * u[0] + w1 * u[1];
  }
  return 0.5 * u[0] + 0.5 * u[1];
}
static double rec_m_WENOm3_2(int n, const double *u, int im)
{
  if (im >= 0 && im <= n - 3) return orc_rec1d_m_WENOm3(u, im);
  /* im == n-2 */
  if (g_rec_end == 1) return u[n - 1];
  if (g_rec_end == 2) {
    double w1 = 0.5 * rec_weightfac(n, u, n - 1), w0 = 1. - w1;
    return w0 * u[n - 1] + w1 * u[n - 2];
  }
  return 0.5 * u[n - 2] + 0.5 * u[n - 1];
}
/* WENOT (rec1d_p_WENOT / rec1d_m_WENOT with the uniform weights,
   rec1d.c:1412-1578): WENO5's stencils and ideal weights (1, 6, 3)/10 with
   the non-normalized weights d_k (1 + |b0 - b2|) / (b_k + eps), i.e. the
   smoothness indicators to the first power */
static double wenot(double a, double b, double c, double d, double e, const double qw[3][3],
                    const double optw[3], int mirror)
{
  double b01, b02, b11, b12, b21, b22;
  if (!mirror) {
    b01 = a - 2. * b + c; b02 = a - 4. * b + 3. * c;
    b11 = b - 2. * c + d; b12 = b - d;
    b21 = c - 2. * d + e; b22 = 3. * c - 4. * d + e;
  } else {   /* rec1d_m_WENOT: b2 from the three right points */
    b21 = c - 2. * d + e; b22 = 3. * c - 4. * d + e;
    b11 = b - 2. * c + d; b12 = b - d;
    b01 = a - 2. * b + c; b02 = a - 4. * b + 3. * c;
  }
  double b0 = (1.083333333333333333333333) * b01 * b01 + 0.25 * b02 * b02;
  double b1 = (1.083333333333333333333333) * b11 * b11 + 0.25 * b12 * b12;
  double b2 = (1.083333333333333333333333) * b21 * b21 + 0.25 * b22 * b22;
  double o_p_db02 = 1. + fabs(b0 - b2);
  double w0 = optw[0] * o_p_db02 / (b0 + 1e-40);
  double w1 = optw[1] * o_p_db02 / (b1 + 1e-40);
  double w2 = optw[2] * o_p_db02 / (b2 + 1e-40);
  double inv = 1. / (w0 + w1 + w2);
  double u0, u1, u2;
  if (!mirror) {
    u0 = qw[0][0] * a + qw[0][1] * b + qw[0][2] * c;
    u1 = qw[1][0] * b + qw[1][1] * c + qw[1][2] * d;
    u2 = qw[2][0] * c + qw[2][1] * d + qw[2][2] * e;
  } else {
    u0 = qw[0][0] * a + qw[0][1] * b + qw[0][2] * c;
    u1 = qw[1][0] * b + qw[1][1] * c + qw[1][2] * d;
    u2 = qw[2][0] * c + qw[2][1] * d + qw[2][2] * e;
  }
  return (w0 * inv) * u0 + (w1 * inv) * u1 + (w2 * inv) * u2;
}
static double orc_rec1d_p_WENOT(const double *u, int im)
{
  const double o6 = 1. / 6.;
  const double qw[3][3] = {{2. * o6, -7. * o6, 11. * o6}, {-o6, 5. * o6, 2. * o6}, {2. * o6, 5. * o6, -o6}};
  const double optw[3] = {1., 6., 3.};
  return wenot(u[im - 2], u[im - 1], u[im], u[im + 1], u[im + 2], qw, optw, 0);
}
static double orc_rec1d_m_WENOT(const double *u, int im)
{
  /* points im-1 .. im+3; stencil 2 on the right end, ideal weights mirrored */
  const double o6 = 1. / 6.;
  const double qw[3][3] = {{-o6, 5. * o6, 2. * o6}, {2. * o6, 5. * o6, -o6}, {11. * o6, -7. * o6, 2. * o6}};
  const double optw[3] = {3., 6., 1.};
  return wenot(u[im - 1], u[im], u[im + 1], u[im + 2], u[im + 3], qw, optw, 1);
}
/* WENOmT_2 (rec1d.c:1579-1594): WENOT inside, WENOm3_2 near the faces */
static double rec_p_WENOmT_2(int n, const double *u, int im)
{
  if (im >= 2 && im <= n - 3) return orc_rec1d_p_WENOT(u, im);
  return rec_p_WENOm3_2(n, u, im);
}
static double rec_m_WENOmT_2(int n, const double *u, int im)
{
  if (im >= 1 && im <= n - 4) return orc_rec1d_m_WENOT(u, im);
  return rec_m_WENOm3_2(n, u, im);
}
/* WENOm5_2 (rec1d.c:1354-1370): WENO5 inside, WENOm3_2 near the faces */
static double rec_p_WENOm5_2(int n, const double *u, int im)
{
  if (im >= 2 && im <= n - 3) return orc_rec1d_p_WENO5(u, im);
  return rec_p_WENOm3_2(n, u, im);
}
static double rec_m_WENOm5_2(int n, const double *u, int im)
{
  if (im >= 1 && im <= n - 4) return orc_rec1d_m_WENO5(u, im);
  return rec_m_WENOm3_2(n, u, im);
}

/* WENOmZ_2 (rec1d.c:1782-1798): WENO-Z inside, WENOm3_2 near the faces */
static double rec_p_WENOmZ_2(int n, const double *u, int im)
{
  if (im >= 2 && im <= n - 3) return orc_rec1d_p_WENOZ(u, im);
  return rec_p_WENOm3_2(n, u, im);
}
static double rec_m_WENOmZ_2(int n, const double *u, int im)
{
  if (im >= 1 && im <= n - 4) return orc_rec1d_m_WENOZ(u, im);
  return rec_m_WENOm3_2(n, u, im);
}

/* the FV reconstruction used by fv_elm_rhs: 3 = WENOm3_2, 5 = WENOm5_2,
   7 = WENOmZ_2 */
static int g_fv_rec = 5;
void orc_set_fv_rec(int mode) { g_fv_rec = mode; }

static double fv_rec_p(int n, const double *u, int im)
{
  if (g_fv_rec == 1) return u[im];   /* rec1d_p_1, rec1d.c:16-19 */
  return g_fv_rec == 5   ? rec_p_WENOm5_2(n, u, im)
         : g_fv_rec == 7 ? rec_p_WENOmZ_2(n, u, im)
         : g_fv_rec == 9 ? rec_p_WENOmT_2(n, u, im) : rec_p_WENOm3_2(n, u, im);
}
static double fv_rec_m(int n, const double *u, int im)
{
  if (g_fv_rec == 1) return u[im + 1];   /* rec1d_m_1, rec1d.c:24-27 */
  return g_fv_rec == 5   ? rec_m_WENOm5_2(n, u, im)
         : g_fv_rec == 7 ? rec_m_WENOmZ_2(n, u, im)
         : g_fv_rec == 9 ? rec_m_WENOmT_2(n, u, im) : rec_m_WENOm3_2(n, u, im);
}

/* the reconstructions at the end midpoints: out[0] = rec_p at midpoint 0,
   out[1] = rec_m at midpoint n-2 */
void orc_fv_rec_pm(int n, const double *u, int im, double *out)
{
  out[0] = fv_rec_p(n, u, im);
  out[1] = fv_rec_m(n, u, im);
}
void orc_fv_rec_ends(int n, const double *u, double *out)
{
  out[0] = fv_rec_p(n, u, 0);
  out[1] = fv_rec_m(n, u, n - 2);
}

/* scalarwave1 numflux for normal nrm between ui and ua */
static void sw_fnum(const OSW *sw, const double nrm[3], const double *ui, const double *ua, double *fnum)
{
  double lam[5] = {1.0, -1.0, 0.0, 0.0, 0.0}, fi[5], fa[5];
  sw_flux1d(fi, nrm, ui);
  sw_flux1d(fa, nrm, ua);
  if (sw->numflux == 1) numflux_LLF(NV, lam, lam, ui, ua, fi, fa, fnum);
  else sw_numflux_upwind(nrm, lam, ui, ua, fnum);
}

/* the numerical flux at one face point, all 5 rows (scalarwave1_fluxes_pt,
   scalarwave1.c:139-209): numflux 0 upwind, 1 LLF */
void orc_numflux(int numflux, const double *nrm, const double *ui, const double *ua, double *fnum)
{
  OSW sw = {numflux, {0., 0., 0.}};
  sw_fnum(&sw, nrm, ui, ua, fnum);
}

/* fv_divf_extrap_s1, _s2, _opt (fv.c:147-150; defaults 0, DBL_MAX, 0) */
static double g_ex_s1 = 0., g_ex_s2 = DBL_MAX;
static int g_ex_opt = 0;
void orc_set_fv_divf_extrap(double s1, double s2, int opt)
{
  g_ex_s1 = s1; g_ex_s2 = s2; g_ex_opt = opt;
}

/* the face value at end i0 (0 or n-1) of one line (rec1d_compute_1s1_u
   with forward = 0, rec1d.c:542-609): the end value sits a quarter spacing
   inside the face and the next one a full spacing, so the face value is
   df_end / 0.75 + (1 - 1/0.75) df_next; with s1 > 0 or s2 < DBL_MAX the
   ratio s = |d1| / 0.75 / |d2| of the first two differences inward (and
   opt on their signs) can switch to the end value itself */
static double extrap_end(int n, const double *u, int i0)
{
  int sign = (i0 > 0) ? 1 : -1;
  int i1 = i0 - sign;
  double w0 = 1. / 0.75;
  double w1 = 1. - w0;
  if ((g_ex_s1 > 0. || g_ex_s2 < DBL_MAX) && n > 2) {
    int i2 = i1 - sign;
    double d2 = u[i2] - u[i1];
    double d1 = u[i1] - u[i0];
    double s = 1.3333333333 * fabs(d1) / (fabs(d2) + 1e-100 * 1.0);
    if (s < g_ex_s1 || s >= g_ex_s2) w1 = 0.;
    if (g_ex_opt == -1) w1 = 0.;
    if (g_ex_opt == 1 && d1 * d2 <= 0.) w1 = 0.;
    if (g_ex_opt == 2 && d1 * d2 >= 0.) w1 = 0.;
    w0 = 1. - w1;
  }
  return w0 * u[i0] + w1 * u[i1];
}

/* dnfn_extrap1 on both ends of one line, the left end first
   (rec1d_LR_uface_to_uin_1_Carray, rec1d.c:612-626; fv.c:558-573), an end
   skipped where the reconstruction kept the end point (FV_REC_NO_*_EXTRAP1) */
static void dnfn_extrap1_sides(int n, double *df, int left, int right)
{
  if (n <= 2) return;
  if (left) df[0] = extrap_end(n, df, 0);
  if (right) df[n - 1] = extrap_end(n, df, n - 1);
}
void orc_dnfn_extrap1(int n, double *df) { dnfn_extrap1_sides(n, df, 1, 1); }

/* r of one FV element: divf from fv_divf, r_{pi,c} = -divf, r_phi = pi */
/* unit normal of face 2*dir+side at grid point g: +-row dir of dXdx over
   its length (node_normal_at_ijk via node_normal_from_idXdx_at_ijk,
   utilities.c:48-81) */
static void node_normal(const OMesh *M, long g, int dir, int side, double nrm[3])
{
  long nn = M->nnodes;
  double sig = side ? 1.0 : -1.0;
  for (int d = 0; d < 3; d++) nrm[d] = M->dXdx[(3 * dir + d) * nn + g];
  double smag = sig * sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
  if (smag == 0.0) smag = sig;
  for (int d = 0; d < 3; d++) nrm[d] /= smag;
}

/* coordinates_midpoint_data: 0 = no (the default: midpoint metric and
   normal are those of the grid point left of the midpoint), 1 = yes
   (computed at the midpoints, coordinates.c:284-415) */
static int g_fv_midpt = 0;
void orc_set_fv_midpoint(int on) { g_fv_midpt = on; }

/* metric of element E at element point Xb for direction dir, as
   coordinates_init_node sets it at a midpoint (coordinates.c:284-415):
   sg = dXbdX[dir] sqrt(sum_l (dX^dir/dx^l)^2) (flat metric,
   coordinates_set_sqrtgdiag_array coordinates.c:728-802), det =
   det(dXb/dx), nrm = row dir of dX/dx over its length
   (node_normal_from_idXdx_at_ijk, utilities.c:48-81) */
static void point_metric(const OMesh *M, const OElm *E, const double Xb[3], int dir, double *sg,
                         double *det, double nrm[3])
{
  const OPat *P = &M->pat[E->p];
  double dXbdX[3], X[3], x[3], D[3][3] = {{0}};
  elm_dXbdX(E, dXbdX);
  elm_X_of_Xb(E, Xb, X);
  if (P->type == PAT_CUBEDSHELL) {
    cubsph_dXdx(P, X, x, D);
  } else {
    for (int d = 0; d < 3; d++) D[d][d] = 1.;
  }
  double detb = dXbdX[0] * dXbdX[1] * dXbdX[2];
  double det3 = D[0][0] * (D[1][1] * D[2][2] - D[1][2] * D[2][1]) -
                D[0][1] * (D[1][0] * D[2][2] - D[1][2] * D[2][0]) +
                D[0][2] * (D[1][0] * D[2][1] - D[1][1] * D[2][0]);
  *det = (P->type == PAT_CUBEDSHELL) ? detb * det3 : detb;
  double gam = 0.0;
  for (int l = 0; l < 3; l++) gam += D[dir][l] * D[dir][l];
  *sg = dXbdX[dir] * sqrt(gam);
  double smag = sqrt(gam);
  if (smag == 0.0) smag = 1.0;
  for (int l = 0; l < 3; l++) nrm[l] = D[dir][l] / smag;
}

static void fv_elm_rhs(const OMesh *M, const OSW *sw, const OElm *E, double t, const double *u,
                       const double *planes, double *r)
{
  const int n = E->n, np = elm_np(E);
  const long nn = M->nnodes, o = E->off;
  const double *Xb = basis(n)->Xu;
  double (*divf)[ORC_MAXN * ORC_MAXN * ORC_MAXN] = calloc(4, sizeof(*divf));
  double (*aj)[NV][ORC_MAXN * ORC_MAXN] = malloc(6 * sizeof(*aj));
  int have[6];
  /* adjacent face values: neighbour planes, or the exterior profile */
  for (int f = 0; f < 6; f++) {
    have[f] = (E->nfnb[f] > 0);
    if (have[f]) {
      for (int v = 0; v < NV; v++) build_ajsurf(M, E, f, planes, v, aj[f][v]);
    } else {
      int onbound = elm_on_patface(M, E, f) && M->pat[E->p].bound[f];
      int dir = f / 2, p = (n - 1) * (f % 2), d1 = (dir == 0), d2 = (dir == 2) ? 1 : 2;
      for (int b = 0; b < n; b++)
        for (int a = 0; a < n; a++) {
          int I[3];
          I[dir] = p; I[d1] = a; I[d2] = b;
          long g = o + I[0] + n * (I[1] + n * I[2]);
          double ua[5] = {0, 0, 0, 0, 0};
          if (onbound) {
            double xyz[3] = {M->x[g], M->x[nn + g], M->x[2 * nn + g]};
            sw_profile_pt(sw->k, xyz, t, ua);
          }
          for (int v = 0; v < NV; v++) aj[f][v][a + n * b] = ua[v];
        }
    }
  }
  /* midpoints and the distances between them (get_coords.c:1749-1860) */
  double Xbm[ORC_MAXN] = {0}, dXb[ORC_MAXN] = {0};
  for (int i = 0; i < n - 1; i++) Xbm[i] = 0.5 * (Xb[i] + Xb[i + 1]);
  dXb[0] = fabs(Xbm[0] - (-1.));
  dXb[n - 1] = fabs(Xbm[n - 2] - 1.);
  for (int i = 1; i < n - 1; i++) dXb[i] = Xbm[i] - Xbm[i - 1];
  for (int dir = 0; dir < 3; dir++) {
    const int d1 = (dir == 0), d2 = (dir == 2) ? 1 : 2;
    for (int b = 0; b < n; b++)
      for (int a = 0; a < n; a++) {
        double qc[NV][ORC_MAXN], df[4][ORC_MAXN];
        long gl[ORC_MAXN];
        for (int i0 = 0; i0 < n; i0++) {
          int I[3];
          I[dir] = i0; I[d1] = a; I[d2] = b;
          gl[i0] = o + I[0] + n * (I[1] + n * I[2]);
          for (int v = 0; v < NV; v++) qc[v][i0] = u[v * nn + gl[i0]];
        }
        double fnumR[5] = {0, 0, 0, 0, 0}, fnumL[5];
        int noL[4] = {0, 0, 0, 0}, noR[4] = {0, 0, 0, 0};
        const int JK = a + n * b;
        /* metric factor sqrtgdiag / det(dXb/dx) and unit normal at each
           midpoint i0+1/2 (fv.c:347-401, node_normal_from_DGinfo
           dg.c:468-479, node_normal_at_midpt_nextto_ijk utilities.c:
           105-181).  With coordinates_midpoint_data = no these are node
           i0's (coordinates.c:467-483: the midpoint variables alias the
           node's); with yes they are computed at the midpoint. */
        double SoJm[ORC_MAXN], nm_[ORC_MAXN][3];
        for (int i0 = 0; i0 < n - 1; i0++) {
          if (g_fv_midpt) {
            const double *xb = elm_Xb1(E);
            double Xbp[3], sg, det;
            Xbp[dir] = Xbm[i0]; Xbp[d1] = xb[a]; Xbp[d2] = xb[b];
            point_metric(M, E, Xbp, dir, &sg, &det, nm_[i0]);
            SoJm[i0] = sg / det;
          } else {
            const long g = gl[i0];
            SoJm[i0] = M->sqrtgdiag[dir * nn + g] / M->det[g];
            node_normal(M, g, dir, 1, nm_[i0]);
          }
        }
        for (int i0 = 0; i0 < n; i0++) {
          const long g = gl[i0];
          /* Jgdow_R from the midpoint right of i0 (the node itself at the
             right end), Jgdow_L from the midpoint left of i0 (the node
             itself at the left end), both over wc of i0; the end faces
             take the end node's normal */
          const double wc = dXb[i0];
          const double SoJn = M->sqrtgdiag[dir * nn + g] / M->det[g];
          const double JR = (i0 < n - 1 ? SoJm[i0] : SoJn) / wc;
          const double JL = (i0 > 0 ? SoJm[i0 - 1] : SoJn) / wc;
          double nR[3];
          if (i0 < n - 1) memcpy(nR, nm_[i0], sizeof(nR));
          else node_normal(M, g, dir, 1, nR);
          if (i0 > 0) {
            for (int l = 0; l < NV; l++) fnumL[l] = -fnumR[l];
          } else {
            double ui[5], ua[5], nL[3];
            node_normal(M, g, dir, 0, nL);
            for (int v = 0; v < NV; v++) { ui[v] = qc[v][0]; ua[v] = aj[2 * dir][v][JK]; }
            sw_fnum(sw, nL, ui, ua, fnumL);
          }
          if (i0 < n - 1) {
            double ui[5], ua[5];
            for (int v = 0; v < NV; v++) {
              ui[v] = fv_rec_p(n, qc[v], i0);
              ua[v] = fv_rec_m(n, qc[v], i0);
              /* fv_stat_WENOm_1or2 (fv.c:45-79): the end point kept */
              if (rec_limit_on() && v < 4) {
                if (i0 == 0) { if (ui[v] == qc[v][0]) noL[v] = 1; }
                else if (i0 == n - 2) { if (ua[v] == qc[v][n - 1]) noR[v] = 1; }
              }
            }
            sw_fnum(sw, nR, ui, ua, fnumR);
          } else {
            double ui[5], ua[5];
            for (int v = 0; v < NV; v++) { ui[v] = qc[v][n - 1]; ua[v] = aj[2 * dir + 1][v][JK]; }
            sw_fnum(sw, nR, ui, ua, fnumR);
          }
          for (int l = 0; l < 4; l++) df[l][i0] = JR * fnumR[l] + JL * fnumL[l];
        }
        for (int l = 0; l < 4; l++) dnfn_extrap1_sides(n, df[l], !noL[l], !noR[l]);
        for (int i0 = 0; i0 < n; i0++)
          for (int l = 0; l < 4; l++) divf[l][gl[i0] - o] += df[l][i0] * M->det[gl[i0]];
      }
  }
  for (int q = 0; q < np; q++) {
    r[4 * nn + o + q] = u[0 * nn + o + q];
    for (int l = 0; l < 4; l++) r[l * nn + o + q] = -divf[l][q];
  }
  free(divf);
  free(aj);
}

/* r = RHS(u, t) (src/main/evolve/evolve.c:182-232) */
void orc_rhs(const OMesh *M, int numflux, const double *k, const double *u, double t, double *r)
{
  OSW sw = sw_from(numflux, k);
  long nn = M->nnodes;
  int maxnp = 0;
  for (int e = 0; e < M->nelem; e++) if (elm_np(&M->e[e]) > maxnp) maxnp = elm_np(&M->e[e]);
  double *divf = malloc(sizeof(double) * 4 * nn);
  double *planes = malloc(sizeof(double) * NV * M->nface);
  /* face planes of the state (MPIexchange_set_all_myln_localdata) */
#pragma omp parallel for schedule(dynamic, 8)
  for (int e = 0; e < M->nelem; e++)
    for (int f = 0; f < 6; f++)
      for (int v = 0; v < NV; v++)
        set_mysurf(M, &M->e[e], f, u, v, planes + v * M->nface + M->foff[6 * e + f]);
#pragma omp parallel
  {
    double *work = malloc(sizeof(double) * 15 * maxnp);
#pragma omp for schedule(dynamic, 8)
    for (int e = 0; e < M->nelem; e++) {
      const OElm *E = &M->e[e];
      if (E->use_fv) fv_elm_rhs(M, &sw, E, t, u, planes, r);
      else sw_vol_rhs(M, E, u, r, divf + 4 * E->off, work);
    }
#pragma omp for schedule(dynamic, 8)
    for (int e = 0; e < M->nelem; e++) {
      const OElm *E = &M->e[e];
      if (!E->use_fv) sw_surf_rhs(M, &sw, E, t, u, planes, r, divf + 4 * E->off);
    }
    free(work);
  }
  free(planes);
  free(divf);
}

/* ------------------------------------------------------------------ */
/* Runge-Kutta steppers (src/main/evolve/RungeKutta.c:19-188,          */
/* src/main/main/variables.c:831-1041)                                 */
/* ------------------------------------------------------------------ */

static void vladd(long N, double *r, double ca, const double *a, double cb, const double *b)
{
  if (ca == 0 && cb == 0) { for (long i = 0; i < N; i++) r[i] = 0; }
  else if (ca == 0) { for (long i = 0; i < N; i++) r[i] = cb * b[i]; }
  else if (cb == 0) { for (long i = 0; i < N; i++) r[i] = ca * a[i]; }
  else { for (long i = 0; i < N; i++) r[i] = ca * a[i] + cb * b[i]; }
}

static void vladdto(long N, double *r, double ca, const double *a)
{
  if (ca == 0) return;
  if (ca == 1) { for (long i = 0; i < N; i++) r[i] += a[i]; }
  else if (ca == -1) { for (long i = 0; i < N; i++) r[i] -= a[i]; }
  else { for (long i = 0; i < N; i++) r[i] += ca * a[i]; }
}

/* method: 0 RK4, 1 sspRK3, 2 Euler; u advanced in place; returns final t */
/* ------------------------------------------------------------------ */
/* MRS limiter (src/main/limiter/limiter.c:56-196, arXiv:1507.03024) as  */
/* scalarwave1 registers it with scalarwave1_limiter = MRS               */
/* (scalarwave1.c:687-691), applied by evolve_limiter_mesh               */
/* (evolve.c:430-541)                                                    */
/* ------------------------------------------------------------------ */

static double g_lim_alpha = -1.;   /* < 0: no limiter */
void orc_set_limiter_MRS(double alpha) { g_lim_alpha = alpha; }
/* minmodB (limiter.c:200-380, arXiv:1506.06140): alpha < 0 is off */
static double g_mm_alpha = -1., g_mm_beta = 0.5;
static int g_mm_scale = 0;
void orc_set_limiter_minmodB(double alpha, double beta, int scaleBound)
{
  g_mm_alpha = alpha; g_mm_beta = beta; g_mm_scale = scaleBound;
}
static void lgl_at_st(int n, const double *x, const double *w, double *AT, double *ST);
static void contract1(int n, const double *MT, int dir, const double *in, double *out);

static double minmod3(double a, double b, double c)
{
  if (a * b >= 0 && b * c >= 0)
    return (a > 0. ? 1. : -1.) * fmin(fabs(a), fmin(fabs(b), fabs(c)));
  return 0.;
}

static double minmod3B(double a, double b, double c, double aBound)
{
  if (fabs(a) <= aBound) return a;
  return minmod3(a, b, c);
}

#define ORC_DEQUAL(a, b) (!((a) < (b) - 1e-10 || (b) < (a) - 1e-10))

/* limdata_c000_100_010_001 on every element (the normalized Legendre
   analysis along x, y, z: LGL_AT_ST_matrices), then limiter_minmodB on
   every element: the three slopes limited against the neighbours' face
   averages; a variable whose slopes change is replaced by its mean and
   limited slopes */
void orc_limit_minmodB(const OMesh *M, double *u)
{
  long nn = M->nnodes;
  double *ld = malloc(sizeof(double) * 20 * M->nelem);
  for (int e = 0; e < M->nelem; e++) {
    const OElm *E = &M->e[e];
    int n = E->n, np = elm_np(E);
    OBasis *b = basis(n);
    double AT[ORC_MAXN * ORC_MAXN], ST[ORC_MAXN * ORC_MAXN];
    lgl_at_st(n, b->Xb, b->Wq, AT, ST);
    double *c = malloc(sizeof(double) * np), *t = malloc(sizeof(double) * np);
    for (int v = 0; v < NV; v++) {
      contract1(n, AT, 0, u + v * nn + E->off, t);
      contract1(n, AT, 1, t, c);
      contract1(n, AT, 2, c, t);
      ld[20 * e + 4 * v + 0] = t[0];
      ld[20 * e + 4 * v + 1] = n > 1 ? t[1] : 0.;
      ld[20 * e + 4 * v + 2] = n > 1 ? t[n] : 0.;
      ld[20 * e + 4 * v + 3] = n > 1 ? t[n * n] : 0.;
    }
    free(c); free(t);
  }
  const double bos3 = g_mm_beta / sqrt(3.);
  for (int e = 0; e < M->nelem; e++) {
    const OElm *E = &M->e[e];
    int n = E->n, np = elm_np(E);
    const double *bb = E->bbox;
    double h = fmax(bb[1] - bb[0], fmax(bb[3] - bb[2], bb[5] - bb[4]));
    double Mt_h = g_mm_alpha * h;
    for (int v = 0; v < NV; v++) {
      const double *my = ld + 20 * e + 4 * v;
      double w0[6], w0L1 = 0.;
      for (int f = 0; f < 6; f++) {
        double av = 0.;
        for (int ni = 0; ni < E->nfnb[f]; ni++) av += ld[20 * E->fnb[f][ni] + 4 * v];
        if (E->nfnb[f] > 0) av = av / E->nfnb[f];
        w0[f] = av;
        w0L1 += fabs(av) * 0.166666666666667;
      }
      double bound = g_mm_scale ? Mt_h * w0L1 : Mt_h;
      double w0c = my[0];
      double l1 = minmod3B(my[1], bos3 * (w0c - w0[0]), bos3 * (w0[1] - w0c), bound);
      double l2 = minmod3B(my[2], bos3 * (w0c - w0[2]), bos3 * (w0[3] - w0c), bound);
      double l3 = minmod3B(my[3], bos3 * (w0c - w0[4]), bos3 * (w0[5] - w0c), bound);
      if (ORC_DEQUAL(l1, my[1]) && ORC_DEQUAL(l2, my[2]) && ORC_DEQUAL(l3, my[3])) continue;
      OBasis *b = basis(n);
      double AT[ORC_MAXN * ORC_MAXN], ST[ORC_MAXN * ORC_MAXN];
      lgl_at_st(n, b->Xb, b->Wq, AT, ST);
      double *c = calloc(np, sizeof(double)), *t = malloc(sizeof(double) * np);
      c[0] = w0c;
      if (n > 1) { c[1] = l1; c[n] = l2; c[n * n] = l3; }
      double *q = u + v * nn + E->off;
      contract1(n, ST, 0, c, t);
      contract1(n, ST, 1, t, c);
      contract1(n, ST, 2, c, q);
      free(c); free(t);
    }
  }
  free(ld);
}

/* var_nodeaverage (basis.c:462-510): the LGL quadrature along x, then y,
   then z, over the Xb volume 8 */
static double elm_average(const OElm *E, const double *q)
{
  const int n = E->n;
  const double *w = basis(n)->Wq;
  double s2[ORC_MAXN];
  for (int k = 0; k < n; k++) {
    double s1[ORC_MAXN];
    for (int j = 0; j < n; j++) {
      double s = 0.;
      for (int i = 0; i < n; i++) s += w[i] * q[i + n * (j + n * k)];
      s1[j] = s;
    }
    double s = 0.;
    for (int j = 0; j < n; j++) s += w[j] * s1[j];
    s2[k] = s;
  }
  double s = 0.;
  for (int k = 0; k < n; k++) s += w[k] * s2[k];
  return s * 0.125;
}

static double mrs_phi(double y)
{
  if (y < 0.) return 0.;
  return fmin(y / 1.1, 1.);
}

static double mrs_theta(double Mi, double wbar, double wMi)
{
  double den = wMi - wbar;
  if (den == 0.) return 1.;
  return mrs_phi((Mi - wbar) / den);
}

/* limdata_MRS on every element, then limiter_MRS on every element */
void orc_limit_MRS(const OMesh *M, double *u)
{
  long nn = M->nnodes;
  double *ld = malloc(sizeof(double) * 15 * M->nelem);
  for (int e = 0; e < M->nelem; e++) {
    const OElm *E = &M->e[e];
    int np = elm_np(E);
    for (int v = 0; v < NV; v++) {
      const double *q = u + v * nn + E->off;
      double mn = q[0], mx = q[0];
      for (int i = 1; i < np; i++) { if (q[i] < mn) mn = q[i]; if (q[i] > mx) mx = q[i]; }
      ld[15 * e + 3 * v + 0] = elm_average(E, q);
      ld[15 * e + 3 * v + 1] = mn;
      ld[15 * e + 3 * v + 2] = mx;
    }
  }
  for (int e = 0; e < M->nelem; e++) {
    const OElm *E = &M->e[e];
    const double *bb = E->bbox;
    double h = fmax(bb[1] - bb[0], fmax(bb[3] - bb[2], bb[5] - bb[4]));
    double ah = g_lim_alpha * pow(h, 1.5);
    double thM = 1.7976931348623157e308, thm = 1.7976931348623157e308;
    for (int v = 0; v < NV; v++) {
      double wbar = ld[15 * e + 3 * v], wmi = ld[15 * e + 3 * v + 1], wMi = ld[15 * e + 3 * v + 2];
      double Mi = -1.7976931348623157e308, mi = 1.7976931348623157e308;
      for (int f = 0; f < 6; f++)
        for (int ni = 0; ni < E->nfnb[f]; ni++) {
          int nb = E->fnb[f][ni];
          double ma = ld[15 * nb + 3 * v + 1], Ma = ld[15 * nb + 3 * v + 2];
          if (ma < mi) mi = ma;
          if (Ma > Mi) Mi = Ma;
        }
      mi = fmin(wbar - ah, mi);
      Mi = fmax(wbar + ah, Mi);
      double a = mrs_theta(Mi, wbar, wMi), b = mrs_theta(mi, wbar, wmi);
      if (a < thM) thM = a;
      if (b < thm) thm = b;
    }
    double th = fmin(1., fmin(thm, thM));
    int np = elm_np(E);
    for (int v = 0; v < NV; v++) {
      double *q = u + v * nn + E->off;
      double qbar = elm_average(E, q);
      for (int i = 0; i < np; i++) q[i] = qbar + th * (q[i] - qbar);
    }
  }
  free(ld);
}

static void limit_if_on(const OMesh *M, double *u)
{
  if (g_lim_alpha >= 0.) orc_limit_MRS(M, u);
  if (g_mm_alpha >= 0.) orc_limit_minmodB(M, u);
}

double orc_evolve(const OMesh *M, int numflux, const double *k, int method, double dt,
                  int nsteps, double t, double *u)
{
  long N = NV * M->nnodes;
  double *up = malloc(sizeof(double) * N), *w = malloc(sizeof(double) * N);
  double *r = malloc(sizeof(double) * N);
  /* time = Te + te accumulated in long double (src/main/main/main.c:588-624) */
  long double Te = t, te = 0.;
  for (int s = 0; s < nsteps; s++) {
    memcpy(up, u, sizeof(double) * N);
    memcpy(w, u, sizeof(double) * N);
    if (method == 0) {
      orc_rhs(M, numflux, k, w, t, r);
      memcpy(u, w, sizeof(double) * N);
      vladdto(N, u, dt / 6., r);
      vladd(N, w, 1., up, dt / 2., r);
      limit_if_on(M, w);
      orc_rhs(M, numflux, k, w, t + 0.5 * dt, r);
      vladdto(N, u, dt / 3., r);
      vladd(N, w, 1., up, dt / 2., r);
      limit_if_on(M, w);
      orc_rhs(M, numflux, k, w, t + 0.5 * dt, r);
      vladdto(N, u, dt / 3., r);
      vladd(N, w, 1., up, dt, r);
      limit_if_on(M, w);
      orc_rhs(M, numflux, k, w, t + dt, r);
      vladdto(N, u, dt / 6., r);
    } else if (method == 1) {
      orc_rhs(M, numflux, k, w, t, r);
      memcpy(u, w, sizeof(double) * N);
      vladdto(N, u, dt / 6., r);
      vladd(N, w, 1., up, dt, r);
      limit_if_on(M, w);
      orc_rhs(M, numflux, k, w, t + dt, r);
      vladdto(N, u, dt / 6., r);
      vladdto(N, w, dt, r);
      vladd(N, w, 0.75, up, 0.25, w);
      limit_if_on(M, w);
      orc_rhs(M, numflux, k, w, t + 0.5 * dt, r);
      vladdto(N, u, dt * 2. / 3., r);
    } else {
      orc_rhs(M, numflux, k, w, t, r);
      memcpy(u, w, sizeof(double) * N);
      vladdto(N, u, dt, r);
    }
    /* the final u is limited in evolve_myln (evolve.c:143-144) */
    limit_if_on(M, u);
    te += dt;
    t = (double)(Te + te);
  }
  free(up); free(w); free(r);
  return t;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* exponential filter (FILTER bin, run after each full step)           */
/* ------------------------------------------------------------------ */

/* P_l(x) by the recurrence P_l = (x (2l-1) P_{l-1} - (l-1) P_{l-2}) / l
   (src/main/basis/Legendre.c:14-38) */
static double legendre_P(int l, double x)
{
  double P0 = 1.0, P1 = x, Pk = 0.0;
  if (l == 0) return P0;
  if (l == 1) return P1;
  for (int k = 2; k <= l; k++) {
    Pk = (x * (2 * k - 1) * P1 - (k - 1) * P0) / k;
    P0 = P1;
    P1 = Pk;
  }
  return Pk;
}

/* normalized Legendre analysis and synthesis on LGL nodes, stored
   transposed: coefficient i = sum_j AT[j + i n] u_j, value j = sum_i
   ST[i + j n] c_i, with c_i = 2/(2i+1) except c_N = 2/N, the discrete norm
   of the top mode (LGL_AT_ST_matrices, src/main/basis/Legendre.c:257-307) */
static void lgl_at_st(int n, const double *x, const double *w, double *AT, double *ST)
{
  int N = n - 1;
  for (int i = 0; i < n; i++) {
    double ooci = 0.5 * (2 * i + 1);
    if (i == N && N != 0) ooci = 0.5 * i;
    double oosci = sqrt(ooci);
    for (int j = 0; j < n; j++) {
      ST[i + j * n] = legendre_P(i, x[j]) * oosci;
      AT[j + i * n] = w[j] * ST[i + j * n];
    }
  }
}

/* out_i = sum_l MT[l + n i] in_l along direction dir (mm_array_indir,
   src/main/amr/array.c:100-300) */
static void contract1(int n, const double *MT, int dir, const double *in, double *out)
{
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        int ii = (dir == 0) ? i : (dir == 1) ? j : k;
        double sum = 0.0;
        for (int l = 0; l < n; l++) {
          int src = (dir == 0) ? l + n * (j + n * k) : (dir == 1) ? i + n * (l + n * k)
                                                                  : i + n * (j + n * l);
          sum += MT[l + n * ii] * in[src];
        }
        out[i + n * (j + n * k)] = sum;
      }
}

/* Filter the variables in varmask (bit v: pi, cx, cy, cz, phi) of every
   element: u -> u J^jpow, coefficients c = A_z A_y A_x u, c_ijk *= f(i) f(j)
   f(k) with f(i) = exp(-alp (i/N)^s), N = n-1-dn (1 if not positive), then
   u = S_z S_y S_x c and u -> u J^-jpow; J = 1/det(dXb/dx)
   (expfilter_coeff_array, expfilter_array, expfilter_var,
   src/main/basis/filter.c:20-92; var_to_var_times_JtoPower,
   src/main/coordinates/utilities.c:885-909; basis_array_analysis3 and
   synthesis3, src/main/basis/basis.c:183-196). */
void orc_filter(const OMesh *M, double *u, int varmask, double alp, double s, int dn, int jpow)
{
  long nn = M->nnodes;
#pragma omp parallel for schedule(dynamic, 4)
  for (int e = 0; e < M->nelem; e++) {
    const OElm *E = &M->e[e];
    if (E->use_fv) continue;   /* basis_filter_fv = no (filter.c:79-80) */
    int n = E->n, np = elm_np(E);
    OBasis *b = basis(n);
    double AT[ORC_MAXN * ORC_MAXN], ST[ORC_MAXN * ORC_MAXN], f[ORC_MAXN];
    lgl_at_st(n, b->Xb, b->Wq, AT, ST);
    double N = n - 1 - dn;
    if (N <= 0.) N = 1.;
    for (int i = 0; i < n; i++) f[i] = exp(-alp * pow(i / N, s));
    double *c = malloc(sizeof(double) * np), *t = malloc(sizeof(double) * np);
    const double *ooJ = M->det + E->off;
    for (int v = 0; v < 5; v++) {
      if (!(varmask >> v & 1)) continue;
      double *uv = u + v * nn + E->off;
      if (jpow == 1) for (int q = 0; q < np; q++) uv[q] = uv[q] / ooJ[q];
      else if (jpow == -1) for (int q = 0; q < np; q++) uv[q] = uv[q] * ooJ[q];
      contract1(n, AT, 0, uv, c);
      contract1(n, AT, 1, c, t);
      contract1(n, AT, 2, t, c);
      for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
          for (int i = 0; i < n; i++) {
            int q = i + n * (j + n * k);
            c[q] = c[q] * f[i] * f[j] * f[k];
          }
      contract1(n, ST, 0, c, t);
      contract1(n, ST, 1, t, c);
      contract1(n, ST, 2, c, uv);
      if (jpow == 1) for (int q = 0; q < np; q++) uv[q] = uv[q] * ooJ[q];
      else if (jpow == -1) for (int q = 0; q < np; q++) uv[q] = uv[q] / ooJ[q];
    }
    free(c); free(t);
  }
}

/* ------------------------------------------------------------------ */
/* Checkpoint: CRC-64 and variables.bin (src/main/main/crc64.c,        */
/* src/main/main/norms.c:331-520, src/main/checkpoint/                 */
/* checkpoint_save.c:337-475)                                          */
/* ------------------------------------------------------------------ */

/* reflected CRC-64 with the ECMA-182 polynomial: 0x42f0e1eba9ea3693
   bit-reversed is 0xc96c5795d7870f42 (crc64.c:9-17) */
#define ORC_CRC64_POLY 0xc96c5795d7870f42ULL
static uint64_t g_crc_tab[256];
static int g_crc_tab_set = 0;

/* table of the byte-wise reflected CRC: entry b is b shifted through 8
   steps of the polynomial division */
static void crc_table(void)
{
  if (g_crc_tab_set) return;
  for (int b = 0; b < 256; b++) {
    uint64_t r = (uint64_t)b;
    for (int k = 0; k < 8; k++) r = (r & 1) ? (r >> 1) ^ ORC_CRC64_POLY : (r >> 1);
    g_crc_tab[b] = r;
  }
  g_crc_tab_set = 1;
}

/* crc64_continue (crc64.c:153-165): pre- and post-inverted */
uint64_t orc_crc64(uint64_t crc, const void *buf, size_t nbytes)
{
  crc_table();
  const uint8_t *d = (const uint8_t *)buf;
  crc = ~crc;
  while (nbytes--) crc = g_crc_tab[(crc ^ *d++) & 0xff] ^ (crc >> 8);
  return ~crc;
}

/* a*b mod P in the reflected representation (bit 63 is x^0) */
static uint64_t gf2_mulmod(uint64_t a, uint64_t b)
{
  uint64_t r = 0;
  for (int i = 63; i >= 0; i--) {
    if ((a >> i) & 1) r ^= b;
    b = (b & 1) ? (b >> 1) ^ ORC_CRC64_POLY : (b >> 1);
  }
  return r;
}

/* x^n mod P, reflected, by square-and-multiply */
static uint64_t gf2_xpow(uint64_t n)
{
  uint64_t r = 1ULL << 63, sq = 1ULL << 62;   /* x^0, x^1 */
  while (n) {
    if (n & 1) r = gf2_mulmod(r, sq);
    sq = gf2_mulmod(sq, sq);
    n >>= 1;
  }
  return r;
}

/* crc64_0start_combine (crc64.c:251-256): the CRC of the concatenation of
   two buffers whose CRCs both started from 0, the second nbytes2 long */
uint64_t orc_crc64_combine(uint64_t crc1, uint64_t crc2, uint64_t nbytes2)
{
  return crc2 ^ gf2_mulmod(crc1, gf2_xpow(8 * nbytes2));
}

/* scalarwave1's evolved variables in the reference's variable-index order
   (AddEvoVar order, nmesh_scalarwave1.c:20-22): phi, pi, cx, cy, cz; the
   state here is indexed pi, cx, cy, cz, phi */
static const int g_ckpt_var[5] = {4, 0, 1, 2, 3};
static const char *g_ckpt_name[5] = {"scalarwave1_phi", "scalarwave1_pi", "scalarwave1_cx",
                                     "scalarwave1_cy", "scalarwave1_cz"};

/* the "vars" CRC of nmesh_CRCs (norms.c:440-458 crc_VarList_local): every
   element in order, each variable of the list in order, its doubles as
   stored */
uint64_t orc_vars_crc(const OMesh *M, const double *u)
{
  long nn = M->nnodes;
  uint64_t crc = 0;
  for (int e = 0; e < M->nelem; e++) {
    const OElm *E = &M->e[e];
    for (int k = 0; k < 5; k++)
      crc = orc_crc64(crc, u + g_ckpt_var[k] * nn + E->off, sizeof(double) * elm_np(E));
  }
  return crc;
}

/* variables.bin as checkpoint_write_vl writes it (checkpoint_save.c:
   412-468): the variable list, then per element "{", its name
   "<patch>_<location digits>" (elmname, connections.c:819-866), np, and per
   variable its list index and np little-endian doubles; "EndOfFile." */
int orc_write_variables_bin(const OMesh *M, const double *u, const char *path)
{
  FILE *fp = fopen(path, "wb");
  if (!fp) return -1;
  long nn = M->nnodes;
  fprintf(fp, "number of variables and variable list in this file:\n");
  fprintf(fp, "%d\n", 5);
  for (int k = 0; k < 5; k++) fprintf(fp, "%s\n", g_ckpt_name[k]);
  fprintf(fp, "\n\n");
  fprintf(fp, "variable-list-data:\n\n");
  for (int e = 0; e < M->nelem; e++) {
    const OElm *E = &M->e[e];
    int np = elm_np(E);
    fprintf(fp, "{\n");
    fprintf(fp, "%d_%.*s\n", E->p, E->l, E->loc);
    fprintf(fp, "%d\n", np);
    for (int k = 0; k < 5; k++) {
      fprintf(fp, "%d\n", k);
      const double *v = u + g_ckpt_var[k] * nn + E->off;
      for (int q = 0; q < np; q++) {
        uint64_t bits;
        memcpy(&bits, v + q, 8);
        unsigned char b[8];
        for (int i = 0; i < 8; i++) b[i] = (unsigned char)(bits >> (8 * i));
        fwrite(b, 1, 8, fp);
      }
      fprintf(fp, "\n");
    }
    fprintf(fp, "}\n");
    fprintf(fp, "\n");
  }
  fprintf(fp, "EndOfFile.\n");
  return fclose(fp) ? -1 : 0;
}

/* ------------------------------------------------------------------ */
/* Mesh adaptation with the state interpolated onto the new elements    */
/* (src/main/amr/refine.c:156-537, storage.c:493-1000)                  */
/* ------------------------------------------------------------------ */

/* a copy of M (elements, patches; geometry rebuilt) */
OMesh *orc_mesh_clone(const OMesh *M)
{
  OMesh *C = calloc(1, sizeof(OMesh));
  C->npat = M->npat;
  memcpy(C->pat, M->pat, sizeof(M->pat));
  C->nelem = M->nelem;
  C->e = malloc(sizeof(OElm) * M->nelem);
  memcpy(C->e, M->e, sizeof(OElm) * M->nelem);
  mesh_finish(C);
  return C;
}

/* is the run of 8 elements from e the 8 children of one parent, in child
   order (amr_elms_are_siblings, child0 first) */
static int are_siblings(const OMesh *M, int e)
{
  if (e + 8 > M->nelem) return 0;
  const OElm *E0 = &M->e[e];
  if (E0->l < 1) return 0;
  for (int c = 0; c < 8; c++) {
    const OElm *E = &M->e[e + c];
    if (E->p != E0->p || E->l != E0->l) return 0;
    if (memcmp(E->loc, E0->loc, E0->l - 1) != 0) return 0;
    if (E->loc[E0->l - 1] != '0' + c) return 0;
  }
  return 1;
}

/* h-adaptation by flag (hp_refine_elms_if_rflag with PARENT_n, refine.c:
   174-276; remove_elms_if_rflag, refine.c:359-506): flag > 0 replaces an
   element by its 8 children of the same n; 8 siblings in a row that all have
   flag < 0 are replaced by their parent, whose n is that of the sibling with
   the most points (refine_sib_with_max_np_pt_typ, refine.c:314-357).
   Refinement first, then coarsening, as two passes. */
void orc_mesh_hadapt(OMesh *M, const int *flag)
{
  int *f = malloc(sizeof(int) * M->nelem);
  int *ref = calloc(M->nelem, sizeof(int));
  for (int e = 0; e < M->nelem; e++) { f[e] = flag[e]; ref[e] = flag[e] > 0; }
  /* coarsening flags survive refinement only on elements not refined */
  int old_n = M->nelem;
  refine_flagged(M, ref);
  int *cf = calloc(M->nelem, sizeof(int));
  for (int e = 0, k = 0; e < old_n; e++) {
    if (ref[e]) { k += 8; continue; }
    cf[k++] = f[e] < 0;
  }
  /* coarsen */
  OElm *ne = malloc(sizeof(OElm) * M->nelem);
  int k = 0;
  for (int e = 0; e < M->nelem;) {
    int all = are_siblings(M, e);
    for (int c = 0; all && c < 8; c++) all = cf[e + c];
    if (!all) { ne[k++] = M->e[e]; e++; continue; }
    OElm P;
    memset(&P, 0, sizeof(P));
    P.p = M->e[e].p;
    P.l = M->e[e].l - 1;
    memcpy(P.loc, M->e[e].loc, sizeof(P.loc));
    P.loc[P.l] = 0;
    int nmax = 0;
    for (int c = 0; c < 8; c++) if (M->e[e + c].n > nmax) nmax = M->e[e + c].n;
    P.n = nmax;
    P.use_fv = M->e[e].use_fv;
    elm_bbox_from_loc(&M->pat[P.p], &P);
    ne[k++] = P;
    e += 8;
  }
  free(M->e);
  M->e = ne;
  M->nelem = k;
  free(f); free(ref); free(cf);
  mesh_finish(M);
}

/* 1D rows A[i*ns + s] = l_s(x_i) of the Lagrange basis on the ns points of
   an element: LGL, or uniform (uni) for an FV element (basis_interp_toIpoints
   with Lagrange_of_x at amr_Lagrange_interp_order n, interpolate.c:540-600;
   between the point types of one n, LGL_to_nUNIt and its PseudoInv inverse
   UNI_to_nLGLt, gridpoints.c:158-290, which for a square matrix is the
   Lagrange interpolation back) */
static void lagrange_rows_on(int nd, const double *x, int ns, int uni, double *A)
{
  OBasis *b = basis(ns);
  const double *xs = uni ? b->Xu : b->Xb;
  double w[ORC_MAXN];
  lagrange_winterp(ns, xs, w);
  for (int i = 0; i < nd; i++)
    for (int s = 0; s < ns; s++) A[i * ns + s] = lagrange_of_x(s, x[i], ns, xs, w);
}

static void lagrange_rows(int nd, const double *x, int ns, double *A)
{
  lagrange_rows_on(nd, x, ns, 0, A);
}

/* out[k][j][i] (nd^3) = sum A2[k][c] A1[j][b] A0[i][a] in[c][b][a] (ns^3) */
static void tensor_apply(int nd, int ns, const double *A0, const double *A1, const double *A2,
                         const double *in, double *out)
{
  double *t1 = malloc(sizeof(double) * ns * ns * nd), *t2 = malloc(sizeof(double) * ns * nd * nd);
  for (int c = 0; c < ns; c++)
    for (int b = 0; b < ns; b++)
      for (int i = 0; i < nd; i++) {
        double s = 0.;
        for (int a = 0; a < ns; a++) s += A0[i * ns + a] * in[a + ns * (b + ns * c)];
        t1[i + nd * (b + ns * c)] = s;
      }
  for (int c = 0; c < ns; c++)
    for (int j = 0; j < nd; j++)
      for (int i = 0; i < nd; i++) {
        double s = 0.;
        for (int b = 0; b < ns; b++) s += A1[j * ns + b] * t1[i + nd * (b + ns * c)];
        t2[i + nd * (j + nd * c)] = s;
      }
  for (int k = 0; k < nd; k++)
    for (int j = 0; j < nd; j++)
      for (int i = 0; i < nd; i++) {
        double s = 0.;
        for (int c = 0; c < ns; c++) s += A2[k * ns + c] * t2[i + nd * (j + nd * c)];
        out[i + nd * (j + nd * k)] = s;
      }
  free(t1); free(t2);
}

static int find_loc(const OMesh *M, int p, int l, const char *loc)
{
  for (int e = 0; e < M->nelem; e++)
    if (M->e[e].p == p && M->e[e].l == l && memcmp(M->e[e].loc, loc, l) == 0) return e;
  return -1;
}

/* The state of Mnew from that of Mold (update_node_n_pt_typ_return_node_old
   storage.c:858-931, make_child_elm storage.c:493-585, make_parent_elm
   storage.c:627-760): an element at the same place keeps its values (same
   n) or is Lagrange-interpolated to its new points; a child is interpolated
   from its old parent; a parent takes at each of its points the average
   over the old children that hold the point of their interpolants.  DG
   elements (LGL points) only. */
int orc_adapt_state(const OMesh *Mold, const double *uold, const OMesh *Mnew, double *unew)
{
  const long nno = Mold->nnodes, nnn = Mnew->nnodes;
  for (int e = 0; e < Mnew->nelem; e++) {
    const OElm *E = &Mnew->e[e];
    const int nd = E->n, npd = elm_np(E);
    const double *xd = E->use_fv ? basis(nd)->Xu : basis(nd)->Xb;
    double A[3][ORC_MAXN * ORC_MAXN];
    int s = find_loc(Mold, E->p, E->l, E->loc);
    if (s >= 0) {                                  /* same place */
      const OElm *S = &Mold->e[s];
      if (S->n == nd && S->use_fv == E->use_fv) {
        for (int v = 0; v < NV; v++) memcpy(unew + v * nnn + E->off, uold + v * nno + S->off, sizeof(double) * npd);
        continue;
      }
      if (S->use_fv && S->n != nd) return -1;    /* WENO p-interpolation: not built */
      for (int d = 0; d < 3; d++) lagrange_rows_on(nd, xd, S->n, S->use_fv, A[d]);
      for (int v = 0; v < NV; v++)
        tensor_apply(nd, S->n, A[0], A[1], A[2], uold + v * nno + S->off, unew + v * nnn + E->off);
      continue;
    }
    if (E->use_fv) return -1;                    /* FV h-adaptation: not built */
    if (E->l > 0 && (s = find_loc(Mold, E->p, E->l - 1, E->loc)) >= 0) {   /* child of an old parent */
      const OElm *S = &Mold->e[s];
      const int c = E->loc[E->l - 1] - '0';
      for (int d = 0; d < 3; d++) {
        double xp[ORC_MAXN];
        for (int i = 0; i < nd; i++) xp[i] = 0.5 * xd[i] + (((c >> d) & 1) ? 0.5 : -0.5);
        lagrange_rows(nd, xp, S->n, A[d]);
      }
      for (int v = 0; v < NV; v++)
        tensor_apply(nd, S->n, A[0], A[1], A[2], uold + v * nno + S->off, unew + v * nnn + E->off);
      continue;
    }
    /* parent of 8 old children */
    char cl[32];
    memcpy(cl, E->loc, sizeof(cl));
    double *sum = calloc((size_t)NV * npd, sizeof(double));
    int *cnt = calloc(npd, sizeof(int));
    double *tmp = malloc(sizeof(double) * npd);
    for (int c = 0; c < 8; c++) {
      cl[E->l] = '0' + c;
      int sc = find_loc(Mold, E->p, E->l + 1, cl);
      if (sc < 0) { free(sum); free(cnt); free(tmp); return -2; }
      const OElm *S = &Mold->e[sc];
      /* my points in the child's Xb; a point is the child's if it lies in
         [-1, 1] there (array_find_XYZ_in_node), boundaries included */
      int in[3][ORC_MAXN];
      for (int d = 0; d < 3; d++) {
        double xc[ORC_MAXN];
        for (int i = 0; i < nd; i++) {
          xc[i] = 2. * xd[i] - (((c >> d) & 1) ? 1. : -1.);
          in[d][i] = (xc[i] >= -1. - 1e-10 && xc[i] <= 1. + 1e-10);
        }
        lagrange_rows(nd, xc, S->n, A[d]);
      }
      for (int v = 0; v < NV; v++) {
        tensor_apply(nd, S->n, A[0], A[1], A[2], uold + v * nno + S->off, tmp);
        for (int k = 0; k < nd; k++)
          for (int j = 0; j < nd; j++)
            for (int i = 0; i < nd; i++)
              if (in[0][i] && in[1][j] && in[2][k]) {
                int q = i + nd * (j + nd * k);
                sum[v * npd + q] += tmp[q];
                if (v == 0) cnt[q]++;
              }
      }
    }
    for (int v = 0; v < NV; v++)
      for (int q = 0; q < npd; q++)
        unew[v * nnn + E->off + q] = cnt[q] > 1 ? sum[v * npd + q] / cnt[q] : sum[v * npd + q];
    free(sum); free(cnt); free(tmp);
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Troubled elements (src/main/evolve/trouble.c)                       */
/* ------------------------------------------------------------------ */

/* evolve_set_trouble_score (trouble.c:27-110) on every element from its
   TROUBLE value tr[e] (1 trouble, 0 keep, -1 all good): 1 starts the score
   at 1 or raises it, 0 sets it to 0, -1 starts it at -1 or lowers it,
   clamped to +-2^30.  tr == NULL reads the scores only.  Returns the
   mesh's score (evolve_read_trouble_score_mesh, trouble.c:129-161): the
   maximum when it is positive, else the minimum. */
int orc_trouble_score(int nelem, const int *tr, int *score)
{
  const int max_trouble = 1073741824;
  int mx = INT_MIN, mn = INT_MAX;
  for (int e = 0; e < nelem; e++) {
    int s = score[e];
    if (tr) {
      if (tr[e] > 0) s = (s <= 0) ? 1 : s + 1;
      else if (tr[e] == 0) s = 0;
      else s = (s >= 0) ? -1 : s - 1;
      if (s < -max_trouble) s = -max_trouble;
      if (s > max_trouble) s = max_trouble;
      score[e] = s;
    }
    if (s > mx) mx = s;
    if (s < mn) mn = s;
  }
  return mx > 0 ? mx : mn;
}

/* The library's RDMP indicator on the data evolve_collect_u_p_data_mesh
   gathers (trouble.c:355-400: limdata_MRS of u_p, the minimum and maximum
   over the element's nodes): m and M are the extremes of u_p over the
   element and its face neighbours, d = max(delta0, eps (M - m)) per
   variable; a value of u outside [m - d, M + d] (or NaN) puts the element
   out of bounds.  tr = 1 for a DG element out of bounds, -1 for an FV
   element in bounds, 0 otherwise. */
void orc_trouble_rdmp(const OMesh *M, const double *u, const double *up, double delta0, double eps, int *tr)
{
  const long nn = M->nnodes;
  double *mm = malloc(sizeof(double) * 2 * NV * M->nelem);
  for (int e = 0; e < M->nelem; e++) {
    const OElm *E = &M->e[e];
    const int np = elm_np(E);
    for (int v = 0; v < NV; v++) {
      const double *q = up + v * nn + E->off;
      double lo = q[0], hi = q[0];
      for (int i = 1; i < np; i++) { if (q[i] < lo) lo = q[i]; if (q[i] > hi) hi = q[i]; }
      mm[2 * (NV * e + v)] = lo;
      mm[2 * (NV * e + v) + 1] = hi;
    }
  }
  for (int e = 0; e < M->nelem; e++) {
    const OElm *E = &M->e[e];
    const int np = elm_np(E);
    int out = 0;
    for (int v = 0; v < NV && !out; v++) {
      double m = mm[2 * (NV * e + v)], Mx = mm[2 * (NV * e + v) + 1];
      for (int f = 0; f < 6; f++)
        for (int ni = 0; ni < E->nfnb[f]; ni++) {
          int nb = E->fnb[f][ni];
          m = fmin(m, mm[2 * (NV * nb + v)]);
          Mx = fmax(Mx, mm[2 * (NV * nb + v) + 1]);
        }
      double d = fmax(delta0, eps * (Mx - m));
      const double *q = u + v * nn + E->off;
      for (int i = 0; i < np; i++)
        if (!(q[i] >= m - d && q[i] <= Mx + d)) { out = 1; break; }
    }
    tr[e] = E->use_fv ? (out ? 0 : -1) : (out ? 1 : 0);
  }
  free(mm);
}

/* the FV flags of all elements (node->dat->use_fv) */
void orc_mesh_fv_flags(const OMesh *M, int *fv)
{
  for (int e = 0; e < M->nelem; e++) fv[e] = M->e[e].use_fv;
}
