// mesh.cpp -- mesh generation, geometry, neighbour and face tables of libnmx.
//
// Builds nmesh's BoxMesh and Shell meshes (src/main/amr/mesh.c:181-669) with
// this library's own data structures: elements carry integer octant
// coordinates, neighbours are found by locating sample points of each face
// in the octree of the same or the adjacent patch, orientations by matching
// face corners, and non-conforming faces get separable 1D Lagrange
// interpolation matrices.  Element order, neighbour lists and the ajsurf
// case of each face follow the reference (surface.c:799-1080).
#include "mesh.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <exception>
#include <mutex>
#include <thread>
#include <unordered_map>

namespace nmx {

// ------------------------------------------------------------------------
// LGL basis
// ------------------------------------------------------------------------

// Legendre P_N and the derivative of Q = P_{N+1} - P_{N-1}: the interior
// LGL nodes are the roots of Q (Q is proportional to (x^2-1) P_N').
static void legendre_q(int N, double x, double *PN, double *Q, double *dQ) {
  double pm1 = 1.0, p = x, dpm1 = 0.0, dp = 1.0;
  double pk = 0, dpk = 0;
  for (int k = 2; k <= N + 1; ++k) {
    pk = (p * x * (2 * k - 1) - pm1 * (k - 1)) / k;
    dpk = p * (2 * k - 1) + dpm1;
    if (k == N + 1) break;
    pm1 = p; p = pk; dpm1 = dp; dp = dpk;
  }
  *PN = p;
  *Q = pk - pm1;
  *dQ = dpk - dpm1;
}

static void make_basis(int n, Basis &b) {
  b.n = n;
  int N = n - 1;
  if (N == 0) {
    b.x[0] = 0; b.w[0] = 2;
  } else {
    b.x[0] = -1; b.x[N] = 1;
    b.w[0] = b.w[N] = 2.0 / (N * (N + 1.0));
    for (int j = 1; j <= N / 2; ++j) {
      double y, PN, Q, dQ;
      if (2 * j == N) {
        y = 0.0;
      } else {
        // Chebyshev-like starting guess, then Newton on Q
        y = -std::cos(M_PI * (j + 0.25) / N - 3 / (8 * N * M_PI * (j + 0.25)));
        for (int it = 0; it < 100; ++it) {
          legendre_q(N, y, &PN, &Q, &dQ);
          double dy = -Q / dQ;
          y += dy;
          if (std::fabs(dy) < std::fabs(1e-15 * y)) break;
        }
      }
      legendre_q(N, y, &PN, &Q, &dQ);
      b.x[j] = y; b.x[N - j] = -y;
      b.w[j] = b.w[N - j] = 2.0 / ((N * (N + 1)) * PN * PN);
    }
  }
  for (int i = 0; i < n; ++i) {
    double d = 1.0;
    for (int m = 0; m < n; ++m)
      if (m != i) d *= (b.x[i] - b.x[m]);
    b.wb[i] = 1.0 / d;
  }
  for (int i = 0; i < n; ++i) {
    double dii = 0.0;
    for (int l = 0; l < n; ++l) {
      if (l == i) continue;
      double dil = (b.wb[l] / b.wb[i]) / (b.x[i] - b.x[l]);
      dii -= dil;
      b.D[i * n + l] = dil;
    }
    b.D[i * n + i] = dii;
  }
}

// uniform points with both faces: ends at -+1, -1 + j h on the left half,
// mirrored, 0 exactly at the centre (uniform_x, Legendre.c:626-663)
static void make_uniform(int n, Basis &b) {
  b.n = n;
  int N = n - 1;
  if (N == 0) { b.x[0] = 0; b.wb[0] = 1; return; }
  b.x[0] = -1; b.x[N] = 1;
  double h = 2. / N;
  for (int j = 1; j <= N / 2 && N > 1; ++j) {
    double y = (j == N / 2.0) ? 0. : -1. + j * h;
    b.x[N - j] = -y;
    b.x[j] = y;
  }
  // barycentric weights 1 / prod_{m != q} (x_q - x_m) for Lagrange rows on
  // these points (fv_surface_interp = ... Lagrange_if_dg)
  for (int q = 0; q < n; ++q) {
    double p = 1.0;
    for (int m = 0; m < n; ++m)
      if (m != q) p *= b.x[q] - b.x[m];
    b.wb[q] = 1.0 / p;
  }
}

// The bases are built once per n under a lock; afterwards a lookup only
// reads the flag (geometry() calls these from several threads).
const Basis &uniform_basis(int n) {
  static Basis cache[kMaxN + 1];
  static std::atomic<bool> done[kMaxN + 1];
  static std::mutex mu;
  if (!done[n].load(std::memory_order_acquire)) {
    std::lock_guard<std::mutex> lk(mu);
    if (!done[n].load(std::memory_order_relaxed)) {
      make_uniform(n, cache[n]);
      done[n].store(true, std::memory_order_release);
    }
  }
  return cache[n];
}

const Basis &lgl_basis(int n) {
  static Basis cache[kMaxN + 1];
  static std::atomic<bool> done[kMaxN + 1];
  static std::mutex mu;
  if (!done[n].load(std::memory_order_acquire)) {
    std::lock_guard<std::mutex> g(mu);
    if (!done[n].load(std::memory_order_relaxed)) {
      make_basis(n, cache[n]);
      done[n].store(true, std::memory_order_release);
    }
  }
  return cache[n];
}

// The exponential filter along one direction on n LGL nodes, F = S diag(f)
// A: the normalized Legendre analysis A_il = w_l P_i(x_l)/sqrt(c_i) and
// synthesis S_ji = P_i(x_j)/sqrt(c_i), c_i = 2/(2i+1) except c_N = 2/N
// (LGL_AT_ST_matrices, src/main/basis/Legendre.c:257-307, with P_i by the
// recurrence of Legendre.c:14-38), around f(i) = exp(-alp (i/N)^s),
// N = n-1-dn, 1 if not positive (expfilter_coeff_array, filter.c:20-55).
// F[j*n + l].
void legendre_analysis_synthesis(int n, double *A, double *S) {
  const Basis &b = lgl_basis(n);
  const int N = n - 1;
  for (int i = 0; i < n; ++i) {
    double ooci = 0.5 * (2 * i + 1);
    if (i == N && N != 0) ooci = 0.5 * i;
    const double oosci = std::sqrt(ooci);
    for (int j = 0; j < n; ++j) {
      const double x = b.x[j];
      double P0 = 1.0, P1 = x, Pk = (i == 0) ? 1.0 : x;
      for (int k = 2; k <= i; ++k) {
        Pk = (x * (2 * k - 1) * P1 - (k - 1) * P0) / k;
        P0 = P1;
        P1 = Pk;
      }
      S[j * n + i] = Pk * oosci;
      A[i * n + j] = b.w[j] * S[j * n + i];
    }
  }
}

void filter_matrix(int n, double alp, double s, int dn, double *F) {
  std::vector<double> A(n * n), S(n * n), f(n);
  legendre_analysis_synthesis(n, A.data(), S.data());
  double Nf = n - 1 - dn;
  if (Nf <= 0.) Nf = 1.;
  for (int i = 0; i < n; ++i) f[i] = std::exp(-alp * std::pow(i / Nf, s));
  for (int j = 0; j < n; ++j)
    for (int l = 0; l < n; ++l) {
      double sum = 0.0;
      for (int i = 0; i < n; ++i) sum += S[j * n + i] * f[i] * A[i * n + l];
      F[j * n + l] = sum;
    }
}

static double lagrange(const Basis &b, int q, double x) {
  double prod = 1.0;
  for (int m = 0; m < b.n; ++m)
    if (m != q) prod *= (x - b.x[m]);
  return prod * b.wb[q];
}

// piecewise linear basis function q on the points of b (basis_pw_linear,
// src/main/basis/interpolate.c:53-102): the end intervals reach past the
// ends; a point on an inner node belongs to the interval on its right
static double pw_linear(const Basis &b, int q, double x) {
  const int np = b.n;
  const double *xp = b.x;
  if (np <= 1) return 1.0;
  if (x < xp[1]) {
    if (q == 0) return (x - xp[1]) / (xp[0] - xp[1]);
    if (q == 1) return (x - xp[0]) / (xp[1] - xp[0]);
    return 0.0;
  }
  if (x >= xp[np - 2]) {
    const int n0 = np - 2, n1 = np - 1;
    if (q == n0) return (x - xp[n1]) / (xp[n0] - xp[n1]);
    if (q == n1) return (x - xp[n0]) / (xp[n1] - xp[n0]);
    return 0.0;
  }
  if (q < np - 1 && x >= xp[q] && x < xp[q + 1]) return (x - xp[q + 1]) / (xp[q] - xp[q + 1]);
  if (q > 0 && x >= xp[q - 1] && x < xp[q]) return (x - xp[q - 1]) / (xp[q] - xp[q - 1]);
  return 0.0;
}

// piecewise parabolic basis function q on the points of b (basis_pw_parab,
// src/main/basis/interpolate.c:104-224): between the midpoints around node
// m the parabola through m-1, m, m+1 (so basis q lives on the intervals
// around q-1, q and q+1); the parabola through the first (last) three points
// below the first (above the last) midpoint; one parabola for 3 points
static double parab3(const double *xp, int a, int b, int c, double x) {
  return (x - xp[b]) * (x - xp[c]) / ((xp[a] - xp[b]) * (xp[a] - xp[c]));
}

static double pw_parab(const Basis &b, int q, double x) {
  const int np = b.n;
  const double *xp = b.x;
  if (np < 3) return pw_linear(b, q, x);
  if (np == 3 || x < 0.5 * (xp[0] + xp[1])) {
    switch (q) {
      case 0: return parab3(xp, 0, 1, 2, x);
      case 1: return parab3(xp, 1, 0, 2, x);
      case 2: return parab3(xp, 2, 0, 1, x);
      default: return 0.0;
    }
  }
  if (x >= 0.5 * (xp[np - 2] + xp[np - 1])) {
    const int a = np - 3, l = q - a;
    switch (l) {
      case 0: return parab3(xp, a, a + 1, a + 2, x);
      case 1: return parab3(xp, a + 1, a, a + 2, x);
      case 2: return parab3(xp, a + 2, a, a + 1, x);
      default: return 0.0;
    }
  }
  // interior x: basis q is nonzero on the cells around q - 1, q, q + 1,
  // whose bounds at the ends of the point set are mirrored outwards
  const int m = q <= 0 ? 0 : (q >= np - 1 ? np - 1 : q);
  double lo, hi;   // cell of node m
  if (m == 0) {
    hi = 0.5 * (xp[0] + xp[1]);
    lo = xp[0] - (hi - xp[0]);
  } else if (m == np - 1) {
    lo = 0.5 * (xp[m - 1] + xp[m]);
    hi = xp[m] + (xp[m] - lo);
  } else {
    lo = 0.5 * (xp[m - 1] + xp[m]);
    hi = 0.5 * (xp[m] + xp[m + 1]);
  }
  const double llo = (m >= 2) ? 0.5 * (xp[m - 2] + xp[m - 1]) : lo - (hi - lo);
  const double hhi = (m + 2 < np) ? 0.5 * (xp[m + 1] + xp[m + 2]) : hi + (hi - lo);
  if (x >= lo && x < hi) {
    // the parabola of cell m: through m-1, m, m+1 (one-sided at the ends)
    if (m == 0) return parab3(xp, 0, 1, 2, x);
    if (m == np - 1) return parab3(xp, m, m - 2, m - 1, x);
    return parab3(xp, m, m - 1, m + 1, x);
  }
  if (x >= llo && x < lo) return (m >= 2) ? parab3(xp, m, m - 2, m - 1, x) : 0.0;
  if (x >= hi && x < hhi) return (m + 2 < np) ? parab3(xp, m, m + 1, m + 2, x) : 0.0;
  return 0.0;
}

// ------------------------------------------------------------------------
// coordinate maps: Cartesian box and CubedShell with constant radii
// (src/main/coordinates/coordtrans_CubedSphere.c:38-509)
// ------------------------------------------------------------------------

static inline void shell_a(const Patch &P, double A, double B, double *a0, double *a1) {
  double pm = 2 * (P.dom % 2) - 1;
  double s = std::sqrt(1.0 + A * A + B * B);
  *a0 = pm * P.s0 / s;
  *a1 = pm * P.s1 / s;
}

void patch_xyz(const Patch &P, const double X[3], double x[3]) {
  if (P.type != PAT_CUBEDSHELL) { x[0] = X[0]; x[1] = X[1]; x[2] = X[2]; return; }
  double a0, a1;
  shell_a(P, X[1], X[2], &a0, &a1);
  double r = (a1 - a0) * X[0] + a0;
  int dir = P.dom / 2;
  // the radial coordinate lies along axis dir; A and B scale the other two
  int ia = (dir == 0) ? 1 : (dir == 1) ? 0 : 1;
  int ib = (dir == 2) ? 0 : 2;
  x[dir] = P.xc[dir] + r;
  x[ia] = P.xc[ia] + X[1] * r;
  x[ib] = P.xc[ib] + X[2] * r;
}

void patch_XYZ(const Patch &P, const double x[3], double X[3]) {
  if (P.type != PAT_CUBEDSHELL) { X[0] = x[0]; X[1] = x[1]; X[2] = x[2]; return; }
  int dir = P.dom / 2;
  int ia = (dir == 0) ? 1 : (dir == 1) ? 0 : 1;
  int ib = (dir == 2) ? 0 : 2;
  double r = x[dir] - P.xc[dir];
  double A = (x[ia] - P.xc[ia]) / r, B = (x[ib] - P.xc[ib]) / r;
  double a0, a1;
  shell_a(P, A, B, &a0, &a1);
  X[0] = (r - a0) / (a1 - a0);
  X[1] = A;
  X[2] = B;
}

// dX^a/dx^b for the CubedShell; same expressions as the reference
// (coordtrans_CubedSphere.c:381-507) with dsigma = 0.
void patch_dXdx(const Patch &P, const double X[3], double x[3], double M[3][3]) {
  if (P.type != PAT_CUBEDSHELL) {
    for (int a = 0; a < 3; ++a) {
      x[a] = X[a];
      for (int b = 0; b < 3; ++b) M[a][b] = (a == b) ? 1.0 : 0.0;
    }
    return;
  }
  const double lam = X[0];
  double a0, a1;
  shell_a(P, X[1], X[2], &a0, &a1);
  patch_xyz(P, X, x);
  const int dir = P.dom / 2;
  double rv[3] = {x[0] - P.xc[0], x[1] - P.xc[1], x[2] - P.xc[2]};
  double rr = rv[dir];
  double rc2 = rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
  double m[3];
  for (int d = 0; d < 3; ++d) m[d] = -rv[d] / rc2;
  m[dir] = 1.0 / rr - rr / rc2;
  for (int d = 0; d < 3; ++d) {
    double da0 = a0 * (m[d] + 0.0), da1 = a1 * (m[d] + 0.0);
    if (d == dir) M[0][d] = (1.0 - da0 - (da1 - da0) * lam) / (a1 - a0);
    else M[0][d] = (-da0 - (da1 - da0) * lam) / (a1 - a0);
  }
  double rr2 = rr * rr;
  int ia = (dir == 0) ? 1 : (dir == 1) ? 0 : 1;
  int ib = (dir == 2) ? 0 : 2;
  // A = r_ia / r_dir, B = r_ib / r_dir
  for (int d = 0; d < 3; ++d) { M[1][d] = 0.0; M[2][d] = 0.0; }
  M[1][dir] = -rv[ia] / rr2;
  M[1][ia] = 1.0 / rr;
  M[2][dir] = -rv[ib] / rr2;
  M[2][ib] = 1.0 / rr;
}

// ------------------------------------------------------------------------
// octree
// ------------------------------------------------------------------------

static void elm_bbox(const Patch &P, Elm &E) {
  double LX[3];
  for (int f = 0; f < 6; ++f) E.bbox[f] = P.bbox[f];
  for (int d = 0; d < 3; ++d) LX[d] = E.bbox[2 * d + 1] - E.bbox[2 * d];
  for (int ll = 0; ll < E.l; ++ll) {
    int c = E.path[ll];
    for (int d = 0; d < 3; ++d) {
      LX[d] /= 2.;
      if (c & (1 << d)) E.bbox[2 * d] = E.bbox[2 * d + 1] - LX[d];
      else E.bbox[2 * d + 1] = E.bbox[2 * d] + LX[d];
    }
  }
}

static void refine(Mesh &M, const std::vector<char> &flag) {
  std::vector<Elm> ne;
  ne.reserve(M.e.size() * 2);
  for (size_t i = 0; i < M.e.size(); ++i) {
    if (!flag[i]) { ne.push_back(M.e[i]); continue; }
    const Elm &P = M.e[i];
    for (int c = 0; c < 8; ++c) {
      Elm C;
      C.p = P.p; C.l = P.l + 1; C.n = P.n;
      std::memcpy(C.path, P.path, sizeof(C.path));
      C.path[P.l] = (uint8_t)c;
      for (int d = 0; d < 3; ++d) C.ic[d] = 2 * P.ic[d] + ((c >> d) & 1);
      elm_bbox(M.pat[C.p], C);
      ne.push_back(C);
    }
  }
  M.e.swap(ne);
}

static void add_roots(Mesh &M, int n) {
  M.e.clear();
  for (size_t p = 0; p < M.pat.size(); ++p) {
    Elm E;
    E.p = (int)p; E.l = 0; E.n = n;
    elm_bbox(M.pat[p], E);
    M.e.push_back(E);
  }
}

static uint64_t okey(int p, int l, uint32_t i, uint32_t j, uint32_t k) {
  return ((uint64_t)p << 60) | ((uint64_t)l << 55) | ((uint64_t)i << 36) |
         ((uint64_t)j << 18) | (uint64_t)k;
}

struct Locator {
  const Mesh &M;
  std::unordered_map<uint64_t, int> leaf;
  explicit Locator(const Mesh &m) : M(m) {
    for (size_t i = 0; i < M.e.size(); ++i) {
      const Elm &E = M.e[i];
      leaf[okey(E.p, E.l, E.ic[0], E.ic[1], E.ic[2])] = (int)i;
    }
  }
  // leaf of patch p containing patch point X (clamped into the patch)
  int find(int p, const double X[3]) const {
    const Patch &P = M.pat[p];
    for (int l = 0; l <= M.maxlevel; ++l) {
      uint32_t c[3];
      uint32_t nl = 1u << l;
      for (int d = 0; d < 3; ++d) {
        double s = (X[d] - P.bbox[2 * d]) / (P.bbox[2 * d + 1] - P.bbox[2 * d]);
        long v = (long)std::floor(s * nl);
        if (v < 0) v = 0;
        if (v >= (long)nl) v = nl - 1;
        c[d] = (uint32_t)v;
      }
      auto it = leaf.find(okey(p, l, c[0], c[1], c[2]));
      if (it != leaf.end()) return it->second;
    }
    return -1;
  }
};

static inline void dirs_of(int dir, int *d1, int *d2) {
  *d1 = (dir == 0) ? 1 : 0;      // Dir1_norm (nmesh_amr_loops.h:295)
  *d2 = (dir == 2) ? 1 : 2;      // Dir2_norm (nmesh_amr_loops.h:296)
}

// patch point on face f of E at in-plane fractions (s,t)
static void face_X(const Elm &E, int f, double s, double t, double X[3]) {
  int dir = f / 2, d1, d2;
  dirs_of(dir, &d1, &d2);
  X[dir] = E.bbox[f];
  X[d1] = E.bbox[2 * d1] + s * (E.bbox[2 * d1 + 1] - E.bbox[2 * d1]);
  X[d2] = E.bbox[2 * d2] + t * (E.bbox[2 * d2 + 1] - E.bbox[2 * d2]);
}

static bool on_patface(const Mesh &M, const Elm &E, int f) {
  return E.bbox[f] == M.pat[E.p].bbox[f];
}

// find the patch q != p and its face that contains physical point x
static bool other_patch(const Mesh &M, int p, const double x[3], int *q, int *fq, double Xq[3]) {
  for (size_t qq = 0; qq < M.pat.size(); ++qq) {
    if ((int)qq == p) continue;
    const Patch &Q = M.pat[qq];
    if (Q.type == PAT_CUBEDSHELL) {
      int dir = Q.dom / 2;
      double pm = 2 * (Q.dom % 2) - 1;
      if ((x[dir] - Q.xc[dir]) * pm <= 0) continue;  // wrong half space
    }
    double X[3];
    patch_XYZ(Q, x, X);
    bool inside = true;
    for (int d = 0; d < 3 && inside; ++d) {
      double L = Q.bbox[2 * d + 1] - Q.bbox[2 * d];
      if (X[d] < Q.bbox[2 * d] - 1e-10 * L || X[d] > Q.bbox[2 * d + 1] + 1e-10 * L) inside = false;
    }
    if (!inside) continue;
    for (int f = 0; f < 6; ++f) {
      int d = f / 2;
      double L = Q.bbox[2 * d + 1] - Q.bbox[2 * d];
      if (std::fabs(X[d] - Q.bbox[f]) <= 1e-9 * L) {
        *q = (int)qq; *fq = f;
        for (int k = 0; k < 3; ++k) Xq[k] = X[k];
        return true;
      }
    }
  }
  return false;
}

// face neighbours of elements [e0, e1): each element writes only its own
// lists; the locator and the other elements' boxes are read only
static void find_neighbours_range(Mesh &M, const Locator &loc, size_t e0, size_t e1) {
  const int K = 8;   // samples per face edge: resolves level jumps up to 3
  for (size_t ei = e0; ei < e1; ++ei) {
    Elm &E = M.e[ei];
    for (int f = 0; f < 6; ++f) {
      std::vector<std::pair<int, int>> found;
      int dir = f / 2;
      double sig = 2 * (f % 2) - 1;
      bool onpf = on_patface(M, E, f);
      for (int sb = 0; sb < K; ++sb)
        for (int sa = 0; sa < K; ++sa) {
          double X[3];
          face_X(E, f, (sa + 0.5) / K, (sb + 0.5) / K, X);
          int nb = -1, nf = -1;
          if (!onpf) {
            double L = E.bbox[2 * dir + 1] - E.bbox[2 * dir];
            X[dir] += sig * 1e-6 * L;
            nb = loc.find(E.p, X);
            nf = f ^ 1;
          } else {
            double x[3], Xq[3];
            int q, fq;
            patch_xyz(M.pat[E.p], X, x);
            if (other_patch(M, E.p, x, &q, &fq, Xq)) {
              int dq = fq / 2;
              double L = M.pat[q].bbox[2 * dq + 1] - M.pat[q].bbox[2 * dq];
              double sq = 2 * (fq % 2) - 1;
              Xq[dq] -= sq * 1e-9 * L;
              nb = loc.find(q, Xq);
              nf = fq;
            }
          }
          if (nb < 0) continue;
          bool have = false;
          for (auto &pr : found) have |= (pr.first == nb);
          if (!have) found.push_back({nb, nf});
        }
      std::sort(found.begin(), found.end());
      E.nfnb[f] = (int)found.size();
      for (size_t i = 0; i < found.size(); ++i) {
        E.fnb[f][i] = found[i].first;
        E.nbf[f][i] = found[i].second;
      }
    }
  }
}

static void find_neighbours(Mesh &M) {
  const Locator loc(M);
  for_element_ranges(M.e.size(), [&M, &loc](unsigned, size_t e0, size_t e1) { find_neighbours_range(M, loc, e0, e1); });
  for (auto &P : M.pat)
    for (int f = 0; f < 6; ++f) P.bound[f] = 1;
  for (auto &E : M.e)
    for (int f = 0; f < 6; ++f)
      if (on_patface(M, E, f) && E.nfnb[f] > 0) M.pat[E.p].bound[f] = 0;
}

// ------------------------------------------------------------------------
// geometry
// ------------------------------------------------------------------------

static void node_X(const Elm &E, int i, int j, int k, double X[3]) {
  const Basis &b = E.fv ? uniform_basis(E.n) : lgl_basis(E.n);
  int I[3] = {i, j, k};
  for (int d = 0; d < 3; ++d)
    X[d] = 0.5 * ((E.bbox[2 * d + 1] - E.bbox[2 * d]) * b.x[I[d]] + (E.bbox[2 * d + 1] + E.bbox[2 * d]));
}

static void node_x(const Mesh &M, const Elm &E, int i, int j, int k, double x[3]) {
  double X[3];
  node_X(E, i, j, k, X);
  patch_xyz(M.pat[E.p], X, x);
}

static double dist3(const double a[3], const double b[3]) {
  double s = 0;
  for (int d = 0; d < 3; ++d) s += (a[d] - b[d]) * (a[d] - b[d]);
  return std::sqrt(s);
}

// smallest axial node distance at the 8 corners, starting from the space
// diagonal / sqrt(3) (src/main/coordinates/utilities.c:1028-1124)
static double find_hmin(const Mesh &M, const Elm &E, bool *pair) {
  const Patch &P = M.pat[E.p];
  double X0[3] = {E.bbox[0], E.bbox[2], E.bbox[4]}, X1[3] = {E.bbox[1], E.bbox[3], E.bbox[5]};
  double x0[3], x1[3];
  patch_xyz(P, X0, x0);
  patch_xyz(P, X1, x1);
  double h = dist3(x0, x1) / std::sqrt(3.);
  *pair = false;
  int n = E.n;
  for (int c = 0; c < 8; ++c) {
    int I[3] = {(c & 1) ? n - 1 : 0, (c & 2) ? n - 1 : 0, (c & 4) ? n - 1 : 0};
    double xa[3];
    node_x(M, E, I[0], I[1], I[2], xa);
    for (int d = 0; d < 3; ++d)
      for (int s = -1; s <= 1; s += 2) {
        int J[3] = {I[0], I[1], I[2]};
        J[d] += s;
        if (J[d] < 0 || J[d] >= n) continue;
        double xb[3];
        node_x(M, E, J[0], J[1], J[2], xb);
        double dd = dist3(xa, xb);
        if (dd < h) { h = dd; *pair = true; }
      }
  }
  return h;
}

// coordinates, dX/dx, sqrtgdiag and hmin of elements [e0, e1)
static void geometry_range(Mesh &M, size_t e0, size_t e1) {
  const long nn = M.nnodes;
  for (size_t ei = e0; ei < e1; ++ei) {
    Elm &E = M.e[ei];
    const Patch &P = M.pat[E.p];
    const Basis &bas = E.fv ? uniform_basis(E.n) : lgl_basis(E.n);
    int n = E.n;
    double dXbdX[3];
    for (int d = 0; d < 3; ++d) dXbdX[d] = 2. / (E.bbox[2 * d + 1] - E.bbox[2 * d]);
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          long g = E.off + i + n * (j + n * k);
          double X[3], x[3], D[3][3];
          const int I[3] = {i, j, k};
          for (int d = 0; d < 3; ++d)   // node_X with the element's basis looked up once
            X[d] = 0.5 * ((E.bbox[2 * d + 1] - E.bbox[2 * d]) * bas.x[I[d]] + (E.bbox[2 * d + 1] + E.bbox[2 * d]));
          patch_dXdx(P, X, x, D);
          for (int d = 0; d < 3; ++d) M.x[d * nn + g] = x[d];
          for (int a = 0; a < 3; ++a) {
            double gam = 0.0;
            for (int b = 0; b < 3; ++b) {
              M.dXdx[(3 * a + b) * nn + g] = D[a][b];
              gam += D[a][b] * D[a][b];
            }
            M.sqrtgdiag[a * nn + g] = dXbdX[a] * std::sqrt(gam);
          }
        }
    E.hmin = find_hmin(M, E, &E.hmin_pair);
  }
}

// Every element's geometry; elements are independent, so ranges of them run
// on up to 16 host threads.
static void geometry(Mesh &M) {
  long off = 0;
  for (auto &E : M.e) { E.off = off; off += (long)E.n * E.n * E.n; }
  M.nnodes = off;
  long nn = off;
  // every entry is written below, so no fill
  M.x.clear();
  M.dXdx.clear();
  M.sqrtgdiag.clear();
  M.x.resize(3 * nn);
  M.dXdx.resize(9 * nn);
  M.sqrtgdiag.resize(3 * nn);
  const size_t ne = M.e.size();
  for (auto &E : M.e) (void)(E.fv ? uniform_basis(E.n) : lgl_basis(E.n));   // fill the caches first
  for_element_ranges(ne, [&M](unsigned, size_t e0, size_t e1) { geometry_range(M, e0, e1); });
}

// ------------------------------------------------------------------------
// face tables
// ------------------------------------------------------------------------

static long node_of(const Elm &E, const int I[3]) {
  return E.off + I[0] + (long)E.n * (I[1] + (long)E.n * I[2]);
}

// face point (a,b) of face f as i,j,k
static void face_ijk(const Elm &E, int f, int a, int b, int I[3]) {
  int dir = f / 2, d1, d2;
  dirs_of(dir, &d1, &d2);
  I[dir] = (E.n - 1) * (f % 2);
  I[d1] = a;
  I[d2] = b;
}

// neighbour face corner (c1,c2 in {0,1}) closest to physical point x0
static double closest_corner(const Mesh &M, const Elm &N, int nf, const double x0[3], int *c1, int *c2) {
  double best = 1e300;
  for (int q2 = 0; q2 < 2; ++q2)
    for (int q1 = 0; q1 < 2; ++q1) {
      int I[3];
      face_ijk(N, nf, q1 * (N.n - 1), q2 * (N.n - 1), I);
      double x[3];
      node_x(M, N, I[0], I[1], I[2], x);
      double d = dist3(x, x0);
      if (d < best) { best = d; *c1 = q1; *c2 = q2; }
    }
  return best;
}

// the face tables of elements [e0, e1): each element writes only its own
// entries and reads its neighbours' boxes, orders and point types
static std::string face_tables_range(Mesh &M, size_t e0, size_t e1) {
  for (size_t ei = e0; ei < e1; ++ei) {
    Elm &E = M.e[ei];
    int n = E.n;
    for (int f = 0; f < 6; ++f) {
      E.ajmode[f] = 0; E.orient[f] = 0;
      if (E.nfnb[f] == 0) continue;
      E.ajmode[f] = 3;
      if (E.nfnb[f] == 1) {
        const Elm &N = M.e[E.fnb[f][0]];
        int nf = E.nbf[f][0];
        bool copy = false;
        if (N.fv != E.fv) {
          copy = false;
        } else if (N.p == E.p) {
          copy = (N.l == E.l && N.n == E.n);
        } else if (N.nfnb[nf] == 1 && N.n == E.n) {
          copy = true;   // if the corners coincide (checked below)
        }
        if (copy) {
          // my corners (0,0), (n-1,0), (0,n-1) against the neighbour's corners
          int c[3][2];
          bool ok = true;
          int pts[3][2] = {{0, 0}, {n - 1, 0}, {0, n - 1}};
          for (int q = 0; q < 3; ++q) {
            int I[3];
            face_ijk(E, f, pts[q][0], pts[q][1], I);
            double x0[3];
            node_x(M, E, I[0], I[1], I[2], x0);
            double d = closest_corner(M, N, nf, x0, &c[q][0], &c[q][1]);
            double mx = std::sqrt(x0[0] * x0[0] + x0[1] * x0[1] + x0[2] * x0[2]);
            if (mx > 0) d /= mx;
            if (!(d <= 1e-10)) ok = false;
          }
          if (ok) {
            // affine map of my (a,b) onto the neighbour's (q1,q2)
            int e1[2] = {c[1][0] - c[0][0], c[1][1] - c[0][1]};   // direction of my a
            int e2[2] = {c[2][0] - c[0][0], c[2][1] - c[0][1]};   // direction of my b
            int I0[3];
            face_ijk(N, nf, c[0][0] * (N.n - 1), c[0][1] * (N.n - 1), I0);
            int od = nf / 2, od1, od2;
            dirs_of(od, &od1, &od2);
            long st[3] = {1, N.n, (long)N.n * N.n};
            FaceCopy &fc = M.fcopy[6 * ei + f];
            fc.nbase = node_of(N, I0);
            fc.s1 = (int)(e1[0] * st[od1] + e1[1] * st[od2]);
            fc.s2 = (int)(e2[0] * st[od1] + e2[1] * st[od2]);
            bool same_patch = (N.p == E.p);
            E.ajmode[f] = same_patch ? 1 : 2;
            if (!same_patch) {
              // orientation code of copy_ajsurf_from_nbsurf0 (surface.c:943-1072)
              int a00 = c[0][0], a02 = c[0][1], a10 = c[1][0], a12 = c[1][1], a20 = c[2][0], a22 = c[2][1];
              int code = -1;
              if (a00 == 0 && a02 == 0) code = (a10 == 1 && a12 == 0) ? 0 : 4;
              else if (a00 == 1 && a02 == 0) code = (a10 == 0 && a12 == 0) ? 2 : 6;
              else if (a00 == 0 && a02 == 1) code = (a10 == 1 && a12 == 1) ? 1 : 5;
              else code = (a10 == 0 && a12 == 1) ? 3 : 7;
              (void)a20; (void)a22;
              E.orient[f] = code;
            }
            continue;
          }
        }
      }
      // interpolation from every neighbour: separable 1D matrices, Lagrange
      // rows from a DG neighbour, piecewise linear or parabolic ones from an
      // FV neighbour on uniform points (fv_surface_interp = linear or
      // parabolic, surface.c:1182-1245)
      int dir = f / 2, d1, d2;
      dirs_of(dir, &d1, &d2);
      std::vector<InterpSlot> &slots = M.fint[6 * ei + f];
      for (int ni = 0; ni < E.nfnb[f]; ++ni) {
        const Elm &N = M.e[E.fnb[f][ni]];
        int nf = E.nbf[f][ni];
        int od = nf / 2, od1, od2;
        dirs_of(od, &od1, &od2);
        const Basis &bn = N.fv ? uniform_basis(N.n) : lgl_basis(N.n);
        auto row = [&](int q, double x) {
          if (!N.fv) return lagrange(bn, q, x);
          // Lagrange_if_dg: full Lagrange on the uniform points into a DG
          // element (surface.c:1198-1238; its 12-point stencil covers n <= 10)
          if (M.fv_lagrange_if_dg && !E.fv) return lagrange(bn, q, x);
          return M.fv_surface_interp == 1 ? pw_parab(bn, q, x) : pw_linear(bn, q, x);
        };
        // neighbour in-plane Xb coordinates of all my face points
        std::vector<double> C0(n * n), C1(n * n);
        std::vector<char> in(n * n);
        for (int b = 0; b < n; ++b)
          for (int a = 0; a < n; ++a) {
            int I[3];
            face_ijk(E, f, a, b, I);
            double X[3], x[3], XN[3];
            node_X(E, I[0], I[1], I[2], X);
            if (N.p == E.p) {
              for (int d = 0; d < 3; ++d) XN[d] = X[d];
            } else {
              patch_xyz(M.pat[E.p], X, x);
              patch_XYZ(M.pat[N.p], x, XN);
            }
            double tol = 1e-10;
            bool inside = !(XN[od1] < N.bbox[2 * od1] - tol || XN[od1] > N.bbox[2 * od1 + 1] + tol ||
                            XN[od2] < N.bbox[2 * od2] - tol || XN[od2] > N.bbox[2 * od2 + 1] + tol);
            in[a + n * b] = inside;
            C0[a + n * b] = (2. * XN[od1] - (N.bbox[2 * od1 + 1] + N.bbox[2 * od1])) /
                            (N.bbox[2 * od1 + 1] - N.bbox[2 * od1]);
            C1[a + n * b] = (2. * XN[od2] - (N.bbox[2 * od2 + 1] + N.bbox[2 * od2])) /
                            (N.bbox[2 * od2 + 1] - N.bbox[2 * od2]);
          }
        InterpSlot s;
        s.nb = E.fnb[f][ni]; s.nbface = nf; s.m = N.n;
        int I0[3];
        face_ijk(N, nf, 0, 0, I0);
        long st[3] = {1, N.n, (long)N.n * N.n};
        s.nbase = node_of(N, I0);
        s.t1 = (int)st[od1];
        s.t2 = (int)st[od2];
        // covered rectangle
        s.a0 = n; s.a1 = -1; s.b0 = n; s.b1 = -1;
        for (int b = 0; b < n; ++b)
          for (int a = 0; a < n; ++a)
            if (in[a + n * b]) {
              s.a0 = std::min(s.a0, a); s.a1 = std::max(s.a1, a);
              s.b0 = std::min(s.b0, b); s.b1 = std::max(s.b1, b);
            }
        if (s.a1 < 0) continue;
        for (int b = s.b0; b <= s.b1; ++b)
          for (int a = s.a0; a <= s.a1; ++a)
            if (!in[a + n * b]) return "face coverage of a neighbour is not a rectangle";
        // separability: C0 depends on a only (swap=0) or b only (swap=1)
        auto depends_only_on_a = [&](const std::vector<double> &C) {
          for (int b = s.b0; b <= s.b1; ++b)
            for (int a = s.a0; a <= s.a1; ++a)
              if (std::fabs(C[a + n * b] - C[a + n * s.b0]) > 1e-11) return false;
          return true;
        };
        auto depends_only_on_b = [&](const std::vector<double> &C) {
          for (int b = s.b0; b <= s.b1; ++b)
            for (int a = s.a0; a <= s.a1; ++a)
              if (std::fabs(C[a + n * b] - C[s.a0 + n * b]) > 1e-11) return false;
          return true;
        };
        int m = N.n;
        s.Ba.assign((size_t)n * m, 0.0);
        s.Bb.assign((size_t)n * m, 0.0);
        if (depends_only_on_a(C0) && depends_only_on_b(C1)) {
          s.swap = 0;
          for (int a = s.a0; a <= s.a1; ++a)
            for (int q = 0; q < m; ++q) s.Ba[a * m + q] = row(q, C0[a + n * s.b0]);
          for (int b = s.b0; b <= s.b1; ++b)
            for (int q = 0; q < m; ++q) s.Bb[b * m + q] = row(q, C1[s.a0 + n * b]);
        } else if (depends_only_on_b(C0) && depends_only_on_a(C1)) {
          s.swap = 1;   // my a indexes the neighbour's q2 axis
          for (int a = s.a0; a <= s.a1; ++a)
            for (int q = 0; q < m; ++q) s.Ba[a * m + q] = row(q, C1[a + n * s.b0]);
          for (int b = s.b0; b <= s.b1; ++b)
            for (int q = 0; q < m; ++q) s.Bb[b * m + q] = row(q, C0[s.a0 + n * b]);
        } else {
          return "non-separable face interpolation is not supported";
        }
        slots.push_back(std::move(s));
      }
    }
  }
  return "";
}

// Every element's face tables, ranges of elements on up to 16 host threads;
// an error is the first element's in element order, as one thread would
// report it.
static std::string face_tables(Mesh &M) {
  size_t ne = M.e.size();
  M.fcopy.assign(6 * ne, FaceCopy());
  M.fint.assign(6 * ne, std::vector<InterpSlot>());
  for (int n = 2; n <= kMaxN; ++n) { (void)uniform_basis(n); (void)lgl_basis(n); }
  std::vector<std::string> err(16);
  const unsigned T = for_element_ranges(
      ne, [&M, &err](unsigned t, size_t e0, size_t e1) { err[t] = face_tables_range(M, e0, e1); });
  for (unsigned t = 0; t < T; ++t)
    if (!err[t].empty()) return err[t];
  return "";
}

Mesh mesh_skeleton(const Mesh &M) {
  Mesh S;
  S.pat = M.pat;
  S.e = M.e;
  S.nnodes = M.nnodes;
  S.maxlevel = M.maxlevel;
  S.fv_surface_interp = M.fv_surface_interp;
  S.fv_lagrange_if_dg = M.fv_lagrange_if_dg;
  return S;
}

static std::string finish(Mesh &M) {
  M.maxlevel = 0;
  for (auto &E : M.e) M.maxlevel = std::max(M.maxlevel, E.l);
  find_neighbours(M);
  geometry(M);
  return face_tables(M);
}

std::string build_box(Mesh &M, int luni, int n, double dout, const double xc[3]) {
  if (n < 2 || n > kMaxN) return "amr_n0 out of range";
  M = Mesh();
  Patch P;
  P.type = PAT_BOX;
  for (int d = 0; d < 3; ++d) {
    P.xc[d] = xc[d];
    P.bbox[2 * d] = xc[d] - dout;
    P.bbox[2 * d + 1] = xc[d] + dout;
  }
  M.pat.push_back(P);
  add_roots(M, n);
  for (int l = 0; l < luni; ++l) {
    std::vector<char> flag(M.e.size());
    for (size_t i = 0; i < M.e.size(); ++i) flag[i] = M.e[i].l < luni;
    refine(M, flag);
  }
  return finish(M);
}

std::string build_shell(Mesh &M, int luni, int n, double rin, double rout, double sph_radius,
                        int sph_levels) {
  if (n < 2 || n > kMaxN) return "amr_n0 out of range";
  if (!(rin > 0 && rout > rin)) return "need 0 < amr_Shell_rin < amr_Shell_rout";
  M = Mesh();
  for (int p = 0; p < 6; ++p) {
    Patch P;
    P.type = PAT_CUBEDSHELL;
    P.dom = p;
    P.s0 = rin; P.s1 = rout;
    double bb[6] = {0., 1., -1., 1., -1., 1.};
    std::memcpy(P.bbox, bb, sizeof(bb));
    M.pat.push_back(P);
  }
  add_roots(M, n);
  for (int l = 0; l < luni; ++l) {
    std::vector<char> flag(M.e.size());
    for (size_t i = 0; i < M.e.size(); ++i) flag[i] = M.e[i].l < luni;
    refine(M, flag);
  }
  // nested spheres about the origin (src/main/amr/refine.c:1025-1100)
  double r = sph_radius * std::pow(2., sph_levels - 1);
  for (int i = 0; i < sph_levels; ++i) {
    std::vector<char> flag(M.e.size());
    for (size_t k = 0; k < M.e.size(); ++k) {
      const Elm &E = M.e[k];
      double Xc[3], xcn[3];
      for (int d = 0; d < 3; ++d) Xc[d] = 0.5 * (E.bbox[2 * d + 1] + E.bbox[2 * d]);
      patch_xyz(M.pat[E.p], Xc, xcn);
      double rr = std::sqrt(xcn[0] * xcn[0] + xcn[1] * xcn[1] + xcn[2] * xcn[2]);
      flag[k] = rr <= r;
    }
    refine(M, flag);
    r *= 0.5;
  }
  return finish(M);
}

std::string import_mesh(Mesh &M, const std::vector<ImportPatch> &pat, const std::vector<ImportElm> &elm) {
  Mesh N;
  for (const ImportPatch &q : pat) {
    Patch P;
    if (q.type != PAT_BOX && q.type != PAT_CUBEDSHELL) return "patch type must be box (0) or CubedShell (1)";
    P.type = q.type;
    P.dom = q.dom;
    P.s0 = q.s0;
    P.s1 = q.s1;
    for (int f = 0; f < 6; ++f) P.bbox[f] = q.bbox[f];
    for (int d = 0; d < 3; ++d) P.xc[d] = q.xc[d];
    if (P.type == PAT_CUBEDSHELL) {
      if (q.dom < 0 || q.dom > 5) return "cubed-sphere dom must be 0..5";
      if (!(q.s0 > 0 && q.s1 > q.s0)) return "cubed-sphere radii need 0 < s0 < s1";
      const double bb[6] = {0., 1., -1., 1., -1., 1.};
      for (int f = 0; f < 6; ++f)
        if (q.bbox[f] != bb[f]) return "cubed-sphere patch box must be lam [0,1], A, B [-1,1]";
    } else {
      for (int d = 0; d < 3; ++d)
        if (!(q.bbox[2 * d + 1] > q.bbox[2 * d])) return "empty patch box";
    }
    N.pat.push_back(P);
  }
  for (size_t i = 0; i < elm.size(); ++i) {
    const ImportElm &q = elm[i];
    if (q.p < 0 || q.p >= (int)N.pat.size()) return "element " + std::to_string(i) + ": no such patch";
    if (q.l < 0 || q.l >= (int)sizeof(q.path)) return "element " + std::to_string(i) + ": level out of range";
    if (q.n < 2 || q.n > kMaxN) return "element " + std::to_string(i) + ": n out of range";
    Elm E;
    E.p = q.p; E.l = q.l; E.n = q.n;
    for (int d = 0; d < 3; ++d) E.ic[d] = 0;
    for (int k = 0; k < q.l; ++k) {
      if (q.path[k] > 7) return "element " + std::to_string(i) + ": octant digit above 7";
      E.path[k] = q.path[k];
      for (int d = 0; d < 3; ++d) E.ic[d] = 2 * E.ic[d] + ((q.path[k] >> d) & 1);
    }
    elm_bbox(N.pat[E.p], E);
    for (int f = 0; f < 6; ++f)
      if (E.bbox[f] != q.bbox[f]) return "element " + std::to_string(i) + ": box differs from the location's";
    N.e.push_back(E);
  }
  // a forest in nmesh's order: patches in order, depth-first octants inside,
  // each patch covered exactly once
  for (size_t i = 1; i < N.e.size(); ++i) {
    const Elm &A = N.e[i - 1], &B = N.e[i];
    bool ok = A.p < B.p;
    if (A.p == B.p) {
      // B follows A in depth-first order: at the first level where they
      // differ, B's octant is the larger
      int k = 0;
      while (k < A.l && k < B.l && A.path[k] == B.path[k]) ++k;
      ok = k < A.l && k < B.l && B.path[k] > A.path[k];
    }
    if (!ok) return "elements " + std::to_string(i - 1) + " and " + std::to_string(i) + " are not in depth-first order";
  }
  // volumes in units of the deepest possible octant (8^-23 of a patch)
  std::vector<unsigned __int128> vol(N.pat.size(), 0);
  for (const Elm &E : N.e) vol[E.p] += (unsigned __int128)1 << (3 * (23 - E.l));
  for (size_t p = 0; p < vol.size(); ++p)
    if (vol[p] != ((unsigned __int128)1 << 69)) return "the elements of patch " + std::to_string(p) + " do not cover it";
  std::string err = finish(N);
  if (!err.empty()) return err;
  for (size_t i = 0; i < N.e.size(); ++i)
    for (int f = 0; f < 6; ++f) {
      const Elm &E = N.e[i];
      const ImportElm &q = elm[i];
      bool same = E.nfnb[f] == q.nfnb[f];
      for (int k = 0; same && k < E.nfnb[f]; ++k) same = E.fnb[f][k] == q.fnb[f][k] && E.nbf[f][k] == q.nbf[f][k];
      if (!same)
        return "element " + std::to_string(i) + " face " + std::to_string(f) +
               ": neighbour list differs from the one built here";
    }
  M = std::move(N);
  return "";
}

std::string set_elm_n(Mesh &M, const std::vector<int> &n) {
  if (n.size() != M.e.size()) return "need one n per element";
  for (size_t i = 0; i < n.size(); ++i) {
    if (n[i] < 2 || n[i] > kMaxN) return "n out of range";
    M.e[i].n = n[i];
  }
  geometry(M);
  return face_tables(M);
}

// the run of 8 elements from e is the 8 children of one parent in child
// order (amr_elms_are_siblings, child0 first)
static bool are_siblings(const Mesh &M, size_t e) {
  if (e + 8 > M.e.size()) return false;
  const Elm &E0 = M.e[e];
  if (E0.l < 1) return false;
  for (int c = 0; c < 8; ++c) {
    const Elm &E = M.e[e + c];
    if (E.p != E0.p || E.l != E0.l || E.path[E0.l - 1] != c) return false;
    if (std::memcmp(E.path, E0.path, E0.l - 1) != 0) return false;
  }
  return true;
}

std::string hadapt(Mesh &M, const std::vector<int> &flag) {
  if (flag.size() != M.e.size()) return "need one flag per element";
  // refinement: children of the same n in place of the parent
  std::vector<char> ref(M.e.size());
  std::vector<char> cf;
  for (size_t i = 0; i < M.e.size(); ++i) {
    ref[i] = flag[i] > 0;
    if (ref[i] && M.e[i].l + 1 >= (int)sizeof(M.e[i].path)) return "refinement level too deep";
  }
  for (size_t i = 0; i < M.e.size(); ++i) {
    if (ref[i]) cf.insert(cf.end(), 8, 0);
    else cf.push_back(flag[i] < 0);
  }
  refine(M, ref);
  // coarsening: 8 flagged siblings in a row become their parent, with the
  // largest sibling n (refine_sib_with_max_np_pt_typ)
  std::vector<Elm> ne;
  ne.reserve(M.e.size());
  for (size_t e = 0; e < M.e.size();) {
    bool all = are_siblings(M, e);
    for (int c = 0; all && c < 8; ++c) all = cf[e + c] != 0;
    if (!all) { ne.push_back(M.e[e]); ++e; continue; }
    const Elm &C0 = M.e[e];
    Elm P;
    P.p = C0.p;
    P.l = C0.l - 1;
    std::memcpy(P.path, C0.path, sizeof(P.path));
    P.path[P.l] = 0;
    for (int d = 0; d < 3; ++d) P.ic[d] = C0.ic[d] >> 1;
    P.n = 0;
    for (int c = 0; c < 8; ++c) P.n = std::max(P.n, M.e[e + c].n);
    P.fv = C0.fv;
    elm_bbox(M.pat[P.p], P);
    ne.push_back(P);
    e += 8;
  }
  M.e = std::move(ne);
  return finish(M);
}

std::string set_fv_elems(Mesh &M, const std::vector<int> &fv) {
  if (fv.size() != M.e.size()) return "need one flag per element";
  for (size_t i = 0; i < fv.size(); ++i) M.e[i].fv = fv[i] != 0;
  geometry(M);
  return face_tables(M);
}

std::string set_fv_patches(Mesh &M, const std::vector<int> &patches) {
  for (auto &E : M.e) {
    E.fv = 0;
    for (int p : patches)
      if (E.p == p) E.fv = 1;
  }
  geometry(M);
  return face_tables(M);
}

std::string set_fv_surface_interp(Mesh &M, int mode, bool lagrange_if_dg) {
  if (mode != 0 && mode != 1) return "fv_surface_interp must be linear or parabolic";
  M.fv_surface_interp = mode;
  M.fv_lagrange_if_dg = lagrange_if_dg;
  return face_tables(M);
}

// (hmin_is_in_uniform_direction and adapt_node_dt_and_mesh_dt,
// src/main/coordinates/utilities.c:1218-1304: the pair at hmin differs in
// one index, and an FV element is uniform along every direction)
double mesh_dt(const Mesh &M, double dtfac, double uniform_dtfac) {
  double dt = 0.1 * 1.7976931348623157e308;
  for (auto &E : M.e) dt = std::min(dt, (E.fv && E.hmin_pair ? uniform_dtfac : dtfac) * E.hmin);
  return dt;
}

}  // namespace nmx
