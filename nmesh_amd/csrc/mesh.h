// mesh.h -- host-side mesh, geometry and face tables of libnmx.
//
// The mesh is nmesh's forest of patches whose leaves (elements) are kept in
// depth-first octant order, exactly as the reference's element list
// (src/main/amr/storage.c:586-621): children replace their parent in place,
// child index c has bit d set for the upper half in direction d
// (src/main/amr/connections.c:1175-1206).
#pragma once
#include <cstdint>
#include <exception>
#include <thread>
#include <memory>
#include <new>
#include <utility>
#include <string>
#include <vector>

namespace nmx {

constexpr int kMaxN = 16;     // points per direction supported by the host code
constexpr int kMaxNb = 8;     // face neighbours per face

enum PatchType { PAT_BOX = 0, PAT_CUBEDSHELL = 1 };

struct Patch {
  int type = PAT_BOX;
  int dom = 0;                // cubed-sphere domain 0..5 (dir = dom/2)
  double s0 = 0, s1 = 0;      // inner and outer radius (sigma) of a CubedShell
  double xc[3] = {0, 0, 0};
  double bbox[6] = {0, 0, 0, 0, 0, 0};
  int bound[6] = {1, 1, 1, 1, 1, 1};
};

// One element.  path[l] is the child index taken at level l+1.
struct Elm {
  int p = 0, l = 0;
  uint8_t path[24] = {0};
  uint32_t ic[3] = {0, 0, 0};   // integer position at level l
  double bbox[6] = {0, 0, 0, 0, 0, 0};
  int n = 2;
  int fv = 0;                   // 1: FV element on uniform points (fv_divf)
  long off = 0;                 // first node
  int nfnb[6] = {0, 0, 0, 0, 0, 0};
  int fnb[6][kMaxNb];
  int nbf[6][kMaxNb];
  int ajmode[6] = {0, 0, 0, 0, 0, 0};
  int orient[6] = {0, 0, 0, 0, 0, 0};
  double hmin = 0;
  bool hmin_pair = false;   // hmin is a distance between two points (not the diagonal)
};

struct Basis {
  int n = 0;
  double x[kMaxN], w[kMaxN], wb[kMaxN];   // LGL nodes, quadrature, barycentric
  double D[kMaxN * kMaxN];                // D[i*n + l] = l_l'(x_i)
};
const Basis &lgl_basis(int n);
void filter_matrix(int n, double alp, double s, int dn, double *F);
// normalized Legendre analysis A[i*n + j] = w_j P_i(x_j)/sqrt(c_i) and
// synthesis S[j*n + i] = P_i(x_j)/sqrt(c_i) on the n LGL nodes, c_N = 2/N
// (LGL_AT_ST_matrices, src/main/basis/Legendre.c:257-307)
void legendre_analysis_synthesis(int n, double *A, double *S);
// uniform points incl. both faces (x only)
const Basis &uniform_basis(int n);

// conforming-face gather: neighbour node = nbase + a*s1 + b*s2 for my face
// point (a,b); a runs along Dir1_norm(dir), b along Dir2_norm(dir)
struct FaceCopy {
  long nbase = -1;
  int s1 = 0, s2 = 0;
};

// one neighbour's contribution to an interpolated face
struct InterpSlot {
  int nb = -1, nbface = 0, m = 0;   // neighbour element, its face, its n
  long nbase = 0;                   // neighbour plane: node = nbase + q1*t1 + q2*t2
  int t1 = 0, t2 = 0;
  int swap = 0;                     // 1: my a-axis maps to the neighbour's q2 axis
  int a0 = 0, a1 = 0, b0 = 0, b1 = 0;   // my covered point ranges [a0,a1]x[b0,b1]
  std::vector<double> Ba, Bb;       // Ba[a*m+q]: l_q at my a-th point (covered rows)
};

// An allocator whose resize leaves new doubles uninitialised: the node
// arrays are written in full by the threads that build them, so a serial
// zero fill (and its page faults) before them is wasted.
template <class T>
struct NoInitAlloc : std::allocator<T> {
  template <class U>
  struct rebind { using other = NoInitAlloc<U>; };
  NoInitAlloc() = default;
  template <class U>
  NoInitAlloc(const NoInitAlloc<U> &) {}
  template <class U>
  void construct(U *p) { ::new (static_cast<void *>(p)) U; }
  template <class U, class... A>
  void construct(U *p, A &&...a) { ::new (static_cast<void *>(p)) U(std::forward<A>(a)...); }
};
using NodeArray = std::vector<double, NoInitAlloc<double>>;

struct Mesh {
  std::vector<Patch> pat;
  std::vector<Elm> e;
  long nnodes = 0;
  NodeArray x;                      // x[d*nnodes + node]
  NodeArray dXdx;                   // dXdx[(3a+b)*nnodes + node]
  NodeArray sqrtgdiag;              // sqrtgdiag[a*nnodes + node]
  std::vector<FaceCopy> fcopy;      // 6 per element
  std::vector<std::vector<InterpSlot>> fint;  // 6 per element
  int maxlevel = 0;
  int fv_surface_interp = 0;        // rows from FV neighbours: 0 piecewise linear, 1 parabolic
  bool fv_lagrange_if_dg = false;   // ... but Lagrange rows into DG elements
};

// fn(t, e0, e1) for the T consecutive ranges t of [0, ne), on up to 16 host
// threads (one range when ne is small); a range whose thread cannot be
// started runs on the calling thread.  Returns T.
template <class F>
inline unsigned for_element_ranges(size_t ne, F fn) {
  unsigned T = std::thread::hardware_concurrency();
  T = T < 1 ? 1 : (T > 16 ? 16 : T);
  if (ne < 64 * (size_t)T) T = 1;
  std::vector<std::thread> th;
  for (unsigned t = 1; t < T; ++t) {
    const size_t e0 = ne * t / T, e1 = ne * (t + 1) / T;
    bool started = false;
    try {
      th.emplace_back(fn, t, e0, e1);
      started = true;
    } catch (const std::exception &) {
    }
    if (!started) fn(t, e0, e1);
  }
  fn(0u, (size_t)0, ne / T);
  for (auto &x : th) x.join();
  return T;
}

// A copy of M's patches, elements and settings without the node arrays
// (x, dXdx, sqrtgdiag) and face tables (fcopy, fint), which every rebuilder
// below (set_elm_n, hadapt, set_fv_elems, set_fv_patches) recomputes: the
// base of a rebuild, without copying the large per-node arrays.
Mesh mesh_skeleton(const Mesh &M);

// builders (return "" on success, else an error message)
std::string build_box(Mesh &M, int luni, int n, double dout, const double xc[3]);
std::string build_shell(Mesh &M, int luni, int n, double rin, double rout,
                        double sph_radius, int sph_levels);
std::string set_elm_n(Mesh &M, const std::vector<int> &n);
// h-adaptation (hp_refine_elms_if_rflag with PARENT_n, refine.c:174-276;
// remove_elms_if_rflag, refine.c:359-506): flag > 0 replaces an element by
// its 8 children of the same n, in place; 8 siblings in a row that all have
// flag < 0 become their parent, with the largest sibling n.  Tables and
// geometry are rebuilt.
std::string hadapt(Mesh &M, const std::vector<int> &flag);
// FV on uniform points where fv[e] is set, DG on LGL points elsewhere;
// geometry and tables rebuilt
std::string set_fv_elems(Mesh &M, const std::vector<int> &fv);
// put every element of the listed patches in FV mode on uniform points
// (amr_uniform_p, amr_fv_p; src/main/amr/mesh.c:40-80) and rebuild
std::string set_fv_patches(Mesh &M, const std::vector<int> &patches);

// fv_surface_interp (0 linear, 1 parabolic) for faces taken from FV
// neighbours; tables rebuilt
std::string set_fv_surface_interp(Mesh &M, int mode, bool lagrange_if_dg);

// A mesh handed over from nmesh's own tables (nmesh_amr.h:87-102): patches
// (type, cubed-sphere dom, radii s0 s1, box, centre), elements in eid order
// (patch, level, child octant per level, n), and the caller's boxes and face
// neighbour lists.  The forest is rebuilt from the element locations and
// every table the library derives (element boxes bit for bit, neighbour
// counts, neighbours in nmesh's order, their faces) must equal the caller's;
// the first difference is returned as the error.
struct ImportPatch {
  int type = PAT_BOX, dom = 0;
  double s0 = 0, s1 = 0, bbox[6] = {0, 0, 0, 0, 0, 0}, xc[3] = {0, 0, 0};
};
struct ImportElm {
  int p = 0, l = 0, n = 2;
  uint8_t path[24] = {0};
  double bbox[6] = {0, 0, 0, 0, 0, 0};
  int nfnb[6] = {0, 0, 0, 0, 0, 0};
  int fnb[6][kMaxNb], nbf[6][kMaxNb];
};
std::string import_mesh(Mesh &M, const std::vector<ImportPatch> &pat, const std::vector<ImportElm> &elm);

// dt = min over elements of dtfac hmin, or uniform_dtfac hmin for FV
// elements whose hmin lies between two (uniform) points
double mesh_dt(const Mesh &M, double dtfac, double uniform_dtfac);

// maps
void patch_xyz(const Patch &P, const double X[3], double x[3]);
void patch_XYZ(const Patch &P, const double x[3], double X[3]);
void patch_dXdx(const Patch &P, const double X[3], double x[3], double M[3][3]);

}  // namespace nmx
