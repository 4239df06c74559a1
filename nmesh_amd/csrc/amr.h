// amr.h -- launch interface of the batched (order-bucketed) DG stage kernel.
#pragma once
#include <hip/hip_runtime.h>

namespace nmx {

// face interpolation job, kJobInts ints per (face, neighbour):
//   [0] n (my points per direction), [1] m (neighbour's), [2] offset of the
//   job's region in the face buffer ([v][n*n]), [3] 1, [4] base, [5] t1,
//   [6] t2 (neighbour plane node = base + q1*t1 + q2*t2 inside the
//   neighbour), [7] swap, [8..11] a0,a1,b0,b1 covered point ranges, [12]
//   offset of the rows for my a, [13] for my b, in Btab (rows of m doubles),
//   [14..15] the neighbour's first state double (a long)
constexpr int kJobInts = 16;
// LDS of the face interpolation kernel, in doubles: first-pass results
// (4 * covered rows * m) and the rows of each pass (covered rows or columns
// * m); covered rows and columns <= kMaxFaceN.
constexpr int kFaceT1 = 4 * 10 * 10, kFaceR1 = 10 * 10;
constexpr int kMaxFaceN = 10;
// Faces with several neighbours (a coarse face against finer ones): one
// region per neighbour; mdesc[foff] = ns, then per neighbour {region, a0,
// a1, b0, b1}.
constexpr int kMaxFaceNb = 4;

// One element of an order bucket, everything its stage needs besides the
// state: read with wave-uniform loads, so no load of the kernel waits on
// another table lookup first.
struct SlotRec {
  long noff;            // first node of the element (state at 5 * noff)
  double dXbdX[3];      // 2 / box widths
  double bb[6];         // element box in patch coordinates
  double s0, s1;        // CubedShell radii of the patch
  long foff[6];         // kind 1: the neighbour's first state double + base
                        // node; kind 2: the face's values in the face buffer;
                        // kind 3: the face's entry in mdesc
  int fkind[6];         // 0 outer boundary, 1 conforming copy, 2 interpolated
                        // from one neighbour, 3 from several
  int fs1[6], fs2[6];   // kind 1: neighbour node = base + a*fs1 + b*fs2
  int dom;              // CubedShell patch direction and sign
  int n;                // points per direction
  int jfirst, jcount;   // batched DG path: the element's interpolation jobs
  int pad[6];
};
static_assert(sizeof(SlotRec) == 256, "SlotRec is 256 bytes");

// One (interpolated face, neighbour) job of the batched DG path, done by the
// receiving element's workgroup in its load phase.  An element's jobs are
// grouped in rounds; a round holds whole faces, its lines (4 m per job, one
// thread each) fit the workgroup and its first-pass results and rows fit the
// stage buffer of the own fields, which is free until those are written.
struct RJob {
  long nbase;           // the neighbour's first state double + the plane's base node
  int t1, t2;           // neighbour plane node = base + q1*t1 + q2*t2
  int pk;               // m | f << 4 | sw << 7 | a0 << 8 | a1 << 12 | b0 << 16 | b1 << 20
                        // | (first job of its face) << 24
  int offA, offB;       // rows (m doubles each) for my a and my b points in Btab
  int line0;            // first of this job's 4 m lines in its round
  int toff, r1off, r2off;   // first-pass results, pass-1 and pass-2 rows in the buffer
  int nround;           // jobs in the round that starts here (0 inside a round)
};
static_assert(sizeof(RJob) == 48, "RJob is 48 bytes");
constexpr int kBatchLines = 256;      // lines per round
constexpr int kBatchRows = 1024;      // row values (both passes) per round
constexpr int kBatchScratch = 4000;   // the own-field buffer at n = 10, in doubles
constexpr int kBatchMaxN = 10;
// doubles a round of an element of order n may use: the own-field buffer,
// at least enough for one face of 4 neighbours of up to 10 points
__host__ __device__ constexpr int batch_scratch(int n) {
  return 4 * n * n * n > 240 * n ? (4 * n * n * n < kBatchScratch ? 4 * n * n * n : kBatchScratch) : 240 * n;
}

// One launch per RK stage for every DG element of a batched mesh: items are
// elements in 8 lists (one per XCD, contiguous in element order, by
// decreasing n inside); workgroup q of XCD x takes items lo[x] + q, + g8, ...
struct BatchArgs {
  const double *win = nullptr, *U = nullptr, *Ain = nullptr;
  double *Aout = nullptr, *Wout = nullptr, *R = nullptr;
  const SlotRec *rec = nullptr;
  const RJob *jobs = nullptr;
  const double *Btab = nullptr;
  const double *dXdx = nullptr;   // element blocks of 9*np
  const double *xyz = nullptr;    // element blocks of 3*np
  const double *Dall = nullptr;   // D of order n at batch_doff(n)
  const double *xball = nullptr;  // LGL nodes of order n at batch_xoff(n)
  const double *wend_inv = nullptr;   // [n]: 1 / (end weight of the LGL quadrature)
  const int *xlo = nullptr;       // [9]: items of XCD x in [xlo[x], xlo[x + 1])
  const double *tdev = nullptr;
  int tidx = 0;
  double cA = 0, cW = 0;
  double k[3] = {0, 0, 0};
  double om = 0;
  int wmode = 0, numflux = 0, profile = 0;
  int nt = 0;                     // streaming hints of the epilogue (NodeOut::nt)
};
// offsets of order n (2 <= n <= 10) in the concatenated D and node tables
__host__ __device__ constexpr int batch_doff(int n) { return (n - 1) * n * (2 * n - 1) / 6 - 1; }
__host__ __device__ constexpr int batch_xoff(int n) { return n * (n - 1) / 2 - 1; }
// grid: a multiple of 8 (XCD lists)
hipError_t launch_dg_batch(int geom, const BatchArgs &a, int grid, hipStream_t s);
// one order: items [a.xlo[0], a.xlo[1]); tile 1 or 2: nodes per thread, 3
// or 4: 1 or 2 nodes per thread with the epilogue's operands requested at
// each element's start; grid a multiple of 8
hipError_t launch_dg_bucket(int np, int geom, const BatchArgs &a, int grid, int tile, hipStream_t s);

struct AmrArgs {
  const double *win = nullptr, *U = nullptr, *Ain = nullptr;
  double *Aout = nullptr, *Wout = nullptr, *R = nullptr;
  const SlotRec *rec = nullptr;   // this bucket's elements
  int nlist = 0;
  const double *aj = nullptr;     // interpolated face values
  const int *mdesc = nullptr;     // faces with several neighbours
  const double *dXdx = nullptr;   // element blocks of 9*np
  const double *xyz = nullptr;    // element blocks of 3*np
  const double *D = nullptr;      // this bucket's D[i*n + l]
  const double *xb = nullptr;     // this bucket's LGL nodes
  const double *tdev = nullptr;
  int tidx = 0;
  double cA = 0, cW = 0, wend_inv = 0;
  double k[3] = {0, 0, 0};
  double om = 0;
  int wmode = 0, numflux = 0, profile = 0, xcd_map = 0;
  int nt = 0;                     // streaming hints of the epilogue (NodeOut::nt)
  int fv_mid = 0;                 // FV buckets: metric and normal at the midpoints
  // FV buckets: fv_divf_extrap_s1, _s2, _opt (ex_on: s1 > 0 or s2 < DBL_MAX)
  double ex_s1 = 0, ex_s2 = 1.7976931348623157e308;
  int ex_opt = 0, ex_on = 0;
  // fv_rec's end midpoints: recend 0 (_2 modes), 1 (_1), 2 (WENOm3/5/Z with
  // fv_rec_WENOm_s1/_s2/_opt; rec_on: s1 > 0 or s2 < DBL_MAX)
  int recend = 0, rec_opt = 0, rec_on = 0;
  int rec_first = 0;               // fv_rec = 1 (first order)
  double rec_s1 = 0, rec_s2 = 1.7976931348623157e308;
};

struct FaceArgs {
  const double *win = nullptr;
  const int *jobs = nullptr;      // [job][kJobInts]
  int nfaces = 0;                 // jobs
  const double *Btab = nullptr;
  double *aj = nullptr;
};

// The exponential filter, in place, over a list of elements of one order.
struct FilterArgs {
  double *u = nullptr;            // register filtered in place
  const SlotRec *rec = nullptr;
  int nlist = 0;
  const double *F = nullptr;      // F[i*np + l]: the filter along one direction
  const double *xb = nullptr;     // LGL nodes of this order
  const double *dXdx = nullptr;   // element blocks of 9*np (stored geometry)
  int varmask = 0;                // bit v: pi, cx, cy, cz, phi
  int jpow = 0;                   // filter u J^jpow, J = 1/det(dXb/dx)
};

// ---------------------------------------------------------------------------
// Discrete adjoint of the batched path.  With G = s_f A(n_f) lam at every
// face point of every element (A the symmetric flux Jacobian, s_f the lift
// scale), the transpose of the face gathers hands each element, at the
// points of its face planes, the G of the points that read them:
//   copy faces: the partner's G at the same point;
//   interpolated faces: per (receiving face, neighbour) job the transposed
//   tensor interpolation T = R1^T (w G) R2 on the neighbour's plane, w the
//   averaging weight (1 / count where several neighbours hold a point),
//   summed over the jobs that read the plane in a fixed order.
// ---------------------------------------------------------------------------

// one face of one element, as the sender of adjoint face terms
struct AdjFace {
  int kind;             // 0 outer boundary, 1 copy, 2 interpolated (one or several neighbours)
  int nb, g;            // kind 1: the partner element and its face
  int base, s1, s2;     // kind 1: partner node = base + a*s1 + b*s2 (element-local)
  int jfirst, jcount;   // kind 2: the jobs that read this face's plane, in AdjAmrArgs::sjobs
};
static_assert(sizeof(AdjFace) == 32, "AdjFace is 32 bytes");

// one (receiving face, neighbour) interpolation job, transposed
struct AdjJob {
  int e, g;             // receiving element and face
  int n, m;             // points per direction of the receiver and of the neighbour
  int sw, a0, a1, b0, b1;   // the forward job's axis swap and covered points
  int offA, offB;       // rows for the receiver's a and b points in Btab (m doubles each)
  int md;               // the receiving face's mdesc entry (several neighbours), -1 for one
  int toff;             // this job's transposed plane [4][m*m] in Tbuf
  // neighbour face coordinates (a, b) -> plane indices: q1 = A1 (a - A0) +
  // B1 (b - B0), q2 = A2 (a - A0) + B2 (b - B0)
  int A0, B0, A1, A2, B1, B2;
  int pad;
};
static_assert(sizeof(AdjJob) == 80, "AdjJob is 80 bytes");

struct AdjAmrArgs {
  const double *lam = nullptr;
  const double *Abase = nullptr, *Wbase = nullptr;
  double *Aout = nullptr, *Wout = nullptr, *R = nullptr;
  double cA = 0, cU = 0, cW = 0;
  const int *list = nullptr;      // this launch's elements
  int nlist = 0;
  const long *eoff = nullptr;     // [e]: first node
  const int *en = nullptr;        // [e]: points per direction
  const double *dXbdX = nullptr;  // [e][3]
  const double *egeo = nullptr;   // [e][8]: box in patch coordinates, s0, s1
  const int *edom = nullptr;      // [e]
  const double *dXdx = nullptr;   // element blocks of 9*np (stored geometry)
  const double *Dall = nullptr;   // D of order n at batch_doff(n)
  const double *xball = nullptr;  // LGL nodes of order n at batch_xoff(n)
  const double *wend_inv = nullptr;   // [n]
  const AdjFace *af = nullptr;    // [e][6]
  const long *goff = nullptr;     // [e]: the element's G planes [6][4][n*n] in Gbuf
  double *G = nullptr;
  const AdjJob *jobs = nullptr;
  int njobs = 0;
  const int *sjobs = nullptr;     // job indices, grouped by the plane they read
  const int *mdesc = nullptr;
  const double *Btab = nullptr;
  double *T = nullptr;
  int numflux = 0;
};

// G at every face point of the listed elements (any orders)
hipError_t launch_adj_gflux(int geom, const AdjAmrArgs &a, hipStream_t s);
// the transposed interpolation jobs into T
hipError_t launch_adj_jobs(const AdjAmrArgs &a, hipStream_t s);
// the transposed RHS and adjoint stage combination of one order bucket
hipError_t launch_dg_adjoint_amr(int np, int geom, const AdjAmrArgs &a, int grid, hipStream_t s);

// div: nodes per thread (1, 2 or 4)
hipError_t launch_dg_amr(int np, int geom, const AmrArgs &a, int grid, int div, hipStream_t s);
hipError_t launch_face_interp(const FaceArgs &a, hipStream_t s);
hipError_t launch_filter(int np, int geom, const FilterArgs &a, hipStream_t s);
// an FV order bucket (fv_divf on uniform points): rec 5 = WENOm5_2, 7 =
// WENOmZ_2, 3 = WENOm3_2; geom 0 Cartesian, 2 CubedShell; wc[np] the
// distances between the midpoints around each point; a.xb uniform points
hipError_t launch_fvelm_amr(int np, int rec, int geom, const AmrArgs &a, const double *wc, int grid,
                            hipStream_t s);

}  // namespace nmx
