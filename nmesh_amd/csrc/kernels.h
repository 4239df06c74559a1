// kernels.h -- launch interface between the libnmx host code and its kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nmx {

// Arguments of one fused DG RHS + RK stage launch over elements [e0, e1).
// Output combinations (each optional):
//   R    = r                                      (RHS only)
//   Aout = (Ain ? Ain : win) + cA * r              (RK accumulator "u")
//   Wout = U + cW * r                   (wmode 1)  (next stage input "w")
//   Wout = 0.75 U + 0.25 (win + cW r)   (wmode 2)  (sspRK3 third stage input)
struct StageArgs {
  const double *win = nullptr;   // stage input state
  const double *U = nullptr;     // state at the start of the step (u_p)
  const double *Ain = nullptr;
  double *Aout = nullptr;
  double *Wout = nullptr;
  double *R = nullptr;
  const double *dXdx = nullptr;  // [e][9][np]
  const double *dXbdX = nullptr; // [e][3]
  const double *xyz = nullptr;   // [e][3][np]
  const int4 *face = nullptr;    // [e][6]: {nb elm (-1 none), local base, s1, s2}
  const double *D = nullptr;     // D[i*np + l]
  const double *tdev = nullptr;  // stage times on the device; t = tdev[tidx]
  const double *egeo = nullptr;  // [e][8]: patch bbox of the element (6), s0, s1
  const int *edom = nullptr;     // [e]: cubed-sphere domain of the element's patch
  const double *xb = nullptr;    // LGL nodes on [-1,1]
  int tidx = 0;
  int xcd_map = 0;               // 1: contiguous element ranges per XCD
  double cA = 0, cW = 0, wend_inv = 0;
  double k[3] = {0, 0, 0};
  double om = 0;
  int e0 = 0, e1 = 0;
  int wmode = 0, numflux = 0, profile = 0;
  int nt = 0;                    // streaming hints of the epilogue (NodeOut::nt)
  int fv_mid = 0;                // FV elements: metric and normal at the midpoints
  // FV elements: fv_divf_extrap_s1, _s2, _opt (ex_on: s1 > 0 or s2 < DBL_MAX)
  double ex_s1 = 0, ex_s2 = 1.7976931348623157e308;
  int ex_opt = 0, ex_on = 0;
  // fv_rec's end midpoints: recend 0 (_2 modes), 1 (_1), 2 (WENOm3/5/Z with
  // fv_rec_WENOm_s1/_s2/_opt; rec_on: s1 > 0 or s2 < DBL_MAX)
  int recend = 0, rec_opt = 0, rec_on = 0;
  int rec_first = 0;               // fv_rec = 1 (first order)
  double rec_s1 = 0, rec_s2 = 1.7976931348623157e308;
};

// Arguments of one fused adjoint RHS + adjoint RK stage launch: y = L^T lam
// (L the Jacobian of the scalarwave1 DG RHS with respect to u), then
//   R    = y
//   Aout = Abase + cA * y
//   Wout = cU * Wbase + cW * y
// (each optional).  Wout must not alias lam: neighbours read lam.
struct AdjArgs {
  const double *lam = nullptr;
  const double *Abase = nullptr, *Wbase = nullptr;
  double *Aout = nullptr, *Wout = nullptr, *R = nullptr;
  double cA = 0, cU = 0, cW = 0;
  const double *dXdx = nullptr, *dXbdX = nullptr, *egeo = nullptr, *xb = nullptr, *D = nullptr;
  const int *edom = nullptr;
  const int4 *face = nullptr;    // as StageArgs::face
  const int *nbface = nullptr;   // [e][6]: the neighbour's face touching face f
  double wend_inv = 0;
  int numflux = 0, xcd_map = 0, e0 = 0, e1 = 0;
};

// the mesh's dt from the device's node coordinates (update_elm_eid_dt with
// find_hmin): per-workgroup minima in part[grid], the mesh's in out[0]
struct DtArgs {
  const double *xyz = nullptr;   // element blocks [e][3][np]
  const long *eoff = nullptr;
  const int *en = nullptr;
  const int *efv = nullptr;      // [e]: 1 FV element
  double dtfac = 0, uniform_dtfac = 0;
  double *part = nullptr, *out = nullptr;
  int nelem = 0;
};
int dt_grid(int nelem);
hipError_t launch_dt(const DtArgs &a, int grid, hipStream_t s);

// geom: 0 Cartesian, 1 stored dXdx, 2 CubedShell recomputed; nt from dg_stage_threads
hipError_t launch_dg_stage(int np, int geom, const StageArgs &a, int grid, int nt, hipStream_t s);
int dg_stage_threads(int np, int div);   // div: nodes per thread 1, 2 or 4
hipError_t launch_dg_adjoint(int np, int geom, const AdjArgs &a, int grid, hipStream_t s);
// scalarwave1 on FV elements (fv_divf); rec: the interior scheme, 5 WENO5,
// 7 WENO-Z, 9 WENOT, 3 WENOm3 (the end midpoints by StageArgs::recend);
// wc[np]: distances between the midpoints around each uniform point
// geom: 0 Cartesian, 2 CubedShell (metric recomputed at nodes and midpoints
// from egeo, edom and the uniform points in xb)
hipError_t launch_fvelm_stage(int np, int rec, int geom, const StageArgs &a, const double *wc, int grid,
                              hipStream_t s);

// The checkpoint's "vars" CRC-64 of a register (element blocks [e][5][np_e]).
struct CrcArgs {
  const double *u = nullptr;
  const long *eoff = nullptr;      // first node per element
  const int *en = nullptr;         // points per direction per element
  int nelem = 0;
  long nnodes = 0;
  const uint64_t *tab = nullptr;   // [8][256] slicing-by-8 tables: S(w) = w x^64 mod P
  const uint64_t *vtab = nullptr;  // [8][256]: (byte b at position j) x^4096 mod P
  const uint64_t *klane = nullptr; // [n - 2][64]: x^(64 (5 n^3 - 1 - lane l's last word)) mod P
  const uint64_t *aw = nullptr;    // [n - 2]: (~0) x^(64 * 5 n^3) mod P
  const uint64_t *ypow = nullptr;  // [4][256]: x^(320 * d * 256^j) mod P
  uint64_t *part = nullptr;        // [nelem] element contributions
};
// out[0] = CRC (device memory)
hipError_t launch_vars_crc(const CrcArgs &a, uint64_t *out, hipStream_t s);

// The static scratch of nmx_vars_crc_noalloc, in 64-bit words: the tables
// of CrcArgs (tab, vtab, klane, aw, ypow) at fixed offsets, the result word,
// then one contribution per element, for at most kCrcNoAllocMaxElems
// elements (256 KiB plus 45 KiB per device).
constexpr long kCrcNoAllocMaxElems = 1L << 15;
constexpr long kCrcScrTab = 0, kCrcScrVtab = 2048, kCrcScrKlane = 4096, kCrcScrAw = 4096 + 576,
               kCrcScrYpow = 4096 + 576 + 16, kCrcScrOut = kCrcScrYpow + 1024, kCrcScrPart = kCrcScrOut + 32;
constexpr long kCrcScratchWords = kCrcScrPart + kCrcNoAllocMaxElems;
// once per device, with the device current: record the array's device
// address and copy the tables (the first kCrcScrPart words) into it
hipError_t crc_scratch_setup(int device, const uint64_t *tables, long nwords);
// the array's device address on this device (nullptr before the set-up)
uint64_t *crc_scratch(int device);

// One source of a new element's state after mesh adaptation: the element's
// points in [lo, hi] (per direction) take A2 (x) A1 (x) A0 applied to the
// source element (rows of ns doubles at a0, a1, a2 in atab).  a0 < 0: copy
// the source unchanged (the only job of its element).
struct AdaptJob {
  long dst, src;        // first node of the new / old element
  int nd, ns;           // points per direction, new / old
  int a0, a1, a2;
  int lo[3], hi[3];
};
struct AdaptArgs {
  const double *uold = nullptr;  // old register, element blocks
  double *unew = nullptr;        // new register, element blocks
  const AdaptJob *jobs = nullptr;
  const int *jfirst = nullptr;   // [nelem + 1]: jobs of new element e
  const double *atab = nullptr;
  int nelem = 0;                 // new elements
};
hipError_t launch_adapt(const AdaptArgs &a, hipStream_t s);

// The MRS or minmodB limiter in place on a register (element blocks).
struct LimArgs {
  double *u = nullptr;
  const long *eoff = nullptr;
  const int *en = nullptr;
  const int *nbr = nullptr;        // [e][6][4]: face neighbours in order (-1 none)
  const double *h15 = nullptr;     // [e]: largest box width h in patch coordinates ^1.5 (MRS)
  const double *hw = nullptr;      // [e]: h (minmodB)
  const double *wq = nullptr;      // [n - 2][10]: LGL quadrature weights
  const double *amat = nullptr;    // [n - 2][100]: normalized Legendre analysis A[i*n + j]
  const double *smat = nullptr;    // [n - 2][100]: synthesis S[j*n + i]
  double *ldat = nullptr;          // [e][5][4]: MRS average, min, max; minmodB c000, c100, c010, c001
  double alpha = 0, beta = 0.5;
  int kind = 0;                    // 0 MRS, 1 minmodB
  int scale_bound = 0;             // minmodB: bound scaled by the neighbours' mean |c000|
  int nelem = 0;
};
hipError_t launch_limiter(const LimArgs &a, hipStream_t s);
hipError_t launch_lim_data(const LimArgs &a, hipStream_t s);

// Troubled elements (evolve_set_trouble_score, trouble.c:27-113): the
// library's RDMP indicator writes each element's TROUBLE value tr (1, 0,
// -1) from u and the element data of u_p (ldat as the MRS data pass
// leaves it, [e][5][4]: average, min, max); the score kernel moves the
// scores and writes the mesh's score; the fix kernel redoes the stage
// epilogue on elements whose score is positive.
struct TroubleArgs {
  const double *u = nullptr;
  const double *ldat = nullptr;
  const long *eoff = nullptr;
  const int *en = nullptr;
  const int *efv = nullptr;        // [e]: 1 FV element
  const int *nbr = nullptr;        // [e][6][4] face neighbours (-1 none)
  int *tr = nullptr;
  int nelem = 0;
  double delta0 = 1e-4, eps = 1e-3;
};
hipError_t launch_rdmp(const TroubleArgs &a, hipStream_t s);
hipError_t launch_trouble_score(const int *tr, int *score, int n, int *out, hipStream_t s);
struct TroubleFix {
  double *y = nullptr;
  const double *x = nullptr, *x2 = nullptr, *r = nullptr;
  double s = 0;
  int mode = 0;                    // 0: y = x + s r; 1: y = x; 2: y = 0.75 x + 0.25 (x2 + s r)
  const int *score = nullptr;
  const long *eoff = nullptr;
  const int *en = nullptr;
  int nelem = 0;
};
hipError_t launch_trouble_fix(const TroubleFix &f, hipStream_t s);

// flat [v][node] <-> element blocks [e][v][np_e] on the device
// (to_blocks: 1 flat -> blocks, 0 blocks -> flat)
hipError_t launch_layout(const double *src, double *dst, const long *eoff, const int *en, int nelem,
                         long nnodes, int to_blocks, hipStream_t s);

}  // namespace nmx
