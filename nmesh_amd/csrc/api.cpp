// api.cpp -- the C ABI of libnmx (include/nmx.h): context, parameters under
// nmesh's names, device mirror of the mesh, state transfer and the RK
// drivers that issue the fused stage kernels.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <mutex>
#include <map>
#include <sstream>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/nmx.h"
#include "amr.h"
#include "ckpt.h"
#include "errors.h"
#include "kernels.h"
#include "mesh.h"

using namespace nmx;

static thread_local std::string g_err;

int nmx::set_error(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

static int fail(int code, const std::string &msg) { return set_error(code, msg); }

#define HIPCHK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess)                                                           \
      return fail(NMX_EHIP, std::string(#x) + ": " + hipGetErrorString(e_));        \
  } while (0)

struct DevBuf {
  double *p = nullptr;
  size_t n = 0;
};

struct nmx_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  std::map<std::string, std::string> par;
  bool have_mesh = false;
  Mesh mesh;
  // derived
  int np1 = 0;           // uniform points per direction (0: mixed)
  bool curv = false;
  long nstate = 0;       // doubles per register
  double dt = 0, time = 0;
  long double Te = 0, te = 0;   // time = Te + te (main.c:588-624)
  long iteration = 0;           // full steps taken (mesh->iteration, main.c:617)
  // device mirror
  double *dU = nullptr, *dA = nullptr, *dW0 = nullptr, *dW1 = nullptr, *dR = nullptr, *dP = nullptr;
  double *dXdx = nullptr, *dXbdX = nullptr, *dxyz = nullptr, *dD = nullptr, *dT = nullptr;
  double *degeo = nullptr, *dxb = nullptr;
  int *dedom = nullptr;
  int4 *dface = nullptr;
  int geom = 0;          // 0 Cartesian, 1 stored dXdx, 2 recomputed CubedShell
  // batched (order-bucketed) mirror
  bool amr = false;
  int *djobs1 = nullptr, *dmdesc = nullptr;
  // exponential filter: matrices per order, the plain path's element list,
  // the variables of its passes (evolve_filter_all, then the varlist)
  double *dFilt = nullptr;      // order n at (n-2)*kMaxN*kMaxN: F[j*n + l]
  SlotRec *dfrec = nullptr;
  std::string filt_key;
  int filt_mask[2] = {0, 0};
  int filt_jpow = 0;
  bool filt_on = false;
  double *dBtab = nullptr, *dDall = nullptr, *dxball = nullptr, *daj = nullptr;
  // persistent batched DG path: every DG element as an item in 8 XCD lists,
  // the items' interpolation jobs, D and nodes of every order
  SlotRec *dbrec = nullptr;
  RJob *drjobs = nullptr;
  double *dbD = nullptr, *dbxb = nullptr, *dbW = nullptr;
  int *dbxlo = nullptr;
  // the order-per-launch shape: records by decreasing n, each order's range,
  // element count and tuned tile (nodes per thread, 0: the nmx_block parameter)
  SlotRec *dbrec2 = nullptr;
  int *dbnr = nullptr;
  std::vector<int> bcount, bdiv;
  int ncu = 0;
  int bgrid = 0;          // workgroups of the persistent launch
  bool anydg = false;     // the batched mesh has DG elements
  // troubled elements: TROUBLE values, scores, the mesh score (device),
  // the scores kept across a re-mirror, the per-element FV flags, a
  // caller's TROUBLE hook
  int *dtr = nullptr, *dscore = nullptr, *dtrout = nullptr, *defv = nullptr;
  std::vector<int> hscore;
  nmx_trouble_fn trouble_hook = nullptr;
  void *trouble_user = nullptr;
  long trouble_redos = 0, trouble_switches = 0;
  bool batch_fixed = false;   // nmx_batch set by the caller (the autotuner leaves it)
  // div: nodes per thread of this bucket's tile (0: the nmx_block parameter)
  // f1first/f1count: the face interpolation jobs of its elements in djobs1
  struct Bucket {
    int n; int count; SlotRec *drec; long doff; long xoff; int div;
    int f1first, f1count;
    float ms;   // tuned stage time alone (0: not tuned) of the shape in use
    float ms_legacy, ms_order;   // ... with the face kernel, and in the order's own launch
    int fv;     // 1: FV elements on uniform points (fv_divf)
    long woff;  // FV: the bucket's midpoint distances in dwcall
  };
  std::vector<Bucket> buckets;
  std::vector<long> eoff;   // node offset per element
  // graph cache
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_steps = 0;
  double graph_dt = 0;
  int graph_method = -1;
  // with troubled elements: one graph per RK stage (the stage's launches
  // only; the scores are read on the host between stages), and the stage
  // arguments each was captured with
  hipGraphExec_t tgexec[4] = {nullptr, nullptr, nullptr, nullptr};
  StageArgs tgargs[4];
  // the u_p whose element bounds dldat holds in this step (RDMP scores
  // every substep against the same u_p); nullptr: not computed
  const double *ldat_up = nullptr;
  double *hT = nullptr;     // pinned stage-time staging
  double *hstage = nullptr; // pinned state staging (element blocks) for host transfers
  long hstage_n = 0;
  int tcap = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // helper streams for concurrent order buckets (fork/join inside graphs)
  static constexpr int kHelpers = 9;
  static constexpr int kLanes = 4;     // stream lanes of the concurrent buckets
  std::vector<int> border;             // buckets by decreasing node count
  std::vector<int> blane;              // lane of each bucket
  hipStream_t helper[kHelpers] = {};
  hipEvent_t evfork = nullptr, evjoin[kHelpers] = {};
  // the faces with several neighbours run beside the single-neighbour ones
  double last_ms = 0;
  bool tuned = false;
  // adjoint: lam register, neighbour face index per element face (plain
  // path), its own graph cache, and the checkpoint pool
  double *dL = nullptr;
  int *dnbf = nullptr;
  // adjoint of the batched path: the interpolation jobs as the forward
  // tables made them (host), then the device tables, G and T buffers and
  // one element list per order (plus one of all elements)
  std::vector<AdjJob> hadjj;
  std::vector<int> hadjsend;   // per job: sender element, sender face
  AdjFace *dadjf = nullptr;
  AdjJob *dadjj = nullptr;
  int *dadjs = nullptr, *dadjl = nullptr;
  long *dadjg = nullptr;
  double *dadjG = nullptr, *dadjT = nullptr;
  std::vector<std::pair<int, int>> adjlists;   // (n, first) in dadjl; all elements first
  std::vector<int> adjcount;
  int adjnjobs = 0;
  bool adj_ready = false;
  // FV elements (fv_divf) on uniform points: every element (fvelm, plain
  // path) or some (anyfv; FV order buckets of the batched path)
  bool fvelm = false, anyfv = false;
  double *dwc = nullptr;    // distances between midpoints around each point
  double *dwcall = nullptr; // the same per FV bucket
  // MRS limiter: neighbour lists, h^1.5, weights, per-element data
  int *dnbr = nullptr;
  double *dlimh = nullptr, *dlimhw = nullptr, *dwqall = nullptr, *dldat = nullptr, *damat = nullptr,
         *dsmat = nullptr;
  bool lim_on = false;
  int lim_kind = 0, lim_scale = 0;
  double lim_alpha = 0, lim_beta = 0.5;
  // checkpoint CRC: device tables, per-element contributions, result
  uint64_t *dcrctab = nullptr, *dcrcvtab = nullptr, *dcrcklane = nullptr, *dcrcaw = nullptr,
           *dcrcypow = nullptr, *dcrcpart = nullptr, *dcrcout = nullptr;
  long *deoff = nullptr;
  int *den = nullptr;
  hipGraph_t agraph = nullptr;
  hipGraphExec_t agexec = nullptr;
  int agraph_steps = 0;
  double agraph_dt = 0;
  int agraph_method = -1;
  double *dCk = nullptr;
  int nck = 0;
  std::vector<long double> ckTe, ckte;
  std::vector<char> ckset;
  double last_adj_ms = 0;
};

extern "C" const char *nmx_last_error(void) { return g_err.c_str(); }

static void set_defaults(nmx_ctx *c) {
  // defaults of the reference parameters (nmesh_amr.c:31-121, nmesh_main.c:29-35,
  // nmesh_evolve.c:26,35-39, nmesh_scalarwave1.c:31-34, nmesh_basis.c:22-24)
  c->par = {
      {"amr_mesh_type", "BoxMesh"},
      {"amr_n0", "5"},
      {"amr_n1", "5"},
      {"amr_n2", "5"},
      {"amr_luni", "0"},
      {"amr_BoxMesh_dout", "1"},
      {"amr_mesh_xc", "0 0 0"},
      {"amr_Shell_rin", "0.5"},
      {"amr_Shell_rout", "1"},
      {"amr_hrefine_sphere_radius", "10"},
      {"amr_hrefine_sphere_levels", "0"},
      {"amr_uniform_p", ""},
      {"amr_fv_p", ""},
      {"amr_fv_if_uniform", "no"},
      {"amr_nmax", "55"},
      {"amr_Shell_nr", "1"},
      {"amr_Shell_r1", "0"},
      {"amr_Shell_r2", "0"},
      {"amr_dir_active0", "yes"},
      {"amr_dir_active1", "yes"},
      {"amr_dir_active2", "yes"},
      {"coordinates_surface_metric", "sqrtgdiag"},
      {"dg_outerBC_flux_fac", "1 1 1"},
      {"evolve_redo_troubled", "substep"},
      {"evolve_u_surfaces", "no"},
      {"fv_rec", "WENOm3_2"},
      {"fv_flux", "fnum"},
      {"fv_divf_extrap", "dnfn_extrap1"},
      {"fv_divf_use_only_right_flux", "yes"},
      {"fv_WENOm3_opt_weightratio", "2"},
      {"fv_surface_interp", "linear"},
      {"fv_rec_WENOm_s1", "0"},
      {"fv_rec_WENOm_s2", "DBL_MAX"},
      {"fv_rec_WENOm_opt", "0"},
      {"fv_divf_extrap_s1", "0"},
      {"fv_divf_extrap_s2", "DBL_MAX"},
      {"fv_divf_extrap_opt", "0"},
      {"fv2dg_interp_use_extrap1", "no"},
      {"dt", "auto"},
      {"dtfac", "0.25"},
      {"iterations", "0"},
      {"finaltime", "0"},
      {"uniform_dtfac", "0.125"},
      {"evolve_method", "RK4"},
      {"scalarwave1_k", "0.8 -0.5 0.33166247903554"},
      {"scalarwave1_numflux", "upwind"},
      {"scalarwave1_profile", "sin"},
      {"evolve_filter_all", "no"},
      {"evolve_filter_varlist", ""},
      {"evolve_filter_alp", "36"},
      {"evolve_filter_s", "32"},
      {"evolve_filter_dn", "0"},
      {"basis_expfilter_JacobianPower", "1"},
      {"basis_filter_fv", "yes"},
      {"coordinates_midpoint_data", "no"},
      {"scalarwave1_limiter", "none"},
      {"limiter_alpha", "5"},
      {"limiter_beta", "0.5"},
      {"limiter_scaleBound", "no"},
      {"nmx_path", "plain"},
      {"nmx_graph", "yes"},
      {"nmx_graph_steps", "8"},
      {"nmx_geometry", "analytic"},
      {"nmx_xcd", "yes"},
      {"nmx_block", "half"},
      {"nmx_autotune", "yes"},
      {"nmx_concurrent", "yes"},
      {"nmx_batch", "legacy"},
      {"nmx_stream_hints", "0"},
      {"nmx_trouble", "none"},
      {"nmx_trouble_delta0", "1e-4"},
      {"nmx_trouble_eps", "1e-3"},
      {"evolve_redo_troubled", "substep"},
  };
}

static std::string P(nmx_ctx *c, const char *k) { return c->par[k]; }
// numerical value, with DBL_MAX and -DBL_MAX as the reference's parameter
// database translates them (translatevalue, parameters.c:406-428)
static double Pd(nmx_ctx *c, const char *k) {
  const std::string &v = c->par[k];
  if (v == "DBL_MAX") return 1.7976931348623157e308;
  if (v == "-DBL_MAX") return -1.7976931348623157e308;
  return std::atof(v.c_str());
}
static int Pi(nmx_ctx *c, const char *k) { return std::atoi(c->par[k].c_str()); }

// fv_rec: the interior scheme (3 WENOm3, 5 WENO5, 7 WENO-Z) and the end
// midpoints (dg.c:295-326, rec1d.c:1032-1186): recend 0 for the _2 modes, 1
// for _1, 2 for WENOm3 / WENOm5 / WENOmZ with fv_rec_WENOm_s1/_s2/_opt
// (dg.c:333-335)
static int fv_rec_pars(nmx_ctx *c, int &recend, double &s1, double &s2, int &opt, int &on) {
  const std::string &fr = P(c, "fv_rec");
  const int rec = fr.compare(0, 6, "WENOm5") == 0   ? 5
                  : fr.compare(0, 6, "WENOmZ") == 0 ? 7
                  : fr.compare(0, 6, "WENOmT") == 0 ? 9 : 3;
  const size_t u = fr.find('_');
  recend = u == std::string::npos ? 2 : (fr.substr(u) == "_1" ? 1 : 0);
  if (fr == "1") recend = 0;
  s1 = Pd(c, "fv_rec_WENOm_s1");
  s2 = Pd(c, "fv_rec_WENOm_s2");
  opt = (int)Pd(c, "fv_rec_WENOm_opt");
  on = (recend == 2 && (s1 > 0. || s2 < 1.7976931348623157e308)) ? 1 : 0;
  if (recend == 2 && !on) recend = 0;   // the average at both ends, as _2
  return rec;
}

// fv_divf_extrap_s1, _s2, _opt for dnfn_extrap1 (dg.c:393-395; the limit is
// on when s1 > 0 or s2 < DBL_MAX, rec1d.c:560)
static void fv_extrap_pars(nmx_ctx *c, double &s1, double &s2, int &opt, int &on) {
  s1 = Pd(c, "fv_divf_extrap_s1");
  s2 = Pd(c, "fv_divf_extrap_s2");
  opt = Pi(c, "fv_divf_extrap_opt");
  on = (s1 > 0. || s2 < 1.7976931348623157e308) ? 1 : 0;
}

// the static scratch of nmx_vars_crc_noalloc on this device (current):
// its address recorded and the CRC tables copied in, once per device
static int noalloc_setup(int device) {
  static std::mutex mu;
  std::lock_guard<std::mutex> lock(mu);
  if (crc_scratch(device)) return NMX_OK;
  std::vector<uint64_t> tab, vtab, klane, aw, ypow;
  crc_device_tables(tab, vtab, klane, aw, ypow);
  std::vector<uint64_t> t(kCrcScrPart, 0);
  std::copy(tab.begin(), tab.end(), t.begin() + kCrcScrTab);
  std::copy(vtab.begin(), vtab.end(), t.begin() + kCrcScrVtab);
  std::copy(klane.begin(), klane.end(), t.begin() + kCrcScrKlane);
  std::copy(aw.begin(), aw.end(), t.begin() + kCrcScrAw);
  std::copy(ypow.begin(), ypow.end(), t.begin() + kCrcScrYpow);
  hipError_t e = crc_scratch_setup(device, t.data(), (long)t.size());
  if (e != hipSuccess) return fail(NMX_EHIP, std::string("static CRC scratch: ") + hipGetErrorString(e));
  return NMX_OK;
}

extern "C" int nmx_ctx_create(nmx_ctx **out, int device) {
  if (!out) return fail(NMX_EINVAL, "ctx pointer is NULL");
  if (device == -1) {   // host-only context: mesh and tables, no device mirror
    nmx_ctx *c = new nmx_ctx();
    c->device = -1;
    set_defaults(c);
    *out = c;
    return NMX_OK;
  }
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(NMX_EINVAL, "no such HIP device");
  nmx_ctx *c = new nmx_ctx();
  c->device = device;
  set_defaults(c);
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking));
  c->stream = c->own_stream;
  HIPCHK(hipEventCreate(&c->ev0));
  HIPCHK(hipEventCreate(&c->ev1));
  HIPCHK(hipEventCreateWithFlags(&c->evfork, hipEventDisableTiming));
  for (int i = 0; i < nmx_ctx::kHelpers; ++i) {
    HIPCHK(hipStreamCreateWithFlags(&c->helper[i], hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&c->evjoin[i], hipEventDisableTiming));
  }
  if (int r = noalloc_setup(device)) return r;
  *out = c;
  return NMX_OK;
}

static void drop_stage_graphs(nmx_ctx *c) {
  for (auto &g : c->tgexec)
    if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
}

static void free_mesh_dev(nmx_ctx *c) {
  c->ldat_up = nullptr;
  double **bufs[] = {&c->dU, &c->dA, &c->dW0, &c->dW1, &c->dR, &c->dP,
                     &c->dXdx, &c->dXbdX, &c->dxyz, &c->dD, &c->dT, &c->degeo, &c->dxb};
  for (auto b : bufs)
    if (*b) { (void)hipFree(*b); *b = nullptr; }
  if (c->dedom) { (void)hipFree(c->dedom); c->dedom = nullptr; }
  {
    void **ptrs[] = {(void **)&c->djobs1, (void **)&c->dmdesc, (void **)&c->dFilt, (void **)&c->dfrec,
                     (void **)&c->dBtab, (void **)&c->dDall, (void **)&c->dxball, (void **)&c->daj,
                     (void **)&c->dbrec, (void **)&c->drjobs, (void **)&c->dbD, (void **)&c->dbxb,
                     (void **)&c->dbW, (void **)&c->dbxlo, (void **)&c->dbrec2, (void **)&c->dbnr,
                     (void **)&c->dtr, (void **)&c->dscore, (void **)&c->dtrout, (void **)&c->defv};
    for (auto q : ptrs)
      if (*q) { (void)hipFree(*q); *q = nullptr; }
    for (auto &b : c->buckets)
      if (b.drec) (void)hipFree(b.drec);
    c->buckets.clear();
    c->border.clear();
    c->blane.clear();
    c->filt_key.clear();
  }
  if (c->dface) { (void)hipFree(c->dface); c->dface = nullptr; }
  drop_stage_graphs(c);
  if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
  if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
  c->graph_steps = 0;
  {
    void **ptrs[] = {(void **)&c->dL, (void **)&c->dnbf, (void **)&c->deoff, (void **)&c->den,
                     (void **)&c->dCk, (void **)&c->dwc, (void **)&c->dwcall,
                     (void **)&c->dcrctab, (void **)&c->dcrcvtab, (void **)&c->dcrcklane,
                     (void **)&c->dcrcaw, (void **)&c->dcrcypow,
                     (void **)&c->dcrcpart, (void **)&c->dcrcout, (void **)&c->dnbr,
                     (void **)&c->dlimh, (void **)&c->dwqall, (void **)&c->dldat, (void **)&c->dlimhw,
                     (void **)&c->damat, (void **)&c->dsmat, (void **)&c->dadjf, (void **)&c->dadjj,
                     (void **)&c->dadjs, (void **)&c->dadjl, (void **)&c->dadjg, (void **)&c->dadjG,
                     (void **)&c->dadjT};
    for (auto q : ptrs)
      if (*q) { (void)hipFree(*q); *q = nullptr; }
    c->adj_ready = false;
    c->adjlists.clear();
    c->adjcount.clear();
  }
  if (c->agexec) { (void)hipGraphExecDestroy(c->agexec); c->agexec = nullptr; }
  if (c->agraph) { (void)hipGraphDestroy(c->agraph); c->agraph = nullptr; }
  c->agraph_steps = 0;
  c->nck = 0;
  c->ckTe.clear(); c->ckte.clear(); c->ckset.clear();
}

extern "C" int nmx_ctx_destroy(nmx_ctx *c) {
  if (!c) return NMX_OK;
  if (c->device < 0) { delete c; return NMX_OK; }
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  free_mesh_dev(c);
  if (c->hT) (void)hipHostFree(c->hT);
  if (c->hstage) (void)hipHostFree(c->hstage);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->evfork) (void)hipEventDestroy(c->evfork);
  for (int i = 0; i < nmx_ctx::kHelpers; ++i) {
    if (c->evjoin[i]) (void)hipEventDestroy(c->evjoin[i]);
    if (c->helper[i]) (void)hipStreamDestroy(c->helper[i]);
  }
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
  return NMX_OK;
}

extern "C" int nmx_set_stream(nmx_ctx *c, void *s) {
  if (!c) return fail(NMX_EINVAL, "ctx is NULL");
  c->stream = s ? (hipStream_t)s : c->own_stream;
  return NMX_OK;
}

extern "C" int nmx_synchronize(nmx_ctx *c) {
  if (!c) return fail(NMX_EINVAL, "ctx is NULL");
  if (c->device < 0) return NMX_OK;
  HIPCHK(hipStreamSynchronize(c->stream));
  return NMX_OK;
}

static int remirror_with_state(nmx_ctx *c, const Mesh &old);

// fv_surface_interp: "linear" or "parabolic", optionally with the word
// "Lagrange_if_dg" (dg.c:369-379); false for anything else
static bool surface_interp_mode(const std::string &v, int *mode, bool *lag) {
  std::istringstream in(v);
  int m = -1;
  bool l = false;
  for (std::string w; in >> w;) {
    if (w == "linear" && m < 0) m = 0;
    else if (w == "parabolic" && m < 0) m = 1;
    else if (w == "Lagrange_if_dg" && !l) l = true;
    else return false;
  }
  if (m < 0) return false;
  if (mode) *mode = m;
  if (lag) *lag = l;
  return true;
}

extern "C" int nmx_set_param(nmx_ctx *c, const char *name, const char *value) {
  if (!c || !name || !value) return fail(NMX_EINVAL, "NULL argument");
  std::string k(name), v(value);
  if (!c->par.count(k)) return fail(NMX_EINVAL, "unknown parameter " + k);
  if (k == "evolve_method" && v != "RK4" && v != "sspRK3" && v != "Euler")
    return fail(NMX_EINVAL, "evolve_method must be RK4, sspRK3 or Euler");
  if (k == "scalarwave1_numflux" && v != "upwind" && v != "LLF")
    return fail(NMX_EINVAL, "scalarwave1_numflux must be upwind or LLF");
  if (k == "scalarwave1_profile" && v != "sin")
    return fail(NMX_EUNSUP, "only scalarwave1_profile = sin is supported");
  if (k == "nmx_trouble" && v != "none" && v != "RDMP" && v != "hook")
    return fail(NMX_EINVAL, "nmx_trouble must be none, RDMP or hook");
  if (k == "evolve_redo_troubled" && v != "substep" && v != "step")
    return fail(NMX_EINVAL, "evolve_redo_troubled must be step or substep");
  if (k == "nmx_batch" && v != "persistent" && v != "buckets" && v != "legacy")
    return fail(NMX_EINVAL, "nmx_batch must be persistent, buckets or legacy");
  if (k == "nmx_path" && v != "plain" && v != "perelement" && v != "batched")
    return fail(NMX_EINVAL, "nmx_path must be plain, perelement or batched");
  if (k == "nmx_geometry" && v != "analytic" && v != "stored")
    return fail(NMX_EINVAL, "nmx_geometry must be analytic or stored");
  if (k == "nmx_geometry" && c->have_mesh)
    return fail(NMX_ESTATE, "set nmx_geometry before amr_setup_mesh");
  if (k == "fv_rec" && v != "WENOm3_2" && v != "WENOm5_2" && v != "WENOmZ_2" && v != "WENOm3_1" &&
      v != "WENOm5_1" && v != "WENOmZ_1" && v != "WENOm3" && v != "WENOm5" && v != "WENOmZ" &&
      v != "WENOmT_2" && v != "1")
    return fail(NMX_EUNSUP, "fv_rec: 1, WENOm3, WENOm5, WENOmZ (and _1, _2) and WENOmT_2 are built");
  if ((k == "fv_flux" && v != "fnum") || (k == "fv_divf_extrap" && v != "dnfn_extrap1") ||
      (k == "fv_divf_use_only_right_flux" && v != "yes") || (k == "fv_WENOm3_opt_weightratio" && v != "2") ||
      (k == "fv2dg_interp_use_extrap1" && v != "no"))
    return fail(NMX_EUNSUP, k + ": only nmesh's default is built");
  if (k == "fv_surface_interp" && !surface_interp_mode(v, nullptr, nullptr))
    return fail(NMX_EUNSUP,
                "fv_surface_interp: linear or parabolic, each optionally with Lagrange_if_dg, is built (not WENO)");
  if ((k == "amr_dir_active0" || k == "amr_dir_active1" || k == "amr_dir_active2") && v != "yes")
    return fail(NMX_EUNSUP, k + ": only 3d meshes are built");
  if (k == "coordinates_surface_metric" && v != "sqrtgdiag")
    return fail(NMX_EUNSUP, "coordinates_surface_metric: only sqrtgdiag is built");
  if (k == "amr_Shell_nr" && v != "1")
    return fail(NMX_EUNSUP, "amr_Shell_nr: one radial shell of 6 patches is built");
  if ((k == "amr_fv_if_uniform" || k == "evolve_u_surfaces") && v != "yes" && v != "no")
    return fail(NMX_EINVAL, k + " must be yes or no");
  if (k == "dg_outerBC_flux_fac") {
    double f[3];
    if (std::sscanf(v.c_str(), "%lg %lg %lg", &f[0], &f[1], &f[2]) != 3)
      return fail(NMX_EINVAL, "dg_outerBC_flux_fac must contain 3 numbers");
  }
  if (k == "scalarwave1_limiter" && v != "none" && v != "MRS" && v != "minmodB")
    return fail(NMX_EINVAL, "scalarwave1_limiter must be none, MRS or minmodB");
  if (k == "limiter_scaleBound" && v != "yes" && v != "no")
    return fail(NMX_EINVAL, "limiter_scaleBound must be yes or no");
  if (k == "coordinates_midpoint_data" && v != "no" && v != "yes" && v != "yes all")
    return fail(NMX_EINVAL, "coordinates_midpoint_data must be no, yes or yes all");
  if ((k == "evolve_filter_all" || k == "basis_filter_fv") && v != "yes" && v != "no")
    return fail(NMX_EINVAL, k + " must be yes or no");
  if (k == "basis_expfilter_JacobianPower" && v != "1" && v != "0" && v != "-1")
    return fail(NMX_EINVAL, "Jpower must be 1, -1 or 0");
  if (k == "nmx_block" && v != "full" && v != "half" && v != "quarter")
    return fail(NMX_EINVAL, "nmx_block must be full, half or quarter");
  const std::string prev = c->par[k];
  c->par[k] = v;
  if (k == "nmx_block" || k == "nmx_xcd") c->tuned = true;   // explicit choice wins
  if (k == "nmx_batch") c->batch_fixed = true;
  if (k == "nmx_block")
    for (auto &b : c->buckets) b.div = 0;
  drop_stage_graphs(c);
  if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
  if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
  c->graph_steps = 0;
  if (c->agexec) { (void)hipGraphExecDestroy(c->agexec); c->agexec = nullptr; }
  if (c->agraph) { (void)hipGraphDestroy(c->agraph); c->agraph = nullptr; }
  c->agraph_steps = 0;
  if (k == "fv_surface_interp" && c->have_mesh) {
    // new face tables for faces taken from FV neighbours; the state stays
    int mode = 0;
    bool lag = false;
    surface_interp_mode(v, &mode, &lag);
    if (mode != c->mesh.fv_surface_interp || lag != c->mesh.fv_lagrange_if_dg) {
      Mesh old = c->mesh;
      std::string err = set_fv_surface_interp(c->mesh, mode, lag);
      if (!err.empty()) { c->mesh = old; c->par[k] = prev; return fail(NMX_EINVAL, err); }
      int r = remirror_with_state(c, old);
      if (r) { c->par[k] = prev; return r; }
    }
  }
  if ((k == "dtfac" || k == "uniform_dtfac") && c->have_mesh && P(c, "dt") == "auto")
    c->dt = mesh_dt(c->mesh, Pd(c, "dtfac"), Pd(c, "uniform_dtfac"));
  if (k == "dt" && c->have_mesh) {
    if (v == "auto") c->dt = mesh_dt(c->mesh, Pd(c, "dtfac"), Pd(c, "uniform_dtfac"));
    else c->dt = std::atof(v.c_str());
  }
  return NMX_OK;
}

extern "C" int nmx_get_param(nmx_ctx *c, const char *name, char *buf, int len) {
  if (!c || !name || !buf || len <= 0) return fail(NMX_EINVAL, "NULL argument");
  auto it = c->par.find(name);
  if (it == c->par.end()) return fail(NMX_EINVAL, std::string("unknown parameter ") + name);
  std::snprintf(buf, len, "%s", it->second.c_str());
  return NMX_OK;
}

template <class T>
static int dev_alloc(T **p, size_t n) {
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) return fail(NMX_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  return NMX_OK;
}

template <class T>
static int dev_upload(T **p, const std::vector<T> &h) {
  int r = dev_alloc(p, h.size());
  if (r) return r;
  HIPCHK(hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return NMX_OK;
}

// An element's state offset and geometry in its SlotRec (faces left empty).
static void slot_geometry(const Mesh &M, int e, SlotRec &R) {
  const Elm &E = M.e[e];
  std::memset(&R, 0, sizeof R);
  R.noff = E.off;
  for (int d = 0; d < 3; ++d) R.dXbdX[d] = 2. / (E.bbox[2 * d + 1] - E.bbox[2 * d]);
  for (int q = 0; q < 6; ++q) R.bb[q] = E.bbox[q];
  R.s0 = M.pat[E.p].s0;
  R.s1 = M.pat[E.p].s1;
  R.dom = M.pat[E.p].dom;
  R.n = E.n;
}

// Launch order and lanes of the concurrent buckets, largest first.  Running
// together, the largest buckets share the GPU and end near the same time,
// so each of the kLanes-1 largest gets a lane of its own and the rest queue
// on the last lane, unless together they outweigh the largest bucket: then
// each goes to the lane with the least work so far.  Work is the bucket's
// tuned stage time once the autotuner has run, its node count before.
static void assign_lanes(nmx_ctx *c) {
  const size_t nb = c->buckets.size();
  bool timed = nb > 0;
  for (const auto &b : c->buckets) timed = timed && b.ms > 0;
  auto work = [c, timed](int i) {
    const nmx_ctx::Bucket &b = c->buckets[i];
    const double n = b.n;
    return timed ? (double)b.ms : (double)b.count * n * n * n;
  };
  c->border.resize(nb);
  for (size_t i = 0; i < nb; ++i) c->border[i] = (int)i;
  std::stable_sort(c->border.begin(), c->border.end(), [&](int x, int y) { return work(x) > work(y); });
  c->blane.assign(nb, 0);
  constexpr int L = nmx_ctx::kLanes;
  double load[L] = {};
  double rest = 0;
  for (size_t k = L - 1; k < nb; ++k) rest += work(c->border[k]);
  const bool last_lane = nb >= (size_t)L && rest <= work(c->border[0]);
  for (size_t k = 0; k < nb; ++k) {
    const int bi = c->border[k];
    int lane = 0;
    if (k < (size_t)L - 1) {
      lane = (int)k;
    } else if (last_lane) {
      lane = L - 1;
    } else {
      for (int l = 1; l < L; ++l)
        if (load[l] < load[lane]) lane = l;
    }
    c->blane[bi] = lane;
    load[lane] += work(bi);
  }
}

// Distances between the midpoints around each of n uniform points, from the
// face to the first midpoint at the ends (get_coords.c:1749-1860).
static std::vector<double> fv_cell_widths(int n) {
  const Basis &ub = uniform_basis(n);
  std::vector<double> xm(n), wc(n);
  for (int i = 0; i < n - 1; ++i) xm[i] = 0.5 * (ub.x[i] + ub.x[i + 1]);
  wc[0] = std::fabs(xm[0] - (-1.));
  wc[n - 1] = std::fabs(xm[n - 2] - 1.);
  for (int i = 1; i < n - 1; ++i) wc[i] = xm[i] - xm[i - 1];
  return wc;
}

// Tables of the persistent batched DG launch.  Items: the DG elements in 8
// lists, one per XCD, each a contiguous range of the element order holding
// about an eighth of the work, with decreasing n inside (stable), so a
// workgroup walks its elements order by order.  Each item's interpolation
// jobs are split into rounds of whole faces whose lines fit the workgroup and
// whose rows and first-pass results fit the own-field buffer.
static int upload_batch_tables(nmx_ctx *c, const std::vector<int> &fd,
                               std::vector<std::vector<RJob>> &erj) {
  const Mesh &M = c->mesh;
  const size_t ne = M.e.size();
  std::vector<int> dg;
  for (size_t i = 0; i < ne; ++i)
    if (!M.e[i].fv) dg.push_back((int)i);
  c->anydg = !dg.empty();
  if (dg.empty()) return NMX_OK;
  auto cost = [&](int e) {
    const double n = M.e[e].n;
    return n * n * n + 0.25 * (double)erj[e].size() * n * n * n / 2.;
  };
  double total = 0;
  for (int e : dg) total += cost(e);
  std::vector<int> items, xlo(9, 0);
  items.reserve(dg.size());
  size_t k = 0;
  double acc = 0;
  for (int x = 0; x < 8; ++x) {
    xlo[x] = (int)items.size();
    const size_t k0 = k;
    while (k < dg.size() && (x == 7 || acc + 0.5 * cost(dg[k]) <= total * (x + 1) / 8.)) acc += cost(dg[k++]);
    std::vector<int> part(dg.begin() + k0, dg.begin() + k);
    std::stable_sort(part.begin(), part.end(), [&](int a, int b) { return M.e[a].n > M.e[b].n; });
    items.insert(items.end(), part.begin(), part.end());
  }
  xlo[8] = (int)items.size();
  std::vector<SlotRec> recs(items.size());
  std::vector<RJob> jobs;
  for (size_t it = 0; it < items.size(); ++it) {
    const int e = items[it];
    const int n = M.e[e].n;
    if (n < 2 || n > kBatchMaxN) return fail(NMX_EUNSUP, "batched DG elements need 2 to 10 points per direction");
    SlotRec &R = recs[it];
    slot_geometry(M, e, R);
    for (int f = 0; f < 6; ++f) {
      const int *d = &fd[48 * e + 8 * f];
      R.fkind[f] = d[0];
      if (d[0] == 1) {
        R.foff[f] = 5 * M.e[d[1]].off + d[2];
        R.fs1[f] = d[3];
        R.fs2[f] = d[4];
      }
    }
    R.jfirst = (int)jobs.size();
    R.jcount = (int)erj[e].size();
    // rounds of whole faces
    std::vector<RJob> &J = erj[e];
    const int cap = batch_scratch(n);
    size_t j = 0;
    while (j < J.size()) {
      const size_t r0 = j;
      int lines = 0, scr = 0, rows = 0;
      while (j < J.size()) {
        size_t je = j + 1;
        while (je < J.size() && !((J[je].pk >> 24) & 1)) ++je;
        int fl = 0, fs = 0, fr = 0;
        for (size_t q = j; q < je; ++q) {
          const int m = J[q].pk & 15, sw = (J[q].pk >> 7) & 1;
          const int a0 = (J[q].pk >> 8) & 15, a1 = (J[q].pk >> 12) & 15;
          const int b0 = (J[q].pk >> 16) & 15, b1 = (J[q].pk >> 20) & 15;
          const int nx1 = (sw ? b1 - b0 : a1 - a0) + 1, nx2 = (sw ? a1 - a0 : b1 - b0) + 1;
          fl += 4 * m;
          fs += 4 * nx1 * m + nx1 * m + nx2 * m;
          fr += nx1 * m + nx2 * m;
        }
        if (fl > kBatchLines || fs > cap || fr > kBatchRows)
          return fail(NMX_EUNSUP, "face too large for one round");
        if (j > r0 && (lines + fl > kBatchLines || scr + fs > cap || rows + fr > kBatchRows)) break;
        rows += fr;
        for (size_t q = j; q < je; ++q) {
          const int m = J[q].pk & 15, sw = (J[q].pk >> 7) & 1;
          const int a0 = (J[q].pk >> 8) & 15, a1 = (J[q].pk >> 12) & 15;
          const int b0 = (J[q].pk >> 16) & 15, b1 = (J[q].pk >> 20) & 15;
          const int nx1 = (sw ? b1 - b0 : a1 - a0) + 1, nx2 = (sw ? a1 - a0 : b1 - b0) + 1;
          if (a1 >= n || b1 >= n) return fail(NMX_EINVAL, "interpolation coverage out of range");
          J[q].line0 = lines;
          J[q].toff = scr;
          J[q].r1off = scr + 4 * nx1 * m;
          J[q].r2off = J[q].r1off + nx1 * m;
          J[q].nround = 0;
          lines += 4 * m;
          scr += 4 * nx1 * m + nx1 * m + nx2 * m;
        }
        j = je;
      }
      J[r0].nround = (int)(j - r0);
    }
    jobs.insert(jobs.end(), J.begin(), J.end());
  }
  if (jobs.empty()) jobs.resize(1);
  std::vector<double> D, xb, wi(kBatchMaxN + 1, 0.0);
  for (int n = 2; n <= kBatchMaxN; ++n) {
    const Basis &b = lgl_basis(n);
    if ((int)D.size() != batch_doff(n) || (int)xb.size() != batch_xoff(n))
      return fail(NMX_EINVAL, "batched D table layout");
    D.insert(D.end(), b.D, b.D + n * n);
    xb.insert(xb.end(), b.x, b.x + n);
    wi[n] = 1.0 / b.w[0];
  }
  // the order-per-launch shape: the same records by decreasing n, element
  // order inside; order n holds [nrange[2n], nrange[2n + 1])
  std::vector<int> perm(items.size());
  for (size_t i = 0; i < perm.size(); ++i) perm[i] = (int)i;
  std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) {
    if (recs[x].n != recs[y].n) return recs[x].n > recs[y].n;
    return items[x] < items[y];
  });
  std::vector<SlotRec> recs2(perm.size());
  std::vector<int> nrange(2 * (kBatchMaxN + 1), 0);
  for (size_t i = 0; i < perm.size(); ++i) recs2[i] = recs[perm[i]];
  for (size_t i = 0; i < recs2.size(); ++i) {
    const int n = recs2[i].n;
    if (i == 0 || recs2[i - 1].n != n) nrange[2 * n] = (int)i;
    nrange[2 * n + 1] = (int)i + 1;
  }
  c->bcount.assign(kBatchMaxN + 1, 0);
  for (int n = 2; n <= kBatchMaxN; ++n) c->bcount[n] = nrange[2 * n + 1] - nrange[2 * n];
  c->bdiv.assign(kBatchMaxN + 1, 0);
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, c->device));
  c->ncu = prop.multiProcessorCount;
  c->bgrid = 8 * ((3 * prop.multiProcessorCount + 7) / 8);
  int r;
  if ((r = dev_upload(&c->dbrec, recs)) || (r = dev_upload(&c->drjobs, jobs)) || (r = dev_upload(&c->dbD, D)) ||
      (r = dev_upload(&c->dbxb, xb)) || (r = dev_upload(&c->dbW, wi)) || (r = dev_upload(&c->dbxlo, xlo)) ||
      (r = dev_upload(&c->dbrec2, recs2)) || (r = dev_upload(&c->dbnr, nrange)))
    return r;
  return NMX_OK;
}

// Tables of the batched path: face descriptors, face interpolation jobs
// (single-neighbour faces and faces with several neighbours in separate
// lists, each grouped by the receiving element's order bucket) with
// deduplicated 1D Lagrange rows, order buckets with one SlotRec per element.
// Every gather a kernel will make is bounds-checked here.
static int upload_amr_tables(nmx_ctx *c) {
  Mesh &M = c->mesh;
  const size_t ne = M.e.size();
  std::vector<long> soff(ne);
  for (size_t i = 0; i < ne; ++i) soff[i] = 5 * M.e[i].off;
  std::vector<int> fd(48 * ne, 0);
  // jobs of each element's interpolated faces, one per (face, neighbour),
  // in the face kernel's form (ej1) and the persistent DG kernel's (erj)
  std::vector<std::vector<int>> ej1(ne);
  std::vector<std::vector<RJob>> erj(ne);
  std::vector<int> mdesc;
  std::vector<double> btab;
  std::map<std::vector<double>, int> bdedup;
  long ajlen = 0;
  auto brow = [&](const std::vector<double> &rows) {
    auto it = bdedup.find(rows);
    if (it != bdedup.end()) return it->second;
    int o = (int)btab.size();
    btab.insert(btab.end(), rows.begin(), rows.end());
    bdedup[rows] = o;
    return o;
  };
  c->hadjj.clear();
  c->hadjsend.clear();
  for (size_t i = 0; i < ne; ++i) {
    const Elm &E = M.e[i];
    const int n = E.n;
    for (int f = 0; f < 6; ++f) {
      int *d = &fd[48 * i + 8 * f];
      if (E.ajmode[f] == 0) { d[0] = 0; continue; }
      if (E.ajmode[f] == 1 || E.ajmode[f] == 2) {
        const FaceCopy &fc = M.fcopy[6 * i + f];
        const int nb = E.fnb[f][0];
        const Elm &N = M.e[nb];
        if (N.n != n) return fail(NMX_EINVAL, "copy face between different orders");
        d[0] = 1; d[1] = nb; d[2] = (int)(fc.nbase - N.off); d[3] = fc.s1; d[4] = fc.s2;
        const long np3 = (long)n * n * n;
        for (int cb = 0; cb < 2; ++cb)
          for (int ca = 0; ca < 2; ++ca) {
            long q = d[2] + (long)ca * (n - 1) * d[3] + (long)cb * (n - 1) * d[4];
            if (q < 0 || q >= np3) return fail(NMX_EINVAL, "face gather table out of range");
          }
        continue;
      }
      const std::vector<InterpSlot> &sl = M.fint[6 * i + f];
      const int ns = (int)sl.size();
      if (ns == 0 || ns > kMaxFaceNb) return fail(NMX_EUNSUP, "face with no or more than 4 interpolation neighbours");
      // the face kernel's LDS holds 4 * covered rows * m first-pass results
      // and covered rows (or columns) * m doubles of each pass's rows
      for (const InterpSlot &q : sl) {
        const int nx1 = q.swap ? (q.b1 - q.b0 + 1) : (q.a1 - q.a0 + 1);
        const int nx2 = q.swap ? (q.a1 - q.a0 + 1) : (q.b1 - q.b0 + 1);
        if (nx1 > kMaxFaceN || nx2 > kMaxFaceN || q.m > kMaxFaceN ||
            4 * nx1 * q.m > kFaceT1 || nx1 * q.m > kFaceR1 || nx2 * q.m > kFaceR1)
          return fail(NMX_EUNSUP, "face order above 10");
      }
      // one region of 4*n*n values per neighbour; a face with several
      // neighbours is averaged by its element from their regions
      if (ajlen + 4L * n * n * ns > (long)INT32_MAX) return fail(NMX_EUNSUP, "face buffer too large");
      if (ns == 1) {
        d[0] = 2; d[2] = (int)ajlen;
      } else {
        d[0] = 3; d[2] = (int)mdesc.size();
        mdesc.push_back(ns);
      }
      for (int si = 0; si < ns; ++si) {
        const InterpSlot &q = sl[si];
        const long region = ajlen + 4L * n * n * si;
        const Elm &N = M.e[q.nb];
        const int m = q.m;
        const long base = q.nbase - N.off;
        const long m3 = (long)m * m * m;
        for (int c2 = 0; c2 < 2; ++c2)
          for (int c1 = 0; c1 < 2; ++c1) {
            long idx = base + (long)c1 * (m - 1) * q.t1 + (long)c2 * (m - 1) * q.t2;
            if (idx < 0 || idx >= m3) return fail(NMX_EINVAL, "interpolation plane out of range");
          }
        if (q.a0 < 0 || q.a1 >= n || q.b0 < 0 || q.b1 >= n || q.a0 > q.a1 || q.b0 > q.b1)
          return fail(NMX_EINVAL, "interpolation coverage out of range");
        if ((int)q.Ba.size() != n * m || (int)q.Bb.size() != n * m)
          return fail(NMX_EINVAL, "interpolation rows have the wrong size");
        int oa = brow(q.Ba), ob = brow(q.Bb);
        int rec[kJobInts] = {n, m, (int)region, 1, (int)base, q.t1, q.t2, q.swap,
                             q.a0, q.a1, q.b0, q.b1, oa, ob, 0, 0};
        const long so = soff[q.nb];
        std::memcpy(rec + 14, &so, sizeof(long));
        ej1[i].insert(ej1[i].end(), rec, rec + kJobInts);
        RJob rj;
        std::memset(&rj, 0, sizeof rj);
        rj.nbase = so + base;
        rj.t1 = q.t1; rj.t2 = q.t2;
        rj.pk = m | (f << 4) | (q.swap << 7) | (q.a0 << 8) | (q.a1 << 12) | (q.b0 << 16) | (q.b1 << 20) |
                ((si == 0 ? 1 : 0) << 24);
        rj.offA = oa; rj.offB = ob;
        erj[i].push_back(rj);
        if (ns > 1) {
          const int md[5] = {(int)region, q.a0, q.a1, q.b0, q.b1};
          mdesc.insert(mdesc.end(), md, md + 5);
        }
        {
          // the job as the adjoint transposes it: the neighbour's face that
          // holds the plane, and its face coordinates as a map of (q1, q2)
          AdjJob aj;
          std::memset(&aj, 0, sizeof aj);
          aj.e = (int)i; aj.g = f; aj.n = n; aj.m = m; aj.sw = q.swap;
          aj.a0 = q.a0; aj.a1 = q.a1; aj.b0 = q.b0; aj.b1 = q.b1; aj.offA = oa; aj.offB = ob;
          aj.md = ns > 1 ? d[2] : -1;
          int ijk[3][3];
          const long pts[3] = {base, base + q.t1, base + q.t2};
          for (int c3 = 0; c3 < 3; ++c3) {
            ijk[c3][0] = (int)(pts[c3] % m); ijk[c3][1] = (int)((pts[c3] / m) % m); ijk[c3][2] = (int)(pts[c3] / ((long)m * m));
          }
          int nd = -1;
          for (int dd = 0; dd < 3; ++dd)
            if (ijk[1][dd] == ijk[0][dd] && ijk[2][dd] == ijk[0][dd]) nd = dd;
          if (nd < 0 || (ijk[0][nd] != 0 && ijk[0][nd] != m - 1))
            return fail(NMX_EINVAL, "interpolation plane is not a face");
          const int sf = 2 * nd + (ijk[0][nd] ? 1 : 0);
          auto fab = [&](const int *I, int &fa, int &fb) {
            fa = (nd == 0) ? I[1] : I[0];
            fb = (nd == 2) ? I[1] : I[2];
          };
          int fa0, fb0, fa1, fb1, fa2, fb2;
          fab(ijk[0], fa0, fb0); fab(ijk[1], fa1, fb1); fab(ijk[2], fa2, fb2);
          aj.A0 = fa0; aj.B0 = fb0;
          aj.A1 = fa1 - fa0; aj.B1 = fb1 - fb0;   // d(a, b) / d q1
          aj.A2 = fa2 - fa0; aj.B2 = fb2 - fb0;   // d(a, b) / d q2
          // a signed permutation: its inverse is its transpose
          if (std::abs(aj.A1) + std::abs(aj.B1) != 1 || std::abs(aj.A2) + std::abs(aj.B2) != 1 ||
              aj.A1 * aj.A2 + aj.B1 * aj.B2 != 0)
            return fail(NMX_EINVAL, "interpolation plane map is not a signed permutation");
          c->hadjj.push_back(aj);
          c->hadjsend.push_back(q.nb);
          c->hadjsend.push_back(sf);
        }
      }
      ajlen += 4L * n * n * ns;
    }
  }
  if (btab.empty()) btab.assign(1, 0.0);
  if (mdesc.empty()) mdesc.assign(1, 0);
  if (ajlen == 0) ajlen = 1;
  int r;
  if ((r = dev_upload(&c->dBtab, btab)) || (r = dev_alloc(&c->daj, ajlen)) ||
      (r = dev_upload(&c->dmdesc, mdesc)))
    return r;
  std::vector<int> jobs1;
  // order buckets with their D matrices and nodes
  std::vector<double> dall, xball, wcall;
  for (int n = 2; n <= 10; ++n)
  for (int fv = 0; fv <= 1; ++fv) {
    std::vector<int> list;
    for (size_t i = 0; i < ne; ++i)
      if (M.e[i].n == n && (M.e[i].fv != 0) == (fv != 0)) list.push_back((int)i);
    if (list.empty()) continue;
    nmx_ctx::Bucket b;
    b.n = n;
    b.count = (int)list.size();
    b.drec = nullptr;
    b.div = fv ? 1 : 0;
    b.ms = b.ms_legacy = b.ms_order = 0.f;
    b.fv = fv;
    b.woff = 0;
    if (fv) {
      b.woff = (long)wcall.size();
      const std::vector<double> wc = fv_cell_widths(n);
      wcall.insert(wcall.end(), wc.begin(), wc.end());
    }
    std::vector<SlotRec> recs(list.size());
    for (size_t it = 0; it < list.size(); ++it) {
      const int e = list[it];
      SlotRec &R = recs[it];
      slot_geometry(M, e, R);
      for (int f = 0; f < 6; ++f) {
        const int *d = &fd[48 * e + 8 * f];
        R.fkind[f] = d[0];
        if (d[0] == 1) {
          R.foff[f] = soff[d[1]] + d[2];
          R.fs1[f] = d[3];
          R.fs2[f] = d[4];
        } else if (d[0] >= 2) {
          R.foff[f] = d[2];
        }
      }
    }
    // the face jobs of this bucket's elements, in element order
    b.f1first = (int)(jobs1.size() / kJobInts);
    for (int e : list) jobs1.insert(jobs1.end(), ej1[e].begin(), ej1[e].end());
    b.f1count = (int)(jobs1.size() / kJobInts) - b.f1first;
    c->buckets.push_back(b);
    if ((r = dev_upload(&c->buckets.back().drec, recs))) return r;
    const Basis &bs = lgl_basis(n);
    c->buckets.back().doff = (long)dall.size();
    dall.insert(dall.end(), bs.D, bs.D + n * n);
    c->buckets.back().xoff = (long)xball.size();
    const Basis &bx = fv ? uniform_basis(n) : bs;   // the bucket's points
    xball.insert(xball.end(), bx.x, bx.x + n);
  }
  if (wcall.empty()) wcall.assign(1, 0.0);
  if ((r = dev_upload(&c->dDall, dall)) || (r = dev_upload(&c->dxball, xball)) ||
      (r = dev_upload(&c->dwcall, wcall)))
    return r;
  assign_lanes(c);
  if (jobs1.empty()) jobs1.assign(kJobInts, 0);
  if ((r = dev_upload(&c->djobs1, jobs1))) return r;
  return upload_batch_tables(c, fd, erj);
}

// host -> device mirror of the mesh (element-block layout)
static int upload_mesh(nmx_ctx *c) {
  free_mesh_dev(c);
  c->tuned = false;
  Mesh &M = c->mesh;
  const long nn = M.nnodes;
  const size_t ne = M.e.size();
  c->np1 = M.e[0].n;
  for (auto &E : M.e)
    if (E.n != c->np1) c->np1 = 0;
  c->curv = false;
  for (auto &Pt : M.pat)
    if (Pt.type == PAT_CUBEDSHELL) c->curv = true;
  c->amr = (c->np1 == 0) || P(c, "nmx_path") == "batched";
  for (size_t i = 0; i < ne; ++i)
    for (int f = 0; f < 6; ++f)
      if (M.e[i].ajmode[f] == 3) c->amr = true;
  for (auto &E : M.e)
    if (E.n > 10) return fail(NMX_EUNSUP, "the device kernels take at most 10 points per direction");
  c->nstate = 5L * nn;
  c->eoff.resize(ne);
  for (size_t i = 0; i < ne; ++i) c->eoff[i] = M.e[i].off;
  int r;
  if ((r = dev_alloc(&c->dU, c->nstate)) || (r = dev_alloc(&c->dA, c->nstate)) ||
      (r = dev_alloc(&c->dW0, c->nstate)) || (r = dev_alloc(&c->dW1, c->nstate)) ||
      (r = dev_alloc(&c->dR, c->nstate)) || (r = dev_alloc(&c->dP, c->nstate)))
    return r;
  if ((r = dev_alloc(&c->dL, c->nstate))) return r;
  HIPCHK(hipMemset(c->dL, 0, c->nstate * sizeof(double)));
  {
    std::vector<long> eo(ne);
    std::vector<int> en(ne);
    for (size_t i = 0; i < ne; ++i) { eo[i] = M.e[i].off; en[i] = M.e[i].n; }
    if ((r = dev_upload(&c->deoff, eo)) || (r = dev_upload(&c->den, en))) return r;
  }
  HIPCHK(hipMemset(c->dU, 0, c->nstate * sizeof(double)));
  HIPCHK(hipMemset(c->dR, 0, c->nstate * sizeof(double)));
  HIPCHK(hipMemset(c->dP, 0, c->nstate * sizeof(double)));
  // geometry in element blocks
  c->geom = c->curv ? (P(c, "nmx_geometry") == "stored" ? 1 : 2) : 0;
  if (c->geom == 2) {
    for (auto &Pt : M.pat)
      if (Pt.type == PAT_CUBEDSHELL && (Pt.xc[0] != 0 || Pt.xc[1] != 0 || Pt.xc[2] != 0))
        c->geom = 1;   // recomputation assumes the shell centre at the origin
  }
  // element-block copies of the node arrays (every entry written, so no
  // fill), built over element ranges on several threads
  NodeArray h9(c->geom == 1 ? 9 * nn : 1), hx(3 * nn);
  std::vector<double> h3(3 * ne), hg(8 * ne);
  std::vector<int> hdom(ne);
  const int geom = c->geom;
  for_element_ranges(ne, [&](unsigned, size_t e0, size_t e1) {
    for (size_t i = e0; i < e1; ++i) {
      const Elm &E = M.e[i];
      long np = (long)E.n * E.n * E.n;
      for (int d = 0; d < 3; ++d) h3[3 * i + d] = 2. / (E.bbox[2 * d + 1] - E.bbox[2 * d]);
      for (int q = 0; q < 6; ++q) hg[8 * i + q] = E.bbox[q];
      hg[8 * i + 6] = M.pat[E.p].s0;
      hg[8 * i + 7] = M.pat[E.p].s1;
      hdom[i] = M.pat[E.p].dom;
      for (long q = 0; q < np; ++q) {
        long g = E.off + q;
        if (geom == 1)
          for (int m = 0; m < 9; ++m) h9[9 * E.off + m * np + q] = M.dXdx[m * nn + g];
        for (int d = 0; d < 3; ++d) hx[3 * E.off + d * np + q] = M.x[d * nn + g];
      }
    }
  });
  if (geom != 1) h9[0] = 0.0;
  if ((r = dev_alloc(&c->dXdx, h9.size())) || (r = dev_alloc(&c->dXbdX, h3.size())) ||
      (r = dev_alloc(&c->dxyz, hx.size())))
    return r;
  HIPCHK(hipMemcpy(c->dXdx, h9.data(), h9.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->dXbdX, h3.data(), h3.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->dxyz, hx.data(), hx.size() * 8, hipMemcpyHostToDevice));
  if ((r = dev_alloc(&c->degeo, hg.size())) || (r = dev_alloc(&c->dedom, hdom.size()))) return r;
  HIPCHK(hipMemcpy(c->degeo, hg.data(), hg.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->dedom, hdom.data(), hdom.size() * sizeof(int), hipMemcpyHostToDevice));
  c->tcap = 4 * 64;
  if ((r = dev_alloc(&c->dT, c->tcap))) return r;
  if (!c->hT) HIPCHK(hipHostMalloc((void **)&c->hT, c->tcap * sizeof(double), hipHostMallocDefault));
  c->fvelm = true;
  c->anyfv = false;
  for (auto &E : M.e) {
    if (E.fv) c->anyfv = true;
    else c->fvelm = false;
    if (E.fv && E.n < 3) return fail(NMX_EUNSUP, "FV elements need n >= 3 points per direction");
  }
  if (c->anyfv && P(c, "nmx_path") == "perelement")
    return fail(NMX_EUNSUP, "FV elements run on the plain or batched path");
  if (c->fvelm && c->amr) c->fvelm = false;   // all FV, but in order buckets
  if (c->fvelm) {
    if ((r = dev_upload(&c->dwc, fv_cell_widths(c->np1)))) return r;
  }
  if (c->amr) return upload_amr_tables(c);
  {
    // every element, for the filter
    std::vector<SlotRec> fr(ne);
    for (size_t i = 0; i < ne; ++i) slot_geometry(M, (int)i, fr[i]);
    if ((r = dev_upload(&c->dfrec, fr))) return r;
  }
  // face gather descriptors {nb, local base, s1, s2}
  std::vector<int4> hf(6 * ne);
  for (size_t i = 0; i < ne; ++i)
    for (int f = 0; f < 6; ++f) {
      const Elm &E = M.e[i];
      int4 d = {-1, 0, 0, 0};
      if (E.ajmode[f] == 1 || E.ajmode[f] == 2) {
        const FaceCopy &fc = M.fcopy[6 * i + f];
        int nb = E.fnb[f][0];
        d.x = nb;
        d.y = (int)(fc.nbase - M.e[nb].off);
        d.z = fc.s1;
        d.w = fc.s2;
        // every gathered node must lie inside the neighbour (the map is affine,
        // so checking the four corners covers the face)
        const long np3 = (long)c->np1 * c->np1 * c->np1;
        for (int cb = 0; cb < 2; ++cb)
          for (int ca = 0; ca < 2; ++ca) {
            long q = d.y + (long)ca * (c->np1 - 1) * d.z + (long)cb * (c->np1 - 1) * d.w;
            if (nb < 0 || nb >= (int)ne || q < 0 || q >= np3)
              return fail(NMX_EINVAL, "face gather table out of range");
          }
      }
      hf[6 * i + f] = d;
    }
  if ((r = dev_alloc(&c->dface, hf.size()))) return r;
  HIPCHK(hipMemcpy(c->dface, hf.data(), hf.size() * sizeof(int4), hipMemcpyHostToDevice));
  // the neighbour's face on each conforming face (adjoint gathers)
  std::vector<int> hnbf(6 * ne, 0);
  for (size_t i = 0; i < ne; ++i)
    for (int f = 0; f < 6; ++f)
      if (hf[6 * i + f].x >= 0) {
        int nf = M.e[i].nbf[f][0];
        if (nf < 0 || nf > 5) return fail(NMX_EINVAL, "neighbour face index out of range");
        hnbf[6 * i + f] = nf;
      }
  if ((r = dev_upload(&c->dnbf, hnbf))) return r;
  const Basis &b = lgl_basis(c->np1);
  std::vector<double> hD(b.D, b.D + c->np1 * c->np1);
  // xb: the element's points per direction (uniform ones on FV elements)
  const double *xbp = c->fvelm ? uniform_basis(c->np1).x : b.x;
  if ((r = dev_alloc(&c->dD, hD.size()))) return r;
  HIPCHK(hipMemcpy(c->dD, hD.data(), hD.size() * 8, hipMemcpyHostToDevice));
  if ((r = dev_alloc(&c->dxb, c->np1))) return r;
  HIPCHK(hipMemcpy(c->dxb, xbp, c->np1 * 8, hipMemcpyHostToDevice));
  return NMX_OK;
}

// the tail of a mesh set-up: time, dt and the device mirror
static int mesh_ready(nmx_ctx *c) {
  c->have_mesh = true;
  c->time = 0; c->Te = 0; c->te = 0; c->iteration = 0;
  if (P(c, "dt") == "auto") c->dt = mesh_dt(c->mesh, Pd(c, "dtfac"), Pd(c, "uniform_dtfac"));
  else c->dt = Pd(c, "dt");
  if (c->device < 0) return NMX_OK;                    // host-only context
  return upload_mesh(c);
}

extern "C" int nmx_mesh_import(nmx_ctx *c, int npat, const int *pat_type, const double *pat_geo, int nelem,
                               const int *elm_pat, const int *elm_level, const unsigned char *elm_loc,
                               const int *elm_n, const double *elm_bbox, const int *nfnb, const int *fnb,
                               const int *nbface) {
  if (!c || !pat_type || !pat_geo || !elm_pat || !elm_level || !elm_loc || !elm_n || !elm_bbox || !nfnb || !fnb ||
      !nbface)
    return fail(NMX_EINVAL, "NULL argument");
  if (npat < 1 || nelem < 1) return fail(NMX_EINVAL, "need patches and elements");
  if (c->device >= 0) HIPCHK(hipSetDevice(c->device));
  std::vector<ImportPatch> pat(npat);
  for (int p = 0; p < npat; ++p) {
    const double *g = pat_geo + 12 * p;
    pat[p].type = pat_type[p];
    pat[p].dom = (int)g[0];
    pat[p].s0 = g[1];
    pat[p].s1 = g[2];
    for (int f = 0; f < 6; ++f) pat[p].bbox[f] = g[3 + f];
    for (int d = 0; d < 3; ++d) pat[p].xc[d] = g[9 + d];
  }
  std::vector<ImportElm> elm(nelem);
  for (int e = 0; e < nelem; ++e) {
    ImportElm &q = elm[e];
    q.p = elm_pat[e];
    q.l = elm_level[e];
    const int *n3 = elm_n + 3 * e;
    if (n3[0] != n3[1] || n3[0] != n3[2]) return fail(NMX_EUNSUP, "elements need the same n in all directions");
    q.n = n3[0];
    if (q.n > Pi(c, "amr_nmax")) return fail(NMX_EINVAL, "element n is above amr_nmax");
    if (q.l < 0 || q.l >= 24) return fail(NMX_EINVAL, "element level out of range");
    for (int k = 0; k < q.l; ++k) {
      const unsigned char ch = elm_loc[24 * e + k];
      q.path[k] = (unsigned char)(ch >= '0' ? ch - '0' : ch);   // nmesh's '0' + octant, or the octant
    }
    for (int f = 0; f < 6; ++f) q.bbox[f] = elm_bbox[6 * e + f];
    for (int f = 0; f < 6; ++f) {
      q.nfnb[f] = nfnb[6 * e + f];
      if (q.nfnb[f] < 0 || q.nfnb[f] > kMaxNb) return fail(NMX_EUNSUP, "more than 8 neighbours on a face");
      for (int k = 0; k < q.nfnb[f]; ++k) {
        q.fnb[f][k] = fnb[48 * e + 8 * f + k];
        q.nbf[f][k] = nbface[48 * e + 8 * f + k];
      }
    }
  }
  Mesh M;
  std::string err = import_mesh(M, pat, elm);
  if (!err.empty()) return fail(NMX_EINVAL, "nmx_mesh_import: " + err);
  for (auto &Pt : M.pat)
    if (Pt.type == PAT_CUBEDSHELL && (Pt.xc[0] != 0 || Pt.xc[1] != 0 || Pt.xc[2] != 0))
      return fail(NMX_EUNSUP, "cubed-sphere patches need their centre at the origin");
  surface_interp_mode(P(c, "fv_surface_interp"), &M.fv_surface_interp, &M.fv_lagrange_if_dg);
  c->mesh = std::move(M);
  return mesh_ready(c);
}

extern "C" int nmx_mesh_build(nmx_ctx *c) {
  if (!c) return fail(NMX_EINVAL, "ctx is NULL");
  if (c->device >= 0) HIPCHK(hipSetDevice(c->device));
  std::string type = P(c, "amr_mesh_type");
  int n0 = Pi(c, "amr_n0"), n1 = Pi(c, "amr_n1"), n2 = Pi(c, "amr_n2");
  if (n0 != n1 || n0 != n2) return fail(NMX_EUNSUP, "amr_n0, amr_n1, amr_n2 must be equal");
  int luni = Pi(c, "amr_luni");
  if (luni < 0 || luni > 8) return fail(NMX_EINVAL, "amr_luni out of range");
  if (n0 > Pi(c, "amr_nmax")) return fail(NMX_EINVAL, "amr_n0 is above amr_nmax");
  std::string err;
  if (type.find("BoxMesh") != std::string::npos) {
    double xc[3] = {0, 0, 0};
    std::sscanf(P(c, "amr_mesh_xc").c_str(), "%lg %lg %lg", &xc[0], &xc[1], &xc[2]);
    err = build_box(c->mesh, luni, n0, Pd(c, "amr_BoxMesh_dout"), xc);
  } else if (type.find("Shell") != std::string::npos) {
    if (Pd(c, "amr_Shell_r1") > Pd(c, "amr_Shell_rin"))   // setup_Shell_mesh, mesh.c:656-664
      return fail(NMX_EUNSUP, "amr_Shell_r1 > amr_Shell_rin (two or three radial shells) is not built");
    err = build_shell(c->mesh, luni, n0, Pd(c, "amr_Shell_rin"), Pd(c, "amr_Shell_rout"),
                      (double)Pi(c, "amr_hrefine_sphere_radius"), Pi(c, "amr_hrefine_sphere_levels"));
  } else {
    return fail(NMX_EUNSUP, "amr_mesh_type must be BoxMesh or Shell");
  }
  if (!err.empty()) return fail(NMX_EINVAL, err);
  surface_interp_mode(P(c, "fv_surface_interp"), &c->mesh.fv_surface_interp, &c->mesh.fv_lagrange_if_dg);
  {
    // FV elements: amr_fv_p patches on uniform points (amr_uniform_p)
    std::vector<int> up, fp;
    std::istringstream su(P(c, "amr_uniform_p")), sf(P(c, "amr_fv_p"));
    for (int q; su >> q;) up.push_back(q);
    for (int q; sf >> q;) fp.push_back(q);
    if (P(c, "amr_fv_if_uniform") == "yes") fp = up;   // use_fv iff P_UNIFORM (mesh.c:86-91)
    std::sort(up.begin(), up.end());
    std::sort(fp.begin(), fp.end());
    if (up != fp)
      return fail(NMX_EUNSUP, "amr_uniform_p and amr_fv_p must name the same patches (DG on uniform "
                              "points and FV on LGL points are not built)");
    if (!fp.empty()) {
      err = set_fv_patches(c->mesh, fp);
      if (!err.empty()) return fail(NMX_EINVAL, err);
      for (auto &Pt : c->mesh.pat)
        if (Pt.type == PAT_CUBEDSHELL && (Pt.xc[0] != 0 || Pt.xc[1] != 0 || Pt.xc[2] != 0))
          return fail(NMX_EUNSUP, "FV elements need the shell centre at the origin");
    }
  }
  return mesh_ready(c);
}

// ---------------------------------------------------------------------------
// mesh adaptation with the evolved state interpolated onto the new elements
// ---------------------------------------------------------------------------

// 1D Lagrange rows l_s(x_i) on the ns points of an element (LGL, or
// uniform for an FV element) at nd points
static void lagrange_rows(int ns, const std::vector<double> &x, std::vector<double> &tab, int *off,
                          bool uni = false) {
  const Basis &b = uni ? uniform_basis(ns) : lgl_basis(ns);
  double w[kMaxN];
  for (int s = 0; s < ns; ++s) {
    if (!uni) { w[s] = b.wb[s]; continue; }
    double prod = 1.0;
    for (int m = 0; m < ns; ++m)
      if (m != s) prod *= (b.x[s] - b.x[m]);
    w[s] = 1.0 / prod;
  }
  *off = (int)tab.size();
  for (double xi : x)
    for (int s = 0; s < ns; ++s) {
      double prod = 1.0;
      for (int m = 0; m < ns; ++m)
        if (m != s) prod *= (xi - b.x[m]);
      tab.push_back(prod * w[s]);
    }
}

// jobs moving the state of Mold (register uold) onto c->mesh, then the launch
static int adapt_state(nmx_ctx *c, const Mesh &Mold, const double *uold, double *unew = nullptr) {
  const Mesh &M = c->mesh;
  std::map<std::string, int> where;
  for (size_t e = 0; e < Mold.e.size(); ++e) where[elm_name(Mold.e[e])] = (int)e;
  std::vector<AdaptJob> jobs;
  std::vector<int> jfirst;
  std::vector<double> atab;
  for (const Elm &E : M.e) {
    jfirst.push_back((int)jobs.size());
    const int nd = E.n;
    const Basis &bd = E.fv ? uniform_basis(nd) : lgl_basis(nd);
    std::vector<double> xd(bd.x, bd.x + nd);
    AdaptJob J;
    J.dst = E.off; J.nd = nd;
    for (int d = 0; d < 3; ++d) { J.lo[d] = 0; J.hi[d] = nd - 1; }
    const std::string name = elm_name(E);
    auto it = where.find(name);
    if (it != where.end()) {                        // same place
      const Elm &S = Mold.e[it->second];
      J.src = S.off; J.ns = S.n;
      if (S.n == nd && S.fv == E.fv) {
        J.a0 = J.a1 = J.a2 = -1;
      } else {
        if (S.fv && S.n != nd) return fail(NMX_EUNSUP, "p-changes of FV elements are not built");
        lagrange_rows(S.n, xd, atab, &J.a0, S.fv != 0);
        J.a1 = J.a2 = J.a0;
      }
      jobs.push_back(J);
      continue;
    }
    if (E.fv) return fail(NMX_EUNSUP, "h-adaptation of FV elements is not built");
    if (E.l > 0 && (it = where.find(name.substr(0, name.size() - 1))) != where.end()) {   // child
      const Elm &S = Mold.e[it->second];
      J.src = S.off; J.ns = S.n;
      const int ch = E.path[E.l - 1];
      int *ao[3] = {&J.a0, &J.a1, &J.a2};
      for (int d = 0; d < 3; ++d) {
        std::vector<double> xp(nd);
        for (int i = 0; i < nd; ++i) xp[i] = 0.5 * xd[i] + (((ch >> d) & 1) ? 0.5 : -0.5);
        lagrange_rows(S.n, xp, atab, ao[d]);
      }
      jobs.push_back(J);
      continue;
    }
    for (int ch = 0; ch < 8; ++ch) {               // parent of 8 old children
      it = where.find(name + std::string(1, (char)('0' + ch)));
      if (it == where.end()) return fail(NMX_EINVAL, "adapted element " + name + " has no source");
      const Elm &S = Mold.e[it->second];
      J.src = S.off; J.ns = S.n;
      int *ao[3] = {&J.a0, &J.a1, &J.a2};
      for (int d = 0; d < 3; ++d) {
        std::vector<double> xc(nd);
        int lo = nd, hi = -1;
        for (int i = 0; i < nd; ++i) {
          xc[i] = 2. * xd[i] - (((ch >> d) & 1) ? 1. : -1.);
          if (xc[i] >= -1. - 1e-10 && xc[i] <= 1. + 1e-10) { lo = std::min(lo, i); hi = std::max(hi, i); }
        }
        J.lo[d] = lo; J.hi[d] = hi;
        lagrange_rows(S.n, xc, atab, ao[d]);
      }
      jobs.push_back(J);
    }
  }
  jfirst.push_back((int)jobs.size());
  // every job reads inside the old register and writes inside the new one
  const long oldlen = 5 * Mold.nnodes, newlen = 5 * M.nnodes;
  for (const AdaptJob &J : jobs) {
    if (J.nd < 2 || J.nd > 10 || J.ns < 2 || J.ns > 10) return fail(NMX_EINVAL, "adapt: order out of range");
    if (J.src < 0 || 5 * J.src + 5L * J.ns * J.ns * J.ns > oldlen ||
        J.dst < 0 || 5 * J.dst + 5L * J.nd * J.nd * J.nd > newlen)
      return fail(NMX_EINVAL, "adapt: job out of range");
    if (J.a0 >= 0 && (J.a0 + J.nd * J.ns > (long)atab.size() || J.a1 + J.nd * J.ns > (long)atab.size() ||
                      J.a2 + J.nd * J.ns > (long)atab.size()))
      return fail(NMX_EINVAL, "adapt: rows out of range");
    if (J.a0 < 0 && J.ns != J.nd) return fail(NMX_EINVAL, "adapt: copy between orders");
  }
  if (atab.empty()) atab.assign(1, 0.0);
  AdaptJob *djobs = nullptr;
  int *djf = nullptr;
  double *dat = nullptr;
  int r;
  if ((r = dev_upload(&djobs, jobs)) || (r = dev_upload(&djf, jfirst)) || (r = dev_upload(&dat, atab))) {
    (void)hipFree(djobs); (void)hipFree(djf); (void)hipFree(dat);
    return r;
  }
  AdaptArgs a;
  a.uold = uold; a.unew = unew ? unew : c->dU; a.jobs = djobs; a.jfirst = djf; a.atab = dat;
  a.nelem = (int)M.e.size();
  hipError_t err = launch_adapt(a, c->stream);
  if (err == hipSuccess) err = hipStreamSynchronize(c->stream);
  (void)hipFree(djobs); (void)hipFree(djf); (void)hipFree(dat);
  if (err != hipSuccess) return fail(NMX_EHIP, std::string("adapt launch: ") + hipGetErrorString(err));
  return NMX_OK;
}

// the mesh changed on the host: new device mirror, the state interpolated.
// If the new mirror cannot be made, the old mesh and state are restored.
// A replaced mesh is freed on a detached host thread: its node arrays and
// face tables (over a GB on large meshes) take tens of milliseconds to
// release, and nothing waits for them.  Without a thread it is freed here.
static void release_mesh(Mesh &&m) {
  try {
    std::thread([](Mesh) {}, std::move(m)).detach();
  } catch (const std::exception &) {
  }
}

static int remirror_with_state(nmx_ctx *c, const Mesh &old) {
  const double dt_old = c->dt;
  if (P(c, "dt") == "auto") c->dt = mesh_dt(c->mesh, Pd(c, "dtfac"), Pd(c, "uniform_dtfac"));
  if (c->device < 0) return NMX_OK;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  double *uold = c->dU;
  c->dU = nullptr;                 // kept across the re-upload
  int r = upload_mesh(c);
  if (!r) {
    r = adapt_state(c, old, uold);
    (void)hipFree(uold);
    return r;
  }
  const std::string msg = nmx_last_error();
  c->mesh = old;
  c->dt = dt_old;
  if (upload_mesh(c) == NMX_OK)
    (void)hipMemcpy(c->dU, uold, c->nstate * sizeof(double), hipMemcpyDeviceToDevice);
  (void)hipFree(uold);
  return fail(r, msg);
}

extern "C" int nmx_mesh_hadapt(nmx_ctx *c, const int *flag, int nelem) {
  if (!c || !flag) return fail(NMX_EINVAL, "NULL argument");
  if (!c->have_mesh) return fail(NMX_ESTATE, "no mesh");
  if (nelem != (int)c->mesh.e.size()) return fail(NMX_EINVAL, "need one flag per element");
  for (const Elm &E : c->mesh.e)
    if (E.fv) return fail(NMX_EUNSUP, "adapting meshes with FV elements is not built");
  Mesh nw = mesh_skeleton(c->mesh);
  std::string err = hadapt(nw, std::vector<int>(flag, flag + nelem));
  if (!err.empty()) return fail(NMX_EINVAL, err);
  Mesh old = std::exchange(c->mesh, std::move(nw));
  const int r = remirror_with_state(c, old);
  release_mesh(std::move(old));
  return r;
}

extern "C" int nmx_mesh_padapt(nmx_ctx *c, const int *n, int nelem) {
  if (!c || !n) return fail(NMX_EINVAL, "NULL argument");
  if (!c->have_mesh) return fail(NMX_ESTATE, "no mesh");
  if (nelem != (int)c->mesh.e.size()) return fail(NMX_EINVAL, "need one n per element");
  for (const Elm &E : c->mesh.e)
    if (E.fv) return fail(NMX_EUNSUP, "adapting meshes with FV elements is not built");
  for (int i = 0; i < nelem; ++i)
    if (n[i] < 2 || n[i] > 10) return fail(NMX_EUNSUP, "the device kernels take 2 to 10 points per direction");
  Mesh nw = mesh_skeleton(c->mesh);
  std::string err = set_elm_n(nw, std::vector<int>(n, n + nelem));
  if (!err.empty()) return fail(NMX_EINVAL, err);
  Mesh old = std::exchange(c->mesh, std::move(nw));
  const int r = remirror_with_state(c, old);
  release_mesh(std::move(old));
  return r;
}

extern "C" int nmx_mesh_switch_fv(nmx_ctx *c, const int *to_fv, int nelem) {
  if (!c || !to_fv) return fail(NMX_EINVAL, "NULL argument");
  if (!c->have_mesh) return fail(NMX_ESTATE, "no mesh");
  if (nelem != (int)c->mesh.e.size()) return fail(NMX_EINVAL, "need one flag per element");
  for (int i = 0; i < nelem; ++i)
    if (to_fv[i] && c->mesh.e[i].n < 3) return fail(NMX_EUNSUP, "FV elements need n >= 3 points per direction");
  Mesh nw = mesh_skeleton(c->mesh);
  std::string err = set_fv_elems(nw, std::vector<int>(to_fv, to_fv + nelem));
  if (!err.empty()) return fail(NMX_EINVAL, err);
  Mesh old = std::exchange(c->mesh, std::move(nw));
  const int r = remirror_with_state(c, old);
  release_mesh(std::move(old));
  return r;
}

extern "C" int nmx_mesh_set_elm_n(nmx_ctx *c, const int *n, int nelem) {
  if (!c || !n) return fail(NMX_EINVAL, "NULL argument");
  if (!c->have_mesh) return fail(NMX_ESTATE, "no mesh");
  std::vector<int> v(n, n + nelem);
  std::string err = set_elm_n(c->mesh, v);
  if (!err.empty()) return fail(NMX_EINVAL, err);
  if (P(c, "dt") == "auto") c->dt = mesh_dt(c->mesh, Pd(c, "dtfac"), Pd(c, "uniform_dtfac"));
  if (c->device < 0) return NMX_OK;
  return upload_mesh(c);
}

#define NEED_MESH(c)                                        \
  do {                                                      \
    if (!(c)) return fail(NMX_EINVAL, "ctx is NULL");      \
    if (!(c)->have_mesh) return fail(NMX_ESTATE, "no mesh"); \
  } while (0)

#define NEED_DEV(c)                                                              \
  do {                                                                           \
    NEED_MESH(c);                                                                \
    if ((c)->device < 0) return fail(NMX_ESTATE, "host-only context has no device"); \
    if (!(c)->dU) return fail(NMX_ESTATE, "no device mirror");                   \
  } while (0)

extern "C" int nmx_mesh_nelem(nmx_ctx *c, int *ne) {
  NEED_MESH(c);
  *ne = (int)c->mesh.e.size();
  return NMX_OK;
}

extern "C" int nmx_mesh_nnodes(nmx_ctx *c, long *nn) {
  NEED_MESH(c);
  *nn = c->mesh.nnodes;
  return NMX_OK;
}

extern "C" int nmx_mesh_elm_info(nmx_ctx *c, int e, long *info) {
  NEED_MESH(c);
  if (e < 0 || e >= (int)c->mesh.e.size()) return fail(NMX_EINVAL, "element out of range");
  const Elm &E = c->mesh.e[e];
  info[0] = E.p; info[1] = E.l; info[2] = E.n; info[3] = E.off;
  for (int f = 0; f < 6; ++f) {
    info[4 + f] = E.nfnb[f];
    info[10 + f] = E.ajmode[f];
    info[16 + f] = E.orient[f];
  }
  return NMX_OK;
}

extern "C" int nmx_mesh_dt_device(nmx_ctx *c, double *dt) {
  NEED_DEV(c);
  if (!dt) return fail(NMX_EINVAL, "NULL argument");
  const int ne = (int)c->mesh.e.size();
  std::vector<int> fv(ne);
  for (int e = 0; e < ne; ++e) fv[e] = c->mesh.e[e].fv != 0;
  const int grid = dt_grid(ne);
  int *dfv = nullptr;
  double *dpart = nullptr;
  int r;
  if ((r = dev_upload(&dfv, fv)) || (r = dev_alloc(&dpart, grid + 1))) {
    (void)hipFree(dfv);
    return r;
  }
  DtArgs a;
  a.xyz = c->dxyz; a.eoff = c->deoff; a.en = c->den; a.efv = dfv;
  a.dtfac = Pd(c, "dtfac"); a.uniform_dtfac = Pd(c, "uniform_dtfac");
  a.part = dpart; a.out = dpart + grid; a.nelem = ne;
  hipError_t err = launch_dt(a, grid, c->stream);
  if (err == hipSuccess) err = hipMemcpyAsync(dt, dpart + grid, sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (err == hipSuccess) err = hipStreamSynchronize(c->stream);
  (void)hipFree(dfv);
  (void)hipFree(dpart);
  if (err != hipSuccess) return fail(NMX_EHIP, std::string("dt reduction: ") + hipGetErrorString(err));
  return NMX_OK;
}

extern "C" int nmx_mesh_fv_flags(nmx_ctx *c, int *fv, int nelem) {
  NEED_MESH(c);
  if (!fv) return fail(NMX_EINVAL, "NULL argument");
  if (nelem != (int)c->mesh.e.size()) return fail(NMX_EINVAL, "need one flag per element");
  for (int e = 0; e < nelem; ++e) fv[e] = c->mesh.e[e].fv != 0;
  return NMX_OK;
}

extern "C" int nmx_mesh_elm_bbox(nmx_ctx *c, int e, double *bb) {
  NEED_MESH(c);
  if (e < 0 || e >= (int)c->mesh.e.size()) return fail(NMX_EINVAL, "element out of range");
  std::memcpy(bb, c->mesh.e[e].bbox, 6 * sizeof(double));
  return NMX_OK;
}

extern "C" int nmx_mesh_elm_fnb(nmx_ctx *c, int e, int f, int *nb, int *nbf, int *nnb) {
  NEED_MESH(c);
  if (e < 0 || e >= (int)c->mesh.e.size() || f < 0 || f > 5) return fail(NMX_EINVAL, "out of range");
  const Elm &E = c->mesh.e[e];
  for (int i = 0; i < E.nfnb[f]; ++i) {
    nb[i] = E.fnb[f][i];
    if (nbf) nbf[i] = E.nbf[f][i];
  }
  *nnb = E.nfnb[f];
  return NMX_OK;
}

extern "C" int nmx_mesh_coords(nmx_ctx *c, double *x) {
  NEED_MESH(c);
  std::memcpy(x, c->mesh.x.data(), c->mesh.x.size() * sizeof(double));
  return NMX_OK;
}

extern "C" int nmx_mesh_geometry(nmx_ctx *c, double *dXdx, double *sg) {
  NEED_MESH(c);
  if (dXdx) std::memcpy(dXdx, c->mesh.dXdx.data(), c->mesh.dXdx.size() * sizeof(double));
  if (sg) std::memcpy(sg, c->mesh.sqrtgdiag.data(), c->mesh.sqrtgdiag.size() * sizeof(double));
  return NMX_OK;
}

extern "C" int nmx_get_dt(nmx_ctx *c, double *dt) { NEED_MESH(c); *dt = c->dt; return NMX_OK; }
// an explicit dt: recorded as the parameter, so a later mesh change or a
// change of dtfac keeps it (the auto rule applies only while dt = auto)
extern "C" int nmx_set_dt(nmx_ctx *c, double dt) {
  NEED_MESH(c);
  if (!(dt > 0.)) return fail(NMX_EINVAL, "dt must be positive");
  char buf[32];
  std::snprintf(buf, sizeof buf, "%.17g", dt);
  c->par["dt"] = buf;
  c->dt = dt;
  return NMX_OK;
}
extern "C" int nmx_get_time(nmx_ctx *c, double *t) { NEED_MESH(c); *t = c->time; return NMX_OK; }
extern "C" int nmx_set_time(nmx_ctx *c, double t) {
  NEED_MESH(c);
  c->time = t; c->Te = t; c->te = 0;
  return NMX_OK;
}

static double *reg_ptr(nmx_ctx *c, int reg) {
  switch (reg) {
    case NMX_REG_U: return c->dU;
    case NMX_REG_R: return c->dR;
    case NMX_REG_P: return c->dP;
    case NMX_REG_ADJ: return c->dL;
    default: return nullptr;
  }
}

// flat [var][node] <-> element blocks [e][var][np]
static void flat_to_blocks(const Mesh &M, const double *u, double *b) {
  long nn = M.nnodes;
  for (const Elm &E : M.e) {
    long np = (long)E.n * E.n * E.n;
    for (int v = 0; v < 5; ++v) std::memcpy(b + 5 * E.off + v * np, u + v * nn + E.off, np * 8);
  }
}
static void blocks_to_flat(const Mesh &M, const double *b, double *u) {
  long nn = M.nnodes;
  for (const Elm &E : M.e) {
    long np = (long)E.n * E.n * E.n;
    for (int v = 0; v < 5; ++v) std::memcpy(u + v * nn + E.off, b + 5 * E.off + v * np, np * 8);
  }
}

// pinned host staging of one register in element blocks: one copy moves
// the whole state
static int host_stage(nmx_ctx *c, double **h) {
  if (c->hstage_n < c->nstate) {
    if (c->hstage) (void)hipHostFree(c->hstage);
    c->hstage = nullptr;
    c->hstage_n = 0;
    HIPCHK(hipHostMalloc((void **)&c->hstage, c->nstate * sizeof(double), hipHostMallocDefault));
    c->hstage_n = c->nstate;
  }
  *h = c->hstage;
  return NMX_OK;
}

extern "C" int nmx_state_upload(nmx_ctx *c, int reg, const double *u) {
  NEED_DEV(c);
  double *d = reg_ptr(c, reg), *b = nullptr;
  if (!d || !u) return fail(NMX_EINVAL, "bad register or NULL state");
  if (int r = host_stage(c, &b)) return r;
  HIPCHK(hipStreamSynchronize(c->stream));   // the staging buffer is free
  flat_to_blocks(c->mesh, u, b);
  HIPCHK(hipMemcpyAsync(d, b, c->nstate * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return NMX_OK;
}

extern "C" int nmx_state_download(nmx_ctx *c, int reg, double *u) {
  NEED_DEV(c);
  double *d = reg_ptr(c, reg), *b = nullptr;
  if (!d || !u) return fail(NMX_EINVAL, "bad register or NULL state");
  if (int r = host_stage(c, &b)) return r;
  HIPCHK(hipMemcpyAsync(b, d, c->nstate * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  blocks_to_flat(c->mesh, b, u);
  return NMX_OK;
}

// every variable of every element through per-element pointers, as nmesh
// holds them (Vard(node, vi), nmesh_amr_loops.h:181): gathered into the
// pinned staging in element-block order, then one copy (and the reverse)
extern "C" int nmx_state_upload_ptrs(nmx_ctx *c, int reg, const double *const *ptrs, int nelem, int nvar) {
  NEED_DEV(c);
  double *d = reg_ptr(c, reg), *b = nullptr;
  if (!d || !ptrs) return fail(NMX_EINVAL, "bad register or NULL pointer table");
  if (nelem != (int)c->mesh.e.size() || nvar != 5) return fail(NMX_EINVAL, "need nelem elements of 5 variables");
  for (long i = 0; i < 5L * nelem; ++i)
    if (!ptrs[i]) return fail(NMX_EINVAL, "NULL variable pointer at element " + std::to_string(i / 5));
  if (int r = host_stage(c, &b)) return r;
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int e = 0; e < nelem; ++e) {
    const Elm &E = c->mesh.e[e];
    const long np = (long)E.n * E.n * E.n;
    for (int v = 0; v < 5; ++v) std::memcpy(b + 5 * E.off + v * np, ptrs[5L * e + v], np * 8);
  }
  HIPCHK(hipMemcpyAsync(d, b, c->nstate * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return NMX_OK;
}

extern "C" int nmx_state_download_ptrs(nmx_ctx *c, int reg, double *const *ptrs, int nelem, int nvar) {
  NEED_DEV(c);
  double *d = reg_ptr(c, reg), *b = nullptr;
  if (!d || !ptrs) return fail(NMX_EINVAL, "bad register or NULL pointer table");
  if (nelem != (int)c->mesh.e.size() || nvar != 5) return fail(NMX_EINVAL, "need nelem elements of 5 variables");
  for (long i = 0; i < 5L * nelem; ++i)
    if (!ptrs[i]) return fail(NMX_EINVAL, "NULL variable pointer at element " + std::to_string(i / 5));
  if (int r = host_stage(c, &b)) return r;
  HIPCHK(hipMemcpyAsync(b, d, c->nstate * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int e = 0; e < nelem; ++e) {
    const Elm &E = c->mesh.e[e];
    const long np = (long)E.n * E.n * E.n;
    for (int v = 0; v < 5; ++v) std::memcpy(ptrs[5L * e + v], b + 5 * E.off + v * np, np * 8);
  }
  return NMX_OK;
}

extern "C" int nmx_state_upload_elm(nmx_ctx *c, int reg, int e, int v, const double *d, int np) {
  NEED_DEV(c);
  double *p = reg_ptr(c, reg);
  if (!p || e < 0 || e >= (int)c->mesh.e.size() || v < 0 || v > 4) return fail(NMX_EINVAL, "out of range");
  const Elm &E = c->mesh.e[e];
  long enp = (long)E.n * E.n * E.n;
  if (np != enp) return fail(NMX_EINVAL, "np does not match the element");
  HIPCHK(hipMemcpyAsync(p + 5 * E.off + v * enp, d, enp * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return NMX_OK;
}

extern "C" int nmx_state_download_elm(nmx_ctx *c, int reg, int e, int v, double *d, int np) {
  NEED_DEV(c);
  double *p = reg_ptr(c, reg);
  if (!p || e < 0 || e >= (int)c->mesh.e.size() || v < 0 || v > 4) return fail(NMX_EINVAL, "out of range");
  const Elm &E = c->mesh.e[e];
  long enp = (long)E.n * E.n * E.n;
  if (np != enp) return fail(NMX_EINVAL, "np does not match the element");
  HIPCHK(hipMemcpyAsync(d, p + 5 * E.off + v * enp, enp * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return NMX_OK;
}

static void kvec(nmx_ctx *c, double k[3]) {
  k[0] = k[1] = k[2] = 0;
  std::sscanf(P(c, "scalarwave1_k").c_str(), "%lg %lg %lg", &k[0], &k[1], &k[2]);
}

extern "C" int nmx_state_set_profile(nmx_ctx *c, int reg, double t) {
  NEED_MESH(c);
  double k[3];
  kvec(c, k);
  double om = std::sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
  long nn = c->mesh.nnodes;
  std::vector<double> u(5 * nn);
  for (long g = 0; g < nn; ++g) {
    double arg = k[0] * c->mesh.x[g] + k[1] * c->mesh.x[nn + g] + k[2] * c->mesh.x[2 * nn + g] - om * t;
    double s = std::sin(arg), co = std::cos(arg);
    u[4 * nn + g] = s;
    u[0 * nn + g] = -om * co;
    u[1 * nn + g] = k[0] * co;
    u[2 * nn + g] = k[1] * co;
    u[3 * nn + g] = k[2] * co;
  }
  return nmx_state_upload(c, reg, u.data());
}

// ---------------------------------------------------------------------------
// stage launches
// ---------------------------------------------------------------------------

static StageArgs base_args(nmx_ctx *c) {
  StageArgs a;
  a.dXdx = c->dXdx;
  a.dXbdX = c->dXbdX;
  a.xyz = c->dxyz;
  a.face = c->dface;
  a.D = c->dD;
  a.tdev = c->dT;
  a.egeo = c->degeo;
  a.edom = c->dedom;
  a.xb = c->dxb;
  a.xcd_map = (P(c, "nmx_xcd") == "yes") ? 1 : 0;
  a.wend_inv = c->np1 ? 1.0 / lgl_basis(c->np1).w[0] : 0.0;
  kvec(c, a.k);
  a.om = std::sqrt(a.k[0] * a.k[0] + a.k[1] * a.k[1] + a.k[2] * a.k[2]);
  a.profile = (a.k[0] != 0 || a.k[1] != 0 || a.k[2] != 0);
  a.numflux = (P(c, "scalarwave1_numflux") == "LLF") ? 1 : 0;
  a.nt = Pi(c, "nmx_stream_hints") & 7;
  a.e0 = 0;
  a.e1 = (int)c->mesh.e.size();
  return a;
}

static AmrArgs amr_args(nmx_ctx *c, const StageArgs &sa) {
  AmrArgs a;
  a.win = sa.win; a.U = sa.U; a.Ain = sa.Ain; a.Aout = sa.Aout; a.Wout = sa.Wout; a.R = sa.R;
  a.aj = c->daj;
  a.mdesc = c->dmdesc;
  a.dXdx = c->dXdx; a.xyz = c->dxyz; a.tdev = sa.tdev; a.tidx = sa.tidx;
  a.cA = sa.cA; a.cW = sa.cW;
  for (int d = 0; d < 3; ++d) a.k[d] = sa.k[d];
  a.om = sa.om; a.wmode = sa.wmode; a.numflux = sa.numflux; a.profile = sa.profile; a.nt = sa.nt;
  return a;
}

// one order bucket's stage kernel with tile div (1, 2 or 4 nodes per thread)
static int launch_bucket(nmx_ctx *c, AmrArgs a, const nmx_ctx::Bucket &b, int div, int xcd_map,
                         hipStream_t st) {
  a.D = c->dDall + b.doff;
  a.xb = c->dxball + b.xoff;
  a.wend_inv = 1.0 / lgl_basis(b.n).w[0];
  a.rec = b.drec;
  a.nlist = b.count;
  a.xcd_map = xcd_map;
  int grid = a.xcd_map ? 8 * ((b.count + 7) / 8) : b.count;
  hipError_t err;
  if (b.fv) {
    const int rec = fv_rec_pars(c, a.recend, a.rec_s1, a.rec_s2, a.rec_opt, a.rec_on);
    a.rec_first = P(c, "fv_rec") == "1";
    a.fv_mid = P(c, "coordinates_midpoint_data") != "no";
    fv_extrap_pars(c, a.ex_s1, a.ex_s2, a.ex_opt, a.ex_on);
    err = launch_fvelm_amr(b.n, rec, c->geom == 0 ? 0 : 2, a, c->dwcall + b.woff, grid, st);
  } else {
    err = launch_dg_amr(b.n, c->geom, a, grid, div, st);
  }
  if (err != hipSuccess) return fail(NMX_EHIP, std::string("stage launch: ") + hipGetErrorString(err));
  return NMX_OK;
}

static BatchArgs batch_args(nmx_ctx *c, const StageArgs &sa) {
  BatchArgs b;
  b.win = sa.win; b.U = sa.U; b.Ain = sa.Ain; b.Aout = sa.Aout; b.Wout = sa.Wout; b.R = sa.R;
  b.rec = c->dbrec; b.jobs = c->drjobs; b.Btab = c->dBtab;
  b.dXdx = c->dXdx; b.xyz = c->dxyz; b.Dall = c->dbD; b.xball = c->dbxb;
  b.tdev = sa.tdev; b.tidx = sa.tidx;
  b.wend_inv = c->dbW; b.xlo = c->dbxlo;
  b.cA = sa.cA; b.cW = sa.cW;
  for (int d = 0; d < 3; ++d) b.k[d] = sa.k[d];
  b.om = sa.om; b.wmode = sa.wmode; b.numflux = sa.numflux; b.profile = sa.profile; b.nt = sa.nt;
  return b;
}

// the persistent launch of every DG element of a batched mesh (one per stage)
static int launch_batch(nmx_ctx *c, const StageArgs &sa) {
  hipError_t err = launch_dg_batch(c->geom, batch_args(c, sa), c->bgrid, c->stream);
  if (err != hipSuccess) return fail(NMX_EHIP, std::string("batch launch: ") + hipGetErrorString(err));
  return NMX_OK;
}

// the DG elements of order n in one launch of the order's own kernel, with
// tile (launch_dg_bucket)
static int launch_order(nmx_ctx *c, const StageArgs &sa, int n, int div, hipStream_t st) {
  BatchArgs b = batch_args(c, sa);
  b.rec = c->dbrec2;
  b.xlo = c->dbnr + 2 * n;
  const int count = c->bcount[n];
  const int grid = 8 * ((std::min(count, 8 * c->ncu) + 7) / 8);
  hipError_t err = launch_dg_bucket(n, c->geom, b, grid, div, st);
  if (err != hipSuccess) return fail(NMX_EHIP, std::string("bucket launch: ") + hipGetErrorString(err));
  return NMX_OK;
}

static int launch_amr(nmx_ctx *c, const StageArgs &sa) {
  const bool perelm = (P(c, "nmx_path") == "perelement");
  AmrArgs a = amr_args(c, sa);
  if (!perelm && P(c, "nmx_batch") == "buckets") {
    // one launch per order, side by side on the stream lanes (FV buckets
    // with their face kernel)
    std::string blk = P(c, "nmx_block");
    const int div0 = blk == "half" ? 2 : blk == "quarter" ? 4 : 1;
    const bool fork = c->buckets.size() > 1 && P(c, "nmx_concurrent") == "yes";
    auto lane_stream = [c](int lane) { return lane == 0 ? c->stream : c->helper[lane - 1]; };
    if (fork) {
      HIPCHK(hipEventRecord(c->evfork, c->stream));
      for (int l = 1; l < nmx_ctx::kLanes; ++l) HIPCHK(hipStreamWaitEvent(lane_stream(l), c->evfork, 0));
    }
    for (size_t k = 0; k < c->buckets.size(); ++k) {
      const int bi = c->border[k];
      const auto &b = c->buckets[bi];
      hipStream_t st = fork ? lane_stream(c->blane[bi]) : c->stream;
      if (b.fv) {
        FaceArgs fa;
        fa.win = sa.win; fa.Btab = c->dBtab; fa.aj = c->daj;
        fa.jobs = c->djobs1 + (long)kJobInts * b.f1first; fa.nfaces = b.f1count;
        hipError_t err = launch_face_interp(fa, st);
        if (err != hipSuccess) return fail(NMX_EHIP, std::string("face launch: ") + hipGetErrorString(err));
        if (int r = launch_bucket(c, a, b, 1, sa.xcd_map, st)) return r;
      } else {
        const int tile = c->bdiv[b.n] ? c->bdiv[b.n] : (div0 >= 2 ? 2 : 1);
        if (int r = launch_order(c, sa, b.n, tile, st)) return r;
      }
    }
    if (fork)
      for (int l = 1; l < nmx_ctx::kLanes; ++l) {
        HIPCHK(hipEventRecord(c->evjoin[l - 1], lane_stream(l)));
        HIPCHK(hipStreamWaitEvent(c->stream, c->evjoin[l - 1], 0));
      }
    return NMX_OK;
  }
  if (!perelm && P(c, "nmx_batch") == "persistent") {
    // FV buckets (faces from the face kernel), then every DG element at once
    std::string blk = P(c, "nmx_block");
    for (size_t k = 0; k < c->buckets.size(); ++k) {
      const auto &b = c->buckets[c->border[k]];
      if (!b.fv) continue;
      FaceArgs fa;
      fa.win = sa.win; fa.Btab = c->dBtab; fa.aj = c->daj;
      fa.jobs = c->djobs1 + (long)kJobInts * b.f1first; fa.nfaces = b.f1count;
      hipError_t err = launch_face_interp(fa, c->stream);
      if (err != hipSuccess) return fail(NMX_EHIP, std::string("face launch: ") + hipGetErrorString(err));
      if (int r = launch_bucket(c, a, b, 1, sa.xcd_map, c->stream)) return r;
    }
    return c->anydg ? launch_batch(c, sa) : NMX_OK;
  }
  std::string blk = P(c, "nmx_block");
  const int div0 = blk == "half" ? 2 : blk == "quarter" ? 4 : 1;
  // interpolated faces of one bucket's elements, just before its stage kernel
  auto faces = [&](const nmx_ctx::Bucket &b, hipStream_t st) -> int {
    FaceArgs fa;
    fa.win = sa.win; fa.Btab = c->dBtab; fa.aj = c->daj;
    fa.jobs = c->djobs1 + (long)kJobInts * b.f1first; fa.nfaces = b.f1count;
    hipError_t err = launch_face_interp(fa, st);
    if (err != hipSuccess) return fail(NMX_EHIP, std::string("face launch: ") + hipGetErrorString(err));
    return NMX_OK;
  };
  // buckets run concurrently on kLanes stream lanes: lane 0 is the context's
  // stream, the others helpers forked from it (parallel branches when
  // captured into a graph); the longest buckets start first, one per lane,
  // the rest queue on the least loaded lane.  A bucket's faces read only the
  // stage input, so each lane computes its own buckets' faces and the face
  // work overlaps other lanes' buckets.
  const bool fork = !perelm && c->buckets.size() > 1 && P(c, "nmx_concurrent") == "yes";
  auto lane_stream = [c](int lane) { return lane == 0 ? c->stream : c->helper[lane - 1]; };
  if (fork) {
    HIPCHK(hipEventRecord(c->evfork, c->stream));
    for (int l = 1; l < nmx_ctx::kLanes; ++l) HIPCHK(hipStreamWaitEvent(lane_stream(l), c->evfork, 0));
  }
  for (size_t k = 0; k < c->buckets.size(); ++k) {
    const int bi = perelm ? (int)k : c->border[k];
    const auto &b = c->buckets[bi];
    hipStream_t st = fork ? lane_stream(c->blane[bi]) : c->stream;
    a.D = c->dDall + b.doff;
    a.xb = c->dxball + b.xoff;
    a.wend_inv = 1.0 / lgl_basis(b.n).w[0];
    const int div = b.div ? b.div : div0;
    if (int r = faces(b, st)) return r;
    if (perelm) {
      if (b.fv) return fail(NMX_EUNSUP, "FV elements run on the plain or batched path");
      a.xcd_map = 0;
      a.nlist = 1;
      for (int q = 0; q < b.count; ++q) {
        a.rec = b.drec + q;
        hipError_t err = launch_dg_amr(b.n, c->geom, a, 1, div, c->stream);
        if (err != hipSuccess) return fail(NMX_EHIP, std::string("stage launch: ") + hipGetErrorString(err));
      }
      continue;
    }
    int r = launch_bucket(c, a, b, div, sa.xcd_map, st);
    if (r) return r;
  }
  if (fork)
    for (int l = 1; l < nmx_ctx::kLanes; ++l) {
      HIPCHK(hipEventRecord(c->evjoin[l - 1], lane_stream(l)));
      HIPCHK(hipStreamWaitEvent(c->stream, c->evjoin[l - 1], 0));
    }
  return NMX_OK;
}

static int launch(nmx_ctx *c, StageArgs a) {
  if (c->amr) return launch_amr(c, a);
  if (c->fvelm) {
    int nel = a.e1 - a.e0;
    int grid = a.xcd_map ? 8 * ((nel + 7) / 8) : nel;
    const int rec = fv_rec_pars(c, a.recend, a.rec_s1, a.rec_s2, a.rec_opt, a.rec_on);
    a.rec_first = P(c, "fv_rec") == "1";
    a.fv_mid = P(c, "coordinates_midpoint_data") != "no";
    fv_extrap_pars(c, a.ex_s1, a.ex_s2, a.ex_opt, a.ex_on);
    hipError_t err = launch_fvelm_stage(c->np1, rec, c->geom == 0 ? 0 : 2, a, c->dwc, grid, c->stream);
    if (err != hipSuccess) return fail(NMX_EHIP, std::string("fv stage launch: ") + hipGetErrorString(err));
    return NMX_OK;
  }
  bool perelm = (P(c, "nmx_path") == "perelement");
  std::string blk = P(c, "nmx_block");
  int nt = dg_stage_threads(c->np1, blk == "half" ? 2 : blk == "quarter" ? 4 : 1);
  if (perelm) {
    int ne = a.e1;
    a.xcd_map = 0;
    for (int e = 0; e < ne; ++e) {
      a.e0 = e; a.e1 = e + 1;
      hipError_t err = launch_dg_stage(c->np1, c->geom, a, 1, nt, c->stream);
      if (err != hipSuccess) return fail(NMX_EHIP, std::string("stage launch: ") + hipGetErrorString(err));
    }
    return NMX_OK;
  }
  int nel = a.e1 - a.e0;
  int grid = a.xcd_map ? 8 * ((nel + 7) / 8) : nel;
  hipError_t err = launch_dg_stage(c->np1, c->geom, a, grid, nt, c->stream);
  if (err != hipSuccess) return fail(NMX_EHIP, std::string("stage launch: ") + hipGetErrorString(err));
  return NMX_OK;
}

extern "C" int nmx_setrhs(nmx_ctx *c, int reg_out, int reg_in, double t) {
  NEED_DEV(c);
  double *out = reg_ptr(c, reg_out), *in = reg_ptr(c, reg_in);
  if (!out || !in || out == in) return fail(NMX_EINVAL, "bad registers");
  c->hT[0] = t;
  HIPCHK(hipMemcpyAsync(c->dT, c->hT, sizeof(double), hipMemcpyHostToDevice, c->stream));
  StageArgs a = base_args(c);
  a.win = in;
  a.R = out;
  a.tidx = 0;
  return launch(c, a);
}

static int method_id(nmx_ctx *c) {
  std::string m = P(c, "evolve_method");
  return m == "RK4" ? 0 : m == "sspRK3" ? 1 : 2;
}

// The FILTER bin's exponential filter (evolve_filter_evosys_mesh,
// evolve.c:656-710): evolve_filter_all filters every evolved variable, then
// evolve_filter_varlist the ones it names (first components; a variable in
// both is filtered twice, as in the reference).  Names of variables this
// build does not hold are skipped, as MeshVarIndLax does for unknown names.
// Builds the per-order matrices when the parameters change; on = false when
// nothing is filtered.
static int filter_prepare(nmx_ctx *c, bool &on) {
  on = false;
  c->filt_mask[0] = P(c, "evolve_filter_all") == "yes" ? 31 : 0;
  c->filt_mask[1] = 0;
  std::istringstream vl(P(c, "evolve_filter_varlist"));
  for (std::string name; vl >> name;) {
    if (name == "scalarwave1_pi") c->filt_mask[1] |= 1;
    else if (name == "scalarwave1_cx") c->filt_mask[1] |= 2 | 4 | 8;
    else if (name == "scalarwave1_phi") c->filt_mask[1] |= 16;
  }
  if (!c->filt_mask[0] && !c->filt_mask[1]) return NMX_OK;
  if (c->anyfv) {
    if (P(c, "basis_filter_fv") == "yes")
      return fail(NMX_EUNSUP, "the filter on FV elements is not built (set basis_filter_fv = no)");
    if (c->fvelm) return NMX_OK;   // FV elements are left unfiltered (filter.c:79-80)
  }
  c->filt_jpow = Pi(c, "basis_expfilter_JacobianPower");
  const double alp = Pd(c, "evolve_filter_alp"), sf = Pd(c, "evolve_filter_s");
  const int dn = Pi(c, "evolve_filter_dn");
  char key[96];
  std::snprintf(key, sizeof key, "%.17g %.17g %d", alp, sf, dn);
  if (c->filt_key != key) {
    constexpr int kF = kMaxN * kMaxN;
    std::vector<double> F(9 * kF, 0.0);
    for (int n = 2; n <= 10; ++n) filter_matrix(n, alp, sf, dn, F.data() + (n - 2) * kF);
    if (c->dFilt) { (void)hipFree(c->dFilt); c->dFilt = nullptr; }
    int r = dev_upload(&c->dFilt, F);
    if (r) return r;
    c->filt_key = key;
  }
  on = true;
  return NMX_OK;
}

// the filter passes on register reg (after filter_prepare)
static int launch_filters(nmx_ctx *c, double *reg) {
  constexpr int kF = kMaxN * kMaxN;
  for (int p = 0; p < 2; ++p) {
    if (!c->filt_mask[p]) continue;
    FilterArgs fa;
    fa.u = reg; fa.dXdx = c->dXdx; fa.varmask = c->filt_mask[p]; fa.jpow = c->filt_jpow;
    hipError_t err = hipSuccess;
    if (c->amr) {
      for (const auto &b : c->buckets) {
        if (b.fv) continue;   // left unfiltered (basis_filter_fv = no, filter.c:79-80)
        fa.rec = b.drec; fa.nlist = b.count;
        fa.F = c->dFilt + (b.n - 2) * kF; fa.xb = c->dxball + b.xoff;
        err = launch_filter(b.n, c->geom, fa, c->stream);
        if (err != hipSuccess) break;
      }
    } else {
      fa.rec = c->dfrec; fa.nlist = (int)c->mesh.e.size();
      fa.F = c->dFilt + (c->np1 - 2) * kF; fa.xb = c->dxb;
      err = launch_filter(c->np1, c->geom, fa, c->stream);
    }
    if (err != hipSuccess) return fail(NMX_EHIP, std::string("filter launch: ") + hipGetErrorString(err));
  }
  return NMX_OK;
}

extern "C" int nmx_filter(nmx_ctx *c, int reg) {
  NEED_DEV(c);
  double *u = reg_ptr(c, reg);
  if (!u) return fail(NMX_EINVAL, "bad register");
  bool on = false;
  int r = filter_prepare(c, on);
  if (r || !on) return r;
  return launch_filters(c, u);
}

// ---------------------------------------------------------------------------
// checkpoint: variables.bin and the "vars" CRC of nmesh_CRCs
// ---------------------------------------------------------------------------

extern "C" uint64_t nmx_crc64(uint64_t crc, const void *buf, size_t nbytes) {
  return crc64(crc, buf, nbytes);
}

extern "C" uint64_t nmx_crc64_combine(uint64_t crc1, uint64_t crc2, uint64_t nbytes2) {
  return crc64_combine(crc1, crc2, nbytes2);
}

extern "C" int nmx_vars_crc(nmx_ctx *c, int reg, uint64_t *crc) {
  NEED_DEV(c);
  double *u = reg_ptr(c, reg);
  if (!u || !crc) return fail(NMX_EINVAL, "bad register or NULL result");
  if (c->mesh.nnodes >= (1L << 32)) return fail(NMX_EUNSUP, "the device CRC takes fewer than 2^32 nodes");
  HIPCHK(hipSetDevice(c->device));
  int r;
  if (!c->dcrctab) {
    std::vector<uint64_t> tab, vtab, klane, aw, ypow;
    crc_device_tables(tab, vtab, klane, aw, ypow);
    if ((r = dev_upload(&c->dcrctab, tab)) || (r = dev_upload(&c->dcrcvtab, vtab)) ||
        (r = dev_upload(&c->dcrcklane, klane)) || (r = dev_upload(&c->dcrcaw, aw)) ||
        (r = dev_upload(&c->dcrcypow, ypow)) || (r = dev_alloc(&c->dcrcpart, c->mesh.e.size())) ||
        (r = dev_alloc(&c->dcrcout, 1)))
      return r;
  }
  CrcArgs a;
  a.u = u; a.eoff = c->deoff; a.en = c->den; a.nelem = (int)c->mesh.e.size(); a.nnodes = c->mesh.nnodes;
  a.tab = c->dcrctab; a.vtab = c->dcrcvtab; a.klane = c->dcrcklane; a.aw = c->dcrcaw;
  a.ypow = c->dcrcypow; a.part = c->dcrcpart;
  hipError_t err = launch_vars_crc(a, c->dcrcout, c->stream);
  if (err != hipSuccess) return fail(NMX_EHIP, std::string("crc launch: ") + hipGetErrorString(err));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(crc, c->dcrcout, sizeof(uint64_t), hipMemcpyDeviceToHost));
  return NMX_OK;
}

// The same CRC as nmx_vars_crc with no device allocation: the tables,
// element contributions and result live in the static scratch set up by
// nmx_ctx_create.  Calls that overlap on one GPU race on that scratch.
extern "C" int nmx_vars_crc_noalloc(nmx_ctx *c, int reg, uint64_t *crc) {
  NEED_DEV(c);
  double *u = reg_ptr(c, reg);
  if (!u || !crc) return fail(NMX_EINVAL, "bad register or NULL result");
  if (c->mesh.nnodes >= (1L << 32)) return fail(NMX_EUNSUP, "the device CRC takes fewer than 2^32 nodes");
  if ((long)c->mesh.e.size() > kCrcNoAllocMaxElems)
    return fail(NMX_EUNSUP, "nmx_vars_crc_noalloc takes at most " + std::to_string(kCrcNoAllocMaxElems) +
                                " elements (its static scratch); use nmx_vars_crc");
  HIPCHK(hipSetDevice(c->device));
  uint64_t *p = crc_scratch(c->device);
  if (!p) return fail(NMX_ESTATE, "the static CRC scratch of this device is not set up");
  CrcArgs a;
  a.u = u; a.eoff = c->deoff; a.en = c->den; a.nelem = (int)c->mesh.e.size(); a.nnodes = c->mesh.nnodes;
  a.tab = p + kCrcScrTab; a.vtab = p + kCrcScrVtab; a.klane = p + kCrcScrKlane; a.aw = p + kCrcScrAw;
  a.ypow = p + kCrcScrYpow; a.part = p + kCrcScrPart;
  hipError_t err = launch_vars_crc(a, p + kCrcScrOut, c->stream);
  if (err != hipSuccess) return fail(NMX_EHIP, std::string("crc launch: ") + hipGetErrorString(err));
  HIPCHK(hipMemcpyAsync(crc, p + kCrcScrOut, sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return NMX_OK;
}

extern "C" int nmx_checkpoint_save_vars(nmx_ctx *c, int reg, const char *path) {
  NEED_DEV(c);
  double *u = reg_ptr(c, reg);
  if (!u || !path) return fail(NMX_EINVAL, "bad register or NULL path");
  HIPCHK(hipSetDevice(c->device));
  std::vector<double> h(c->nstate);
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(h.data(), u, c->nstate * sizeof(double), hipMemcpyDeviceToHost));
  std::string err = write_variables_bin(c->mesh, h.data(), path);
  if (!err.empty()) return fail(NMX_EIO, err);
  return NMX_OK;
}

extern "C" int nmx_checkpoint_load_vars(nmx_ctx *c, int reg, const char *path) {
  NEED_DEV(c);
  double *u = reg_ptr(c, reg);
  if (!u || !path) return fail(NMX_EINVAL, "bad register or NULL path");
  HIPCHK(hipSetDevice(c->device));
  std::vector<double> h(c->nstate, 0.0);
  std::string err = read_variables_bin(c->mesh, h.data(), path);
  if (!err.empty()) return fail(NMX_EINVAL, err);   // the register is left as it was
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(u, h.data(), c->nstate * sizeof(double), hipMemcpyHostToDevice));
  return NMX_OK;
}

// ---------------------------------------------------------------------------
// MRS limiter (scalarwave1_limiter = MRS)
// ---------------------------------------------------------------------------

// limiter on or off from the parameters; its tables on first use
static int nbr_tables(nmx_ctx *c);

static int limiter_prepare(nmx_ctx *c) {
  const std::string &lim = P(c, "scalarwave1_limiter");
  c->lim_on = lim == "MRS" || lim == "minmodB";
  if (!c->lim_on) return NMX_OK;
  if (c->anyfv) return fail(NMX_EUNSUP, "the limiters on FV elements are not built");
  c->lim_kind = lim == "minmodB" ? 1 : 0;
  c->lim_alpha = Pd(c, "limiter_alpha");
  c->lim_beta = Pd(c, "limiter_beta");
  c->lim_scale = P(c, "limiter_scaleBound") == "yes";
  return nbr_tables(c);
}

// face-neighbour lists, h^1.5, quadrature and Legendre tables, element data
// (the limiters' and the trouble indicator's), made once per mirror
static int nbr_tables(nmx_ctx *c) {
  if (c->dnbr) return NMX_OK;
  const Mesh &M = c->mesh;
  const size_t ne = M.e.size();
  std::vector<int> nbr(24 * ne, -1);
  std::vector<double> h15(ne), hw(ne), wq(9 * 10, 0.0), am(9 * 100, 0.0), sm(9 * 100, 0.0);
  for (size_t e = 0; e < ne; ++e) {
    const Elm &E = M.e[e];
    for (int f = 0; f < 6; ++f) {
      if (E.nfnb[f] > 4) return fail(NMX_EUNSUP, "more than 4 neighbours on a face");
      for (int ni = 0; ni < E.nfnb[f]; ++ni) nbr[24 * e + 4 * f + ni] = E.fnb[f][ni];
    }
    const double *bb = E.bbox;
    const double h = std::max(bb[1] - bb[0], std::max(bb[3] - bb[2], bb[5] - bb[4]));
    h15[e] = std::pow(h, 1.5);
    hw[e] = h;
  }
  for (int n = 2; n <= 10; ++n) {
    const Basis &b = lgl_basis(n);
    for (int i = 0; i < n; ++i) wq[(n - 2) * 10 + i] = b.w[i];
    legendre_analysis_synthesis(n, am.data() + (n - 2) * 100, sm.data() + (n - 2) * 100);
  }
  int r;
  if ((r = dev_upload(&c->dnbr, nbr)) || (r = dev_upload(&c->dlimh, h15)) || (r = dev_upload(&c->dlimhw, hw)) ||
      (r = dev_upload(&c->dwqall, wq)) || (r = dev_upload(&c->damat, am)) || (r = dev_upload(&c->dsmat, sm)) ||
      (r = dev_alloc(&c->dldat, 20 * ne)))
    return r;
  return NMX_OK;
}

// ---------------------------------------------------------------------------
// troubled elements: scores, switches and redone substeps (trouble.c,
// RungeKutta.c:19-188, evolve.c:25-176)
// ---------------------------------------------------------------------------

static bool trouble_on(nmx_ctx *c) { return P(c, "nmx_trouble") != "none"; }

// device arrays of the trouble machinery for this mirror; the scores come
// back from the host copy kept across re-mirrors
static int trouble_prepare(nmx_ctx *c) {
  if (int r = nbr_tables(c)) return r;
  const size_t ne = c->mesh.e.size();
  if (c->hscore.size() != ne) c->hscore.assign(ne, 0);
  if (c->dtr) return NMX_OK;
  std::vector<int> fv(ne);
  for (size_t e = 0; e < ne; ++e) fv[e] = c->mesh.e[e].fv != 0;
  int r;
  if ((r = dev_alloc(&c->dtr, ne)) || (r = dev_upload(&c->dscore, c->hscore)) || (r = dev_alloc(&c->dtrout, 1)) ||
      (r = dev_upload(&c->defv, fv)))
    return r;
  return NMX_OK;
}

// evolve_set_trouble_score_mesh (trouble.c:112-127) on state u against u_p
// (the TROUBLE values from the library's indicator or the caller's hook),
// or evolve_read_trouble_score_mesh when update is false: the mesh's score
static int trouble_score(nmx_ctx *c, const double *u, const double *up, bool update, int *score) {
  if (int r = trouble_prepare(c)) return r;
  const int ne = (int)c->mesh.e.size();
  if (update) {
    if (c->trouble_hook) {
      if (P(c, "nmx_trouble") != "hook") return fail(NMX_EINVAL, "a TROUBLE hook is set: nmx_trouble must be hook");
      if (int r = c->trouble_hook(c->trouble_user, u, up, c->dtr, ne, (void *)c->stream))
        return fail(NMX_EINVAL, "the TROUBLE hook returned " + std::to_string(r));
    } else {
      if (P(c, "nmx_trouble") != "RDMP") return fail(NMX_ESTATE, "nmx_trouble = hook needs nmx_set_trouble_hook");
      hipError_t err = hipSuccess;
      if (c->ldat_up != up) {
        LimArgs la;
        la.u = const_cast<double *>(up); la.eoff = c->deoff; la.en = c->den; la.wq = c->dwqall;
        la.ldat = c->dldat; la.nelem = ne;
        err = launch_lim_data(la, c->stream);
        if (err != hipSuccess) return fail(NMX_EHIP, std::string("trouble data: ") + hipGetErrorString(err));
        c->ldat_up = up;
      }
      TroubleArgs ta;
      ta.u = u; ta.ldat = c->dldat; ta.eoff = c->deoff; ta.en = c->den; ta.efv = c->defv; ta.nbr = c->dnbr;
      ta.tr = c->dtr; ta.nelem = ne;
      ta.delta0 = Pd(c, "nmx_trouble_delta0");
      ta.eps = Pd(c, "nmx_trouble_eps");
      err = launch_rdmp(ta, c->stream);
      if (err != hipSuccess) return fail(NMX_EHIP, std::string("trouble indicator: ") + hipGetErrorString(err));
    }
  }
  hipError_t err = launch_trouble_score(update ? c->dtr : nullptr, c->dscore, ne, c->dtrout, c->stream);
  if (err != hipSuccess) return fail(NMX_EHIP, std::string("trouble score: ") + hipGetErrorString(err));
  HIPCHK(hipMemcpyAsync(score, c->dtrout, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(c->hscore.data(), c->dscore, ne * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return NMX_OK;
}

// registers a switch carries, besides U (which the re-mirror adapts)
enum { CARRY_A = 1, CARRY_W0 = 2, CARRY_W1 = 4 };

// elements to FV on uniform points (to_fv 1) or back to DG on LGL points,
// at the same n (PARENT_n_P_UNIFORM / PARENT_n_P_LGL; trouble.c:186-330,
// RungeKutta.c:239-312): the tables are rebuilt, U and the carried registers
// interpolated, the scores kept
// (nw: the switched mesh, already built by the caller from the current one)
static int trouble_switch_to(nmx_ctx *c, Mesh nw, int carry) {
  struct Held { int flag; double *tmp; };
  std::vector<Held> held;
  for (int flag : {CARRY_A, CARRY_W0, CARRY_W1}) {
    if (!(carry & flag)) continue;
    double *src = flag == CARRY_A ? c->dA : flag == CARRY_W0 ? c->dW0 : c->dW1;
    double *tmp = nullptr;
    int r = dev_alloc(&tmp, c->nstate);
    if (!r && hipMemcpyAsync(tmp, src, c->nstate * 8, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
      r = fail(NMX_EHIP, "trouble switch: hold a register");
    if (r) {
      (void)hipFree(tmp);
      for (const Held &h : held) (void)hipFree(h.tmp);
      return r;
    }
    held.push_back({flag, tmp});
  }
  const std::vector<int> keep = c->hscore;
  Mesh old = std::exchange(c->mesh, std::move(nw));
  int r = remirror_with_state(c, old);
  for (const Held &h : held) {
    double *dst = h.flag == CARRY_A ? c->dA : h.flag == CARRY_W0 ? c->dW0 : c->dW1;
    if (!r) r = adapt_state(c, old, h.tmp, dst);
    (void)hipFree(h.tmp);
  }
  release_mesh(std::move(old));
  if (r) return r;
  // the step's stage times (the mirror's copy is new)
  HIPCHK(hipMemcpyAsync(c->dT, c->hT, 4 * sizeof(double), hipMemcpyHostToDevice, c->stream));
  c->hscore = keep;
  c->tuned = true;                // mid-run: keep the launch shapes chosen so far
  c->trouble_switches++;
  return trouble_prepare(c);
}

static int trouble_switch(nmx_ctx *c, const std::vector<int> &to_fv, int carry) {
  Mesh nw = mesh_skeleton(c->mesh);
  std::string err = set_fv_elems(nw, to_fv);
  if (!err.empty()) return fail(NMX_EINVAL, err);
  return trouble_switch_to(c, std::move(nw), carry);
}

static int trouble_fix(nmx_ctx *c, double *y, const double *x, const double *x2, const double *r, double sfac,
                       int mode, const int *mask = nullptr) {
  TroubleFix f;
  f.y = y; f.x = x; f.x2 = x2; f.r = r; f.s = sfac; f.mode = mode;
  f.score = mask ? mask : c->dscore; f.eoff = c->deoff; f.en = c->den; f.nelem = (int)c->mesh.e.size();
  hipError_t err = launch_trouble_fix(f, c->stream);
  if (err != hipSuccess) return fail(NMX_EHIP, std::string("trouble fix: ") + hipGetErrorString(err));
  return NMX_OK;
}

static int launch(nmx_ctx *c, StageArgs a);
static void stage_times(int method, double t, double dt, double *ts);

// one stage of evolve_RK4_mesh / evolve_sspRK3_mesh / evolve_Euler_mesh
// (RungeKutta.c:19-188) with u_p in U, u in A, r kept in R
struct TStage { int win, wout, wmode; double cA, cW; };   // win/wout: 0 U, 1 W0, 2 W1, -1 none
static void trouble_stages(int method, double dt, std::vector<TStage> &st) {
  if (method == 0)
    st = {{0, 1, 1, dt / 6., dt / 2.}, {1, 2, 1, dt / 3., dt / 2.}, {2, 1, 1, dt / 3., dt}, {1, -1, 0, dt / 6., 0.}};
  else if (method == 1)
    st = {{0, 1, 1, dt / 6., dt}, {1, 2, 2, dt / 6., dt}, {2, -1, 0, dt * 2. / 3., 0.}};
  else
    st = {{0, -1, 0, dt, 0.}};
}

static double *treg(nmx_ctx *c, int which) { return which == 0 ? c->dU : which == 1 ? c->dW0 : which == 2 ? c->dW1 : nullptr; }

// classes of the redone elements: evolve_trouble_redo_u_step_mesh redoes
// one element at a time, its neighbours as they were before any switch.  A
// face that is a copy face once both sides are FV (same n, level and
// points) hands over the same values either way (the polynomial at the
// uniform face points), so only two redone elements that share another
// face, at least one of them newly switched, go into different classes.
// Greedy colouring in element order; the class of element e in cls[e] (-1:
// not redone), the count returned.
static int redo_classes(const Mesh &M0, const Mesh &Mall, const std::vector<int> &score, std::vector<int> &cls) {
  const size_t ne = M0.e.size();
  cls.assign(ne, -1);
  int ncls = 0;
  for (size_t e = 0; e < ne; ++e) {
    if (score[e] <= 0) continue;
    uint64_t used = 0;
    for (int f = 0; f < 6; ++f)
      for (int ni = 0; ni < M0.e[e].nfnb[f]; ++ni) {
        const int nb = M0.e[e].fnb[f][ni];
        if (cls[nb] < 0 || (M0.e[e].fv && M0.e[nb].fv)) continue;
        const int nf = M0.e[e].nbf[f][ni];
        const bool copy = (Mall.e[e].ajmode[f] == 1 || Mall.e[e].ajmode[f] == 2) &&
                          (Mall.e[nb].ajmode[nf] == 1 || Mall.e[nb].ajmode[nf] == 2);
        if (!copy) used |= 1ull << cls[nb];
      }
    int k = 0;
    while (used >> k & 1) ++k;
    if (k >= 64) return -1;
    cls[e] = k;
    ncls = std::max(ncls, k + 1);
  }
  return ncls;
}

// a new device mirror of c->mesh in the middle of a step: tables, stage
// times, scores; the launch shapes chosen so far kept
static int mirror_again(nmx_ctx *c) {
  if (int r = upload_mesh(c)) return r;
  HIPCHK(hipMemcpyAsync(c->dT, c->hT, 4 * sizeof(double), hipMemcpyHostToDevice, c->stream));
  c->tuned = true;
  return trouble_prepare(c);
}

// the redo of stage S (evolve_trouble_redo_u_step_mesh, RungeKutta.c:
// 223-330) after u -= r rfac: the troubled elements become FV (u, u_p and w
// interpolated), then R of each redone element is the RHS of w with that
// element switched and its neighbours as before.  With one class that is
// the RHS on the switched mesh; with more, each class is switched on its
// own in turn and its R kept.
static int redo_rhs(nmx_ctx *c, const TStage &S, int s) {
  const Mesh M0 = mesh_skeleton(c->mesh);    // the base of each class's rebuild
  const size_t ne = M0.e.size();
  const std::vector<int> score = c->hscore;
  std::vector<int> fv(ne);
  for (size_t e = 0; e < ne; ++e) fv[e] = M0.e[e].fv || score[e] > 0;
  Mesh Mall = M0;
  std::string err = set_fv_elems(Mall, fv);
  if (!err.empty()) return fail(NMX_EINVAL, err);
  std::vector<int> cls;
  const int ncls = redo_classes(M0, Mall, score, cls);
  if (ncls < 0) return fail(NMX_EUNSUP, "more than 64 classes of redone elements");
  const long ns = c->nstate;
  double *win0 = nullptr;
  if (ncls > 1) {
    if (int r = dev_alloc(&win0, ns)) return r;
    HIPCHK(hipMemcpyAsync(win0, treg(c, S.win), ns * 8, hipMemcpyDeviceToDevice, c->stream));
  }
  const int carry = CARRY_A | (S.win == 1 || S.wout == 1 ? CARRY_W0 : 0) | (S.win == 2 || S.wout == 2 ? CARRY_W1 : 0);
  // the switched mesh is Mall: copied in when the class mirrors follow, else moved
  const int rs = ncls > 1 ? trouble_switch_to(c, Mall, carry) : trouble_switch_to(c, std::move(Mall), carry);
  if (rs) { (void)hipFree(win0); return rs; }
  if (ncls <= 1) {
    StageArgs b = base_args(c);
    b.win = treg(c, S.win); b.R = c->dR; b.tidx = s;
    return launch(c, b);
  }
  // registers of the switched mesh held across the class mirrors
  double *held[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};   // U A W0 W1 R
  int *dmask = nullptr;
  int r = NMX_OK;
  for (int q = 0; q < 5 && !r; ++q) r = dev_alloc(&held[q], ns);
  if (!r) {
    double *src[4] = {c->dU, c->dA, c->dW0, c->dW1};
    for (int q = 0; q < 4; ++q)
      if (hipMemcpyAsync(held[q], src[q], ns * 8, hipMemcpyDeviceToDevice, c->stream) != hipSuccess) r = fail(NMX_EHIP, "redo: hold registers");
    if (hipMemsetAsync(held[4], 0, ns * 8, c->stream) != hipSuccess) r = fail(NMX_EHIP, "redo: clear R");
  }
  const double *wall = S.win == 0 ? held[0] : S.win == 1 ? held[2] : held[3];
  for (int k = 0; k < ncls && !r; ++k) {
    std::vector<int> fk(ne), mk(ne);
    for (size_t e = 0; e < ne; ++e) { mk[e] = cls[e] == k; fk[e] = M0.e[e].fv || mk[e]; }
    release_mesh(std::exchange(c->mesh, Mesh(M0)));
    err = set_fv_elems(c->mesh, fk);
    if (!err.empty()) { r = fail(NMX_EINVAL, err); break; }
    if ((r = mirror_again(c))) break;
    (void)hipFree(dmask); dmask = nullptr;
    if ((r = dev_upload(&dmask, mk))) break;
    // w: as before the switch, this class's elements interpolated
    if (hipMemcpyAsync(c->dW0, win0, ns * 8, hipMemcpyDeviceToDevice, c->stream) != hipSuccess) { r = fail(NMX_EHIP, "redo: w"); break; }
    if ((r = trouble_fix(c, c->dW0, wall, nullptr, nullptr, 0., 1, dmask))) break;
    StageArgs b = base_args(c);
    b.win = c->dW0; b.R = c->dR; b.tidx = s;
    if ((r = launch(c, b))) break;
    r = trouble_fix(c, held[4], c->dR, nullptr, nullptr, 0., 1, dmask);
  }
  // back to the switched mesh with its registers, also after a failure
  // (the first error is the one returned)
  const bool held_ok = held[0] && held[1] && held[2] && held[3] && held[4];
  release_mesh(std::exchange(c->mesh, std::move(Mall)));
  int r2 = mirror_again(c);
  if (!r2 && held_ok) {
    double *dst[5] = {c->dU, c->dA, c->dW0, c->dW1, c->dR};
    for (int q = 0; q < 5; ++q)
      if (hipMemcpyAsync(dst[q], held[q], ns * 8, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
        r2 = fail(NMX_EHIP, "redo: restore registers");
    if (hipStreamSynchronize(c->stream) != hipSuccess) r2 = fail(NMX_EHIP, "redo: sync");
  }
  if (!r) r = r2;
  for (double *h : held) (void)hipFree(h);
  (void)hipFree(dmask);
  (void)hipFree(win0);
  return r;
}

// Evolve_mesh with evolve_redo_troubled = substep when redo: after each
// stage the scores of u; if the mesh's score is positive, each troubled
// element takes its substep back, becomes FV (u, u_p and w interpolated),
// and redoes it with the RHS of w (evolve_trouble_redo_u_step_mesh,
// RungeKutta.c:223-330).  The step's times are ts.
// stage s's launches from a graph captured on the first call with these
// arguments (the mirror's pointers and parameters drop the graphs when they
// change); eager launches where steps are not captured
static int launch_stage_graph(nmx_ctx *c, const StageArgs &a, int s) {
  if (P(c, "nmx_graph") != "yes" || P(c, "nmx_path") == "perelement" || s < 0 || s > 3) return launch(c, a);
  const StageArgs &k = c->tgargs[s];
  const bool same = c->tgexec[s] && k.win == a.win && k.U == a.U && k.Ain == a.Ain && k.Aout == a.Aout &&
                    k.Wout == a.Wout && k.R == a.R && k.tidx == a.tidx && k.cA == a.cA && k.cW == a.cW &&
                    k.wmode == a.wmode;
  if (!same) {
    if (c->tgexec[s]) { (void)hipGraphExecDestroy(c->tgexec[s]); c->tgexec[s] = nullptr; }
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int r = launch(c, a);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(c->stream, &g);
    if (r || ec != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      if (r) return r;
      return fail(NMX_EHIP, std::string("stage capture: ") + hipGetErrorString(ec));
    }
    const hipError_t ei = hipGraphInstantiate(&c->tgexec[s], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      c->tgexec[s] = nullptr;
      return fail(NMX_EHIP, std::string("stage graph: ") + hipGetErrorString(ei));
    }
    c->tgargs[s] = a;
  }
  HIPCHK(hipGraphLaunch(c->tgexec[s], c->stream));
  return NMX_OK;
}

static int trouble_evolve_mesh(nmx_ctx *c, int method, double dt, bool redo, int *nredo) {
  std::vector<TStage> st;
  trouble_stages(method, dt, st);
  for (size_t s = 0; s < st.size(); ++s) {
    const TStage &S = st[s];
    StageArgs a = base_args(c);
    a.win = treg(c, S.win); a.U = c->dU; a.Ain = s ? c->dA : nullptr; a.Aout = c->dA;
    a.Wout = treg(c, S.wout); a.wmode = S.wmode; a.cA = S.cA; a.cW = S.cW; a.tidx = (int)s;
    if (int r = launch_stage_graph(c, a, (int)s)) return r;
    if (!redo) continue;
    int score = 0;
    if (int r = trouble_score(c, c->dA, c->dU, true, &score)) return r;
    if (score <= 0) continue;
    ++*nredo;
    c->trouble_redos++;
    // the substep's r again from its w (no stage writes its w input), on
    // the mesh it was taken on: the stages keep no r, which only a redo reads
    {
      StageArgs b = base_args(c);
      b.win = a.win; b.R = c->dR; b.tidx = (int)s;
      if (int r = launch(c, b)) return r;
    }
    // u -= r rfac on the troubled elements, they become FV and get their r
    // again, then u += r rfac and the next w
    if (int r = trouble_fix(c, c->dA, c->dA, nullptr, c->dR, -S.cA, 0)) return r;
    if (int r = redo_rhs(c, S, (int)s)) return r;
    if (int r = trouble_fix(c, c->dA, c->dA, nullptr, c->dR, S.cA, 0)) return r;
    if (S.wout >= 0 && S.wmode == 1)
      if (int r = trouble_fix(c, treg(c, S.wout), c->dU, nullptr, c->dR, S.cW, 0)) return r;
    if (S.wout >= 0 && S.wmode == 2)
      if (int r = trouble_fix(c, treg(c, S.wout), c->dU, treg(c, S.win), c->dR, S.cW, 2)) return r;
  }
  return NMX_OK;
}

// evolve_myln (evolve.c:25-176) with troubled elements: the step, its
// scores (after each substep, or after the step and then the step redone
// from u_p on the switched mesh), then the elements whose score is at most
// -notroubles back to DG.  The state ends in U.
static int trouble_step(nmx_ctx *c, int method, double dt) {
  const bool substep = P(c, "evolve_redo_troubled") == "substep";
  const int nsub = method == 0 ? 4 : method == 1 ? 3 : 1;
  const int notroubles = 11 * (substep ? nsub : 1);   // NOTROUBLES (evolve.h:60), evolve_notroubles
  int nredo = 0, score = 0;
  c->ldat_up = nullptr;           // u_p is new
  if (int r = trouble_evolve_mesh(c, method, dt, substep, &nredo)) return r;
  if (!substep) {
    if (int r = trouble_score(c, c->dA, c->dU, true, &score)) return r;
    if (score > 0) {
      // evolve_prepare_do_over_mesh: back to u_p, troubled elements to FV,
      // the step again
      std::vector<int> fv(c->mesh.e.size());
      for (size_t e = 0; e < fv.size(); ++e) fv[e] = c->mesh.e[e].fv || c->hscore[e] > 0;
      if (int r = trouble_switch(c, fv, 0)) return r;
      c->trouble_redos++;
      if (int r = trouble_evolve_mesh(c, method, dt, false, &nredo)) return r;
    }
  } else {
    if (int r = trouble_score(c, c->dA, c->dU, false, &score)) return r;
  }
  HIPCHK(hipMemcpyAsync(c->dU, c->dA, c->nstate * 8, hipMemcpyDeviceToDevice, c->stream));
  if (score > 0 || score <= -notroubles) {
    // evolve_switch_nontroubled_nodes_mesh: scores <= -notroubles back to DG
    std::vector<int> fv(c->mesh.e.size());
    bool any = false;
    for (size_t e = 0; e < fv.size(); ++e) {
      const bool back = c->mesh.e[e].fv && c->hscore[e] <= -notroubles;
      fv[e] = c->mesh.e[e].fv && !back;
      any = any || back;
    }
    if (any)
      if (int r = trouble_switch(c, fv, 0)) return r;
  }
  return NMX_OK;
}

static int launch_limiter_on(nmx_ctx *c, double *reg) {
  LimArgs a;
  a.u = reg; a.eoff = c->deoff; a.en = c->den; a.nbr = c->dnbr; a.h15 = c->dlimh; a.hw = c->dlimhw;
  a.wq = c->dwqall; a.amat = c->damat; a.smat = c->dsmat; a.ldat = c->dldat; a.alpha = c->lim_alpha;
  a.beta = c->lim_beta; a.kind = c->lim_kind; a.scale_bound = c->lim_scale; a.nelem = (int)c->mesh.e.size();
  hipError_t err = launch_limiter(a, c->stream);
  if (err != hipSuccess) return fail(NMX_EHIP, std::string("limiter launch: ") + hipGetErrorString(err));
  return NMX_OK;
}

// issue the stage launches of one step; stage times at dT[tbase + s]
static int issue_step(nmx_ctx *c, int method, double dt, int tbase, bool swap_parity) {
  double *U = swap_parity ? c->dA : c->dU;
  double *A = swap_parity ? c->dU : c->dA;
  StageArgs a = base_args(c);
  int r;
  // the MRS limiter acts on every later stage's input w (RungeKutta.c:45,
  // 56, 67, 161, 173) and on the final u (evolve.c:143-144)
  auto lim = [&](double *reg) { return c->lim_on ? launch_limiter_on(c, reg) : NMX_OK; };
  if (method == 0) {   // evolve_RK4_mesh (RungeKutta.c:19-82)
    a.win = U; a.U = U; a.Ain = nullptr; a.Aout = A; a.cA = dt / 6.; a.Wout = c->dW0; a.wmode = 1;
    a.cW = dt / 2.; a.tidx = tbase + 0;
    if ((r = launch(c, a)) || (r = lim(c->dW0))) return r;
    a.win = c->dW0; a.Ain = A; a.cA = dt / 3.; a.Wout = c->dW1; a.cW = dt / 2.; a.tidx = tbase + 1;
    if ((r = launch(c, a)) || (r = lim(c->dW1))) return r;
    a.win = c->dW1; a.cA = dt / 3.; a.Wout = c->dW0; a.cW = dt; a.tidx = tbase + 2;
    if ((r = launch(c, a)) || (r = lim(c->dW0))) return r;
    a.win = c->dW0; a.cA = dt / 6.; a.Wout = nullptr; a.wmode = 0; a.tidx = tbase + 3;
    if ((r = launch(c, a))) return r;
  } else if (method == 1) {   // evolve_sspRK3_mesh (RungeKutta.c:130-188)
    a.win = U; a.U = U; a.Aout = A; a.cA = dt / 6.; a.Wout = c->dW0; a.wmode = 1; a.cW = dt;
    a.tidx = tbase + 0;
    if ((r = launch(c, a)) || (r = lim(c->dW0))) return r;
    a.win = c->dW0; a.Ain = A; a.cA = dt / 6.; a.Wout = c->dW1; a.wmode = 2; a.cW = dt; a.tidx = tbase + 1;
    if ((r = launch(c, a)) || (r = lim(c->dW1))) return r;
    a.win = c->dW1; a.cA = dt * 2. / 3.; a.Wout = nullptr; a.wmode = 0; a.tidx = tbase + 2;
    if ((r = launch(c, a))) return r;
  } else {   // evolve_Euler_mesh (RungeKutta.c:85-115)
    a.win = U; a.U = U; a.Aout = A; a.cA = dt; a.wmode = 0; a.tidx = tbase;
    if ((r = launch(c, a))) return r;
  }
  if ((r = lim(A))) return r;
  // FILTER bin after the full step (main.c:628-630)
  if (c->filt_on && (r = launch_filters(c, A))) return r;
  return NMX_OK;
}

static void stage_times(int method, double t, double dt, double *ts) {
  if (method == 0) { ts[0] = t; ts[1] = t + 0.5 * dt; ts[2] = t + 0.5 * dt; ts[3] = t + dt; }
  else if (method == 1) { ts[0] = t; ts[1] = t + dt; ts[2] = t + 0.5 * dt; ts[3] = 0; }
  else { ts[0] = t; ts[1] = ts[2] = ts[3] = 0; }
}

// Autotuner: time the stage kernel's tile size (nodes per thread) and the
// XCD-aware element mapping on this mesh with a few RK steps, keep the
// fastest.  The state and time are restored afterwards.
extern "C" int nmx_set_trouble_hook(nmx_ctx *c, nmx_trouble_fn fn, void *user) {
  if (!c) return fail(NMX_EINVAL, "ctx is NULL");
  c->trouble_hook = fn;
  c->trouble_user = user;
  return NMX_OK;
}

extern "C" int nmx_get_trouble_scores(nmx_ctx *c, int *score, int nelem, long *redos, long *switches) {
  if (!c) return fail(NMX_EINVAL, "ctx is NULL");
  if (!c->have_mesh) return fail(NMX_ESTATE, "no mesh");
  if (score) {
    if (nelem != (int)c->mesh.e.size()) return fail(NMX_EINVAL, "need one score per element");
    for (int e = 0; e < nelem; ++e) score[e] = (size_t)e < c->hscore.size() ? c->hscore[e] : 0;
  }
  if (redos) *redos = c->trouble_redos;
  if (switches) *switches = c->trouble_switches;
  return NMX_OK;
}

extern "C" int nmx_autotune(nmx_ctx *c) {
  NEED_DEV(c);
  HIPCHK(hipSetDevice(c->device));
  if (int r = limiter_prepare(c)) return r;
  if (P(c, "nmx_path") == "perelement" || c->fvelm) { c->tuned = true; return NMX_OK; }
  double *backup = nullptr;
  int r;
  if ((r = dev_alloc(&backup, c->nstate))) return r;
  HIPCHK(hipMemcpyAsync(backup, c->dU, c->nstate * 8, hipMemcpyDeviceToDevice, c->stream));
  const int method = method_id(c);
  stage_times(method, c->time, c->dt, c->hT);
  HIPCHK(hipMemcpyAsync(c->dT, c->hT, 4 * sizeof(double), hipMemcpyHostToDevice, c->stream));
  const char *blocks[] = {"full", "half", "quarter"};
  const char *xcds[] = {"yes", "no"};
  std::string best_b = P(c, "nmx_block"), best_x = P(c, "nmx_xcd");
  float best = 1e30f;
  hipEvent_t t0, t1;
  HIPCHK(hipEventCreate(&t0));
  HIPCHK(hipEventCreate(&t1));
  for (const char *b : blocks)
    for (const char *x : xcds) {
      c->par["nmx_block"] = b;
      c->par["nmx_xcd"] = x;
      if ((r = issue_step(c, method, c->dt, 0, false))) return r;   // warm-up
      HIPCHK(hipEventRecord(t0, c->stream));
      for (int s = 0; s < 4 && !r; ++s) r = issue_step(c, method, c->dt, 0, (s & 1) != 0);
      if (r) return r;
      HIPCHK(hipEventRecord(t1, c->stream));
      HIPCHK(hipEventSynchronize(t1));
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, t0, t1));
      if (ms < best) { best = ms; best_b = b; best_x = x; }
    }
  c->par["nmx_block"] = best_b;
  c->par["nmx_xcd"] = best_x;
  // batched path: each order bucket gets its own tile, timed alone on an
  // RK4 second-stage launch (reads w, u_p, u_acc; writes u_acc, w)
  if (c->amr) {
    StageArgs sa = base_args(c);
    sa.win = c->dW0; sa.U = c->dU; sa.Ain = c->dA; sa.Aout = c->dA; sa.Wout = c->dW1;
    sa.wmode = 1; sa.cA = c->dt / 3.; sa.cW = c->dt / 2.; sa.tidx = 0;
    HIPCHK(hipMemcpyAsync(c->dW0, c->dU, c->nstate * 8, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->dA, c->dU, c->nstate * 8, hipMemcpyDeviceToDevice, c->stream));
    {   // the face buffer of every bucket, for the face-kernel shape
      const std::string shape = P(c, "nmx_batch");
      c->par["nmx_batch"] = "legacy";
      r = launch_amr(c, sa);
      c->par["nmx_batch"] = shape;
      if (r) return r;
    }
    const int xcd = best_x == "yes" ? 1 : 0;
    AmrArgs a = amr_args(c, sa);
    for (auto &b : c->buckets) {
      // each order's tile, timed alone: with the face kernel's buffer
      // (FV buckets and the legacy shape) and in the order's own launch of
      // the receiver-gathering kernel (DG buckets)
      for (int variant = 0; variant < (b.fv ? 1 : 2); ++variant) {
        int best_div = 0;
        float bms = 1e30f;
        const std::vector<int> divs = b.fv ? std::vector<int>{1}
                                      : variant == 0 ? std::vector<int>{1, 2, 4} : std::vector<int>{1, 2, 3, 4};
        auto one = [&](int div) {
          return variant == 0 ? launch_bucket(c, a, b, div, xcd, c->stream)
                              : launch_order(c, sa, b.n, div, c->stream);
        };
        for (int div : divs) {
          if ((r = one(div))) return r;   // warm-up
          HIPCHK(hipEventRecord(t0, c->stream));
          for (int s = 0; s < 5 && !r; ++s) r = one(div);
          if (r) return r;
          HIPCHK(hipEventRecord(t1, c->stream));
          HIPCHK(hipEventSynchronize(t1));
          float ms = 0;
          HIPCHK(hipEventElapsedTime(&ms, t0, t1));
          if (ms < bms) { bms = ms; best_div = div; }
        }
        if (variant == 0) { b.div = best_div; b.ms_legacy = bms; }
        else { c->bdiv[b.n] = best_div; b.ms_order = bms; }
      }
      if (b.fv) b.ms_order = b.ms_legacy;
    }
    auto lanes_for = [c](const std::string &shape) {
      for (auto &b : c->buckets) b.ms = shape == "legacy" ? b.ms_legacy : b.ms_order;
      assign_lanes(c);
    };
    lanes_for(P(c, "nmx_batch"));
    // launch shape: one persistent launch per stage, one launch per order on
    // the stream lanes, or per order with the face kernel first, whichever
    // takes the steps fastest (unless set)
    if (!c->batch_fixed && c->anydg) {
      std::string best_shape = P(c, "nmx_batch");
      float bt = 1e30f;
      for (const char *shape : {"persistent", "buckets", "legacy"}) {
        c->par["nmx_batch"] = shape;
        lanes_for(shape);
        if ((r = issue_step(c, method, c->dt, 0, false))) return r;   // warm-up
        HIPCHK(hipEventRecord(t0, c->stream));
        for (int s = 0; s < 4 && !r; ++s) r = issue_step(c, method, c->dt, 0, (s & 1) != 0);
        if (r) return r;
        HIPCHK(hipEventRecord(t1, c->stream));
        HIPCHK(hipEventSynchronize(t1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, t0, t1));
        if (ms < bt) { bt = ms; best_shape = shape; }
      }
      c->par["nmx_batch"] = best_shape;
      lanes_for(best_shape);
    }
    // with the face kernel's shape, the order buckets' tiles refined on whole
    // steps, where the buckets run beside one another as they will: each DG
    // bucket, longest first, keeps the tile that makes the steps fastest with
    // the others fixed (the tile that is fastest alone need not be)
    if (P(c, "nmx_batch") == "legacy" && c->anydg && c->buckets.size() > 1) {
      auto steps_ms = [&](float &ms) -> int {
        if (int rr = issue_step(c, method, c->dt, 0, false)) return rr;   // warm-up
        HIPCHK(hipEventRecord(t0, c->stream));
        for (int s = 0; s < 6; ++s)
          if (int rr = issue_step(c, method, c->dt, 0, (s & 1) != 0)) return rr;
        HIPCHK(hipEventRecord(t1, c->stream));
        HIPCHK(hipEventSynchronize(t1));
        HIPCHK(hipEventElapsedTime(&ms, t0, t1));
        return NMX_OK;
      };
      float cur = 0.f;
      if ((r = steps_ms(cur))) return r;
      for (size_t k = 0; k < c->buckets.size(); ++k) {
        auto &b = c->buckets[c->border[k]];
        if (b.fv) continue;
        const int keep = b.div;
        for (int div : {1, 2, 4}) {
          if (div == keep) continue;
          const int prev = b.div;
          b.div = div;
          float ms = 0.f;
          if ((r = steps_ms(ms))) return r;
          if (ms < cur) cur = ms;
          else b.div = prev;
        }
      }
    }
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  HIPCHK(hipMemcpyAsync(c->dU, backup, c->nstate * 8, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  (void)hipFree(backup);
  drop_stage_graphs(c);
  if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
  if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
  c->graph_steps = 0;
  c->tuned = true;
  return NMX_OK;
}

// the time after one full step: te accumulates the steps and moves into Te
// every 10^6 iterations (evolve_mesh, main.c:612-624)
static void advance_time(long double &Te, long double &te, long &iteration, double &time, double dt) {
  iteration++;
  te += dt;
  if (iteration % 1000000 == 0) {
    Te += te;
    te = 0;
  }
  time = (double)(Te + te);
}

static void advance_time(nmx_ctx *c, double dt) { advance_time(c->Te, c->te, c->iteration, c->time, dt); }

extern "C" int nmx_evolve(nmx_ctx *c, int nsteps) {
  NEED_DEV(c);
  if (nsteps < 0) return fail(NMX_EINVAL, "nsteps < 0");
  HIPCHK(hipSetDevice(c->device));
  {
    bool on = false;
    int r = filter_prepare(c, on);
    if (r) return r;
    c->filt_on = on;
    if ((r = limiter_prepare(c))) return r;
  }
  if (!c->tuned && P(c, "nmx_autotune") == "yes" && nsteps > 0) {
    int r = nmx_autotune(c);
    if (r) return r;
  }
  const int method = method_id(c);
  const double dt = c->dt;
  if (trouble_on(c)) {
    // troubled elements: scores read on the host after each substep, so one
    // step at a time, each stage's launches from a graph of their own
    if (c->lim_on || c->filt_on) return fail(NMX_EUNSUP, "troubled elements with a limiter or filter are not built");
    if (P(c, "nmx_trouble") == "hook" && !c->trouble_hook) return fail(NMX_ESTATE, "nmx_trouble = hook needs nmx_set_trouble_hook");
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    for (int k = 0; k < nsteps; ++k) {
      HIPCHK(hipStreamSynchronize(c->stream));
      stage_times(method, c->time, dt, c->hT);
      HIPCHK(hipMemcpyAsync(c->dT, c->hT, 4 * sizeof(double), hipMemcpyHostToDevice, c->stream));
      if (int r = trouble_step(c, method, dt)) return r;
      advance_time(c, dt);
    }
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    return NMX_OK;
  }
  const bool use_graph = (P(c, "nmx_graph") == "yes") && P(c, "nmx_path") != "perelement";
  int G = Pi(c, "nmx_graph_steps");
  if (G < 2) G = 2;
  if (G > 64) G = 64;
  G &= ~1;   // even: the U/A ping-pong returns to its start
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  int done = 0;
  bool parity = false;   // false: current state in dU
  if (use_graph && nsteps >= G) {
    if (!c->gexec || c->graph_steps != G || c->graph_dt != dt || c->graph_method != method) {
      if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
      if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
      HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int r = NMX_OK;
      for (int s = 0; s < G && !r; ++s) r = issue_step(c, method, dt, 4 * s, (s & 1) != 0);
      hipGraph_t g = nullptr;
      hipError_t ec = hipStreamEndCapture(c->stream, &g);
      if (r || ec != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (r) return r;
        return fail(NMX_EHIP, std::string("capture: ") + hipGetErrorString(ec));
      }
      c->graph = g;
      HIPCHK(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
      c->graph_steps = G; c->graph_dt = dt; c->graph_method = method;
    }
    while (nsteps - done >= G) {
      // stage times of the G steps (time = Te + te, main.c:588-624)
      HIPCHK(hipStreamSynchronize(c->stream));   // hT is reused
      for (int s = 0; s < G; ++s) {
        stage_times(method, c->time, dt, c->hT + 4 * s);
        advance_time(c, dt);
      }
      HIPCHK(hipMemcpyAsync(c->dT, c->hT, 4 * G * sizeof(double), hipMemcpyHostToDevice, c->stream));
      HIPCHK(hipGraphLaunch(c->gexec, c->stream));
      done += G;
    }
  }
  for (; done < nsteps; ++done) {
    HIPCHK(hipStreamSynchronize(c->stream));
    stage_times(method, c->time, dt, c->hT);
    HIPCHK(hipMemcpyAsync(c->dT, c->hT, 4 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    int r = issue_step(c, method, dt, 0, parity);
    if (r) return r;
    parity = !parity;
    advance_time(c, dt);
  }
  // keep register pointers fixed (captured graphs refer to them)
  if (parity) HIPCHK(hipMemcpyAsync(c->dU, c->dA, c->nstate * 8, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  HIPCHK(hipEventSynchronize(c->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_ms = ms;
  return NMX_OK;
}

// evolve_mesh (main.c:583-666): steps until the mesh time reaches finaltime
// (or iterations * dt when finaltime <= 0) within max(dt 1e-7,
// finaltime 1e-14), Te starting at the mesh time.  The number of steps is
// counted on the host with the same time accumulation, then taken by
// nmx_evolve (graphs included).
extern "C" int nmx_evolve_mesh(nmx_ctx *c) {
  NEED_DEV(c);
  const int itmax = Pi(c, "iterations");
  double timemax = Pd(c, "finaltime");
  const double dt = c->dt;
  const double ttol = std::max(dt * 1e-7, timemax * 1e-14);
  if (timemax <= 0. && itmax > 0) timemax = itmax * dt;
  if (timemax <= 0.) return NMX_OK;
  if (!(dt > 0.)) return fail(NMX_EINVAL, "dt must be positive");
  c->Te = c->time;
  c->te = 0;
  long double Te = c->Te, te = 0;
  long it = c->iteration;
  double t = c->time;
  long k = 0;
  while (t < timemax - ttol) {
    advance_time(Te, te, it, t, dt);
    if (++k > 2000000000L) return fail(NMX_EINVAL, "finaltime needs more than 2e9 steps");
  }
  return nmx_evolve(c, (int)k);
}

extern "C" int nmx_get_iteration(nmx_ctx *c, long *it) {
  if (!c || !it) return fail(NMX_EINVAL, "NULL argument");
  *it = c->iteration;
  return NMX_OK;
}

extern "C" int nmx_last_evolve_ms(nmx_ctx *c, double *ms) {
  if (!c) return fail(NMX_EINVAL, "ctx is NULL");
  *ms = c->last_ms;
  return NMX_OK;
}

extern "C" int nmx_state_device_ptr(nmx_ctx *c, int reg, void **p, long *len) {
  NEED_DEV(c);
  double *d = reg_ptr(c, reg);
  if (!d) return fail(NMX_EINVAL, "bad register");
  *p = d;
  *len = c->nstate;
  return NMX_OK;
}

// ---------------------------------------------------------------------------
// discrete adjoint
// ---------------------------------------------------------------------------

static AdjArgs adj_base_args(nmx_ctx *c) {
  AdjArgs a;
  a.dXdx = c->dXdx; a.dXbdX = c->dXbdX; a.egeo = c->degeo; a.xb = c->dxb; a.D = c->dD;
  a.edom = c->dedom; a.face = c->dface; a.nbface = c->dnbf;
  a.wend_inv = c->np1 ? 1.0 / lgl_basis(c->np1).w[0] : 0.0;
  a.numflux = (P(c, "scalarwave1_numflux") == "LLF") ? 1 : 0;
  a.xcd_map = (P(c, "nmx_xcd") == "yes") ? 1 : 0;
  a.e0 = 0;
  a.e1 = (int)c->mesh.e.size();
  return a;
}

// the adjoint tables of the batched path (made once per mirror, at the
// first adjoint call): per element face its kind, copy partner or the jobs
// that read its plane; per job its transposed plane in T; G planes per
// element; element lists per order
static int adj_amr_prepare(nmx_ctx *c) {
  if (c->adj_ready) return NMX_OK;
  const Mesh &M = c->mesh;
  const int ne = (int)M.e.size();
  std::vector<long> goff(ne);
  long glen = 0;
  for (int e = 0; e < ne; ++e) { goff[e] = glen; glen += 24L * M.e[e].n * M.e[e].n; }
  std::vector<AdjFace> af(6L * ne);
  for (int e = 0; e < ne; ++e) {
    const Elm &E = M.e[e];
    for (int f = 0; f < 6; ++f) {
      AdjFace &F = af[6L * e + f];
      std::memset(&F, 0, sizeof F);
      if (E.ajmode[f] == 0) continue;
      if (E.ajmode[f] == 1 || E.ajmode[f] == 2) {
        const FaceCopy &fc = M.fcopy[6L * e + f];
        const int nb = E.fnb[f][0];
        const Elm &N = M.e[nb];
        F.kind = 1; F.nb = nb; F.g = E.nbf[f][0];
        F.base = (int)(fc.nbase - N.off); F.s1 = fc.s1; F.s2 = fc.s2;
        if (N.n != E.n || F.g < 0 || F.g > 5) return fail(NMX_EINVAL, "adjoint: copy face between different orders");
        const int n = E.n;
        for (int cb = 0; cb < 2; ++cb)
          for (int ca = 0; ca < 2; ++ca) {
            const long p = F.base + (long)ca * (n - 1) * F.s1 + (long)cb * (n - 1) * F.s2;
            if (p < 0 || p >= (long)n * n * n) return fail(NMX_EINVAL, "adjoint: copy table out of range");
          }
        // the partner's point lies on its face g
        const long p0 = F.base;
        const int I0[3] = {(int)(p0 % n), (int)((p0 / n) % n), (int)(p0 / ((long)n * n))};
        if (I0[F.g >> 1] != (F.g & 1) * (n - 1)) return fail(NMX_EINVAL, "adjoint: copy partner face");
        continue;
      }
      F.kind = 2;
    }
  }
  const int nj = (int)c->hadjj.size();
  std::vector<AdjJob> jobs = c->hadjj;
  std::vector<std::vector<int>> byplane(6L * ne);
  long tlen = 0;
  for (int j = 0; j < nj; ++j) {
    AdjJob &J = jobs[j];
    if (tlen > INT32_MAX - 4L * J.m * J.m) return fail(NMX_EUNSUP, "adjoint: transposed planes too large");
    J.toff = (int)tlen;
    tlen += 4L * J.m * J.m;
    const int se = c->hadjsend[2 * j], sf = c->hadjsend[2 * j + 1];
    if (se < 0 || se >= ne || M.e[se].n != J.m || af[6L * se + sf].kind != 2)
      return fail(NMX_EINVAL, "adjoint: a job reads a plane that is not an interpolated face");
    if (J.e < 0 || J.e >= ne || M.e[J.e].n != J.n || J.g < 0 || J.g > 5)
      return fail(NMX_EINVAL, "adjoint: job receiver out of range");
    for (int cb = 0; cb < 2; ++cb)
      for (int ca = 0; ca < 2; ++ca) {
        const int fa = ca * (J.m - 1), fb = cb * (J.m - 1);
        const int q1 = J.A1 * (fa - J.A0) + J.B1 * (fb - J.B0), q2 = J.A2 * (fa - J.A0) + J.B2 * (fb - J.B0);
        if (q1 < 0 || q1 >= J.m || q2 < 0 || q2 >= J.m) return fail(NMX_EINVAL, "adjoint: plane map out of range");
      }
    byplane[6L * se + sf].push_back(j);
  }
  std::vector<int> sj;
  for (long k = 0; k < 6L * ne; ++k) {
    af[k].jfirst = (int)sj.size();
    af[k].jcount = (int)byplane[k].size();
    sj.insert(sj.end(), byplane[k].begin(), byplane[k].end());
  }
  // lists: all elements, then one per order
  std::vector<int> lists;
  c->adjlists.clear();
  c->adjcount.clear();
  c->adjlists.push_back({0, 0});
  for (int e = 0; e < ne; ++e) lists.push_back(e);
  c->adjcount.push_back(ne);
  for (int n = 2; n <= kBatchMaxN; ++n) {
    const int first = (int)lists.size();
    for (int e = 0; e < ne; ++e)
      if (M.e[e].n == n) lists.push_back(e);
    if ((int)lists.size() > first) {
      c->adjlists.push_back({n, first});
      c->adjcount.push_back((int)lists.size() - first);
    }
  }
  if (jobs.empty()) jobs.resize(1);
  if (sj.empty()) sj.assign(1, 0);
  int r;
  if ((r = dev_upload(&c->dadjf, af)) || (r = dev_upload(&c->dadjj, jobs)) || (r = dev_upload(&c->dadjs, sj)) ||
      (r = dev_upload(&c->dadjl, lists)) || (r = dev_upload(&c->dadjg, goff)) ||
      (r = dev_alloc(&c->dadjG, glen)) || (r = dev_alloc(&c->dadjT, tlen > 0 ? tlen : 1)))
    return r;
  c->adjnjobs = nj;
  c->adj_ready = true;
  return NMX_OK;
}

#define NEED_ADJ(c)                                                                      \
  do {                                                                                   \
    NEED_DEV(c);                                                                         \
    if ((c)->fvelm || (c)->anyfv)                                                        \
      return fail(NMX_EUNSUP, "the adjoint of FV elements is not built");               \
    if ((c)->amr ? !(c)->anydg : !(c)->dnbf)                                             \
      return fail(NMX_EUNSUP, "the adjoint needs a DG mesh");                           \
    if ((c)->amr) {                                                                      \
      if (int r_ = adj_amr_prepare(c)) return r_;                                        \
    }                                                                                    \
    bool filt_ = false;                                                                  \
    if (int r_ = filter_prepare((c), filt_)) return r_;                                  \
    if (filt_) return fail(NMX_EUNSUP, "the adjoint of the exponential filter is not built"); \
    if (P((c), "scalarwave1_limiter") != "none")                                          \
      return fail(NMX_EUNSUP, "the adjoint of the limiter is not built");                \
  } while (0)

// the batched path's adjoint stage: G, the transposed jobs, then one
// launch per order, all on the context's stream
static int launch_adj_amr(nmx_ctx *c, const AdjArgs &s) {
  AdjAmrArgs a;
  a.lam = s.lam; a.Abase = s.Abase; a.Wbase = s.Wbase; a.Aout = s.Aout; a.Wout = s.Wout; a.R = s.R;
  a.cA = s.cA; a.cU = s.cU; a.cW = s.cW;
  a.eoff = c->deoff; a.en = c->den; a.dXbdX = c->dXbdX; a.egeo = c->degeo; a.edom = c->dedom;
  a.dXdx = c->dXdx; a.Dall = c->dbD; a.xball = c->dbxb; a.wend_inv = c->dbW;
  a.af = c->dadjf; a.goff = c->dadjg; a.G = c->dadjG; a.jobs = c->dadjj; a.njobs = c->adjnjobs;
  a.sjobs = c->dadjs; a.mdesc = c->dmdesc; a.Btab = c->dBtab; a.T = c->dadjT;
  a.numflux = s.numflux;
  a.list = c->dadjl;
  a.nlist = c->adjcount[0];
  hipError_t err = launch_adj_gflux(c->geom, a, c->stream);
  if (err == hipSuccess) err = launch_adj_jobs(a, c->stream);
  for (size_t b = 1; b < c->adjlists.size() && err == hipSuccess; ++b) {
    a.list = c->dadjl + c->adjlists[b].second;
    a.nlist = c->adjcount[b];
    const int grid = std::min(a.nlist, 8 * c->ncu);
    err = launch_dg_adjoint_amr(c->adjlists[b].first, c->geom, a, grid, c->stream);
  }
  if (err != hipSuccess) return fail(NMX_EHIP, std::string("adjoint launch: ") + hipGetErrorString(err));
  return NMX_OK;
}

static int launch_adj(nmx_ctx *c, const AdjArgs &a) {
  if (c->amr) return launch_adj_amr(c, a);
  int nel = a.e1 - a.e0;
  int grid = a.xcd_map ? 8 * ((nel + 7) / 8) : nel;
  hipError_t err = launch_dg_adjoint(c->np1, c->geom, a, grid, c->stream);
  if (err != hipSuccess) return fail(NMX_EHIP, std::string("adjoint launch: ") + hipGetErrorString(err));
  return NMX_OK;
}

extern "C" int nmx_adjoint_setrhs(nmx_ctx *c, int reg_out, int reg_in) {
  NEED_ADJ(c);
  double *out = reg_ptr(c, reg_out), *in = reg_ptr(c, reg_in);
  if (!out || !in || out == in) return fail(NMX_EINVAL, "bad registers");
  AdjArgs a = adj_base_args(c);
  a.lam = in;
  a.R = out;
  return launch_adj(c, a);
}

// One adjoint RK step lam <- (du_{n+1}/du_n)^T lam: the forward stages of
// issue_step taken in reverse.  With y_s = L^T x_s:
//   RK4:    x4 = lam; acc = lam + dt/6 y4; x3 = dt/3 lam + dt^2/6 y4;
//           acc += y3; x2 = dt/3 lam + dt/2 y3; acc += y2; x1 = dt/6 lam + dt/2 y2;
//           acc += y1
//   sspRK3: x3 = lam; acc = lam + 2dt/3 y3; x2 = dt/6 lam + dt^2/6 y3;
//           acc += y2; x1 = x2 + dt y2; acc += y1
//   Euler:  acc = lam + dt y1
static int issue_adj_step(nmx_ctx *c, int method, double dt, bool swap_parity) {
  double *L = swap_parity ? c->dA : c->dL;
  double *Acc = swap_parity ? c->dL : c->dA;
  AdjArgs a = adj_base_args(c);
  int r;
  if (method == 0) {
    a.lam = L; a.Abase = L; a.Aout = Acc; a.cA = dt / 6.;
    a.Wbase = L; a.Wout = c->dW0; a.cU = dt / 3.; a.cW = dt * dt / 6.;
    if ((r = launch_adj(c, a))) return r;
    a.lam = c->dW0; a.Abase = Acc; a.cA = 1.0; a.Wbase = L; a.Wout = c->dW1; a.cU = dt / 3.; a.cW = dt / 2.;
    if ((r = launch_adj(c, a))) return r;
    a.lam = c->dW1; a.Wout = c->dW0; a.cU = dt / 6.; a.cW = dt / 2.;
    if ((r = launch_adj(c, a))) return r;
    a.lam = c->dW0; a.Wout = nullptr; a.Wbase = nullptr;
    if ((r = launch_adj(c, a))) return r;
  } else if (method == 1) {
    a.lam = L; a.Abase = L; a.Aout = Acc; a.cA = 2. * dt / 3.;
    a.Wbase = L; a.Wout = c->dW0; a.cU = dt / 6.; a.cW = dt * dt / 6.;
    if ((r = launch_adj(c, a))) return r;
    a.lam = c->dW0; a.Abase = Acc; a.cA = 1.0; a.Wbase = c->dW0; a.Wout = c->dW1; a.cU = 1.0; a.cW = dt;
    if ((r = launch_adj(c, a))) return r;
    a.lam = c->dW1; a.Wout = nullptr; a.Wbase = nullptr;
    if ((r = launch_adj(c, a))) return r;
  } else {
    a.lam = L; a.Abase = L; a.Aout = Acc; a.cA = dt;
    if ((r = launch_adj(c, a))) return r;
  }
  return NMX_OK;
}

extern "C" int nmx_adjoint_evolve(nmx_ctx *c, int nsteps) {
  NEED_ADJ(c);
  if (nsteps < 0) return fail(NMX_EINVAL, "nsteps < 0");
  HIPCHK(hipSetDevice(c->device));
  const int method = method_id(c);
  const double dt = c->dt;
  const bool use_graph = (P(c, "nmx_graph") == "yes");
  int G = Pi(c, "nmx_graph_steps");
  if (G < 2) G = 2;
  if (G > 64) G = 64;
  G &= ~1;
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  int done = 0;
  bool parity = false;
  if (use_graph && nsteps >= G) {
    if (!c->agexec || c->agraph_steps != G || c->agraph_dt != dt || c->agraph_method != method) {
      if (c->agexec) { (void)hipGraphExecDestroy(c->agexec); c->agexec = nullptr; }
      if (c->agraph) { (void)hipGraphDestroy(c->agraph); c->agraph = nullptr; }
      HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int r = NMX_OK;
      for (int s = 0; s < G && !r; ++s) r = issue_adj_step(c, method, dt, (s & 1) != 0);
      hipGraph_t g = nullptr;
      hipError_t ec = hipStreamEndCapture(c->stream, &g);
      if (r || ec != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (r) return r;
        return fail(NMX_EHIP, std::string("capture: ") + hipGetErrorString(ec));
      }
      c->agraph = g;
      HIPCHK(hipGraphInstantiate(&c->agexec, c->agraph, nullptr, nullptr, 0));
      c->agraph_steps = G; c->agraph_dt = dt; c->agraph_method = method;
    }
    for (; nsteps - done >= G; done += G) HIPCHK(hipGraphLaunch(c->agexec, c->stream));
  }
  for (; done < nsteps; ++done) {
    int r = issue_adj_step(c, method, dt, parity);
    if (r) return r;
    parity = !parity;
  }
  if (parity) HIPCHK(hipMemcpyAsync(c->dL, c->dA, c->nstate * 8, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  HIPCHK(hipEventSynchronize(c->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_adj_ms = ms;
  return NMX_OK;
}

extern "C" int nmx_last_adjoint_ms(nmx_ctx *c, double *ms) {
  if (!c || !ms) return fail(NMX_EINVAL, "NULL argument");
  *ms = c->last_adj_ms;
  return NMX_OK;
}

// checkpoint pool: nslots device copies of a register plus the mesh time
extern "C" int nmx_checkpoint_alloc(nmx_ctx *c, int nslots) {
  NEED_DEV(c);
  if (nslots < 0) return fail(NMX_EINVAL, "nslots < 0");
  if (c->dCk) { HIPCHK(hipStreamSynchronize(c->stream)); (void)hipFree(c->dCk); c->dCk = nullptr; }
  c->nck = 0;
  if (nslots > 0) {
    int r = dev_alloc(&c->dCk, (size_t)nslots * c->nstate);
    if (r) return r;
  }
  c->nck = nslots;
  c->ckTe.assign(nslots, 0);
  c->ckte.assign(nslots, 0);
  c->ckset.assign(nslots, 0);
  return NMX_OK;
}

extern "C" int nmx_checkpoint_store(nmx_ctx *c, int slot, int reg) {
  NEED_DEV(c);
  double *d = reg_ptr(c, reg);
  if (!d) return fail(NMX_EINVAL, "bad register");
  if (slot < 0 || slot >= c->nck) return fail(NMX_EINVAL, "checkpoint slot out of range");
  HIPCHK(hipMemcpyAsync(c->dCk + (size_t)slot * c->nstate, d, c->nstate * 8, hipMemcpyDeviceToDevice, c->stream));
  c->ckTe[slot] = c->Te;
  c->ckte[slot] = c->te;
  c->ckset[slot] = 1;
  return NMX_OK;
}

extern "C" int nmx_checkpoint_load(nmx_ctx *c, int slot, int reg) {
  NEED_DEV(c);
  double *d = reg_ptr(c, reg);
  if (!d) return fail(NMX_EINVAL, "bad register");
  if (slot < 0 || slot >= c->nck || !c->ckset[slot]) return fail(NMX_EINVAL, "checkpoint slot empty or out of range");
  HIPCHK(hipMemcpyAsync(d, c->dCk + (size_t)slot * c->nstate, c->nstate * 8, hipMemcpyDeviceToDevice, c->stream));
  if (reg == NMX_REG_U) {
    c->Te = c->ckTe[slot];
    c->te = c->ckte[slot];
    c->time = (double)(c->Te + c->te);
  }
  return NMX_OK;
}

// state transfer from / to device memory in the flat layout, ordered on the
// context's stream (no host synchronisation)
extern "C" int nmx_state_copy_in(nmx_ctx *c, int reg, const double *dflat) {
  NEED_DEV(c);
  double *d = reg_ptr(c, reg);
  if (!d || !dflat) return fail(NMX_EINVAL, "bad register or pointer");
  hipError_t err = launch_layout(dflat, d, c->deoff, c->den, (int)c->mesh.e.size(), c->mesh.nnodes, 1, c->stream);
  if (err != hipSuccess) return fail(NMX_EHIP, std::string("layout launch: ") + hipGetErrorString(err));
  return NMX_OK;
}

extern "C" int nmx_state_copy_out(nmx_ctx *c, int reg, double *dflat) {
  NEED_DEV(c);
  double *d = reg_ptr(c, reg);
  if (!d || !dflat) return fail(NMX_EINVAL, "bad register or pointer");
  hipError_t err = launch_layout(d, dflat, c->deoff, c->den, (int)c->mesh.e.size(), c->mesh.nnodes, 0, c->stream);
  if (err != hipSuccess) return fail(NMX_EHIP, std::string("layout launch: ") + hipGetErrorString(err));
  return NMX_OK;
}
