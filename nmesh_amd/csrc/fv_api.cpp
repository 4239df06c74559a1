// fv_api.cpp -- C ABI of the FV WENO5 + LLF path (include/nmx.h, nmx_fv_*):
// a periodic grid of 16^3-cell blocks evolved with special-relativistic
// Euler (ideal gas), c2p + fused RHS/stage kernels per RK stage, RK steps
// captured in hipGraphs.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/nmx.h"
#include "errors.h"
#include "fv.h"

using namespace nmx;

struct nmx_fv {
  int device = 0;
  hipStream_t own_stream = nullptr, stream = nullptr;
  std::map<std::string, std::string> par;
  bool ready = false;
  FvGrid g;
  long ncell = 0, nstate = 0;
  double *dU = nullptr, *dA = nullptr, *dW0 = nullptr, *dW1 = nullptr, *dR = nullptr, *dP = nullptr,
         *dT = nullptr;
  int *dbad = nullptr;
  double dt = 0, time = 0;
  long double Te = 0, te = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_steps = 0, graph_method = -1;
  double graph_dt = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0;
};

static int fail(int code, const std::string &msg) { return set_error(code, msg); }
#define HIPCHK(x) NMX_HIPCHK(x)

static const std::string &P(nmx_fv *f, const char *k) { return f->par[k]; }

static void fv_defaults(nmx_fv *f) {
  f->par = {
      {"fv_n", "256"},          {"fv_block", "16"},      {"fv_gamma", "2"},
      {"fv_c2p_tol", "1e-12"},  {"fv_dtfac", "0.25"},    {"dt", "auto"},
      {"evolve_method", "sspRK3"}, {"fv_rec", "WENO5"},  {"fv_flux", "LLF"},
      {"nmx_graph", "yes"},     {"nmx_graph_steps", "8"}, {"nmx_xcd", "yes"},
  };
}

static void free_grid(nmx_fv *f) {
  double **bufs[] = {&f->dU, &f->dA, &f->dW0, &f->dW1, &f->dR, &f->dP, &f->dT};
  for (auto b : bufs)
    if (*b) { (void)hipFree(*b); *b = nullptr; }
  if (f->dbad) { (void)hipFree(f->dbad); f->dbad = nullptr; }
  if (f->gexec) { (void)hipGraphExecDestroy(f->gexec); f->gexec = nullptr; }
  if (f->graph) { (void)hipGraphDestroy(f->graph); f->graph = nullptr; }
  f->graph_steps = 0;
  f->ready = false;
}

extern "C" int nmx_fv_create(nmx_fv **out, int device) {
  if (!out) return fail(NMX_EINVAL, "NULL argument");
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(NMX_EINVAL, "no such HIP device");
  nmx_fv *f = new nmx_fv;
  f->device = device;
  fv_defaults(f);
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipStreamCreateWithFlags(&f->own_stream, hipStreamNonBlocking));
  f->stream = f->own_stream;
  HIPCHK(hipEventCreate(&f->ev0));
  HIPCHK(hipEventCreate(&f->ev1));
  *out = f;
  return NMX_OK;
}

extern "C" int nmx_fv_destroy(nmx_fv *f) {
  if (!f) return NMX_OK;
  (void)hipSetDevice(f->device);
  if (f->stream) (void)hipStreamSynchronize(f->stream);
  free_grid(f);
  if (f->ev0) (void)hipEventDestroy(f->ev0);
  if (f->ev1) (void)hipEventDestroy(f->ev1);
  if (f->own_stream) (void)hipStreamDestroy(f->own_stream);
  delete f;
  return NMX_OK;
}

extern "C" int nmx_fv_set_stream(nmx_fv *f, void *s) {
  if (!f) return fail(NMX_EINVAL, "handle is NULL");
  f->stream = s ? (hipStream_t)s : f->own_stream;
  return NMX_OK;
}

extern "C" int nmx_fv_synchronize(nmx_fv *f) {
  if (!f) return fail(NMX_EINVAL, "handle is NULL");
  HIPCHK(hipStreamSynchronize(f->stream));
  return NMX_OK;
}

extern "C" int nmx_fv_set_param(nmx_fv *f, const char *name, const char *value) {
  if (!f || !name || !value) return fail(NMX_EINVAL, "NULL argument");
  std::string k(name), v(value);
  if (!f->par.count(k)) return fail(NMX_EINVAL, "unknown parameter " + k);
  if (k == "evolve_method" && v != "RK4" && v != "sspRK3" && v != "Euler")
    return fail(NMX_EINVAL, "evolve_method must be RK4, sspRK3 or Euler");
  if (k == "fv_rec" && v != "WENO5") return fail(NMX_EUNSUP, "fv_rec: only WENO5 is built");
  if (k == "fv_flux" && v != "LLF") return fail(NMX_EUNSUP, "fv_flux: only LLF is built");
  if (k == "fv_block" && v != "16") return fail(NMX_EUNSUP, "fv_block: the kernels take 16^3-cell blocks");
  if ((k == "fv_n" || k == "fv_block" || k == "fv_gamma") && f->ready)
    return fail(NMX_ESTATE, k + " must be set before nmx_fv_setup");
  f->par[k] = v;
  if (f->gexec) { (void)hipGraphExecDestroy(f->gexec); f->gexec = nullptr; }
  if (f->graph) { (void)hipGraphDestroy(f->graph); f->graph = nullptr; }
  f->graph_steps = 0;
  if (f->ready && (k == "dt" || k == "fv_dtfac"))
    f->dt = P(f, "dt") == "auto" ? std::atof(P(f, "fv_dtfac").c_str()) : std::atof(P(f, "dt").c_str());
  return NMX_OK;
}

extern "C" int nmx_fv_get_param(nmx_fv *f, const char *name, char *buf, int buflen) {
  if (!f || !name || !buf || buflen <= 0) return fail(NMX_EINVAL, "NULL argument");
  auto it = f->par.find(name);
  if (it == f->par.end()) return fail(NMX_EINVAL, std::string("unknown parameter ") + name);
  std::snprintf(buf, buflen, "%s", it->second.c_str());
  return NMX_OK;
}

template <class T>
static int dalloc(T **p, size_t n) {
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) return fail(NMX_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  return NMX_OK;
}

extern "C" int nmx_fv_setup(nmx_fv *f) {
  if (!f) return fail(NMX_EINVAL, "handle is NULL");
  HIPCHK(hipSetDevice(f->device));
  free_grid(f);
  const int N = std::atoi(P(f, "fv_n").c_str());
  const int B = std::atoi(P(f, "fv_block").c_str());
  if (B != 16) return fail(NMX_EUNSUP, "fv_block must be 16");
  if (N < B || N % B) return fail(NMX_EINVAL, "fv_n must be a positive multiple of fv_block");
  const double gam = std::atof(P(f, "fv_gamma").c_str());
  if (!(gam > 1.0)) return fail(NMX_EINVAL, "fv_gamma must exceed 1");
  f->g.B = B;
  f->g.nbx = N / B;
  f->g.gam = gam;
  f->ncell = (long)N * N * N;
  f->nstate = 5 * f->ncell;
  int r;
  double **bufs[] = {&f->dU, &f->dA, &f->dW0, &f->dW1, &f->dR, &f->dP, &f->dT};
  for (auto b : bufs)
    if ((r = dalloc(b, f->nstate))) return r;
  const long nb = (long)f->g.nbx * f->g.nbx * f->g.nbx;
  if ((r = dalloc(&f->dbad, nb))) return r;
  HIPCHK(hipMemset(f->dbad, 0, nb * sizeof(int)));
  HIPCHK(hipMemset(f->dU, 0, f->nstate * 8));
  HIPCHK(hipMemset(f->dP, 0, f->nstate * 8));
  f->dt = P(f, "dt") == "auto" ? std::atof(P(f, "fv_dtfac").c_str()) : std::atof(P(f, "dt").c_str());
  f->time = 0; f->Te = 0; f->te = 0;
  f->ready = true;
  return NMX_OK;
}

#define NEED_GRID(f)                                                   \
  do {                                                                 \
    if (!(f)) return fail(NMX_EINVAL, "handle is NULL");              \
    if (!(f)->ready) return fail(NMX_ESTATE, "call nmx_fv_setup first"); \
  } while (0)

extern "C" int nmx_fv_ncells(nmx_fv *f, long *n) { NEED_GRID(f); *n = f->ncell; return NMX_OK; }
extern "C" int nmx_fv_get_dt(nmx_fv *f, double *dt) { NEED_GRID(f); *dt = f->dt; return NMX_OK; }
extern "C" int nmx_fv_set_dt(nmx_fv *f, double dt) { NEED_GRID(f); f->dt = dt; return NMX_OK; }
extern "C" int nmx_fv_get_time(nmx_fv *f, double *t) { NEED_GRID(f); *t = f->time; return NMX_OK; }
extern "C" int nmx_fv_set_time(nmx_fv *f, double t) {
  NEED_GRID(f);
  f->time = t; f->Te = t; f->te = 0;
  return NMX_OK;
}

// host flat [v][cell] -> block layout register (through the R scratch register)
static int upload_flat(nmx_fv *f, double *dst, const double *h) {
  HIPCHK(hipMemcpyAsync(f->dT, h, f->nstate * 8, hipMemcpyHostToDevice, f->stream));
  hipError_t e = launch_fv_layout(f->g, f->dT, dst, 1, f->stream);
  if (e != hipSuccess) return fail(NMX_EHIP, std::string("layout: ") + hipGetErrorString(e));
  HIPCHK(hipStreamSynchronize(f->stream));
  return NMX_OK;
}

static double *fv_reg(nmx_fv *f, int which) {
  switch (which) {
    case NMX_FV_CONS: return f->dU;
    case NMX_FV_PRIM: return f->dP;
    case NMX_FV_RHS: return f->dR;
    default: return nullptr;
  }
}

extern "C" int nmx_fv_upload_prim(nmx_fv *f, const double *prim) {
  NEED_GRID(f);
  if (!prim) return fail(NMX_EINVAL, "NULL argument");
  int r = upload_flat(f, f->dP, prim);
  if (r) return r;
  hipError_t e = launch_fv_p2c(f->g, f->dP, f->dU, f->stream);
  if (e != hipSuccess) return fail(NMX_EHIP, std::string("p2c: ") + hipGetErrorString(e));
  HIPCHK(hipStreamSynchronize(f->stream));
  return NMX_OK;
}

extern "C" int nmx_fv_upload(nmx_fv *f, int which, const double *h) {
  NEED_GRID(f);
  double *d = fv_reg(f, which);
  if (!d || !h) return fail(NMX_EINVAL, "bad register or NULL argument");
  return upload_flat(f, d, h);
}

extern "C" int nmx_fv_download(nmx_fv *f, int which, double *h) {
  NEED_GRID(f);
  double *d = fv_reg(f, which);
  if (!d || !h) return fail(NMX_EINVAL, "bad register or NULL argument");
  hipError_t e = launch_fv_layout(f->g, d, f->dT, 0, f->stream);
  if (e != hipSuccess) return fail(NMX_EHIP, std::string("layout: ") + hipGetErrorString(e));
  HIPCHK(hipMemcpyAsync(h, f->dT, f->nstate * 8, hipMemcpyDeviceToHost, f->stream));
  HIPCHK(hipStreamSynchronize(f->stream));
  return NMX_OK;
}

// c2p of win into P, then the fused stage
static int fv_launch(nmx_fv *f, FvStageArgs a) {
  FvC2pArgs c;
  c.g = f->g;
  c.U = a.win;
  c.P = f->dP;
  c.bad = f->dbad;
  c.tol = std::atof(P(f, "fv_c2p_tol").c_str());
  hipError_t e = launch_fv_c2p(c, f->stream);
  if (e != hipSuccess) return fail(NMX_EHIP, std::string("c2p: ") + hipGetErrorString(e));
  a.g = f->g;
  a.P = f->dP;
  a.xcd_map = P(f, "nmx_xcd") == "yes" ? 1 : 0;
  e = launch_fv_stage(a, f->stream);
  if (e != hipSuccess) return fail(NMX_EHIP, std::string("fv stage: ") + hipGetErrorString(e));
  return NMX_OK;
}

extern "C" int nmx_fv_max_speed(nmx_fv *f, double *amax) {
  NEED_GRID(f);
  if (!amax) return fail(NMX_EINVAL, "NULL argument");
  FvC2pArgs c;
  c.g = f->g;
  c.U = f->dU;
  c.P = f->dP;
  c.bad = f->dbad;
  c.tol = std::atof(P(f, "fv_c2p_tol").c_str());
  hipError_t e = launch_fv_c2p(c, f->stream);
  if (e != hipSuccess) return fail(NMX_EHIP, std::string("c2p: ") + hipGetErrorString(e));
  const int grid = fv_speed_grid(f->g);
  double *part = nullptr;
  if (int r = dalloc(&part, (size_t)grid + 1)) return r;
  e = launch_fv_max_speed(f->g, f->dP, part, part + grid, f->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(amax, part + grid, sizeof(double), hipMemcpyDeviceToHost, f->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(f->stream);
  (void)hipFree(part);
  if (e != hipSuccess) return fail(NMX_EHIP, std::string("speed reduction: ") + hipGetErrorString(e));
  return NMX_OK;
}

extern "C" int nmx_fv_setrhs(nmx_fv *f) {
  NEED_GRID(f);
  FvStageArgs a;
  a.win = f->dU;
  a.R = f->dR;
  return fv_launch(f, a);
}

static int method_id(nmx_fv *f) {
  const std::string &m = P(f, "evolve_method");
  return m == "RK4" ? 0 : m == "sspRK3" ? 1 : 2;
}

// the stage sequences of evolve_RK4_mesh / evolve_sspRK3_mesh /
// evolve_Euler_mesh (RungeKutta.c:19-188), as in the DG driver
static int fv_issue_step(nmx_fv *f, int method, double dt, bool swap) {
  double *U = swap ? f->dA : f->dU;
  double *A = swap ? f->dU : f->dA;
  FvStageArgs a;
  int r;
  if (method == 0) {
    a.win = U; a.U = U; a.Aout = A; a.cA = dt / 6.; a.Wout = f->dW0; a.wmode = 1; a.cW = dt / 2.;
    if ((r = fv_launch(f, a))) return r;
    a.win = f->dW0; a.Ain = A; a.cA = dt / 3.; a.Wout = f->dW1; a.cW = dt / 2.;
    if ((r = fv_launch(f, a))) return r;
    a.win = f->dW1; a.cA = dt / 3.; a.Wout = f->dW0; a.cW = dt;
    if ((r = fv_launch(f, a))) return r;
    a.win = f->dW0; a.cA = dt / 6.; a.Wout = nullptr; a.wmode = 0;
    if ((r = fv_launch(f, a))) return r;
  } else if (method == 1) {
    a.win = U; a.U = U; a.Aout = A; a.cA = dt / 6.; a.Wout = f->dW0; a.wmode = 1; a.cW = dt;
    if ((r = fv_launch(f, a))) return r;
    a.win = f->dW0; a.Ain = A; a.cA = dt / 6.; a.Wout = f->dW1; a.wmode = 2; a.cW = dt;
    if ((r = fv_launch(f, a))) return r;
    a.win = f->dW1; a.cA = dt * 2. / 3.; a.Wout = nullptr; a.wmode = 0;
    if ((r = fv_launch(f, a))) return r;
  } else {
    a.win = U; a.U = U; a.Aout = A; a.cA = dt; a.wmode = 0;
    if ((r = fv_launch(f, a))) return r;
  }
  return NMX_OK;
}

extern "C" int nmx_fv_evolve(nmx_fv *f, int nsteps) {
  NEED_GRID(f);
  if (nsteps < 0) return fail(NMX_EINVAL, "nsteps < 0");
  HIPCHK(hipSetDevice(f->device));
  const int method = method_id(f);
  const double dt = f->dt;
  int G = std::atoi(P(f, "nmx_graph_steps").c_str());
  if (G < 2) G = 2;
  if (G > 64) G = 64;
  G &= ~1;
  HIPCHK(hipEventRecord(f->ev0, f->stream));
  int done = 0;
  bool parity = false;
  if (P(f, "nmx_graph") == "yes" && nsteps >= G) {
    if (!f->gexec || f->graph_steps != G || f->graph_dt != dt || f->graph_method != method) {
      if (f->gexec) { (void)hipGraphExecDestroy(f->gexec); f->gexec = nullptr; }
      if (f->graph) { (void)hipGraphDestroy(f->graph); f->graph = nullptr; }
      HIPCHK(hipStreamBeginCapture(f->stream, hipStreamCaptureModeThreadLocal));
      int r = NMX_OK;
      for (int s = 0; s < G && !r; ++s) r = fv_issue_step(f, method, dt, (s & 1) != 0);
      hipGraph_t g = nullptr;
      hipError_t ec = hipStreamEndCapture(f->stream, &g);
      if (r || ec != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (r) return r;
        return fail(NMX_EHIP, std::string("capture: ") + hipGetErrorString(ec));
      }
      f->graph = g;
      HIPCHK(hipGraphInstantiate(&f->gexec, f->graph, nullptr, nullptr, 0));
      f->graph_steps = G; f->graph_dt = dt; f->graph_method = method;
    }
    for (; nsteps - done >= G; done += G) HIPCHK(hipGraphLaunch(f->gexec, f->stream));
  }
  for (; done < nsteps; ++done) {
    int r = fv_issue_step(f, method, dt, parity);
    if (r) return r;
    parity = !parity;
  }
  if (parity) HIPCHK(hipMemcpyAsync(f->dU, f->dA, f->nstate * 8, hipMemcpyDeviceToDevice, f->stream));
  HIPCHK(hipEventRecord(f->ev1, f->stream));
  HIPCHK(hipEventSynchronize(f->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, f->ev0, f->ev1));
  f->last_ms = ms;
  for (int s = 0; s < nsteps; ++s) f->te += dt;   // time = Te + te (main.c:588-624)
  f->time = (double)(f->Te + f->te);
  return NMX_OK;
}

extern "C" int nmx_fv_last_evolve_ms(nmx_fv *f, double *ms) {
  if (!f || !ms) return fail(NMX_EINVAL, "NULL argument");
  *ms = f->last_ms;
  return NMX_OK;
}

// blocks in which a c2p Newton solve did not converge since the last call
extern "C" int nmx_fv_status(nmx_fv *f, long *nbad) {
  NEED_GRID(f);
  const long nb = (long)f->g.nbx * f->g.nbx * f->g.nbx;
  std::vector<int> h(nb);
  HIPCHK(hipMemcpyAsync(h.data(), f->dbad, nb * sizeof(int), hipMemcpyDeviceToHost, f->stream));
  HIPCHK(hipStreamSynchronize(f->stream));
  long n = 0;
  for (int v : h) n += v != 0;
  HIPCHK(hipMemsetAsync(f->dbad, 0, nb * sizeof(int), f->stream));
  HIPCHK(hipStreamSynchronize(f->stream));
  *nbad = n;
  return NMX_OK;
}
