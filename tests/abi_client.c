/* abi_client.c -- a C caller of libnmx, driving it the way the nmesh-side
 * shim does: parameters by name, nmesh's own mesh tables handed over with
 * nmx_mesh_import, the state moved through per-element arrays (what
 * Vard(node, vi) holds, nmesh_amr_loops.h:181), steps and the checkpoint
 * CRC.  Built with gcc -std=c99 -pedantic against include/nmx.h alone.
 *
 *   abi_client host    no device: a shell is built, h-refined and given
 *                      mixed orders in one host-only context; its tables
 *                      (element info, boxes, face neighbours, and locations
 *                      derived from the boxes) are handed to a second one
 *                      with nmx_mesh_import and must come back identical.
 *                      A swapped neighbour pair and two elements out of
 *                      depth-first order are refused with NMX_EINVAL.
 *   abi_client device  device 0: the same import into a device context; a
 *                      state set flat in the built context and through
 *                      per-element arrays in the imported one; the RHS, 3
 *                      RK4 steps, the time and the vars CRC of both are
 *                      bitwise equal; the per-element arrays read back equal
 *                      the flat download.
 * Prints "abi_client <mode> ok" and exits 0; otherwise prints what failed
 * and exits 1. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "nmx.h"

#define NINFO 22
#define MAXNB 8

static int failed;

#define CHECK(call)                                                              \
  do {                                                                           \
    int rc_ = (call);                                                            \
    if (rc_ != NMX_OK) {                                                         \
      printf("%s:%d: %s -> %d (%s)\n", __FILE__, __LINE__, #call, rc_,           \
             nmx_last_error());                                                  \
      exit(1);                                                                   \
    }                                                                            \
  } while (0)

#define EXPECT(cond, ...)            \
  do {                               \
    if (!(cond)) {                   \
      printf("failed: " __VA_ARGS__); \
      printf("\n");                  \
      failed = 1;                    \
    }                                \
  } while (0)

static const char *common_pars[][2] = {
    {"scalarwave1_k", "0.8 -0.5 0.33166247903554"},
    {"evolve_method", "RK4"},
    {"nmx_autotune", "no"},
    {"nmx_batch", "legacy"},
    {"nmx_path", "batched"},
    {NULL, NULL}};

/* nmesh's tables of one mesh, as the shim collects them from mesh->pat[]
   and the tElm of each element */
typedef struct {
  int npat, ne;
  int *ptype;
  double *pgeo;
  int *epat, *elev, *en, *nfnb, *fnb, *nbf;
  unsigned char *eloc;
  double *ebb;
  long *info;
  long nnodes;
} Tables;

static void set_pars(nmx_ctx *c) {
  for (int i = 0; common_pars[i][0]; i++) CHECK(nmx_set_param(c, common_pars[i][0], common_pars[i][1]));
}

/* a cubed-sphere shell of 6 x 8 elements, 3 of them h-refined, then orders
   4 to 7 */
static nmx_ctx *built_mesh(int device) {
  nmx_ctx *c;
  CHECK(nmx_ctx_create(&c, device));
  set_pars(c);
  CHECK(nmx_set_param(c, "amr_mesh_type", "Shell"));
  CHECK(nmx_set_param(c, "amr_luni", "1"));
  CHECK(nmx_set_param(c, "amr_n0", "5"));
  CHECK(nmx_set_param(c, "amr_n1", "5"));
  CHECK(nmx_set_param(c, "amr_n2", "5"));
  CHECK(nmx_set_param(c, "amr_Shell_rin", "2"));
  CHECK(nmx_set_param(c, "amr_Shell_rout", "6"));
  CHECK(nmx_mesh_build(c));
  int ne;
  CHECK(nmx_mesh_nelem(c, &ne));
  if (ne != 48) {
    printf("expected 48 elements, got %d\n", ne);
    exit(1);
  }
  int flag[48] = {0};
  flag[3] = flag[17] = flag[40] = 1;
  CHECK(nmx_mesh_hadapt(c, flag, ne));
  CHECK(nmx_mesh_nelem(c, &ne));
  int *n = malloc(ne * sizeof(int));
  for (int e = 0; e < ne; e++) n[e] = 4 + (e * 7) % 4;
  CHECK(nmx_mesh_padapt(c, n, ne));
  free(n);
  return c;
}

/* the octant digits of an element from its box inside the patch box
   lam [0,1], A, B [-1,1]: at level k+1 the octant i + 2j + 4k of the
   element's centre, as nmesh's eloc->loc holds it ('0' + digit) */
static void eloc_from_box(const double bb[6], int level, unsigned char *loc) {
  static const double lo[3] = {0., -1., -1.}, width[3] = {1., 2., 2.};
  int ic[3];
  for (int d = 0; d < 3; d++) {
    double centre = 0.5 * (bb[2 * d] + bb[2 * d + 1]);
    ic[d] = (int)((centre - lo[d]) / width[d] * (double)(1 << level));
  }
  for (int k = 0; k < level; k++) {
    int digit = 0;
    for (int d = 0; d < 3; d++) digit |= ((ic[d] >> (level - 1 - k)) & 1) << d;
    loc[k] = (unsigned char)('0' + digit);
  }
}

static void read_tables(nmx_ctx *c, Tables *T) {
  CHECK(nmx_mesh_nelem(c, &T->ne));
  CHECK(nmx_mesh_nnodes(c, &T->nnodes));
  int ne = T->ne;
  T->npat = 6;
  T->ptype = calloc(6, sizeof(int));
  T->pgeo = calloc(6 * 12, sizeof(double));
  for (int p = 0; p < 6; p++) {
    static const double box[6] = {0., 1., -1., 1., -1., 1.};
    T->ptype[p] = 1;
    T->pgeo[12 * p] = p;
    T->pgeo[12 * p + 1] = 2.0;
    T->pgeo[12 * p + 2] = 6.0;
    memcpy(T->pgeo + 12 * p + 3, box, sizeof box);
  }
  T->epat = calloc(ne, sizeof(int));
  T->elev = calloc(ne, sizeof(int));
  T->en = calloc(3 * ne, sizeof(int));
  T->nfnb = calloc(6 * ne, sizeof(int));
  T->fnb = calloc(6 * MAXNB * ne, sizeof(int));
  T->nbf = calloc(6 * MAXNB * ne, sizeof(int));
  T->eloc = calloc(24 * ne, 1);
  T->ebb = calloc(6 * ne, sizeof(double));
  T->info = calloc(NINFO * ne, sizeof(long));
  for (int e = 0; e < ne; e++) {
    long *inf = T->info + NINFO * e;
    CHECK(nmx_mesh_elm_info(c, e, inf));
    CHECK(nmx_mesh_elm_bbox(c, e, T->ebb + 6 * e));
    T->epat[e] = (int)inf[0];
    T->elev[e] = (int)inf[1];
    for (int d = 0; d < 3; d++) T->en[3 * e + d] = (int)inf[2];
    eloc_from_box(T->ebb + 6 * e, T->elev[e], T->eloc + 24 * e);
    for (int f = 0; f < 6; f++) {
      int nnb;
      CHECK(nmx_mesh_elm_fnb(c, e, f, T->fnb + 48 * e + 8 * f, T->nbf + 48 * e + 8 * f, &nnb));
      T->nfnb[6 * e + f] = nnb;
    }
  }
}

static int import_tables(nmx_ctx *c, const Tables *T) {
  return nmx_mesh_import(c, T->npat, T->ptype, T->pgeo, T->ne, T->epat, T->elev, T->eloc, T->en, T->ebb,
                         T->nfnb, T->fnb, T->nbf);
}

static void compare_tables(const Tables *A, const Tables *B) {
  EXPECT(A->ne == B->ne && A->nnodes == B->nnodes, "element or node count differs");
  if (A->ne != B->ne) return;
  EXPECT(!memcmp(A->info, B->info, NINFO * A->ne * sizeof(long)), "element info differs");
  EXPECT(!memcmp(A->ebb, B->ebb, 6 * A->ne * sizeof(double)), "element boxes differ");
  EXPECT(!memcmp(A->nfnb, B->nfnb, 6 * A->ne * sizeof(int)), "neighbour counts differ");
  for (int e = 0; e < A->ne; e++)
    for (int f = 0; f < 6; f++)
      for (int k = 0; k < A->nfnb[6 * e + f]; k++) {
        int i = 48 * e + 8 * f + k;
        EXPECT(A->fnb[i] == B->fnb[i] && A->nbf[i] == B->nbf[i], "element %d face %d neighbour %d differs", e, f, k);
      }
}

static void free_tables(Tables *T) {
  free(T->ptype); free(T->pgeo); free(T->epat); free(T->elev); free(T->en); free(T->nfnb);
  free(T->fnb); free(T->nbf); free(T->eloc); free(T->ebb); free(T->info);
}

/* the refusals: a face with several neighbours listed in another order,
   and two elements of one patch swapped */
static void check_refusals(int device, const Tables *A) {
  nmx_ctx *c;
  CHECK(nmx_ctx_create(&c, device));
  set_pars(c);
  Tables T = *A;
  int *fnb = malloc(6 * MAXNB * A->ne * sizeof(int)), *nbf = malloc(6 * MAXNB * A->ne * sizeof(int));
  memcpy(fnb, A->fnb, 6 * MAXNB * A->ne * sizeof(int));
  memcpy(nbf, A->nbf, 6 * MAXNB * A->ne * sizeof(int));
  int swapped = 0;
  for (int i = 0; i < 6 * A->ne && !swapped; i++)
    if (A->nfnb[i] > 1) {
      int *p = fnb + 8 * i, *q = nbf + 8 * i, t;
      t = p[0]; p[0] = p[1]; p[1] = t;
      t = q[0]; q[0] = q[1]; q[1] = t;
      swapped = 1;
    }
  EXPECT(swapped, "the mesh has no face with several neighbours");
  T.fnb = fnb;
  T.nbf = nbf;
  int rc = import_tables(c, &T);
  EXPECT(rc == NMX_EINVAL && strstr(nmx_last_error(), "neighbour list differs"),
         "a swapped neighbour pair gave %d (%s)", rc, nmx_last_error());
  /* elements 0 and 1 (same patch) exchanged: their boxes and locations */
  unsigned char *eloc = malloc(24 * A->ne);
  double *ebb = malloc(6 * A->ne * sizeof(double));
  memcpy(eloc, A->eloc, 24 * A->ne);
  memcpy(ebb, A->ebb, 6 * A->ne * sizeof(double));
  memcpy(eloc, A->eloc + 24, 24);
  memcpy(eloc + 24, A->eloc, 24);
  memcpy(ebb, A->ebb + 6, 6 * sizeof(double));
  memcpy(ebb + 6, A->ebb, 6 * sizeof(double));
  T = *A;
  T.eloc = eloc;
  T.ebb = ebb;
  rc = import_tables(c, &T);
  EXPECT(rc == NMX_EINVAL && strstr(nmx_last_error(), "depth-first order"),
         "elements out of order gave %d (%s)", rc, nmx_last_error());
  int ne = -1;
  EXPECT(nmx_mesh_nelem(c, &ne) != NMX_OK || ne == 0, "a refused import left a mesh (%d elements)", ne);
  free(fnb); free(nbf); free(eloc); free(ebb);
  CHECK(nmx_ctx_destroy(c));
}

/* per-element arrays of the flat state u[var * nnodes + node] */
static double **element_arrays(const Tables *T, const double *u) {
  double **p = malloc(5 * T->ne * sizeof(double *));
  for (int e = 0; e < T->ne; e++) {
    long off = T->info[NINFO * e + 3], np = T->info[NINFO * e + 2];
    np = np * np * np;
    for (int v = 0; v < 5; v++) {
      p[5 * e + v] = malloc(np * sizeof(double));
      if (u) memcpy(p[5 * e + v], u + v * T->nnodes + off, np * sizeof(double));
      else memset(p[5 * e + v], 0, np * sizeof(double));
    }
  }
  return p;
}

static void free_arrays(const Tables *T, double **p) {
  for (int i = 0; i < 5 * T->ne; i++) free(p[i]);
  free(p);
}

static void device_checks(nmx_ctx *A, nmx_ctx *B, const Tables *T) {
  const long len = 5 * T->nnodes;
  double *u = malloc(len * sizeof(double)), *ra = malloc(len * sizeof(double)),
         *rb = malloc(len * sizeof(double));
  double dta, dtb;
  CHECK(nmx_get_dt(A, &dta));
  CHECK(nmx_get_dt(B, &dtb));
  EXPECT(dta == dtb, "dt differs: %.17g %.17g", dta, dtb);
  /* the profile at t = 0.1 plus a deterministic perturbation */
  CHECK(nmx_state_set_profile(A, NMX_REG_U, 0.1));
  CHECK(nmx_state_download(A, NMX_REG_U, u));
  unsigned long long s = 88172645463325252ULL;
  for (long i = 0; i < len; i++) {
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    u[i] += 0.05 * ((double)(s >> 11) / 9007199254740992.0 - 0.5);
  }
  CHECK(nmx_state_upload(A, NMX_REG_U, u));
  double **arr = element_arrays(T, u);
  CHECK(nmx_state_upload_ptrs(B, NMX_REG_U, (const double *const *)arr, T->ne, 5));
  EXPECT(nmx_state_upload_ptrs(B, NMX_REG_U, (const double *const *)arr, T->ne, 4) == NMX_EINVAL,
         "nvar 4 accepted");
  CHECK(nmx_setrhs(A, NMX_REG_R, NMX_REG_U, 0.1));
  CHECK(nmx_setrhs(B, NMX_REG_R, NMX_REG_U, 0.1));
  CHECK(nmx_state_download(A, NMX_REG_R, ra));
  CHECK(nmx_state_download(B, NMX_REG_R, rb));
  EXPECT(!memcmp(ra, rb, len * sizeof(double)), "RHS differs");
  double amax = 0.0;
  for (long i = 0; i < len; i++) amax = ra[i] > amax ? ra[i] : (-ra[i] > amax ? -ra[i] : amax);
  EXPECT(amax > 0.1, "RHS is (nearly) zero: %g", amax);
  CHECK(nmx_set_time(A, 0.1));
  CHECK(nmx_set_time(B, 0.1));
  CHECK(nmx_evolve(A, 3));
  CHECK(nmx_evolve(B, 3));
  double ta, tb;
  CHECK(nmx_get_time(A, &ta));
  CHECK(nmx_get_time(B, &tb));
  EXPECT(ta == tb && ta > 0.1, "time differs: %.17g %.17g", ta, tb);
  uint64_t ca, cb;
  CHECK(nmx_vars_crc(A, NMX_REG_U, &ca));
  CHECK(nmx_vars_crc(B, NMX_REG_U, &cb));
  EXPECT(ca == cb, "vars CRC differs: %016llx %016llx", (unsigned long long)ca, (unsigned long long)cb);
  CHECK(nmx_state_download(A, NMX_REG_U, ra));
  double **out = element_arrays(T, NULL);
  CHECK(nmx_state_download_ptrs(B, NMX_REG_U, out, T->ne, 5));
  int same = 1, moved = 0;
  for (int e = 0; e < T->ne; e++) {
    long off = T->info[NINFO * e + 3], np = T->info[NINFO * e + 2];
    np = np * np * np;
    for (int v = 0; v < 5; v++) {
      same &= !memcmp(out[5 * e + v], ra + v * T->nnodes + off, np * sizeof(double));
      moved |= memcmp(out[5 * e + v], u + v * T->nnodes + off, np * sizeof(double)) != 0;
    }
  }
  EXPECT(same, "per-element arrays differ from the flat download");
  EXPECT(moved, "the state did not change in 3 steps");
  free_arrays(T, arr);
  free_arrays(T, out);
  free(u); free(ra); free(rb);
}

int main(int argc, char **argv) {
  if (argc != 2 || (strcmp(argv[1], "host") && strcmp(argv[1], "device"))) {
    printf("usage: abi_client host|device\n");
    return 2;
  }
  const int device = strcmp(argv[1], "device") ? -1 : 0;
  nmx_ctx *A = built_mesh(device);
  Tables TA, TB;
  read_tables(A, &TA);
  int maxlev = 0, nmin = 10, nmax = 2;
  for (int e = 0; e < TA.ne; e++) {
    maxlev = TA.elev[e] > maxlev ? TA.elev[e] : maxlev;
    nmin = TA.en[3 * e] < nmin ? TA.en[3 * e] : nmin;
    nmax = TA.en[3 * e] > nmax ? TA.en[3 * e] : nmax;
  }
  EXPECT(TA.ne == 69 && maxlev == 2 && nmin == 4 && nmax == 7, "mesh: %d elements, level %d, n %d..%d", TA.ne,
         maxlev, nmin, nmax);
  nmx_ctx *B;
  CHECK(nmx_ctx_create(&B, device));
  set_pars(B);
  CHECK(import_tables(B, &TA));
  read_tables(B, &TB);
  compare_tables(&TA, &TB);
  check_refusals(device, &TA);
  if (device >= 0) device_checks(A, B, &TA);
  free_tables(&TA);
  free_tables(&TB);
  CHECK(nmx_ctx_destroy(B));
  CHECK(nmx_ctx_destroy(A));
  if (failed) return 1;
  printf("abi_client %s ok\n", argv[1]);
  return 0;
}
