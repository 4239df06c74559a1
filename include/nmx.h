/* nmx.h -- C ABI of the MI355X-native DG/FV evolution library (libnmx.so).
 *
 * The library replaces nmesh's element-local evolution hot path: the DG
 * right-hand side of the first-order scalar wave (scalarwave1), the face
 * numerical flux and lift, the Runge-Kutta stage update, the FV WENO5 + LLF
 * reconstruction, and a discrete adjoint of the RK step.  Everything below is
 * plain C: opaque handles, pointers and sizes.  Every entry returns 0 on
 * success and a negative NMX_E* code on failure; nmx_last_error() then holds a
 * message (per thread).  One host thread drives one context; work is ordered
 * on the context's HIP stream (nmx_set_stream).
 *
 * Reference call surface each group replaces (paths relative to the nmesh
 * source tree):
 *   - parameters by name: AddPar/Par/Get* (src/main/main/nmesh_main.h:89-107);
 *     names are nmesh's own (amr_*, dt, dtfac, evolve_method, scalarwave1_*).
 *   - mesh set-up: amr_setup_mesh (src/main/amr/mesh.c:181-319).
 *   - per-element variable arrays: Vard(node, vi) (src/main/amr/nmesh_amr_loops.h:181).
 *   - RHS: evolve_setrhs_mesh (src/main/evolve/evolve.h:15, evolve.c:182-232).
 *   - steppers: evolve_myln / evolve_RK4_mesh / evolve_sspRK3_mesh /
 *     evolve_Euler_mesh (src/main/evolve/evolve.c:25, RungeKutta.c:19-188).
 *   - time step: update_mesh_myelms_elm_eid_dt (src/main/amr/storage.c:1370-1386).
 *   - checkpoint variables and CRCs: checkpoint_save_EvoVars /
 *     checkpoint_load_Vars (src/main/checkpoint/checkpoint_save.c:337-468,
 *     checkpoint_load.c:396-700), nmesh_CRCs (src/main/main/norms.c:480-520).
 */
#ifndef NMX_H
#define NMX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define NMX_OK 0
#define NMX_EINVAL (-1)   /* bad argument or parameter value */
#define NMX_ESTATE (-2)   /* call out of order (e.g. no mesh yet) */
#define NMX_EHIP (-3)     /* HIP runtime error */
#define NMX_ENOMEM (-4)   /* allocation failed */
#define NMX_EUNSUP (-5)   /* configuration not supported */
#define NMX_EIO (-6)      /* a file could not be opened or written */

/* state registers */
#define NMX_REG_U 0       /* evolved variables u (scalarwave1 vlu) */
#define NMX_REG_R 1       /* RHS r = RHS(u) ("_r" vars, evosys.c:256) */
#define NMX_REG_P 2       /* u at previous time ("_p" vars, evosys.c:257) */
#define NMX_REG_ADJ 3     /* adjoint variable lam = dJ/du (no counterpart in the reference) */

/* variables of scalarwave1's vlu, in order (scalarwave1.c:670-672) */
#define NMX_NVARS 5       /* pi, cx, cy, cz, phi */

typedef struct nmx_ctx nmx_ctx;

const char *nmx_last_error(void);

/* context: device is the HIP device ordinal. */
int nmx_ctx_create(nmx_ctx **ctx, int device);
int nmx_ctx_destroy(nmx_ctx *ctx);
/* stream is a hipStream_t (NULL = the library's own stream) */
int nmx_set_stream(nmx_ctx *ctx, void *stream);
int nmx_synchronize(nmx_ctx *ctx);

/* Parameters under nmesh's names (replaces AddPar/Par, nmesh_main.h:89-107):
 * amr_mesh_type (BoxMesh|Shell), amr_n0/1/2, amr_luni, amr_BoxMesh_dout,
 * amr_mesh_xc, amr_Shell_rin, amr_Shell_rout, amr_hrefine_sphere_radius,
 * amr_hrefine_sphere_levels, dt (auto|number), dtfac, uniform_dtfac (dt on
 * FV elements, nmesh_main.c:32), iterations, finaltime (nmx_evolve_mesh),
 * evolve_method
 * (RK4|sspRK3|Euler), scalarwave1_k, scalarwave1_numflux (upwind|LLF),
 * scalarwave1_profile (sin), the FILTER bin's evolve_filter_all (yes|no),
 * evolve_filter_varlist, evolve_filter_alp, evolve_filter_s, evolve_filter_dn
 * (nmesh_evolve.c:35-39), basis_expfilter_JacobianPower (1|0|-1) and
 * basis_filter_fv (yes|no) (nmesh_basis.c:22-24), the FV elements'
 * amr_uniform_p, amr_fv_p, fv_rec (1|WENOm3|WENOm5|WENOmZ, each WENOm also
 * _1 and _2, and WENOmT_2) with fv_rec_WENOm_s1, _s2, _opt (nmesh_dg.c:35-40), and
 * coordinates_midpoint_data (no|yes|yes all, nmesh_coordinates.c:84),
 * fv_surface_interp (linear|parabolic, optionally with the word
 * Lagrange_if_dg, nmesh_dg.c:53; set on a built mesh
 * it rebuilds the face tables and keeps the state), fv_divf_extrap_s1,
 * fv_divf_extrap_s2 (a number or DBL_MAX), fv_divf_extrap_opt
 * (nmesh_dg.c:41-46),
 * scalarwave1_limiter (none|MRS|minmodB; nmesh_scalarwave1.c:34) with
 * limiter_alpha, limiter_beta, limiter_scaleBound (nmesh_limiter.c:19-23):
 * the limiter on every later RK stage's input
 * and on the final u of each step (RungeKutta.c:45-67,161-173;
 * evolve.c:143-144),
 * accepted with the reference's defaults only: amr_dir_active0/1/2 (yes),
 * coordinates_surface_metric (sqrtgdiag), amr_Shell_nr (1), amr_Shell_r1
 * (<= amr_Shell_rin); amr_nmax bounds amr_n0; amr_fv_if_uniform (yes: FV
 * exactly in amr_uniform_p, mesh.c:86-91); accepted without effect on this
 * path: dg_outerBC_flux_fac (3 numbers; its use in scalarwave1 is commented
 * out, scalarwave1.c:202), evolve_u_surfaces (one device: no surface
 * exchange); the troubled-element machinery below: nmx_trouble
 * (none|RDMP|hook), evolve_redo_troubled (substep|step), nmx_trouble_delta0,
 * nmx_trouble_eps,
 * plus library knobs
 * prefixed nmx_ (nmx_path:
 * plain|batched|perelement, nmx_graph: yes|no, nmx_graph_steps, nmx_geometry:
 * stored|analytic, nmx_block: full|half|quarter, nmx_xcd: yes|no,
 * nmx_autotune: yes|no, nmx_batch: legacy|buckets|persistent (the batched
 * path's launch shape; the autotuner times the three unless it is set),
 * nmx_stream_hints: 0..7 (streaming cache hints of the stage epilogue)). */
int nmx_set_param(nmx_ctx *ctx, const char *name, const char *value);
int nmx_get_param(nmx_ctx *ctx, const char *name, char *buf, int buflen);

/* Build the mesh, geometry, neighbour tables and device mirror from the
 * parameters (replaces amr_setup_mesh, mesh.c:181-319, and coordinates_init,
 * coordinates.c:133-443).  Sets mesh time 0 and dt per the dt parameter. */
int nmx_mesh_build(nmx_ctx *ctx);
/* Take the mesh nmesh already holds (the patch and element tables behind
 * amr_setup_mesh, mesh.c:181-319; tElm, nmesh_amr.h:87-102), instead of
 * building one:
 *   npat patches: pat_type 0 box, 1 CubedShell; pat_geo[12 p ..]: dom, s0,
 *     s1 (CubedSphere radii), bbox[6], xc[3];
 *   nelem elements in eid order: elm_pat, elm_level, elm_loc[24 e + k] the
 *     octant digit of level k+1 (nmesh's eloc->loc, '0'+ijk or ijk),
 *     elm_n[3 e + d] points per direction (equal), elm_bbox[6 e ..];
 *   per face: nfnb[6 e + f] neighbours (node->nfnb), fnb[48 e + 8 f + k] their
 *     element index (node->fnb in eid order), nbface[...] the face of the
 *     neighbour touching f (amr_get_nbface, connections.c:1114-1157).
 * The library rebuilds the forest from the locations and requires every
 * table it derives to equal the caller's: the boxes bit for bit, neighbour
 * counts, neighbours in nmesh's order and their faces
 * (amr_elm_nbinfo_to_elm_fnb, connections.c:2882; locate_facenb_in_fnbs,
 * connections.c:1090).  A difference is NMX_EINVAL naming the first element
 * and face that differ; on success the mesh is taken (as nmx_mesh_build). */
int nmx_mesh_import(nmx_ctx *ctx, int npat, const int *pat_type, const double *pat_geo, int nelem,
                    const int *elm_pat, const int *elm_level, const unsigned char *elm_loc, const int *elm_n,
                    const double *elm_bbox, const int *nfnb, const int *fnb, const int *nbface);
/* Set points per direction of every element (n[nelem]) at set-up and rebuild
 * tables; the registers start from zero (element orders chosen before the
 * initial data, as mesh.c:181-319 does with amr_n). */
int nmx_mesh_set_elm_n(nmx_ctx *ctx, const int *n, int nelem);
/* Adapt the mesh between steps with the evolved state (NMX_REG_U)
 * interpolated onto the new elements; the other registers start from zero,
 * time is kept and dt follows the dt parameter.  DG meshes (no FV elements).
 * nmx_mesh_hadapt: flag > 0 replaces an element by its 8 children of the
 *   same n, interpolated from it (hp_refine_elms_if_rflag with PARENT_n,
 *   src/main/amr/refine.c:174-276; make_child_elm, storage.c:493-585);
 *   8 siblings in a row all with flag < 0 become their parent, whose n is
 *   the largest sibling n and whose points average the interpolants of the
 *   children that hold them (remove_elms_if_rflag, refine.c:359-506;
 *   make_parent_elm, storage.c:627-760).
 * nmx_mesh_padapt: new points per direction n[nelem], the state Lagrange-
 *   interpolated per element (update_node_n_pt_typ, storage.c:858-989). */
int nmx_mesh_hadapt(nmx_ctx *ctx, const int *flag, int nelem);
int nmx_mesh_padapt(nmx_ctx *ctx, const int *n, int nelem);
/* Switch elements between DG on LGL points (to_fv[e] = 0) and FV on uniform
 * points of the same n (to_fv[e] = 1; n >= 3), the state interpolated: LGL
 * to uniform by Lagrange (INTERP_LGL_TO_n_UNIFORM, LGL_to_nUNIt) and back by
 * its inverse, the Lagrange interpolation from the uniform points
 * (INTERP_UNIFORM_TO_n_LGL, UNI_to_nLGLt with basis_BackInterpMatrix =
 * PseudoInv; gridpoints.c:158-290).  The mechanism of
 * evolve_switch_troubled_nodes_mesh / evolve_switch_nontroubled_nodes_mesh
 * (src/main/evolve/trouble.c:186-330), called by hand here; with
 * nmx_trouble set, nmx_evolve switches by the trouble scores itself.
 * Meshes mixing both kinds run on the batched path. */
int nmx_mesh_switch_fv(nmx_ctx *ctx, const int *to_fv, int nelem);
/* The mesh's dt computed on the device from the mirror's node coordinates
 * (update_mesh_myelms_elm_eid_dt, storage.c:1305-1386, with find_hmin,
 * utilities.c:1028-1128): per element the smallest corner-node distance
 * along each direction (from the space diagonal / sqrt 3), times
 * uniform_dtfac on FV elements whose minimum is such a pair and dtfac
 * otherwise, minimised with wavefront shuffles.  A query: the mesh's dt
 * (nmx_get_dt, set up on the host by the same rule) is not changed. */
int nmx_mesh_dt_device(nmx_ctx *ctx, double *dt);
/* fv[e] = 1 for an FV element on uniform points, 0 for DG on LGL points
 * (node->dat->use_fv, refine_set_use_fv_true_iff_P_UNIFORM, trouble.c:246) */
int nmx_mesh_fv_flags(nmx_ctx *ctx, int *fv, int nelem);

int nmx_mesh_nelem(nmx_ctx *ctx, int *nelem);
int nmx_mesh_nnodes(nmx_ctx *ctx, long *nnodes);
/* info[22]: patch, level, n, node offset, nfnb[6], ajsurf mode[6] (0 none,
 * 1 alias, 2 oriented copy, 3 interpolation), orientation code[6]
 * (intrch*4 + rev1*2 + rev2 as in copy_ajsurf_from_nbsurf0, surface.c:1361) */
int nmx_mesh_elm_info(nmx_ctx *ctx, int elm, long *info);
int nmx_mesh_elm_bbox(nmx_ctx *ctx, int elm, double *bbox6);
/* face neighbours of elm on face f (node->fnb[f], nmesh_amr.h:96-98); nb and
 * nbface must hold 8 entries; returns the count in *nnb */
int nmx_mesh_elm_fnb(nmx_ctx *ctx, int elm, int f, int *nb, int *nbface, int *nnb);
/* per-node Cartesian coordinates x,y,z as x[d*nnodes + node] */
int nmx_mesh_coords(nmx_ctx *ctx, double *x);
/* dXdx[(3a+b)*nnodes+node], sqrtgdiag[a*nnodes+node] (coordinates.c:133-224,728-805) */
int nmx_mesh_geometry(nmx_ctx *ctx, double *dXdx, double *sqrtgdiag);
/* time step from the dt/dtfac parameters (storage.c:1305-1386) and time;
 * nmx_set_dt sets the parameter dt to the value, so it stays through mesh
 * changes until dt is set to auto again */
int nmx_get_dt(nmx_ctx *ctx, double *dt);
int nmx_set_dt(nmx_ctx *ctx, double dt);
int nmx_get_time(nmx_ctx *ctx, double *t);
int nmx_set_time(nmx_ctx *ctx, double t);

/* State transfer in the flat layout u[var*nnodes + node], node = element
 * offset + ijk, ijk = i + n*(j + n*k) (nmesh_amr_loops.h:279). */
int nmx_state_upload(nmx_ctx *ctx, int reg, const double *u);
int nmx_state_download(nmx_ctx *ctx, int reg, double *u);
/* one variable of one element, as Vard(node, vi) would hold it */
int nmx_state_upload_elm(nmx_ctx *ctx, int reg, int elm, int var, const double *d, int np);
int nmx_state_download_elm(nmx_ctx *ctx, int reg, int elm, int var, double *d, int np);
/* every variable of every element in one call: ptrs[5*elm + var] is the
 * caller's array of that variable of that element (Vard(node, vi),
 * nmesh_amr_loops.h:181; np doubles in ijk order), var in scalarwave1's vlu
 * order pi, cx, cy, cz, phi (scalarwave1.c:670-672); nvar must be 5.  The
 * arrays are gathered into pinned host staging and moved in one copy (and
 * the reverse), which replaces nelem*5 single transfers. */
int nmx_state_upload_ptrs(nmx_ctx *ctx, int reg, const double *const *ptrs, int nelem, int nvar);
int nmx_state_download_ptrs(nmx_ctx *ctx, int reg, double *const *ptrs, int nelem, int nvar);
/* set u to scalarwave1's profile at time t (scalarwave1_set_profile, scalarwave1.c:574-642) */
int nmx_state_set_profile(nmx_ctx *ctx, int reg, double t);

/* Time the stage kernel variants (nodes per thread, XCD-aware element
 * mapping) on the current mesh and keep the fastest; runs automatically
 * before the first nmx_evolve when nmx_autotune = yes.  No counterpart in the
 * reference (its element loop is an OpenMP for, nmesh_amr_loops.h:61-63). */
int nmx_autotune(nmx_ctx *ctx);

/* Troubled elements (trouble.c:27-400; RungeKutta.c:19-344; evolve.c:
 * 25-176).  With nmx_trouble = RDMP or hook, nmx_evolve scores every element
 * after each substep (evolve_redo_troubled = substep, the default) or after
 * the step (= step): a TROUBLE value 1 / 0 / -1 per element moves its
 * trouble score (evolve_set_trouble_score); when the mesh's score is
 * positive the troubled elements become FV elements on uniform points
 * (same n) and redo their substep (evolve_trouble_redo_u_step_mesh), or the
 * whole step is redone from u_p (evolve_prepare_do_over_mesh); after the
 * step, elements whose score reached -11 per substep go back to DG
 * (evolve_switch_nontroubled_nodes_mesh).  RDMP is the library's indicator:
 * a value of u outside the u_p extremes of the element and its face
 * neighbours widened by max(nmx_trouble_delta0, nmx_trouble_eps (M - m))
 * makes a DG element troubled, an FV element in bounds is good.
 * nmx_trouble_fn is a project's TROUBLE function on the device
 * (evosys->f[TROUBLE]): for every element it writes tr[e] in {1, 0, -1}
 * from u (after the substep) and u_p (the step's start), registers in
 * element blocks (nmx_state_device_ptr), ordered on stream (a hipStream_t);
 * it returns 0 on success.  nmx_get_trouble_scores reads the scores
 * (node->dat->info->trbl_score) and counts of redone steps and switches. */
typedef int (*nmx_trouble_fn)(void *user, const double *u, const double *u_p, int *tr, int nelem, void *stream);
int nmx_set_trouble_hook(nmx_ctx *ctx, nmx_trouble_fn fn, void *user);
int nmx_get_trouble_scores(nmx_ctx *ctx, int *score, int nelem, long *redos, long *switches);

/* r = RHS(u, t) on the device (evolve_setrhs_mesh, evolve.c:182-232) */
int nmx_setrhs(nmx_ctx *ctx, int reg_out, int reg_in, double t);
/* nsteps full time steps with evolve_method (evolve_myln, evolve.c:25-176;
 * RungeKutta.c:19-188).  Advances NMX_REG_U and the mesh time. */
int nmx_evolve(nmx_ctx *ctx, int nsteps);
/* evolve_mesh (main.c:583-666): full steps until the mesh time reaches the
 * parameter finaltime, or iterations * dt when finaltime <= 0, within
 * max(dt 1e-7, finaltime 1e-14); nothing when both are 0 (their defaults,
 * nmesh_main.c:34-35).  Te starts at the mesh time as in evolve_mesh.  The
 * step count is fixed at the call from the current dt, finaltime and
 * iterations (the reference re-reads them after each step): callers that
 * change dt or those parameters during a run call this again in chunks. */
int nmx_evolve_mesh(nmx_ctx *ctx);
/* full steps taken since the mesh was built (mesh->iteration, main.c:617) */
int nmx_get_iteration(nmx_ctx *ctx, long *iteration);
/* The FILTER bin's exponential filter, in place on register reg, with the
 * evolve_filter_* and basis_expfilter_* parameters (evolve_filter_evosys_mesh,
 * evolve.c:656-710; expfilter_var, filter.c:75-92).  nmx_evolve applies it
 * after every full step when evolve_filter_all or evolve_filter_varlist asks
 * for it (main.c:628-630).  No filter requested: does nothing. */
int nmx_filter(nmx_ctx *ctx, int reg);

/* Checkpoint variables file and CRCs.
 * nmx_checkpoint_save_vars writes register reg as nmesh's variables.bin:
 * scalarwave1's evolved variables (phi, pi, cx, cy, cz) of every element,
 * little-endian doubles, under nmesh's element names
 * (checkpoint_save_EvoVars / checkpoint_write_vl, checkpoint_save.c:337-468).
 * nmx_checkpoint_load_vars reads one into register reg
 * (checkpoint_load_Vars, checkpoint_load.c:396-700): every element of the
 * mesh must be present with its five variables and point count; variables
 * of other names are skipped; on any error (NMX_EINVAL) the register is
 * left unchanged.  Only the variables: time, mesh and parameters stay the
 * caller's (nmx_set_time).
 * nmx_vars_crc computes on the device the "vars" CRC of nmesh_CRCs for
 * register reg (norms.c:440-458 crc_VarList_local; what
 * checkpoint_save_CRCs writes as "vars:"), fewer than 2^32 nodes.
 * nmx_crc64 and nmx_crc64_combine are nmesh's crc64_continue and
 * crc64_0start_combine (crc64.c:153-256): reflected CRC-64, ECMA-182
 * polynomial, pre- and post-inverted. */
int nmx_checkpoint_save_vars(nmx_ctx *ctx, int reg, const char *path);
int nmx_checkpoint_load_vars(nmx_ctx *ctx, int reg, const char *path);
int nmx_vars_crc(nmx_ctx *ctx, int reg, uint64_t *crc);
/* nmx_vars_crc_noalloc: no counterpart in the reference.  It is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same result as
 * nmx_vars_crc, but it calls no allocator: the CRC tables, the per-element
 * contributions and the result word live in one static device array per GPU
 * (256 KiB + 45 KiB, set up by nmx_ctx_create), so meshes of at most 32,768
 * elements (else NMX_EUNSUP).  The array keeps the last call's contents;
 * calls that overlap on one GPU race on it. */
int nmx_vars_crc_noalloc(nmx_ctx *ctx, int reg, uint64_t *crc);
uint64_t nmx_crc64(uint64_t crc, const void *buf, size_t nbytes);
uint64_t nmx_crc64_combine(uint64_t crc1, uint64_t crc2, uint64_t nbytes2);

/* device pointer and length (doubles) of a register, for zero-copy use from
 * frameworks; valid until the mesh is rebuilt */
int nmx_state_device_ptr(nmx_ctx *ctx, int reg, void **dptr, long *len);

/* Discrete adjoint (no counterpart in the reference; it transposes the
 * forward operators above, evolve_setrhs_mesh and evolve_RK4_mesh /
 * evolve_sspRK3_mesh / evolve_Euler_mesh, RungeKutta.c:19-188).  The
 * scalarwave1 RHS is linear in u, so its Jacobian L does not depend on the
 * state or the time; the boundary profile enters only as a u-independent
 * source.  DG meshes of every kind: conforming meshes of one order, and
 * h-refined mixed-order meshes on the batched path, where the transpose of
 * the interpolated and averaged faces (surface.c:1081-1337) is applied.
 * Meshes with FV elements, the limiter and the filter: NMX_EUNSUP.
 *
 * reg_out = L^T reg_in */
int nmx_adjoint_setrhs(nmx_ctx *ctx, int reg_out, int reg_in);
/* NMX_REG_ADJ <- (du_{n+1}/du_n)^T NMX_REG_ADJ, nsteps times, for
 * evolve_method and the current dt (the forward stages in reverse).  The
 * mesh time is not changed. */
int nmx_adjoint_evolve(nmx_ctx *ctx, int nsteps);
/* time (ms) of the last nmx_adjoint_evolve on the device */
int nmx_last_adjoint_ms(nmx_ctx *ctx, double *ms);

/* Checkpoints for reverse passes: a pool of nslots device copies of a
 * register; store/load copy reg <-> slot on the stream, and keep the mesh
 * time with the slot (restored by loading into NMX_REG_U). */
int nmx_checkpoint_alloc(nmx_ctx *ctx, int nslots);
int nmx_checkpoint_store(nmx_ctx *ctx, int slot, int reg);
int nmx_checkpoint_load(nmx_ctx *ctx, int slot, int reg);

/* State transfer between a register and DEVICE memory in the flat layout
 * u[var*nnodes + node], ordered on the context's stream with no host
 * synchronisation (for frameworks that hold the state on the GPU). */
int nmx_state_copy_in(nmx_ctx *ctx, int reg, const double *dflat);
int nmx_state_copy_out(nmx_ctx *ctx, int reg, double *dflat);

/* ------------------------------------------------------------------------
 * FV WENO5 + LLF on a periodic grid of 16^3-cell blocks, special-relativistic
 * Euler with an ideal gas (conserved D, S_i, tau; primitive rho, v_i, p).
 * Replaces the reconstruction and flux of fv_divf (src/main/dg/fv.c:115-624)
 * with rec1d_p/m_WENO5 (rec1d.c:1190-1352) and numflux1d_LLF
 * (numflux1d.c:85-112), stepped by evolve_sspRK3_mesh / evolve_RK4_mesh /
 * evolve_Euler_mesh (RungeKutta.c:19-188).  The SR-Euler physics has no
 * counterpart in the reference (its SR projects are absent).  Parameters:
 * fv_n (cells per edge, a multiple of 16), fv_block (16), fv_gamma,
 * fv_c2p_tol (Newton tolerance, relative), dt (auto = fv_dtfac * dx, dx = 1),
 * fv_dtfac, evolve_method, fv_rec (WENO5), fv_flux (LLF), nmx_graph,
 * nmx_graph_steps, nmx_xcd.  Host arrays are flat: u[var*ncell + cell],
 * cell = i + N (j + N k).
 * ------------------------------------------------------------------------ */
#define NMX_FV_CONS 0     /* conserved state (evolved) */
#define NMX_FV_PRIM 1     /* primitives of the last c2p */
#define NMX_FV_RHS 2      /* RHS of nmx_fv_setrhs */

typedef struct nmx_fv nmx_fv;

int nmx_fv_create(nmx_fv **fv, int device);
int nmx_fv_destroy(nmx_fv *fv);
int nmx_fv_set_stream(nmx_fv *fv, void *stream);
int nmx_fv_synchronize(nmx_fv *fv);
int nmx_fv_set_param(nmx_fv *fv, const char *name, const char *value);
int nmx_fv_get_param(nmx_fv *fv, const char *name, char *buf, int buflen);
/* allocate the grid from the parameters; time 0, dt per the dt parameter */
int nmx_fv_setup(nmx_fv *fv);
int nmx_fv_ncells(nmx_fv *fv, long *ncell);
int nmx_fv_get_dt(nmx_fv *fv, double *dt);
int nmx_fv_set_dt(nmx_fv *fv, double dt);
int nmx_fv_get_time(nmx_fv *fv, double *t);
int nmx_fv_set_time(nmx_fv *fv, double t);
/* primitives -> the conserved state (and the c2p guess) */
int nmx_fv_upload_prim(nmx_fv *fv, const double *prim);
int nmx_fv_upload(nmx_fv *fv, int which, const double *u);
int nmx_fv_download(nmx_fv *fv, int which, double *u);
/* NMX_FV_RHS = RHS(conserved state), with c2p first */
int nmx_fv_setrhs(nmx_fv *fv);
/* the CFL reduction: c2p of the conserved state, then the largest
 * characteristic speed |lambda+-| (the SR-Euler eigenvalues of the LLF
 * flux) over every cell and direction, reduced on the device with
 * wavefront shuffles; a caller's CFL dt is fv_dtfac dx / *amax.  No
 * counterpart in the reference (its SR projects are absent); dt itself stays
 * the dt parameter's. */
int nmx_fv_max_speed(nmx_fv *fv, double *amax);
int nmx_fv_evolve(nmx_fv *fv, int nsteps);
int nmx_fv_last_evolve_ms(nmx_fv *fv, double *ms);
/* number of blocks whose c2p Newton solve failed since the last call */
int nmx_fv_status(nmx_fv *fv, long *nbad);

/* profiling helpers: time (ms) of the last nmx_evolve on the device and the
 * number of kernel launches it issued */
int nmx_last_evolve_ms(nmx_ctx *ctx, double *ms);

#ifdef __cplusplus
}
#endif
#endif
