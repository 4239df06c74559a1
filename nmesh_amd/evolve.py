"""Host-side mirror of nmesh's evolution interface over libnmx.

Names follow the reference (paths relative to the nmesh source tree):

- ``Mesh`` stands in for ``tMesh`` plus its parameter database
  (src/main/amr/nmesh_amr.h:227-251, src/main/main/parameters.c); parameters
  are set under nmesh's own names (``amr_mesh_type``, ``amr_luni``, ``dt``,
  ``evolve_method``, ``scalarwave1_k`` ...).
- ``amr_setup_mesh`` (src/main/amr/mesh.c:181-319)
- ``evolve_setrhs_mesh`` (src/main/evolve/evolve.c:182-232)
- ``evolve_RK4_mesh`` / ``evolve_sspRK3_mesh`` / ``evolve_Euler_mesh``
  (src/main/evolve/RungeKutta.c:19-188) and ``evolve_myln`` (evolve.c:25-176)
- ``scalarwave1_set_profile`` (src/tests/scalarwave1/scalarwave1.c:574-642)

Errors raise ``NmxError`` where the reference would call ``errorexit``
(src/main/main/nmesh_main.h:244-252).
"""
import ctypes as C

import numpy as np

from ._lib import check, lib

NVARS = 5
VARS = ("scalarwave1_pi", "scalarwave1_cx", "scalarwave1_cy", "scalarwave1_cz",
        "scalarwave1_phi")
REG_U, REG_R, REG_P, REG_ADJ = 0, 1, 2, 3


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Mesh:
    """A mesh with its device mirror; one context per Mesh."""

    def __init__(self, pars=None, device=0):
        L = lib()
        self._ctx = C.c_void_p()
        check(L.nmx_ctx_create(C.byref(self._ctx), int(device)))
        for k, v in (pars or {}).items():
            self.set_par(k, v)

    def close(self):
        if self._ctx:
            lib().nmx_ctx_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # parameters (AddPar / Par / Gets)
    def set_par(self, name, value):
        check(lib().nmx_set_param(self._ctx, name.encode(), str(value).encode()))

    def par(self, name):
        buf = C.create_string_buffer(256)
        check(lib().nmx_get_param(self._ctx, name.encode(), buf, 256))
        return buf.value.decode()

    @property
    def ctx(self):
        return self._ctx

    @property
    def nelem(self):
        n = C.c_int()
        check(lib().nmx_mesh_nelem(self._ctx, C.byref(n)))
        return n.value

    @property
    def nnodes(self):
        n = C.c_long()
        check(lib().nmx_mesh_nnodes(self._ctx, C.byref(n)))
        return n.value

    @property
    def dt(self):
        d = C.c_double()
        check(lib().nmx_get_dt(self._ctx, C.byref(d)))
        return d.value

    @dt.setter
    def dt(self, v):
        check(lib().nmx_set_dt(self._ctx, float(v)))

    @property
    def time(self):
        d = C.c_double()
        check(lib().nmx_get_time(self._ctx, C.byref(d)))
        return d.value

    @time.setter
    def time(self, v):
        check(lib().nmx_set_time(self._ctx, float(v)))

    @property
    def iteration(self):
        """full steps taken since the mesh was built (mesh->iteration)"""
        it = C.c_long()
        check(lib().nmx_get_iteration(self._ctx, C.byref(it)))
        return it.value

    def elm_info(self, e):
        a = np.zeros(22, dtype=np.int64)
        check(lib().nmx_mesh_elm_info(self._ctx, int(e), _p(a)))
        return a

    def elm_bbox(self, e):
        a = np.zeros(6)
        check(lib().nmx_mesh_elm_bbox(self._ctx, int(e), _p(a)))
        return a

    def elm_fnb(self, e, f):
        nb = np.zeros(8, dtype=np.int32)
        nbf = np.zeros(8, dtype=np.int32)
        n = C.c_int()
        check(lib().nmx_mesh_elm_fnb(self._ctx, int(e), int(f), _p(nb), _p(nbf), C.byref(n)))
        return nb[:n.value].copy(), nbf[:n.value].copy()

    def coords(self):
        x = np.zeros(3 * self.nnodes)
        check(lib().nmx_mesh_coords(self._ctx, _p(x)))
        return x.reshape(3, -1)

    def geometry(self):
        nn = self.nnodes
        dXdx = np.zeros(9 * nn)
        sg = np.zeros(3 * nn)
        check(lib().nmx_mesh_geometry(self._ctx, _p(dXdx), _p(sg)))
        return dXdx.reshape(9, -1), sg.reshape(3, -1)

    def set_elm_n(self, n):
        n = np.ascontiguousarray(n, dtype=np.int32)
        check(lib().nmx_mesh_set_elm_n(self._ctx, _p(n), int(n.size)))

    def hadapt(self, flag):
        """h-adapt between steps, the state interpolated (refine.c:174-506):
        flag > 0 refines an element, 8 siblings in a row with flag < 0
        coarsen into their parent"""
        f = np.ascontiguousarray(flag, dtype=np.int32)
        check(lib().nmx_mesh_hadapt(self._ctx, _p(f), int(f.size)))

    def switch_fv(self, to_fv):
        """elements to FV on uniform points (1) or DG on LGL points (0),
        the state interpolated (evolve_switch_troubled_nodes_mesh and
        evolve_switch_nontroubled_nodes_mesh, trouble.c:186-330)"""
        f = np.ascontiguousarray(to_fv, dtype=np.int32)
        check(lib().nmx_mesh_switch_fv(self._ctx, _p(f), int(f.size)))

    def dt_device(self):
        """the mesh's dt from the device's node coordinates (a query)"""
        d = C.c_double()
        check(lib().nmx_mesh_dt_device(self._ctx, C.byref(d)))
        return d.value

    def fv_flags(self):
        """1 for FV elements on uniform points, 0 for DG (node->dat->use_fv)"""
        f = np.zeros(self.nelem, dtype=np.int32)
        check(lib().nmx_mesh_fv_flags(self._ctx, _p(f), int(f.size)))
        return f

    def set_trouble_hook(self, fn):
        """Install a TROUBLE function (evosys->f[TROUBLE]): fn(u, u_p, tr,
        nelem, stream) with device pointers (element blocks) and a
        hipStream_t, returning 0; it must write tr[e] in {1, 0, -1} for
        every element in stream order.  None removes it.  The callback
        object is kept alive with the mesh."""
        if fn is None:
            self._trouble_cb = None
            check(lib().nmx_set_trouble_hook(self._ctx, None, None))
            return
        proto = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p)
        self._trouble_cb = proto(lambda user, u, up, tr, n, stream: int(fn(u, up, tr, n, stream)))
        check(lib().nmx_set_trouble_hook(self._ctx, C.cast(self._trouble_cb, C.c_void_p), None))

    def trouble_scores(self):
        """(scores per element, redone substeps or steps, switches)
        (node->dat->info->trbl_score, trouble.c:27-113)"""
        s = np.zeros(self.nelem, dtype=np.int32)
        redos, switches = C.c_long(), C.c_long()
        check(lib().nmx_get_trouble_scores(self._ctx, _p(s), int(s.size), C.byref(redos), C.byref(switches)))
        return s, redos.value, switches.value

    def padapt(self, n):
        """new points per direction per element, the state interpolated
        (update_node_n_pt_typ, storage.c:858-989)"""
        n = np.ascontiguousarray(n, dtype=np.int32)
        check(lib().nmx_mesh_padapt(self._ctx, _p(n), int(n.size)))

    # state (Vard over all elements, flat [var][node])
    def upload(self, u, reg=REG_U):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        if u.size != NVARS * self.nnodes:
            raise ValueError("state must hold 5*nnodes values")
        check(lib().nmx_state_upload(self._ctx, int(reg), _p(u)))

    def download(self, reg=REG_U):
        u = np.zeros(NVARS * self.nnodes)
        check(lib().nmx_state_download(self._ctx, int(reg), _p(u)))
        return u

    def Vard(self, e, vi, reg=REG_U):
        info = self.elm_info(e)
        np_ = int(info[2]) ** 3
        d = np.zeros(np_)
        check(lib().nmx_state_download_elm(self._ctx, int(reg), int(e), int(vi), _p(d), np_))
        return d

    def set_Vard(self, e, vi, d, reg=REG_U):
        d = np.ascontiguousarray(d, dtype=np.float64)
        check(lib().nmx_state_upload_elm(self._ctx, int(reg), int(e), int(vi), _p(d), int(d.size)))

    def upload_vars(self, arrays, reg=REG_U):
        """every variable of every element from per-element arrays
        (arrays[e][v], as nmesh's Vard holds them) in one transfer"""
        ptrs = (C.c_void_p * (NVARS * self.nelem))()
        keep = []
        for e in range(self.nelem):
            for v in range(NVARS):
                a = np.ascontiguousarray(arrays[e][v], dtype=np.float64)
                keep.append(a)
                ptrs[NVARS * e + v] = a.ctypes.data
        check(lib().nmx_state_upload_ptrs(self._ctx, int(reg), ptrs, self.nelem, NVARS))

    def download_vars(self, arrays, reg=REG_U):
        """the reverse of upload_vars, into the given float64 arrays"""
        ptrs = (C.c_void_p * (NVARS * self.nelem))()
        for e in range(self.nelem):
            for v in range(NVARS):
                a = arrays[e][v]
                if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
                    raise ValueError("download_vars needs contiguous float64 arrays")
                ptrs[NVARS * e + v] = a.ctypes.data
        check(lib().nmx_state_download_ptrs(self._ctx, int(reg), ptrs, self.nelem, NVARS))

    # device-resident state (flat [var][node] at a device pointer), ordered
    # on the context's stream
    def set_stream(self, stream_handle):
        """Run the library's work on an existing hipStream_t (an integer
        handle such as torch.cuda.current_stream().cuda_stream)."""
        check(lib().nmx_set_stream(self._ctx, C.c_void_p(int(stream_handle))))

    def copy_in(self, dptr, reg=REG_U):
        check(lib().nmx_state_copy_in(self._ctx, int(reg), C.c_void_p(int(dptr))))

    def copy_out(self, dptr, reg=REG_U):
        check(lib().nmx_state_copy_out(self._ctx, int(reg), C.c_void_p(int(dptr))))

    # checkpoints (device pool; see nmx_checkpoint_*)
    def checkpoint_alloc(self, nslots):
        check(lib().nmx_checkpoint_alloc(self._ctx, int(nslots)))

    def checkpoint_store(self, slot, reg=REG_U):
        check(lib().nmx_checkpoint_store(self._ctx, int(slot), int(reg)))

    def checkpoint_load(self, slot, reg=REG_U):
        check(lib().nmx_checkpoint_load(self._ctx, int(slot), int(reg)))

    def last_adjoint_ms(self):
        d = C.c_double()
        check(lib().nmx_last_adjoint_ms(self._ctx, C.byref(d)))
        return d.value

    def synchronize(self):
        check(lib().nmx_synchronize(self._ctx))

    def last_evolve_ms(self):
        d = C.c_double()
        check(lib().nmx_last_evolve_ms(self._ctx, C.byref(d)))
        return d.value


def amr_setup_mesh(mesh):
    """Build patches, elements, tables, geometry and dt (mesh.c:181-319)."""
    check(lib().nmx_mesh_build(mesh.ctx))
    return 0


def amr_import_mesh(mesh, patches, elements):
    """Take a mesh nmesh already holds instead of building one
    (nmx_mesh_import).  patches: dicts with type (0 box, 1 CubedShell), dom,
    s0, s1, bbox (6), xc (3); elements in eid order: dicts with p, loc (the
    octant digits, a string like nmesh's eloc->loc), n, bbox (6), fnb (6 lists
    of neighbour indices) and nbface (6 lists of their faces)."""
    npat, ne = len(patches), len(elements)
    ptype = np.array([q["type"] for q in patches], np.int32)
    pgeo = np.zeros((npat, 12))
    for i, q in enumerate(patches):
        pgeo[i, 0] = q.get("dom", 0)
        pgeo[i, 1] = q.get("s0", 0.0)
        pgeo[i, 2] = q.get("s1", 0.0)
        pgeo[i, 3:9] = q["bbox"]
        pgeo[i, 9:12] = q.get("xc", (0.0, 0.0, 0.0))
    epat = np.array([q["p"] for q in elements], np.int32)
    elev = np.array([len(q["loc"]) for q in elements], np.int32)
    eloc = np.zeros((ne, 24), np.uint8)
    en = np.zeros((ne, 3), np.int32)
    ebb = np.zeros((ne, 6))
    nfnb = np.zeros((ne, 6), np.int32)
    fnb = np.zeros((ne, 6, 8), np.int32)
    nbf = np.zeros((ne, 6, 8), np.int32)
    for e, q in enumerate(elements):
        for k, ch in enumerate(q["loc"]):
            eloc[e, k] = ord(ch)
        en[e] = q["n"]
        ebb[e] = q["bbox"]
        for f in range(6):
            nfnb[e, f] = len(q["fnb"][f])
            fnb[e, f, :nfnb[e, f]] = q["fnb"][f]
            nbf[e, f, :nfnb[e, f]] = q["nbface"][f]
    check(lib().nmx_mesh_import(mesh.ctx, npat, _p(ptype), _p(pgeo), ne, _p(epat), _p(elev), _p(eloc),
                                _p(en), _p(ebb), _p(nfnb), _p(fnb), _p(nbf)))
    return 0


def scalarwave1_set_profile(mesh, t=None, reg=REG_U):
    """u = plane wave of scalarwave1_k at time t (scalarwave1.c:574-642)."""
    check(lib().nmx_state_set_profile(mesh.ctx, int(reg), float(mesh.time if t is None else t)))
    return 0


def evolve_setrhs_mesh(mesh, t=None, reg_out=REG_R, reg_in=REG_U):
    """r = RHS(u, t) on the device; returns r as a flat array (evolve.c:182-232)."""
    check(lib().nmx_setrhs(mesh.ctx, int(reg_out), int(reg_in),
                           float(mesh.time if t is None else t)))
    return mesh.download(reg_out)


def _step(mesh, method, nsteps):
    old = mesh.par("evolve_method")
    if old != method:
        mesh.set_par("evolve_method", method)
    try:
        check(lib().nmx_evolve(mesh.ctx, int(nsteps)))
    finally:
        if old != method:
            mesh.set_par("evolve_method", old)


def evolve_RK4_mesh(mesh, nsteps=1):
    """RK4 step(s) (RungeKutta.c:19-82)."""
    _step(mesh, "RK4", nsteps)


def evolve_sspRK3_mesh(mesh, nsteps=1):
    """SSP-RK3 step(s) (RungeKutta.c:130-188)."""
    _step(mesh, "sspRK3", nsteps)


def evolve_Euler_mesh(mesh, nsteps=1):
    """Euler step(s) (RungeKutta.c:85-115)."""
    _step(mesh, "Euler", nsteps)


def autotune(mesh):
    """Time the stage-kernel variants on this mesh and keep the fastest."""
    check(lib().nmx_autotune(mesh.ctx))


def evolve_filter_evosys_mesh(mesh, reg=REG_U):
    """The FILTER bin's exponential filter of the evolved variables, in place
    (evolve_filter_evosys_mesh, evolve.c:656-710); nmx_evolve applies it after
    every step when evolve_filter_all / evolve_filter_varlist ask for it."""
    check(lib().nmx_filter(mesh.ctx, int(reg)))
    return 0


def evolve_mesh(mesh):
    """evolve_mesh (main.c:583-666): steps until finaltime, or iterations *
    dt when finaltime <= 0 (both parameters of the mesh)"""
    check(lib().nmx_evolve_mesh(mesh.ctx))
    return 0


def evolve_myln(mesh, nsteps=1):
    """Step(s) with the method named by evolve_method (evolve.c:25-176)."""
    check(lib().nmx_evolve(mesh.ctx, int(nsteps)))
    return 0
