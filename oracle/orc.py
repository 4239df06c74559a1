"""ctypes binding of the oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker / CPU baseline.  The product
(nmesh_amd, libnmx.so) never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_L = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp, i, l, d = C.c_void_p, C.c_int, C.c_long, C.c_double
        u64, sz, cp = C.c_uint64, C.c_size_t, C.c_char_p
        sig = {
            "orc_lgl": (None, [i, vp, vp]),
            "orc_dt_matrix": (None, [i, vp]),
            "orc_mesh_box": (vp, [i, i, d, vp]),
            "orc_mesh_shell": (vp, [i, i, d, d]),
            "orc_mesh_amr_shell": (vp, [i, i, d, d, d, i]),
            "orc_mesh_set_n": (None, [vp, vp]),
            "orc_mesh_free": (None, [vp]),
            "orc_mesh_nelem": (i, [vp]),
            "orc_mesh_nnodes": (l, [vp]),
            "orc_mesh_npat": (i, [vp]),
            "orc_mesh_elm_info": (None, [vp, i, vp]),
            "orc_mesh_elm_bbox": (None, [vp, i, vp]),
            "orc_mesh_elm_fnb": (i, [vp, i, i, vp, vp]),
            "orc_mesh_pat_bound": (i, [vp, i, i]),
            "orc_mesh_coords": (None, [vp, vp, vp]),
            "orc_mesh_geom": (None, [vp, vp, vp, vp]),
            "orc_mesh_dt": (d, [vp, d, d]),
            "orc_mesh_elm_hmin": (d, [vp, i]),
            "orc_set_profile": (None, [vp, vp, d, vp, vp]),
            "orc_rhs": (None, [vp, i, vp, vp, d, vp]),
            "orc_numflux": (None, [i, vp, vp, vp, vp]),
            "orc_evolve": (d, [vp, i, vp, i, d, i, d, vp]),
            "orc_filter": (None, [vp, vp, i, d, d, i, i]),
            "orc_num_threads": (i, []),
            "orc_mesh_set_fv": (None, [vp, i]),
            "orc_mesh_set_fv_patches": (None, [vp, i]),
            "orc_face_values": (None, [vp, vp, i, i, i, vp]),
            "orc_crc64": (u64, [u64, vp, sz]),
            "orc_crc64_combine": (u64, [u64, u64, u64]),
            "orc_vars_crc": (u64, [vp, vp]),
            "orc_write_variables_bin": (i, [vp, vp, cp]),
            "orc_mesh_clone": (vp, [vp]),
            "orc_mesh_hadapt": (None, [vp, vp]),
            "orc_adapt_state": (i, [vp, vp, vp, vp]),
            "orc_mesh_set_fv_elems": (None, [vp, vp]),
            "orc_set_limiter_MRS": (None, [d]),
            "orc_limit_MRS": (None, [vp, vp]),
            "orc_set_limiter_minmodB": (None, [d, d, i]),
            "orc_limit_minmodB": (None, [vp, vp]),
            "orc_set_fv_rec": (None, [i]),
            "orc_set_fv_midpoint": (None, [i]),
            "orc_set_fv_surface_interp": (None, [i]),
            "orc_set_fv_divf_extrap": (None, [d, d, i]),
            "orc_set_fv_rec_end": (None, [i, d, d, i]),
            "orc_fv_rec_ends": (None, [i, vp, vp]),
            "orc_fv_rec_pm": (None, [i, vp, i, vp]),
            "orc_pw_parab": (d, [i, d, i, vp]),
            "orc_dnfn_extrap1": (None, [i, vp]),
            "orc_rec1d_p_WENOm3": (d, [vp, i]),
            "orc_rec1d_m_WENOm3": (d, [vp, i]),
            "orc_rec1d_p_WENO5": (d, [vp, i]),
            "orc_rec1d_m_WENO5": (d, [vp, i]),
            "orc_rec1d_p_WENOZ": (d, [vp, i]),
            "orc_rec1d_m_WENOZ": (d, [vp, i]),
            "orc_sr_p2c": (None, [d, vp, vp]),
            "orc_sr_c2p": (i, [d, vp, vp, d]),
            "orc_sr_flux": (None, [d, i, vp, vp, vp, vp]),
            "orc_sr_llf": (None, [d, i, vp, vp, vp]),
            "orc_fv_c2p": (i, [i, d, vp, vp, d]),
            "orc_fv_p2c": (None, [i, d, vp, vp]),
            "orc_fv_rhs_prim": (None, [i, d, vp, vp]),
            "orc_fv_rhs": (i, [i, d, vp, vp, vp, d]),
            "orc_fv_evolve": (i, [i, d, d, i, vp, vp, d]),
            "orc_trouble_score": (i, [i, vp, vp]),
            "orc_trouble_rdmp": (None, [vp, vp, vp, d, d, vp]),
            "orc_mesh_fv_flags": (None, [vp, vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _L = L
    return _L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


NUMFLUX = {"upwind": 0, "LLF": 1}
METHOD = {"RK4": 0, "sspRK3": 1, "Euler": 2}


class OMesh:
    """Oracle mesh built from nmesh parameter values."""

    def __init__(self, kind, luni, n, dout=1.0, rin=2.0, rout=20.0, radius=10.0, levels=0, fv=False,
                 fv_patches=None):
        L = lib()
        if kind == "box":
            self.h = L.orc_mesh_box(int(luni), int(n), float(dout), None)
        elif kind == "shell" and levels == 0:
            self.h = L.orc_mesh_shell(int(luni), int(n), float(rin), float(rout))
        elif kind == "shell":
            self.h = L.orc_mesh_amr_shell(int(luni), int(n), float(rin), float(rout),
                                          float(radius), int(levels))
        else:
            raise ValueError(kind)
        if fv:
            L.orc_mesh_set_fv(self.h, 1)
        elif fv_patches:
            L.orc_mesh_set_fv_patches(self.h, sum(1 << p for p in fv_patches))

    def __del__(self):
        try:
            lib().orc_mesh_free(self.h)
        except Exception:
            pass

    @property
    def nelem(self):
        return lib().orc_mesh_nelem(self.h)

    @property
    def nnodes(self):
        return lib().orc_mesh_nnodes(self.h)

    def set_n(self, n):
        n = np.ascontiguousarray(n, dtype=np.int32)
        lib().orc_mesh_set_n(self.h, _p(n))

    def elm_info(self, e):
        a = np.zeros(22, dtype=np.int64)
        lib().orc_mesh_elm_info(self.h, int(e), _p(a))
        return a

    def elm_bbox(self, e):
        a = np.zeros(6)
        lib().orc_mesh_elm_bbox(self.h, int(e), _p(a))
        return a

    def elm_fnb(self, e, f):
        nb = np.zeros(8, dtype=np.int32)
        nbf = np.zeros(8, dtype=np.int32)
        k = lib().orc_mesh_elm_fnb(self.h, int(e), int(f), _p(nb), _p(nbf))
        return nb[:k].copy(), nbf[:k].copy()

    def pat_bound(self, p, f):
        return lib().orc_mesh_pat_bound(self.h, int(p), int(f))

    def coords(self):
        X = np.zeros(3 * self.nnodes)
        x = np.zeros(3 * self.nnodes)
        lib().orc_mesh_coords(self.h, _p(X), _p(x))
        return X.reshape(3, -1), x.reshape(3, -1)

    def geometry(self):
        nn = self.nnodes
        g = np.zeros(9 * nn)
        sg = np.zeros(3 * nn)
        det = np.zeros(nn)
        lib().orc_mesh_geom(self.h, _p(g), _p(sg), _p(det))
        return g.reshape(9, -1), sg.reshape(3, -1), det

    def dt(self, dtfac=0.25, uniform_dtfac=0.125):
        return lib().orc_mesh_dt(self.h, float(dtfac), float(uniform_dtfac))

    def profile(self, k, t):
        k = np.ascontiguousarray(k, dtype=np.float64)
        u = np.zeros(5 * self.nnodes)
        du = np.zeros(5 * self.nnodes)
        lib().orc_set_profile(self.h, _p(k), float(t), _p(u), _p(du))
        return u, du

    def rhs(self, u, t, k, numflux="upwind"):
        k = np.ascontiguousarray(k, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64)
        r = np.zeros_like(u)
        lib().orc_rhs(self.h, NUMFLUX[numflux], _p(k), _p(u), float(t), _p(r))
        return r

    def evolve(self, u, dt, nsteps, t0, k, method="RK4", numflux="upwind"):
        k = np.ascontiguousarray(k, dtype=np.float64)
        u = np.array(u, dtype=np.float64, copy=True)
        t = lib().orc_evolve(self.h, NUMFLUX[numflux], _p(k), METHOD[method], float(dt),
                             int(nsteps), float(t0), _p(u))
        return u, t

    def adapted(self, u, hflag=None, n=None, fv=None):
        """(new mesh, new state): h-adaptation by hflag (> 0 refine, < 0
        coarsen 8 flagged siblings), then per-element n (p-change), the state
        interpolated as nmesh's refine and update_node_n_pt_typ do"""
        new = OMesh.__new__(OMesh)
        new.h = lib().orc_mesh_clone(self.h)
        if hflag is not None:
            f = np.ascontiguousarray(hflag, dtype=np.int32)
            lib().orc_mesh_hadapt(new.h, _p(f))
        if n is not None:
            new.set_n(n)
        if fv is not None:
            f = np.ascontiguousarray(fv, dtype=np.int32)
            lib().orc_mesh_set_fv_elems(new.h, _p(f))
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros(5 * new.nnodes)
        r = lib().orc_adapt_state(self.h, _p(u), new.h, _p(out))
        if r != 0:
            raise ValueError("orc_adapt_state failed (%d)" % r)
        return new, out

    def fv_flags(self):
        f = np.zeros(self.nelem, dtype=np.int32)
        lib().orc_mesh_fv_flags(self.h, _p(f))
        return f

    def trouble_rdmp(self, u, up, delta0=1e-4, eps=1e-3):
        """TROUBLE values of the library's RDMP indicator for state u
        against u_p (1 DG out of bounds, -1 FV in bounds, 0 otherwise)"""
        u = np.ascontiguousarray(u, dtype=np.float64)
        up = np.ascontiguousarray(up, dtype=np.float64)
        tr = np.zeros(self.nelem, dtype=np.int32)
        lib().orc_trouble_rdmp(self.h, _p(u), _p(up), float(delta0), float(eps), _p(tr))
        return tr

    def limit_MRS(self, u, alpha):
        """one application of the MRS limiter to state u (limdata_MRS then
        limiter_MRS on every element, limiter.c:56-196)"""
        u = np.array(u, dtype=np.float64, copy=True)
        set_limiter_MRS(alpha)
        try:
            lib().orc_limit_MRS(self.h, _p(u))
        finally:
            set_limiter_MRS(None)
        return u

    def limit_minmodB(self, u, alpha, beta=0.5, scale_bound=False):
        """one application of the minmodB limiter (limiter.c:200-380)"""
        u = np.array(u, dtype=np.float64, copy=True)
        set_limiter_minmodB(alpha, beta, scale_bound)
        try:
            lib().orc_limit_minmodB(self.h, _p(u))
        finally:
            set_limiter_minmodB(None)
        return u

    def vars_crc(self, u):
        """the reference's "vars" CRC-64 of state u (norms.c:440-458)"""
        u = np.ascontiguousarray(u, dtype=np.float64)
        return int(lib().orc_vars_crc(self.h, _p(u)))

    def write_variables_bin(self, u, path):
        """variables.bin of state u (checkpoint_save.c:412-468)"""
        u = np.ascontiguousarray(u, dtype=np.float64)
        if lib().orc_write_variables_bin(self.h, _p(u), str(path).encode()) != 0:
            raise OSError("orc_write_variables_bin failed: %s" % path)

    def filter(self, u, varmask=31, alp=36.0, s=32.0, dn=0, jpow=1):
        """exponential filter of the variables in varmask (bit v of pi, cx,
        cy, cz, phi) in every element (filter.c:20-92)"""
        u = np.array(u, dtype=np.float64, copy=True)
        lib().orc_filter(self.h, _p(u), int(varmask), float(alp), float(s), int(dn), int(jpow))
        return u


def set_limiter_MRS(alpha):
    """scalarwave1_limiter = MRS with limiter_alpha = alpha in the oracle's
    RK steps (None: no limiter)"""
    lib().orc_set_limiter_MRS(-1.0 if alpha is None else float(alpha))


def set_limiter_minmodB(alpha, beta=0.5, scale_bound=False):
    """scalarwave1_limiter = minmodB with limiter_alpha, limiter_beta and
    limiter_scaleBound in the oracle's RK steps (alpha None: off)"""
    lib().orc_set_limiter_minmodB(-1.0 if alpha is None else float(alpha), float(beta),
                                  1 if scale_bound else 0)


def set_fv_midpoint(on):
    """coordinates_midpoint_data of the oracle's FV elements: False (the
    default, "no") or True ("yes")"""
    lib().orc_set_fv_midpoint(1 if on else 0)


def pw_parab(k, x, xp):
    """piecewise parabolic basis function k of the points xp at x"""
    xp = np.ascontiguousarray(xp, dtype=np.float64)
    return lib().orc_pw_parab(int(k), float(x), len(xp), _p(xp))


def set_fv_divf_extrap(s1=0.0, s2=None, opt=0):
    """fv_divf_extrap_s1, _s2 (None: DBL_MAX) and _opt of the oracle's
    dnfn_extrap1"""
    lib().orc_set_fv_divf_extrap(float(s1), 1.7976931348623157e308 if s2 is None else float(s2), int(opt))


def set_fv_surface_interp(mode):
    """fv_surface_interp of the oracle's faces taken from FV neighbours:
    "linear" (the default) or "parabolic" """
    words = mode.split()
    lib().orc_set_fv_surface_interp({"linear": 0, "parabolic": 1}[words[0]] +
                                    (2 if "Lagrange_if_dg" in words[1:] else 0))


def set_fv_rec(mode, s1=0.0, s2=None, opt=0):
    """reconstruction of the oracle's FV elements (fv_rec): WENOm3, WENOm5,
    WENOmZ and their _1 and _2 modes, with fv_rec_WENOm_s1 / _s2 (None:
    DBL_MAX) / _opt for the modes without suffix"""
    base = mode.split("_")[0]
    lib().orc_set_fv_rec({"WENOm5": 5, "WENOm3": 3, "WENOmZ": 7, "WENOmT": 9, "1": 1}[base])
    end = {"_2": 0, "_1": 1, "": 2}[mode[len(base):]] if base != "1" else 0
    lib().orc_set_fv_rec_end(end, float(s1), 1.7976931348623157e308 if s2 is None else float(s2), int(opt))


# ---------------------------------------------------------------------------
# FV WENO5 + LLF, SR-Euler (fv_oracle.c)
# ---------------------------------------------------------------------------

def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def rec_p(u, im):
    u = np.ascontiguousarray(u, dtype=np.float64)
    return lib().orc_rec1d_p_WENO5(_ptr(u), int(im))


def rec_m(u, im):
    u = np.ascontiguousarray(u, dtype=np.float64)
    return lib().orc_rec1d_m_WENO5(_ptr(u), int(im))


def rec_pz(u, im):
    u = np.ascontiguousarray(u, dtype=np.float64)
    return lib().orc_rec1d_p_WENOZ(_ptr(u), int(im))


def rec_mz(u, im):
    u = np.ascontiguousarray(u, dtype=np.float64)
    return lib().orc_rec1d_m_WENOZ(_ptr(u), int(im))


def sr_p2c(gam, P):
    P = np.ascontiguousarray(P, dtype=np.float64)
    U = np.zeros(5)
    lib().orc_sr_p2c(float(gam), _ptr(P), _ptr(U))
    return U


def sr_c2p(gam, U, guess_p, tol=1e-12):
    U = np.ascontiguousarray(U, dtype=np.float64)
    P = np.zeros(5)
    P[4] = guess_p
    it = lib().orc_sr_c2p(float(gam), _ptr(U), _ptr(P), float(tol))
    return P, it


def sr_flux(gam, d, P):
    P = np.ascontiguousarray(P, dtype=np.float64)
    F, U, lam = np.zeros(5), np.zeros(5), np.zeros(2)
    lib().orc_sr_flux(float(gam), int(d), _ptr(P), _ptr(F), _ptr(U), _ptr(lam))
    return F, U, lam


def sr_llf(gam, d, PL, PR):
    PL = np.ascontiguousarray(PL, dtype=np.float64)
    PR = np.ascontiguousarray(PR, dtype=np.float64)
    F = np.zeros(5)
    lib().orc_sr_llf(float(gam), int(d), _ptr(PL), _ptr(PR), _ptr(F))
    return F


class OFV:
    """Periodic N^3 grid of the FV oracle; states are flat [var][cell],
    cell = i + N (j + N k)."""

    def __init__(self, N, gam=2.0, tol=1e-12):
        self.N, self.gam, self.tol = int(N), float(gam), float(tol)
        self.ncell = self.N ** 3

    def p2c(self, P):
        P = np.ascontiguousarray(P, dtype=np.float64).ravel()
        u = np.zeros_like(P)
        lib().orc_fv_p2c(self.N, self.gam, _ptr(P), _ptr(u))
        return u

    def c2p(self, u, P_guess):
        u = np.ascontiguousarray(u, dtype=np.float64).ravel()
        P = np.ascontiguousarray(P_guess, dtype=np.float64).ravel().copy()
        bad = lib().orc_fv_c2p(self.N, self.gam, _ptr(u), _ptr(P), self.tol)
        return P, bad

    def rhs_prim(self, P):
        P = np.ascontiguousarray(P, dtype=np.float64).ravel()
        r = np.zeros_like(P)
        lib().orc_fv_rhs_prim(self.N, self.gam, _ptr(P), _ptr(r))
        return r

    def rhs(self, u, P_guess):
        u = np.ascontiguousarray(u, dtype=np.float64).ravel()
        P = np.ascontiguousarray(P_guess, dtype=np.float64).ravel().copy()
        r = np.zeros_like(u)
        bad = lib().orc_fv_rhs(self.N, self.gam, _ptr(u), _ptr(P), _ptr(r), self.tol)
        return r, P, bad

    def evolve(self, u, P_guess, dt, nsteps):
        u = np.ascontiguousarray(u, dtype=np.float64).ravel().copy()
        P = np.ascontiguousarray(P_guess, dtype=np.float64).ravel().copy()
        bad = lib().orc_fv_evolve(self.N, self.gam, float(dt), int(nsteps), _ptr(u), _ptr(P), self.tol)
        return u, P, bad


def trouble_score(score, tr=None):
    """evolve_set_trouble_score on every element (tr None: read only);
    updates score in place, returns the mesh's score"""
    assert score.dtype == np.int32 and score.flags.c_contiguous
    if tr is not None:
        tr = np.ascontiguousarray(tr, dtype=np.int32)
    return int(lib().orc_trouble_score(len(score), None if tr is None else _p(tr), _p(score)))
