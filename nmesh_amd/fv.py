"""FV WENO5 + LLF evolution of special-relativistic Euler on a periodic grid
of 16^3-cell blocks (libnmx nmx_fv_*).

The reconstruction and flux replace those of nmesh's fv_divf
(src/main/dg/fv.c:115-624): rec1d_p_WENO5 / rec1d_m_WENO5 with uniform
weights (src/main/dg/rec1d.c:1190-1352) and numflux1d_LLF
(src/main/dg/numflux1d.c:85-112); the steps are those of evolve_sspRK3_mesh /
evolve_RK4_mesh / evolve_Euler_mesh (src/main/evolve/RungeKutta.c:19-188).
Parameters keep nmesh's names where it has them (fv_rec, fv_flux, dt,
evolve_method).  Host arrays are flat: u[var*ncell + cell], cell = i + N (j +
N k); conserved (D, Sx, Sy, Sz, tau), primitive (rho, vx, vy, vz, p).
"""
import ctypes as C

import numpy as np

from ._lib import check, lib

CONS, PRIM, RHS = 0, 1, 2


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class FvGrid:
    """One periodic N^3 grid on one device."""

    def __init__(self, pars=None, device=0):
        self._h = C.c_void_p()
        check(lib().nmx_fv_create(C.byref(self._h), int(device)))
        for k, v in (pars or {}).items():
            self.set_par(k, v)

    def close(self):
        if self._h:
            lib().nmx_fv_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_par(self, name, value):
        check(lib().nmx_fv_set_param(self._h, name.encode(), str(value).encode()))

    def par(self, name):
        buf = C.create_string_buffer(256)
        check(lib().nmx_fv_get_param(self._h, name.encode(), buf, 256))
        return buf.value.decode()

    def setup(self):
        check(lib().nmx_fv_setup(self._h))

    @property
    def ncell(self):
        n = C.c_long()
        check(lib().nmx_fv_ncells(self._h, C.byref(n)))
        return n.value

    @property
    def dt(self):
        d = C.c_double()
        check(lib().nmx_fv_get_dt(self._h, C.byref(d)))
        return d.value

    @dt.setter
    def dt(self, v):
        check(lib().nmx_fv_set_dt(self._h, float(v)))

    @property
    def time(self):
        d = C.c_double()
        check(lib().nmx_fv_get_time(self._h, C.byref(d)))
        return d.value

    def upload_prim(self, P):
        P = np.ascontiguousarray(P, dtype=np.float64).reshape(-1)
        if P.size != 5 * self.ncell:
            raise ValueError("primitives must hold 5*ncell values")
        check(lib().nmx_fv_upload_prim(self._h, _p(P)))

    def upload(self, u, which=CONS):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        if u.size != 5 * self.ncell:
            raise ValueError("a state holds 5*ncell values")
        check(lib().nmx_fv_upload(self._h, int(which), _p(u)))

    def download(self, which=CONS):
        u = np.zeros(5 * self.ncell)
        check(lib().nmx_fv_download(self._h, int(which), _p(u)))
        return u

    def setrhs(self):
        """RHS of the conserved state (c2p first); returns it flat."""
        check(lib().nmx_fv_setrhs(self._h))
        return self.download(RHS)

    def max_speed(self):
        """largest characteristic speed over cells and directions (the CFL
        reduction, c2p first)"""
        a = C.c_double()
        check(lib().nmx_fv_max_speed(self._h, C.byref(a)))
        return a.value

    def evolve(self, nsteps):
        check(lib().nmx_fv_evolve(self._h, int(nsteps)))

    def last_evolve_ms(self):
        d = C.c_double()
        check(lib().nmx_fv_last_evolve_ms(self._h, C.byref(d)))
        return d.value

    def c2p_failures(self):
        """Blocks whose Newton c2p did not converge since the last call."""
        n = C.c_long()
        check(lib().nmx_fv_status(self._h, C.byref(n)))
        return n.value


def blob_initial_data(N, seed=2026, nblobs=32):
    """Primitives of the FV benchmark workload: rho = 1 + sum of nblobs
    Gaussian blobs (amplitude U[0.1, 1], width U[2, 8] cells, centres
    U[0, N)^3, periodic minimum-image distance), v_i = U[-0.3, 0.3] per cell,
    p = 1; numpy PCG64 with the given seed.  Returns P flat [var][cell]."""
    rng = np.random.default_rng(seed)
    A = rng.uniform(0.1, 1.0, nblobs)
    w = rng.uniform(2.0, 8.0, nblobs)
    xc = rng.uniform(0.0, N, (nblobs, 3))
    x = np.arange(N) + 0.5
    rho = np.ones((N, N, N))          # indexed [k, j, i]
    for b in range(nblobs):
        g = []
        for d in range(3):
            dd = x - xc[b, d]
            dd = dd - N * np.round(dd / N)
            g.append(np.exp(-dd ** 2 / (2 * w[b] ** 2)))
        rho += A[b] * g[2][:, None, None] * g[1][None, :, None] * g[0][None, None, :]
    v = rng.uniform(-0.3, 0.3, (3, N, N, N))
    P = np.empty((5, N, N, N))
    P[0] = rho
    P[1:4] = v
    P[4] = 1.0
    return P.reshape(5, -1)
