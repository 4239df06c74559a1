"""Convergence of the device path under p- and h-refinement.

scalarwave1's own known answer (scalarwave1_analyze, scalarwave1.c:710-751):
the plane wave u = profile(k, t) with the plane wave as exterior state.  The
state after RK4 steps to t = T on the device is compared with the exact
profile at T:
  - p-refinement (box, 2 elements per direction, n = 3..8 points per
    direction, dt = T/256 so the time error stays far below): the error
    falls spectrally;
  - h-refinement (n = 4, i.e. degree 3, with 2, 4, 8 elements per direction):
    the observed order is the DG order N + 1 = 4 (dt shrinks with h, and RK4
    is fourth order, so the time error follows the same rate).
Each error is also compared with the oracle's error for the same mesh, dt
and step count: the device reproduces the reference's convergence, not only
its own.
"""
import math

import numpy as np
import pytest

import nmesh_amd as nm
from nmesh_amd.evolve import REG_P
from oracle.orc import OMesh

pytestmark = pytest.mark.gpu

K = [0.8, -0.5, 0.33166247903554]    # |k| = 1
T = 1.0
DOUT = 2.0


def run(luni, n, nsteps=None):
    pars = {"amr_mesh_type": "BoxMesh", "amr_luni": luni, "amr_n0": n, "amr_n1": n, "amr_n2": n,
            "amr_BoxMesh_dout": DOUT, "scalarwave1_k": "%.17g %.17g %.17g" % tuple(K),
            "scalarwave1_numflux": "upwind", "evolve_method": "RK4"}
    m = nm.Mesh(pars, device=0)
    nm.amr_setup_mesh(m)
    if nsteps is None:
        nsteps = int(math.ceil(T / m.dt))
    assert T / nsteps <= m.dt
    dt = T / nsteps
    m.dt = dt
    m.time = 0.0
    nm.scalarwave1_set_profile(m, t=0.0)
    u0 = m.download()
    nm.evolve_myln(m, nsteps)
    u = m.download()
    nm.scalarwave1_set_profile(m, t=T, reg=REG_P)
    exact = m.download(REG_P)
    err = float(np.abs(u - exact).max())
    om = OMesh("box", luni, n, dout=DOUT)
    uo, _ = om.evolve(u0, dt, nsteps, 0.0, K, "RK4")
    err_o = float(np.abs(uo - exact).max())
    return err, err_o, float(np.abs(exact).max())


def test_p_refinement_spectral():
    print("p-refinement", flush=True)
    errs = []
    for n in (3, 4, 5, 6, 7, 8):
        err, err_o, scale = run(1, n, nsteps=256)     # time error far below
        print("n=%d err %.3e oracle %.3e" % (n, err, err_o), flush=True)
        assert abs(err - err_o) <= 1e-10 * scale + 1e-6 * err_o
        errs.append(err)
    errs = np.array(errs)
    # every two more points per direction gain at least a factor 10
    assert np.all(errs[2:] < 0.1 * errs[:-2])
    assert errs[-1] < 1e-5 * errs[0]


def test_h_refinement_order():
    print("h-refinement", flush=True)
    errs = []
    for luni in (1, 2, 3):
        err, err_o, scale = run(luni, 4)
        print("luni=%d err %.3e oracle %.3e" % (luni, err, err_o), flush=True)
        assert abs(err - err_o) <= 1e-10 * scale + 1e-6 * err_o
        errs.append(err)
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    print("observed orders", orders, flush=True)
    assert np.all(orders > 3.4) and np.all(orders < 5.0)
