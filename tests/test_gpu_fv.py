"""GPU parity of the FV WENO5 + LLF path (libnmx nmx_fv_*) against the FV
oracle (oracle/fv_oracle.c) on the same seeded inputs.

The kernels and the oracle evaluate the same formulas in the same order
(reconstruction, flux, per-direction accumulation x, y, z); they differ only
in FMA contraction, so the RHS agrees to ~1e-14 relative and states after
each RK step to ~1e-15.  The bar is 1e-12 relative to the field maximum.
"""
import numpy as np
import pytest

from nmesh_amd.fv import blob_initial_data
from oracle import orc

pytestmark = pytest.mark.gpu
TOL = 1e-12


def make(N, method="sspRK3", graph="yes", gsteps=2, xcd="yes"):
    from nmesh_amd.fv import FvGrid
    g = FvGrid({"fv_n": N, "evolve_method": method, "nmx_graph": graph,
                "nmx_graph_steps": gsteps, "nmx_xcd": xcd}, device=0)
    g.setup()
    return g


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("N,xcd", [(32, "yes"), (48, "yes"), (48, "no")])
def test_rhs_and_primitives_match_oracle(N, xcd):
    print("fv rhs", N, xcd, flush=True)
    P = blob_initial_data(N, seed=11, nblobs=6)
    g = make(N, xcd=xcd)
    g.upload_prim(P)
    ofv = orc.OFV(N)
    u = ofv.p2c(P)
    assert rel(g.download(), u) <= 1e-15
    r = g.setrhs()
    r_o, P_o, bad = ofv.rhs(u, P)
    assert bad == 0 and g.c2p_failures() == 0
    assert rel(g.download(1), P_o) <= TOL
    assert rel(r, r_o) <= TOL


@pytest.mark.parametrize("method,nsteps", [("sspRK3", 5), ("RK4", 3), ("Euler", 3)])
def test_steps_match_oracle_every_step(method, nsteps):
    print("fv steps", method, nsteps, flush=True)
    N = 32
    P = blob_initial_data(N, seed=12, nblobs=6)
    g = make(N, method=method, gsteps=2)
    g.upload_prim(P)
    ofv = orc.OFV(N)
    u = ofv.p2c(P)
    Pg = P.ravel().copy()
    dt = g.dt
    assert dt == 0.25
    for s in range(nsteps):
        u, Pg, bad = ofv_step(ofv, u, Pg, dt, method)
        g.evolve(1)
        assert rel(g.download(), u) <= TOL, "step %d" % s
    # graph-captured steps give the same states as single steps
    g2 = make(N, method=method, gsteps=2)
    g2.upload_prim(P)
    g2.evolve(nsteps)
    assert np.array_equal(g2.download(), g.download())
    assert g.c2p_failures() == 0


def ofv_step(ofv, u, P, dt, method):
    """one oracle step of the given method (sspRK3 is orc_fv_evolve itself;
    RK4 and Euler are composed from the oracle RHS like RungeKutta.c)"""
    if method == "sspRK3":
        return ofv.evolve(u, P, dt, 1)
    r, P, b0 = ofv.rhs(u, P)
    if method == "Euler":
        return u + dt * r, P, b0
    A = u + dt / 6 * r
    w = u + dt / 2 * r
    r, P, b1 = ofv.rhs(w, P)
    A = A + dt / 3 * r
    w = u + dt / 2 * r
    r, P, b2 = ofv.rhs(w, P)
    A = A + dt / 3 * r
    w = u + dt * r
    r, P, b3 = ofv.rhs(w, P)
    return A + dt / 6 * r, P, b0 + b1 + b2 + b3


def test_constant_state_is_preserved_exactly():
    N = 32
    g = make(N)
    P = np.empty((5, N ** 3))
    P[:] = np.array([1.3, 0.2, -0.1, 0.05, 0.7])[:, None]
    g.upload_prim(P)
    u0 = g.download()
    assert np.abs(g.setrhs()).max() == 0.0
    g.evolve(4)
    assert np.array_equal(g.download(), u0)


def test_periodic_shift_commutes_with_the_step():
    """Shifting the data by a whole block shifts the result bitwise: periodic
    wrap, block addressing and tile mapping are consistent."""
    N = 48
    P = blob_initial_data(N, seed=13, nblobs=4).reshape(5, N, N, N)
    Ps = np.roll(P, shift=(16, -16, 16), axis=(1, 2, 3))
    g1, g2 = make(N), make(N)
    g1.upload_prim(P.reshape(5, -1))
    g2.upload_prim(Ps.reshape(5, -1))
    g1.evolve(2)
    g2.evolve(2)
    u1 = g1.download().reshape(5, N, N, N)
    u2 = g2.download().reshape(5, N, N, N)
    assert np.array_equal(np.roll(u1, shift=(16, -16, 16), axis=(1, 2, 3)), u2)


def test_bad_parameters_are_rejected():
    from nmesh_amd import NmxError
    from nmesh_amd.fv import FvGrid
    g = FvGrid(device=0)
    with pytest.raises(NmxError):
        g.set_par("fv_rec", "WENOZ")
    g.set_par("fv_n", 40)
    with pytest.raises(NmxError):
        g.setup()
