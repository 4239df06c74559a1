"""GPU tests of the outer evolution loop (nmx_evolve_mesh; evolve_mesh,
main.c:583-666) with the parameters finaltime and iterations:

- the number of steps is the one the reference's loop takes: while time <
  finaltime - max(dt 1e-7, finaltime 1e-14), time = Te + te with te
  accumulating dt (the loop restated below in Python);
- iterations = N with finaltime = 0 evolves to N dt, and a second call does
  nothing;
- the state equals nmx_evolve's over the same number of steps bitwise, and
  the oracle's steps to 1e-12 relative to the field maximum.
"""
import numpy as np
import pytest

import nmesh_amd as nm
from oracle.orc import OMesh

pytestmark = pytest.mark.gpu
K = [0.8, -0.5, 0.33166247903554]


def box(extra=None):
    pars = {"amr_mesh_type": "BoxMesh", "amr_luni": 1, "amr_n0": 5, "amr_n1": 5, "amr_n2": 5,
            "scalarwave1_k": "%.17g %.17g %.17g" % tuple(K), "nmx_graph_steps": 4}
    pars.update(extra or {})
    m = nm.Mesh(pars, device=0)
    nm.amr_setup_mesh(m)
    nm.scalarwave1_set_profile(m, 0.0)
    m.time = 0.0
    return m


def loop_steps(t0, dt, finaltime, iterations):
    """the step count of evolve_mesh's loop (main.c:585-624)"""
    ttol = max(dt * 1e-7, finaltime * 1e-14)
    if finaltime <= 0 and iterations > 0:
        finaltime = iterations * dt
    if finaltime <= 0:
        return 0
    te, t, k = 0.0, t0, 0
    while t < finaltime - ttol:
        te += dt
        t = t0 + te
        k += 1
    return k


@pytest.mark.parametrize("frac", [0.0, 0.3, 1e-9])
def test_finaltime_steps_and_state(frac):
    print("evolve_mesh finaltime", frac, flush=True)
    m = box()
    dt = m.dt
    ft = (11 + frac) * dt
    m.set_par("finaltime", "%.17g" % ft)
    nm.evolve_mesh(m)
    k = loop_steps(0.0, dt, ft, 0)
    assert k == (11 if frac < 1e-7 else 12)
    assert m.iteration == k
    assert m.time >= ft - max(dt * 1e-7, ft * 1e-14)
    ref = box()
    nm.evolve_myln(ref, k)
    assert np.array_equal(m.download(), ref.download())
    assert m.time == ref.time
    om = OMesh("box", 1, 5)
    u0, _ = om.profile(K, 0.0)
    v, _ = om.evolve(u0, dt, k, 0.0, K, "RK4")
    assert np.abs(m.download() - v).max() <= 1e-12 * np.abs(v).max()


def test_iterations_then_nothing_more():
    print("evolve_mesh iterations", flush=True)
    m = box({"iterations": 7, "evolve_method": "sspRK3"})
    nm.evolve_mesh(m)
    assert m.iteration == 7 == loop_steps(0.0, m.dt, 0.0, 7)
    u = m.download()
    nm.evolve_mesh(m)              # time already at 7 dt
    assert m.iteration == 7 and np.array_equal(m.download(), u)
    m.set_par("iterations", "0")
    nm.evolve_mesh(m)              # finaltime = iterations = 0: nothing
    assert m.iteration == 7
