"""GPU parity of the limiters (scalarwave1_limiter = MRS: libnmx
lim_data_kernel / lim_apply_kernel; minmodB: mm_data_kernel /
mm_apply_kernel) inside the RK steps against the oracle's steps with the
same limiter (orc_limit_MRS / orc_limit_minmodB in orc_evolve): every later
stage's input and the final u of each step are limited (RungeKutta.c:19-188,
evolve.c:143-144).  The limiter's averages, theta and update are evaluated
in the reference's operation order on both sides; the RHS differs at
rounding, so the bar is 1e-12 relative to the field maximum per step.
"""
import numpy as np
import pytest

import nmesh_amd as nm
from oracle import orc
from oracle.orc import OMesh

pytestmark = pytest.mark.gpu
K = [0.8, -0.5, 0.33166247903554]
TOL = 1e-12


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def make(path, alpha, method, levels=0, extra=None, limiter="MRS"):
    pars = {"amr_mesh_type": "Shell", "amr_luni": 1, "amr_n0": 5, "amr_n1": 5, "amr_n2": 5,
            "amr_Shell_rin": 2.0, "amr_Shell_rout": 6.0, "amr_hrefine_sphere_radius": 4,
            "amr_hrefine_sphere_levels": levels, "scalarwave1_k": "%.17g %.17g %.17g" % tuple(K),
            "scalarwave1_limiter": limiter, "limiter_alpha": alpha, "evolve_method": method,
            "nmx_path": path, "nmx_graph_steps": 2}
    pars.update(extra or {})
    m = nm.Mesh(pars, device=0)
    nm.amr_setup_mesh(m)
    om = OMesh("shell", 1, 5, rin=2.0, rout=6.0, radius=4, levels=levels)
    return m, om


def noisy(om, seed):
    u, _ = om.profile(K, 0.0)
    rng = np.random.default_rng(seed)
    # a step in one region: the limiter has something to do
    _, x = om.coords()
    jump = (x.reshape(3, -1)[0] > 1.0).astype(float)
    return u + 0.5 * np.tile(jump, 5) + 0.05 * rng.uniform(-1, 1, u.shape)


@pytest.mark.parametrize("method", ["RK4", "sspRK3", "Euler"])
@pytest.mark.parametrize("alpha", [0.3, 5.0])
def test_limited_steps_match_oracle(method, alpha):
    print("limiter", method, alpha, flush=True)
    m, om = make("plain", alpha, method)
    u = noisy(om, 1)
    m.upload(u)
    m.time = 0.0
    orc.set_limiter_MRS(alpha)
    try:
        t = 0.0
        v = u.copy()
        for s in range(3):
            v, t = om.evolve(v, m.dt, 1, t, K, method)
            nm.evolve_myln(m, 1)
            assert rel(m.download(), v) <= TOL, "step %d" % s
    finally:
        orc.set_limiter_MRS(None)
    # with the small alpha the limiter acts (the unlimited steps differ);
    # alpha = 5 widens the bounds enough to leave this state alone
    w, _ = om.evolve(u, m.dt, 3, 0.0, K, method)
    if alpha < 1.0:
        assert rel(w, v) > 1e-6


def test_limited_steps_batched_refined_with_graph():
    print("limiter batched", flush=True)
    m, om = make("batched", 0.3, "RK4", levels=1)
    u = noisy(om, 2)
    m.upload(u)
    m.time = 0.0
    orc.set_limiter_MRS(0.3)
    try:
        v, _ = om.evolve(u, m.dt, 4, 0.0, K, "RK4")
    finally:
        orc.set_limiter_MRS(None)
    nm.evolve_myln(m, 4)          # two captured 2-step graphs
    assert rel(m.download(), v) <= TOL


def test_limiter_then_filter():
    """evolve_myln limits the final u, then the FILTER bin filters it"""
    print("limiter and filter", flush=True)
    m, om = make("plain", 0.3, "RK4", extra={"evolve_filter_all": "yes", "evolve_filter_alp": "4",
                                              "evolve_filter_s": "6"})
    u = noisy(om, 3)
    m.upload(u)
    m.time = 0.0
    orc.set_limiter_MRS(0.3)
    try:
        v, _ = om.evolve(u, m.dt, 1, 0.0, K, "RK4")
    finally:
        orc.set_limiter_MRS(None)
    v = om.filter(v, 31, 4.0, 6.0, 0, 1)
    nm.evolve_myln(m, 1)
    assert rel(m.download(), v) <= TOL


def test_limiter_refusals():
    print("limiter refusals", flush=True)
    m, om = make("plain", 0.3, "RK4")
    with pytest.raises(nm.NmxError):
        m.set_par("scalarwave1_limiter", "TVB")
    import torch
    from nmesh_amd import adjoint
    lam = torch.zeros(5 * om.nnodes, dtype=torch.float64, device="cuda")
    with pytest.raises(nm.NmxError):
        adjoint.adjoint_evolve(m, lam, 1)
    fv = nm.Mesh({"amr_mesh_type": "BoxMesh", "amr_luni": 1, "amr_n0": 5, "amr_n1": 5, "amr_n2": 5,
                  "amr_uniform_p": "0", "amr_fv_p": "0", "scalarwave1_limiter": "MRS"}, device=0)
    nm.amr_setup_mesh(fv)
    with pytest.raises(nm.NmxError):
        nm.evolve_myln(fv, 1)


@pytest.mark.parametrize("method", ["RK4", "sspRK3"])
@pytest.mark.parametrize("alpha,beta,scale", [(0.05, 0.5, "no"), (0.5, 1.0, "yes")])
def test_minmodB_steps_match_oracle(method, alpha, beta, scale):
    print("minmodB", method, alpha, beta, scale, flush=True)
    m, om = make("plain", alpha, method, limiter="minmodB",
                 extra={"limiter_beta": beta, "limiter_scaleBound": scale})
    u = noisy(om, 4)
    m.upload(u)
    m.time = 0.0
    orc.set_limiter_minmodB(alpha, beta, scale == "yes")
    try:
        t = 0.0
        v = u.copy()
        for s in range(3):
            v, t = om.evolve(v, m.dt, 1, t, K, method)
            nm.evolve_myln(m, 1)
            assert rel(m.download(), v) <= TOL, "step %d" % s
    finally:
        orc.set_limiter_minmodB(None)
    w, _ = om.evolve(u, m.dt, 3, 0.0, K, method)
    assert rel(w, v) > 1e-6


def test_minmodB_batched_refined():
    print("minmodB batched", flush=True)
    m, om = make("batched", 0.05, "RK4", levels=1, limiter="minmodB")
    u = noisy(om, 5)
    m.upload(u)
    m.time = 0.0
    orc.set_limiter_minmodB(0.05, 0.5, False)
    try:
        v, _ = om.evolve(u, m.dt, 4, 0.0, K, "RK4")
    finally:
        orc.set_limiter_minmodB(None)
    nm.evolve_myln(m, 4)
    assert rel(m.download(), v) <= TOL
