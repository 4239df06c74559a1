"""GPU parity of the exponential filter (FILTER bin) against the oracle's
orc_filter (filter.c:20-92 restated): alone through nmx_filter, and after
every full step inside nmx_evolve (main.c:628-630), on Cartesian and
cubed-sphere meshes, the plain and the batched mixed-order paths.  The
device applies the three directional products S diag(f) A instead of the
reference's six passes, so it agrees with the oracle to rounding: the bar is
1e-12 relative to the field maximum."""
import numpy as np
import pytest

import nmesh_amd as nm
from nmesh_amd import NmxError
from oracle.orc import OMesh

pytestmark = pytest.mark.gpu
K = [0.8, -0.5, 0.33166247903554]
TOL = 1e-12


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def make(kind, luni, n, extra=None, nvec=None, radius=10, levels=0):
    pars = {"amr_mesh_type": "BoxMesh" if kind == "box" else "Shell", "amr_luni": luni,
            "amr_n0": n, "amr_n1": n, "amr_n2": n, "amr_BoxMesh_dout": 2.0,
            "amr_Shell_rin": 2.0, "amr_Shell_rout": 20.0,
            "amr_hrefine_sphere_radius": radius, "amr_hrefine_sphere_levels": levels,
            "scalarwave1_k": "%.17g %.17g %.17g" % tuple(K)}
    pars.update(extra or {})
    m = nm.Mesh(pars, device=0)
    nm.amr_setup_mesh(m)
    om = OMesh(kind, luni, n, dout=2.0, rin=2.0, rout=20.0, radius=radius, levels=levels)
    if nvec is not None:
        m.set_elm_n(nvec)
        om.set_n(nvec)
    return m, om


def fpars(alp, s, dn, jpow, all_="yes", varlist=""):
    return {"evolve_filter_all": all_, "evolve_filter_varlist": varlist,
            "evolve_filter_alp": alp, "evolve_filter_s": s, "evolve_filter_dn": dn,
            "basis_expfilter_JacobianPower": jpow}


@pytest.mark.parametrize("kind,luni,n,geometry,alp,s,dn,jpow", [
    ("box", 1, 6, "analytic", 36.0, 32.0, 0, 1),
    ("box", 1, 5, "analytic", 2.0, 2.0, 1, 0),
    ("shell", 1, 7, "analytic", 36.0, 32.0, 0, 1),
    ("shell", 1, 5, "stored", 3.0, 4.0, 0, 1),
    ("shell", 1, 6, "analytic", 3.0, 4.0, 0, -1),
])
def test_filter_matches_oracle(kind, luni, n, geometry, alp, s, dn, jpow):
    print("filter", kind, n, geometry, alp, s, dn, jpow, flush=True)
    m, om = make(kind, luni, n, dict(fpars(alp, s, dn, jpow), nmx_geometry=geometry))
    rng = np.random.default_rng(n)
    u = rng.standard_normal(5 * om.nnodes)
    m.upload(u)
    nm.evolve_filter_evosys_mesh(m)
    assert rel(m.download(), om.filter(u, 31, alp, s, dn, jpow)) <= TOL


def test_filter_varlist_and_both_passes():
    """evolve_filter_varlist names first components (cx filters cx, cy, cz);
    with evolve_filter_all too, those variables are filtered twice"""
    print("filter varlist", flush=True)
    m, om = make("shell", 1, 5, fpars(3.0, 4.0, 0, 1, all_="no",
                                      varlist="scalarwave1_cx scalarwave1_none"))
    rng = np.random.default_rng(7)
    u = rng.standard_normal(5 * om.nnodes)
    m.upload(u)
    nm.evolve_filter_evosys_mesh(m)
    assert rel(m.download(), om.filter(u, 0b01110, 3.0, 4.0, 0, 1)) <= TOL
    m.set_par("evolve_filter_all", "yes")
    m.set_par("evolve_filter_varlist", "scalarwave1_pi scalarwave1_phi")
    m.upload(u)
    nm.evolve_filter_evosys_mesh(m)
    ref = om.filter(om.filter(u, 31, 3.0, 4.0, 0, 1), 0b10001, 3.0, 4.0, 0, 1)
    assert rel(m.download(), ref) <= TOL


def test_filter_on_mixed_order_refined_shell_batched():
    print("filter batched", flush=True)
    base = OMesh("shell", 1, 4, rin=2.0, rout=20.0, radius=7, levels=1)
    rng = np.random.default_rng(8)
    nvec = rng.integers(3, 11, base.nelem).astype(np.int32)
    m, om = make("shell", 1, 4, fpars(6.0, 8.0, 0, 1), nvec=nvec, radius=7, levels=1)
    u = rng.standard_normal(5 * om.nnodes)
    m.upload(u)
    nm.evolve_filter_evosys_mesh(m)
    assert rel(m.download(), om.filter(u, 31, 6.0, 8.0, 0, 1)) <= TOL


@pytest.mark.parametrize("path", ["plain", "batched"])
def test_steps_with_filter_match_oracle_every_step(path):
    """RK4 steps with the filter after each one (FILTER bin), checked every
    step; then the same steps from a captured multi-step graph, bitwise"""
    print("filtered steps", path, flush=True)
    m, om = make("shell", 1, 6, dict(fpars(4.0, 6.0, 0, 1), nmx_path=path,
                                     nmx_graph_steps=2))
    u, _ = om.profile(K, 0.1)
    u = u + 0.05 * np.random.default_rng(9).uniform(-1, 1, u.shape)
    m.upload(u)
    m.time = 0.1
    dt = m.dt
    t = 0.1
    for s in range(3):
        u, t = om.evolve(u, dt, 1, t, K, "RK4")
        u = om.filter(u, 31, 4.0, 6.0, 0, 1)
        nm.evolve_myln(m, 1)
        assert rel(m.download(), u) <= TOL, "step %d" % s
    m2, _ = make("shell", 1, 6, dict(fpars(4.0, 6.0, 0, 1), nmx_path=path, nmx_graph_steps=2))
    u0, _ = om.profile(K, 0.1)
    m2.upload(u0 + 0.05 * np.random.default_rng(9).uniform(-1, 1, u0.shape))
    m2.time = 0.1
    nm.evolve_myln(m2, 3)
    assert np.array_equal(m2.download(), m.download())


def test_filter_off_and_refusals():
    print("filter refusals", flush=True)
    m, om = make("box", 1, 5)
    u = np.random.default_rng(10).standard_normal(5 * om.nnodes)
    m.upload(u)
    nm.evolve_filter_evosys_mesh(m)            # nothing asked for: unchanged
    assert np.array_equal(m.download(), u)
    with pytest.raises(NmxError):
        m.set_par("basis_expfilter_JacobianPower", "2")
    m.set_par("evolve_filter_all", "yes")
    import torch
    from nmesh_amd import adjoint
    lam = torch.zeros(5 * om.nnodes, dtype=torch.float64, device="cuda")
    with pytest.raises(NmxError):
        adjoint.adjoint_evolve(m, lam, 1)      # the filter's adjoint is not built


def test_fv_elements_are_refused_or_left_alone():
    """basis_filter_fv = yes on FV elements is not built (refused);
    basis_filter_fv = no leaves them unfiltered (filter.c:79-80)"""
    print("filter fv elements", flush=True)
    pars = {"amr_mesh_type": "BoxMesh", "amr_luni": 1, "amr_n0": 5, "amr_n1": 5, "amr_n2": 5,
            "amr_BoxMesh_dout": 2.0, "amr_uniform_p": "0", "amr_fv_p": "0",
            "evolve_filter_all": "yes"}
    m = nm.Mesh(pars, device=0)
    nm.amr_setup_mesh(m)
    u = np.random.default_rng(11).standard_normal(5 * m.nnodes)
    m.upload(u)
    with pytest.raises(NmxError):
        nm.evolve_filter_evosys_mesh(m)
    m.set_par("basis_filter_fv", "no")
    nm.evolve_filter_evosys_mesh(m)
    assert np.array_equal(m.download(), u)
