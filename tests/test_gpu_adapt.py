"""GPU parity of mesh adaptation with the state interpolated on the device
(libnmx adapt_kernel behind nmx_mesh_hadapt / nmx_mesh_padapt /
nmx_mesh_switch_fv) against the
oracle (orc_mesh_hadapt, orc_adapt_state).  The device applies the three
1D Lagrange matrices one direction at a time, the oracle the same products
in its own loop order, so the bar is 1e-12 relative to the field maximum.

- h-refinement, coarsening (with the shared-point average) and p-changes on
  the shell, one after the other, each checked against the oracle;
- refine then coarsen gives back the state;
- RK4 steps on the adapted mesh (batched path, interpolated faces) match the
  oracle's steps on its adapted mesh;
- FV meshes and wrong sizes are refused, the mesh left as it was;
- the DG <-> FV element switch (nmx_mesh_switch_fv): the state after the
  switch, RK4 steps on the mixed mesh, the switch back, and a round trip.
"""
import numpy as np
import pytest

import nmesh_amd as nm
from oracle.orc import OMesh

pytestmark = pytest.mark.gpu
K = [0.8, -0.5, 0.33166247903554]
TOL = 1e-12


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def shell(n=5):
    pars = {"amr_mesh_type": "Shell", "amr_luni": 1, "amr_n0": n, "amr_n1": n, "amr_n2": n,
            "amr_Shell_rin": 2.0, "amr_Shell_rout": 6.0,
            "scalarwave1_k": "%.17g %.17g %.17g" % tuple(K)}
    m = nm.Mesh(pars, device=0)
    nm.amr_setup_mesh(m)
    return m, OMesh("shell", 1, n, rin=2.0, rout=6.0)


def same_tables(m, om):
    assert m.nelem == om.nelem
    for e in range(om.nelem):
        assert np.array_equal(m.elm_info(e), om.elm_info(e)), e


def test_refine_p_change_coarsen_sequence():
    print("adapt sequence", flush=True)
    m, om = shell()
    u, _ = om.profile(K, 0.3)
    u = u + 0.1 * np.random.default_rng(1).uniform(-1, 1, u.shape)
    m.upload(u)
    # h-refine three elements
    flag = np.zeros(om.nelem, np.int32)
    flag[[3, 17, 40]] = 1
    m.hadapt(flag)
    om, u = om.adapted(u, hflag=flag)
    same_tables(m, om)
    assert rel(m.download(), u) <= TOL
    # p-change every element
    nv = np.random.default_rng(2).integers(3, 10, om.nelem).astype(np.int32)
    m.padapt(nv)
    om, u = om.adapted(u, n=nv)
    same_tables(m, om)
    assert rel(m.download(), u) <= TOL
    # coarsen the refined ones (their siblings have different n now)
    cflag = np.array([-1 if om.elm_info(e)[1] == 2 else 0 for e in range(om.nelem)], np.int32)
    m.hadapt(cflag)
    om, u = om.adapted(u, hflag=cflag)
    same_tables(m, om)
    assert rel(m.download(), u) <= TOL


def test_refine_then_coarsen_restores_the_state():
    print("adapt round trip", flush=True)
    m, om = shell(6)
    u = np.random.default_rng(3).standard_normal(5 * om.nnodes)
    m.upload(u)
    flag = np.zeros(om.nelem, np.int32)
    flag[::5] = 1
    m.hadapt(flag)
    ref, _ = om.adapted(u, hflag=flag)
    cflag = np.array([-1 if ref.elm_info(e)[1] == 2 else 0 for e in range(ref.nelem)], np.int32)
    m.hadapt(cflag)
    same_tables(m, om)
    assert rel(m.download(), u) <= TOL


def test_steps_after_adaptation_match_oracle():
    print("adapt then steps", flush=True)
    m, om = shell()
    u, _ = om.profile(K, 0.0)
    m.upload(u)
    m.time = 0.0
    flag = np.zeros(om.nelem, np.int32)
    flag[[1, 30]] = 1
    m.hadapt(flag)
    om, u = om.adapted(u, hflag=flag)
    assert abs(m.dt - om.dt()) == 0.0
    t = 0.0
    for s in range(3):
        u, t = om.evolve(u, m.dt, 1, t, K, "RK4")
        nm.evolve_myln(m, 1)
        assert rel(m.download(), u) <= TOL, "step %d" % s
    assert abs(m.time - t) < 1e-15


def test_refusals_leave_the_mesh():
    print("adapt refusals", flush=True)
    m, om = shell()
    u = np.random.default_rng(4).standard_normal(5 * om.nnodes)
    m.upload(u)
    with pytest.raises(nm.NmxError):
        m.hadapt(np.zeros(om.nelem - 1, np.int32))
    with pytest.raises(nm.NmxError):
        m.padapt(np.full(om.nelem, 11, np.int32))
    assert m.nelem == om.nelem and np.array_equal(m.download(), u)
    fv = nm.Mesh({"amr_mesh_type": "BoxMesh", "amr_luni": 1, "amr_n0": 5, "amr_n1": 5, "amr_n2": 5,
                  "amr_uniform_p": "0", "amr_fv_p": "0"}, device=0)
    nm.amr_setup_mesh(fv)
    with pytest.raises(nm.NmxError):
        fv.hadapt(np.ones(8, np.int32))


@pytest.mark.parametrize("rec", ["WENOm5_2", "WENOm3_2"])
def test_switch_to_fv_and_back_then_steps(rec):
    """DG -> FV switch of some elements with the state interpolated on the
    device, against the oracle; RK4 steps on the mixed mesh; the switch
    back; a round trip restores the state"""
    print("switch fv", rec, flush=True)
    from oracle import orc
    m, om = shell(6)
    m.set_par("fv_rec", rec)
    orc.set_fv_rec(rec)
    orc.set_fv_midpoint(False)
    u, _ = om.profile(K, 0.0)
    m.upload(u)
    m.time = 0.0
    to_fv = np.zeros(om.nelem, np.int32)
    to_fv[np.random.default_rng(5).permutation(om.nelem)[:20]] = 1
    m.switch_fv(to_fv)
    om2, u2 = om.adapted(u, fv=to_fv)
    same_tables(m, om2)
    assert rel(m.download(), u2) <= TOL
    t = 0.0
    for s in range(2):
        u2, t = om2.evolve(u2, m.dt, 1, t, K, "RK4")
        nm.evolve_myln(m, 1)
        assert rel(m.download(), u2) <= TOL, "step %d" % s
    m.switch_fv(np.zeros(om.nelem, np.int32))
    om3, u3 = om2.adapted(u2, fv=np.zeros(om.nelem, np.int32))
    same_tables(m, om3)
    assert rel(m.download(), u3) <= TOL
    # round trip of an arbitrary state
    v = np.random.default_rng(6).standard_normal(5 * om.nnodes)
    m.upload(v)
    m.switch_fv(to_fv)
    m.switch_fv(np.zeros(om.nelem, np.int32))
    assert rel(m.download(), v) <= TOL
