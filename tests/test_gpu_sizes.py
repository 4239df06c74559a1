"""Sizes at the edges and at full scale:

- every element order the kernels take, n = 2 .. 10, on a box and a shell
  (plain path), RHS against the oracle;
- a single-element box (every face on the outer boundary), RHS and RK4
  steps against the oracle;
- n = 11 refused with a message, the mesh unchanged;
- the C2 shell (3072 elements, n = 8, 1,572,864 nodes) and the C4
  mixed-order AMR shell (19,200 elements, 8,213,159 nodes) at full size: one
  RHS against the oracle's at 1e-12, the C4 RHS linear (zero exterior) and
  its launch shapes agreeing, and C2 plain against batched.
"""
import numpy as np
import pytest

import nmesh_amd as nm
from oracle.orc import OMesh

pytestmark = pytest.mark.gpu
K = [0.8, -0.5, 0.33166247903554]
K0 = [0.0, 0.0, 0.0]
TOL = 1e-12


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def mesh(kind, luni, n, k=K, path="plain"):
    pars = {"amr_mesh_type": "BoxMesh" if kind == "box" else "Shell", "amr_luni": luni,
            "amr_n0": n, "amr_n1": n, "amr_n2": n, "amr_BoxMesh_dout": 8.0,
            "amr_Shell_rin": 2.0, "amr_Shell_rout": 20.0,
            "scalarwave1_k": "%.17g %.17g %.17g" % tuple(k), "nmx_path": path}
    m = nm.Mesh(pars, device=0)
    nm.amr_setup_mesh(m)
    return m, OMesh(kind, luni, n, dout=8.0, rin=2.0, rout=20.0)


@pytest.mark.parametrize("n", [2, 3, 4, 7, 9, 10])
@pytest.mark.parametrize("kind,luni", [("box", 1), ("shell", 0)])
def test_every_order(kind, luni, n):
    print("order", kind, n, flush=True)
    m, om = mesh(kind, luni, n)
    u, _ = om.profile(K, 0.2)
    u = u + 0.1 * np.random.default_rng(n).uniform(-1, 1, u.shape)
    m.upload(u)
    assert rel(nm.evolve_setrhs_mesh(m, t=0.2), om.rhs(u, 0.2, K)) <= TOL


def test_single_element_box():
    print("single element", flush=True)
    m, om = mesh("box", 0, 6)
    assert m.nelem == 1
    u, _ = om.profile(K, 0.0)
    u = u + 0.05 * np.random.default_rng(1).uniform(-1, 1, u.shape)
    m.upload(u)
    assert rel(nm.evolve_setrhs_mesh(m, t=0.0), om.rhs(u, 0.0, K)) <= TOL
    m.time = 0.0
    t = 0.0
    for s in range(3):
        u, t = om.evolve(u, m.dt, 1, t, K, "RK4")
        nm.evolve_myln(m)
        assert rel(m.download(), u) <= TOL, s


def test_order_eleven_is_refused():
    print("order 11", flush=True)
    m, _ = mesh("box", 1, 5)
    before = [m.elm_info(e).copy() for e in range(m.nelem)]
    with pytest.raises(nm.NmxError, match="10 points"):
        m.padapt(np.full(m.nelem, 11, np.int32))
    assert all(np.array_equal(m.elm_info(e), b) for e, b in enumerate(before))


def test_c2_full_size_against_oracle_and_batched():
    import bench
    print("C2 full size", flush=True)
    p = bench.c2_params()
    p["scalarwave1_k"] = "%.17g %.17g %.17g" % tuple(K)
    out = {}
    for path in ("plain", "batched"):
        q = dict(p)
        q["nmx_path"] = path
        m = nm.Mesh(q, device=0)
        nm.amr_setup_mesh(m)
        if path == "plain":
            om = OMesh("shell", 3, 8, rin=2.0, rout=20.0)
            u, _ = om.profile(K, 0.1)
            u = u + 0.05 * np.random.default_rng(2).uniform(-1, 1, u.shape)
        m.upload(u)
        out[path] = nm.evolve_setrhs_mesh(m, t=0.1)
        del m
    assert rel(out["plain"], om.rhs(u, 0.1, K)) <= TOL
    assert rel(out["batched"], out["plain"]) <= 1e-13


def test_c4_full_size_against_oracle_linearity_and_shapes():
    import bench
    print("C4 full size", flush=True)
    m = bench.c4_mesh(nm, "batched")
    om = OMesh("shell", 3, 5, rin=2.0, rout=20.0, radius=3, levels=2)
    om.set_n(np.random.default_rng(2026).integers(4, 11, om.nelem).astype(np.int32))
    assert m.nelem == om.nelem == 19200 and m.nnodes == om.nnodes
    X, x = om.coords()
    u = bench.shell_initial_data(x)
    v = np.random.default_rng(3).standard_normal(u.shape)
    m.upload(u)
    ru = nm.evolve_setrhs_mesh(m, t=0.0)
    assert rel(ru, om.rhs(u, 0.0, K0)) <= TOL
    m.upload(v)
    rv = nm.evolve_setrhs_mesh(m, t=0.0)
    m.upload(u + 2.0 * v)
    ruv = nm.evolve_setrhs_mesh(m, t=0.0)
    assert rel(ruv, ru + 2.0 * rv) <= 1e-13
    for shape in ("buckets", "persistent"):
        m.set_par("nmx_batch", shape)
        m.upload(u)
        assert rel(nm.evolve_setrhs_mesh(m, t=0.0), ru) <= 1e-13, shape
