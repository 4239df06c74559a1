"""GPU parity of the batched (order-bucketed) DG path on meshes with mixed
element orders and mixed refinement levels against the oracle, whose face
values on such faces follow nmesh's 2D Lagrange interpolation with
averaging (surface.c:1081-1337).  Tolerance 1e-12 relative, per RHS and per
RK step."""
import numpy as np
import pytest

import nmesh_amd as nm
from oracle.orc import OMesh

K = [0.8, -0.5, 0.33166247903554]
TOL = 1e-12

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def make(kind, luni, n, nvec=None, path="batched", radius=10, levels=0, k=K, numflux="upwind", shape=None):
    pars = {"amr_mesh_type": "BoxMesh" if kind == "box" else "Shell", "amr_luni": luni,
            "amr_n0": n, "amr_n1": n, "amr_n2": n, "amr_BoxMesh_dout": 8.0,
            "amr_Shell_rin": 2.0, "amr_Shell_rout": 20.0,
            "amr_hrefine_sphere_radius": radius, "amr_hrefine_sphere_levels": levels,
            "scalarwave1_k": "%.17g %.17g %.17g" % tuple(k),
            "scalarwave1_numflux": numflux, "nmx_path": path}
    if shape is not None:
        pars["nmx_batch"] = shape
    m = nm.Mesh(pars, device=0)
    nm.amr_setup_mesh(m)
    om = OMesh(kind, luni, n, dout=8.0, rin=2.0, rout=20.0, radius=radius, levels=levels)
    if nvec is not None:
        m.set_elm_n(nvec)
        om.set_n(nvec)
    return m, om


def state(om, seed, amp=0.1):
    u, _ = om.profile(K, 0.1)
    return u + amp * np.random.default_rng(seed).uniform(-1, 1, u.shape)


SHAPES = ["persistent", "buckets", "legacy"]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("kind,luni,n", [("box", 2, 5), ("shell", 1, 6)])
def test_batched_kernel_on_uniform_mesh(kind, luni, n, shape):
    print("batched uniform", kind, shape, flush=True)
    m, om = make(kind, luni, n, shape=shape)
    u = state(om, 1)
    m.upload(u)
    assert rel(nm.evolve_setrhs_mesh(m, t=0.1), om.rhs(u, 0.1, K)) <= TOL


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("numflux", ["upwind", "LLF"])
def test_mixed_order_shell_rhs(numflux, shape):
    print("mixed order", numflux, shape, flush=True)
    rng = np.random.default_rng(12)
    nvec = rng.integers(4, 11, 48).astype(np.int32)     # N in 3..9
    m, om = make("shell", 1, 5, nvec=nvec, numflux=numflux, shape=shape)
    u = state(om, 2)
    m.upload(u)
    assert rel(nm.evolve_setrhs_mesh(m, t=0.1), om.rhs(u, 0.1, K, numflux)) <= TOL


@pytest.mark.parametrize("shape", SHAPES)
def test_refined_mixed_order_shell_rk4_steps(shape):
    """h-refined shell (nested spheres: coarse faces see 4 fine neighbours)
    with random orders; every RK4 step checked."""
    print("amr mixed steps", shape, flush=True)
    base = OMesh("shell", 1, 4, rin=2.0, rout=20.0, radius=7, levels=1)
    rng = np.random.default_rng(13)
    nvec = rng.integers(4, 11, base.nelem).astype(np.int32)
    m, om = make("shell", 1, 4, nvec=nvec, radius=7, levels=1, shape=shape)
    info = np.array([m.elm_info(e) for e in range(m.nelem)])
    assert (info[:, 1] == 2).any() and (info[:, 1] == 1).any()
    u = state(om, 3, amp=0.02)
    m.upload(u)
    m.time = 0.1
    dt = m.dt
    assert dt == om.dt()
    t = 0.1
    for s in range(4):
        u, t = om.evolve(u, dt, 1, t, K, "RK4")
        nm.evolve_myln(m)
        assert rel(m.download(), u) <= TOL, "step %d" % s


@pytest.mark.parametrize("shape", SHAPES)
def test_perelement_and_batched_launches_agree_bitwise(shape):
    print("perelement vs batched", shape, flush=True)
    rng = np.random.default_rng(14)
    nvec = rng.integers(3, 9, 48).astype(np.int32)
    m1, om = make("shell", 1, 5, nvec=nvec, path="batched", shape=shape)
    m2, _ = make("shell", 1, 5, nvec=nvec, path="perelement")
    u = state(om, 4)
    m1.upload(u)
    m2.upload(u)
    nm.evolve_myln(m1, 2)
    nm.evolve_myln(m2, 2)
    assert np.array_equal(m1.download(), m2.download())


@pytest.mark.parametrize("method", ["sspRK3", "Euler"])
def test_refined_mixed_order_shell_other_methods(method):
    """The interpolated faces of each stage's input come from the previous
    stage's senders: sspRK3 (w, w, then the accumulator) and Euler (the
    accumulator only), every step checked."""
    print("amr mixed steps", method, flush=True)
    base = OMesh("shell", 1, 4, rin=2.0, rout=20.0, radius=7, levels=1)
    rng = np.random.default_rng(15)
    nvec = rng.integers(4, 11, base.nelem).astype(np.int32)
    m, om = make("shell", 1, 4, nvec=nvec, radius=7, levels=1)
    m.set_par("evolve_method", method)
    u = state(om, 5, amp=0.02)
    m.upload(u)
    m.time = 0.1
    dt = m.dt
    t = 0.1
    for s in range(3):
        u, t = om.evolve(u, dt, 1, t, K, method)
        nm.evolve_myln(m)
        assert rel(m.download(), u) <= TOL, "step %d" % s


@pytest.mark.parametrize("shape", SHAPES)
def test_graph_steps_match_single_steps_and_rhs_after_evolve(shape):
    """Captured multi-step graphs against one step per call (bitwise), and an
    RHS of the evolved state (its faces recomputed from HBM) against the
    oracle."""
    print("amr graph vs single", shape, flush=True)
    base = OMesh("shell", 1, 4, rin=2.0, rout=20.0, radius=7, levels=1)
    rng = np.random.default_rng(16)
    nvec = rng.integers(4, 11, base.nelem).astype(np.int32)
    m1, om = make("shell", 1, 4, nvec=nvec, radius=7, levels=1, shape=shape)
    m2, _ = make("shell", 1, 4, nvec=nvec, radius=7, levels=1, shape=shape)
    m1.set_par("nmx_graph_steps", 4)
    u = state(om, 6, amp=0.02)
    m1.upload(u)
    m2.upload(u)
    nm.evolve_myln(m1, 9)
    for _ in range(9):
        nm.evolve_myln(m2, 1)
    u1 = m1.download()
    assert np.array_equal(u1, m2.download())
    t = m1.time
    assert rel(nm.evolve_setrhs_mesh(m1, t=t), om.rhs(u1, t, K)) <= TOL


@pytest.mark.parametrize("shape", SHAPES)
def test_every_order_pair_on_a_refined_shell(shape):
    """orders 2..10 on an h-refined shell: the face interpolation of every
    (receiver, neighbour) order pair that the mesh holds (the face pass is
    compiled per neighbour order, its transpose per receiver order), RHS
    against the oracle"""
    print("order pairs", shape, flush=True)
    base = OMesh("shell", 1, 4, rin=2.0, rout=20.0, radius=7, levels=1)
    nvec = np.random.default_rng(0).integers(2, 11, base.nelem).astype(np.int32)
    m, om = make("shell", 1, 4, nvec=nvec, radius=7, levels=1, shape=shape)
    info = np.array([m.elm_info(e) for e in range(m.nelem)])
    pairs = set()
    for e in range(m.nelem):
        for f in range(6):
            nb, _ = om.elm_fnb(e, f)
            for j in nb:
                pairs.add((int(info[e, 2]), int(info[int(j), 2])))
    assert len(pairs) == 81, sorted(pairs)   # every (receiver, neighbour) pair of 2..10
    u = state(om, 21)
    m.upload(u)
    assert rel(nm.evolve_setrhs_mesh(m, t=0.1), om.rhs(u, 0.1, K)) <= TOL
