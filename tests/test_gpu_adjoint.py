"""Discrete adjoint on the GPU (nmx_adjoint_setrhs / nmx_adjoint_evolve and the
torch.autograd wrapper) against the oracle's forward operator.

The scalarwave1 DG RHS with a zero exterior state (scalarwave1_k = 0 0 0) is
linear, so the forward RHS, the RK steps and their adjoints are matrices:
- dot-product tests <J v, w> = <v, J^T w> with J v from the oracle (the CPU
  restatement of the reference forward operator) and J^T w from the device,
  at 1e-12 relative to |J v| |w|;
- on meshes small enough, the whole matrix: J from the oracle applied to unit
  vectors, J^T from the device applied to unit vectors, elementwise;
- the C5 workload (config-1 box, 256 RK4 steps, checkpoints every 16, loss on
  the outer-boundary psi): the autograd gradient against central finite
  differences (h = 1e-6, 64 seeded directions) at 1e-10 relative, and a
  dot-product test of the 256-step Jacobian against the oracle.
"""
import numpy as np
import pytest

import nmesh_amd as nm
from oracle.orc import OMesh

torch = pytest.importorskip("torch")
from nmesh_amd import adjoint as adj  # noqa: E402

pytestmark = pytest.mark.gpu
K0 = [0.0, 0.0, 0.0]


def make(kind, luni, n, numflux="upwind", geometry="analytic", method="RK4", **kw):
    pars = {
        "amr_mesh_type": "BoxMesh" if kind == "box" else "Shell",
        "amr_luni": luni, "amr_n0": n, "amr_n1": n, "amr_n2": n,
        "scalarwave1_k": "0 0 0", "scalarwave1_numflux": numflux,
        "nmx_geometry": geometry, "evolve_method": method,
    }
    if kind == "box":
        pars["amr_BoxMesh_dout"] = kw.get("dout", 8.0)
    else:
        pars["amr_Shell_rin"] = kw.get("rin", 2.0)
        pars["amr_Shell_rout"] = kw.get("rout", 20.0)
    m = nm.Mesh(pars, device=0)
    nm.amr_setup_mesh(m)
    om = OMesh(kind, luni, n, dout=kw.get("dout", 8.0), rin=kw.get("rin", 2.0),
               rout=kw.get("rout", 20.0))
    return m, om


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def host(t):
    return t.detach().cpu().numpy()


RHS_CASES = [("box", 2, 5, "analytic"), ("box", 1, 8, "analytic"), ("shell", 1, 6, "analytic"),
             ("shell", 1, 6, "stored"), ("shell", 2, 4, "analytic")]


@pytest.mark.parametrize("kind,luni,n,geometry", RHS_CASES)
@pytest.mark.parametrize("numflux", ["upwind", "LLF"])
def test_adjoint_rhs_dot_product(kind, luni, n, geometry, numflux):
    print("adjoint rhs", kind, luni, n, geometry, numflux, flush=True)
    m, om = make(kind, luni, n, numflux=numflux, geometry=geometry)
    rng = np.random.default_rng(11)
    for _ in range(3):
        v = rng.standard_normal(5 * om.nnodes)
        w = rng.standard_normal(5 * om.nnodes)
        Jv = om.rhs(v, 0.0, K0, numflux)
        JTw = host(adj.adjoint_setrhs(m, dev(w)))
        lhs, rhs = float(Jv @ w), float(v @ JTw)
        assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(Jv) * np.linalg.norm(w), (lhs, rhs)


@pytest.mark.parametrize("kind,luni,n", [("box", 1, 3), ("shell", 0, 3)])
@pytest.mark.parametrize("numflux", ["upwind", "LLF"])
def test_adjoint_rhs_is_the_transposed_matrix(kind, luni, n, numflux):
    print("adjoint matrix", kind, luni, n, numflux, flush=True)
    m, om = make(kind, luni, n, numflux=numflux)
    N = 5 * om.nnodes
    J = np.zeros((N, N))
    JT = np.zeros((N, N))
    eye = torch.eye(N, dtype=torch.float64, device="cuda")
    for i in range(N):
        e = np.zeros(N)
        e[i] = 1.0
        J[:, i] = om.rhs(e, 0.0, K0, numflux)
        JT[:, i] = host(adj.adjoint_setrhs(m, eye[i]))
    scale = np.abs(J).max()
    assert np.abs(JT - J.T).max() <= 1e-13 * scale


@pytest.mark.parametrize("method", ["RK4", "sspRK3", "Euler"])
@pytest.mark.parametrize("kind,luni,n", [("box", 2, 5), ("shell", 1, 6)])
def test_adjoint_step_dot_product(method, kind, luni, n):
    print("adjoint step", method, kind, luni, n, flush=True)
    m, om = make(kind, luni, n, method=method)
    m.set_par("nmx_graph_steps", 2)
    rng = np.random.default_rng(5)
    for nsteps in (1, 5):      # 5 = two captured 2-step graphs and one eager step
        v = rng.standard_normal(5 * om.nnodes)
        w = rng.standard_normal(5 * om.nnodes)
        Jv, _ = om.evolve(v, m.dt, nsteps, 0.0, K0, method)
        JTw = host(adj.adjoint_evolve(m, dev(w), nsteps))
        lhs, rhs = float(Jv @ w), float(v @ JTw)
        assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(Jv) * np.linalg.norm(w), (nsteps, lhs, rhs)


# ---------------------------------------------------------------------------
# C5: config-1 box, 256 RK4 steps, checkpoints every 16
# ---------------------------------------------------------------------------

def c5_setup():
    m, om = make("box", 2, 5, dout=8.0)
    X, x = om.coords()
    r2 = (x ** 2).sum(0)
    psi = np.exp(-r2)
    u0 = np.zeros((5, om.nnodes))
    u0[4] = psi
    u0[1:4] = -2 * x * psi
    u0 = u0.ravel()
    rng = np.random.default_rng(2027)
    du = 1e-3 * rng.uniform(-1, 1, u0.shape)
    bnd = adj.outer_boundary_nodes(m)
    return m, om, u0, du, bnd


def c5_loss(m, u0_t, target_phi, bnd, nsteps=256, every=16, recompute=False):
    traj = adj.evolve(m, u0_t, nsteps, every, recompute)
    nn = m.nnodes
    phi = traj[-1][4 * nn:][bnd]
    return ((phi - target_phi) ** 2).sum()


def test_c5_gradient_against_finite_differences():
    print("c5 fd", flush=True)
    m, om, u0, du, bnd = c5_setup()
    nn = m.nnodes
    with torch.no_grad():
        target = adj.evolve(m, dev(u0 + du), 256, 256)[-1][4 * nn:][bnd].clone()
    u0_t = dev(u0).requires_grad_(True)
    loss = c5_loss(m, u0_t, target, bnd)
    (g,) = torch.autograd.grad(loss, u0_t)
    g = host(g)
    rng = np.random.default_rng(64)
    h = 1e-6
    errs, dirs = [], []
    with torch.no_grad():
        for _ in range(64):
            d = rng.uniform(-1, 1, u0.shape)
            lp = float(c5_loss(m, dev(u0 + h * d), target, bnd, every=256))
            lm = float(c5_loss(m, dev(u0 - h * d), target, bnd, every=256))
            fd = (lp - lm) / (2 * h)
            dirs.append(float(g @ d))
            errs.append(abs(fd - dirs[-1]))
    rel = max(errs) / max(abs(x) for x in dirs)
    print("c5 fd max relative error", rel, flush=True)
    assert rel <= 1e-10


def test_c5_jacobian_dot_product_with_oracle():
    print("c5 dot", flush=True)
    m, om, u0, du, bnd = c5_setup()
    rng = np.random.default_rng(9)
    v = rng.standard_normal(u0.shape)
    w = rng.standard_normal(u0.shape)
    Jv, _ = om.evolve(v, m.dt, 256, 0.0, K0, "RK4")
    JTw = host(adj.adjoint_evolve(m, dev(w), 256))
    lhs, rhs = float(Jv @ w), float(v @ JTw)
    assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(Jv) * np.linalg.norm(w), (lhs, rhs)


def test_c5_recompute_and_segments_give_the_same_gradient():
    print("c5 recompute", flush=True)
    m, om, u0, du, bnd = c5_setup()
    nn = m.nnodes
    with torch.no_grad():
        target = adj.evolve(m, dev(u0 + du), 256, 256)[-1][4 * nn:][bnd].clone()
        final = adj.evolve(m, dev(u0), 256, 16)[-1].clone()
    grads = []
    for every, recompute in ((16, False), (16, True), (100, False), (256, False)):
        u0_t = dev(u0).requires_grad_(True)
        loss = c5_loss(m, u0_t, target, bnd, every=every, recompute=recompute)
        (g,) = torch.autograd.grad(loss, u0_t)
        grads.append(host(g))
        if recompute:   # the forward's final state is restored after the reverse pass
            u_now = torch.empty_like(final)
            m.copy_out(u_now.data_ptr())
            m.synchronize()
            assert torch.equal(u_now, final)
    assert np.array_equal(grads[0], grads[1])
    for g in grads[2:]:
        assert np.abs(g - grads[0]).max() <= 1e-13 * np.abs(grads[0]).max()


def test_trajectory_loss_gradient():
    """A loss on two checkpointed states: the reverse pass adds the
    gradient of each segment's output at the segment boundary."""
    print("trajectory loss", flush=True)
    m, om, u0, du, bnd = c5_setup()
    nn = m.nnodes
    rng = np.random.default_rng(3)
    a = dev(rng.standard_normal(5 * nn))
    b = dev(rng.standard_normal(5 * nn))

    def loss_of(u0_t):
        traj = adj.evolve(m, u0_t, 48, 16)
        return (traj[0] * a).sum() + (traj[2] * b).sum()

    u0_t = dev(u0).requires_grad_(True)
    (g,) = torch.autograd.grad(loss_of(u0_t), u0_t)
    # the same gradient from two explicit adjoint runs
    ref = host(adj.adjoint_evolve(m, a, 16)) + host(adj.adjoint_evolve(m, b, 48))
    assert np.abs(host(g) - ref).max() <= 1e-13 * np.abs(ref).max()


def test_adjoint_refuses_fv_elements():
    pars = {"amr_mesh_type": "Shell", "amr_luni": 1, "amr_n0": 4, "amr_n1": 4, "amr_n2": 4,
            "amr_Shell_rin": 2.0, "amr_Shell_rout": 20.0, "amr_uniform_p": "0", "amr_fv_p": "0"}
    m = nm.Mesh(pars, device=0)
    nm.amr_setup_mesh(m)
    with pytest.raises(nm.NmxError):
        adj.adjoint_setrhs(m, torch.zeros(5 * m.nnodes, dtype=torch.float64, device="cuda"))
