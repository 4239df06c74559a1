"""Discrete adjoint of the batched path (h-refined, mixed-order meshes):
nmx_adjoint_setrhs / nmx_adjoint_evolve through adj_gflux_kernel,
adj_jobs_kernel and dg_adjoint_amr_kernel, against the oracle's forward
operator (the CPU restatement of the reference RHS with its interpolated and
averaged faces, surface.c:1081-1337).  The RHS with a zero exterior state is
linear, so:

- dot-product tests <J v, w> = <v, J^T w> on the refined shell with orders
  3..7 per element (copy faces, interpolated faces, faces with several
  neighbours), upwind and LLF, recomputed and stored metric, 1e-12
  relative to |J v| |w|;
- the whole matrix on a small refined mixed-order box: J from the oracle on
  unit vectors, J^T from the device on unit vectors, elementwise;
- RK4, sspRK3 and Euler steps (graph-captured and eager) by dot products;
- a uniform mesh on the batched path against the plain adjoint kernel.
"""
import numpy as np
import pytest

import nmesh_amd as nm
from oracle.orc import OMesh

torch = pytest.importorskip("torch")
from nmesh_amd import adjoint as adj  # noqa: E402

pytestmark = pytest.mark.gpu
K0 = [0.0, 0.0, 0.0]


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def host(t):
    return t.detach().cpu().numpy()


def refined_mixed(kind="shell", numflux="upwind", geometry="analytic", method="RK4", seed=4):
    """the refined shell (or box) with per-element orders, both sides"""
    if kind == "shell":
        pars = {"amr_mesh_type": "Shell", "amr_luni": 1, "amr_n0": 5, "amr_n1": 5, "amr_n2": 5,
                "amr_Shell_rin": 2.0, "amr_Shell_rout": 6.0}
        om = OMesh("shell", 1, 5, rin=2.0, rout=6.0)
        hflag_idx, nrange = [3, 17, 40], (3, 8)
    else:
        pars = {"amr_mesh_type": "BoxMesh", "amr_luni": 1, "amr_n0": 3, "amr_n1": 3, "amr_n2": 3,
                "amr_BoxMesh_dout": 1.0}
        om = OMesh("box", 1, 3, dout=1.0)
        hflag_idx, nrange = [0], (3, 5)
    pars.update({"scalarwave1_k": "0 0 0", "scalarwave1_numflux": numflux,
                 "nmx_geometry": geometry, "evolve_method": method, "nmx_path": "batched"})
    m = nm.Mesh(pars, device=0)
    nm.amr_setup_mesh(m)
    flag = np.zeros(om.nelem, np.int32)
    flag[hflag_idx] = 1
    m.hadapt(flag)
    om, _ = om.adapted(np.zeros(5 * om.nnodes), hflag=flag)
    nv = np.random.default_rng(seed).integers(nrange[0], nrange[1], om.nelem).astype(np.int32)
    m.padapt(nv)
    om, _ = om.adapted(np.zeros(5 * om.nnodes), n=nv)
    assert m.nelem == om.nelem and m.nnodes == om.nnodes
    info = np.array([om.elm_info(e) for e in range(om.nelem)])
    assert (info[:, 10:16] == 3).any() and (info[:, 4:10] > 1).any()   # interpolated, several neighbours
    return m, om


@pytest.mark.parametrize("numflux", ["upwind", "LLF"])
@pytest.mark.parametrize("geometry", ["analytic", "stored"])
def test_adjoint_rhs_dot_product_refined_mixed_shell(numflux, geometry):
    print("adjoint amr rhs", numflux, geometry, flush=True)
    m, om = refined_mixed("shell", numflux, geometry)
    rng = np.random.default_rng(12)
    for _ in range(3):
        v = rng.standard_normal(5 * om.nnodes)
        w = rng.standard_normal(5 * om.nnodes)
        Jv = om.rhs(v, 0.0, K0, numflux)
        JTw = host(adj.adjoint_setrhs(m, dev(w)))
        lhs, rhs = float(Jv @ w), float(v @ JTw)
        assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(Jv) * np.linalg.norm(w), (lhs, rhs)


def test_adjoint_rhs_dot_product_every_order_pair():
    """orders 2..10 on an h-refined shell, so the transposed interpolation
    runs for every (receiver, neighbour) order pair (its jobs are compiled
    per receiver order): dot products against the oracle's RHS"""
    print("adjoint amr order pairs", flush=True)
    pars = {"amr_mesh_type": "Shell", "amr_luni": 1, "amr_n0": 4, "amr_n1": 4, "amr_n2": 4,
            "amr_Shell_rin": 2.0, "amr_Shell_rout": 20.0, "amr_hrefine_sphere_radius": 7,
            "amr_hrefine_sphere_levels": 1, "scalarwave1_k": "0 0 0", "nmx_path": "batched"}
    m = nm.Mesh(pars, device=0)
    nm.amr_setup_mesh(m)
    om = OMesh("shell", 1, 4, rin=2.0, rout=20.0, radius=7, levels=1)
    nv = np.random.default_rng(0).integers(2, 11, om.nelem).astype(np.int32)
    m.set_elm_n(nv)
    om.set_n(nv)
    pairs = {(int(nv[e]), int(nv[int(j)])) for e in range(om.nelem) for f in range(6) for j in om.elm_fnb(e, f)[0]}
    assert len(pairs) == 81
    rng = np.random.default_rng(13)
    for _ in range(2):
        v = rng.standard_normal(5 * om.nnodes)
        w = rng.standard_normal(5 * om.nnodes)
        Jv = om.rhs(v, 0.0, K0)
        JTw = host(adj.adjoint_setrhs(m, dev(w)))
        lhs, rhs = float(Jv @ w), float(v @ JTw)
        assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(Jv) * np.linalg.norm(w), (lhs, rhs)


@pytest.mark.parametrize("numflux", ["upwind", "LLF"])
def test_adjoint_rhs_is_the_transposed_matrix_refined_box(numflux):
    print("adjoint amr matrix", numflux, flush=True)
    m, om = refined_mixed("box", numflux)
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
def test_adjoint_step_dot_product_refined_mixed_shell(method):
    print("adjoint amr step", method, flush=True)
    m, om = refined_mixed("shell", method=method)
    m.set_par("nmx_graph_steps", 2)
    rng = np.random.default_rng(6)
    for nsteps in (1, 5):      # 5 = two captured 2-step graphs and one eager step
        v = rng.standard_normal(5 * om.nnodes)
        w = rng.standard_normal(5 * om.nnodes)
        Jv, _ = om.evolve(v, m.dt, nsteps, 0.0, K0, method)
        JTw = host(adj.adjoint_evolve(m, dev(w), nsteps))
        lhs, rhs = float(Jv @ w), float(v @ JTw)
        assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(Jv) * np.linalg.norm(w), (nsteps, lhs, rhs)


def test_batched_adjoint_matches_the_plain_one():
    print("adjoint batched vs plain", flush=True)
    out = {}
    for path in ("plain", "batched"):
        pars = {"amr_mesh_type": "Shell", "amr_luni": 1, "amr_n0": 6, "amr_n1": 6, "amr_n2": 6,
                "amr_Shell_rin": 2.0, "amr_Shell_rout": 20.0, "scalarwave1_k": "0 0 0",
                "nmx_path": path}
        m = nm.Mesh(pars, device=0)
        nm.amr_setup_mesh(m)
        w = np.random.default_rng(8).standard_normal(5 * m.nnodes)
        out[path] = host(adj.adjoint_setrhs(m, dev(w)))
    scale = np.abs(out["plain"]).max()
    assert np.abs(out["batched"] - out["plain"]).max() <= 1e-13 * scale
