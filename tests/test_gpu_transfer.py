"""GPU tests of the drop-in boundary's state transfers and mesh import:

- per element and variable (nmx_state_upload_elm / _download_elm, what
  nmesh's Vard(node, vi) holds, nmesh_amr_loops.h:181) and all at once
  through per-element pointers (nmx_state_upload_ptrs / _download_ptrs),
  both checked against the flat transfer, bitwise, on an h-refined shell
  with FV patches and on the C2 and C4 meshes;
- a mesh taken from nmesh's tables (nmx_mesh_import, here the literal
  fixtures of tests/golden) runs the same RHS as the oracle's mesh, 1e-12.
"""
import json
import os

import numpy as np
import pytest

import nmesh_amd as nm
from oracle.orc import OMesh

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def offsets(m):
    info = np.array([m.elm_info(e) for e in range(m.nelem)])
    return info[:, 3], info[:, 2] ** 3


def per_element(u, m):
    """the flat state [var][node] as arrays[e][v]"""
    off, np_ = offsets(m)
    U = u.reshape(5, -1)
    return [[U[v, off[e]:off[e] + np_[e]].copy() for v in range(5)] for e in range(m.nelem)]


def mixed_mesh():
    pars = {"amr_mesh_type": "Shell", "amr_luni": 1, "amr_n0": 5, "amr_n1": 5, "amr_n2": 5,
            "amr_Shell_rin": 2.0, "amr_Shell_rout": 6.0, "amr_hrefine_sphere_radius": 4,
            "amr_hrefine_sphere_levels": 1, "amr_uniform_p": "0 2 4", "amr_fv_p": "0 2 4"}
    m = nm.Mesh(pars, device=0)
    nm.amr_setup_mesh(m)
    return m


def test_per_element_transfers_match_flat_on_mixed_fv_mesh():
    print("transfers: refined shell with FV patches", flush=True)
    m = mixed_mesh()
    info = np.array([m.elm_info(e) for e in range(m.nelem)])
    assert (info[:, 1] == 2).any()
    rng = np.random.default_rng(40)
    u = rng.standard_normal(5 * m.nnodes)
    m.upload(u)
    A = per_element(u, m)
    for e in range(m.nelem):
        for v in range(5):
            assert np.array_equal(m.Vard(e, v), A[e][v]), (e, v)
    w = rng.standard_normal(5 * m.nnodes)
    B = per_element(w, m)
    for e in range(m.nelem):
        for v in range(5):
            m.set_Vard(e, v, B[e][v])
    assert np.array_equal(m.download(), w)
    # all at once through the pointer table, both ways
    m.upload_vars(A)
    assert np.array_equal(m.download(), u)
    out = [[np.zeros_like(a) for a in row] for row in B]
    m.download_vars(out)
    for e in range(m.nelem):
        for v in range(5):
            assert np.array_equal(out[e][v], A[e][v])
    with pytest.raises(nm.NmxError):
        m.set_Vard(0, 0, np.zeros(7))        # wrong point count


@pytest.mark.parametrize("which", ["C2", "C4"])
def test_pointer_transfers_on_benchmark_meshes(which):
    import bench
    print("transfers:", which, flush=True)
    if which == "C2":
        m = nm.Mesh(bench.c2_params(), device=0)
        nm.amr_setup_mesh(m)
    else:
        m = bench.c4_mesh(nm, "batched")
    u = np.random.default_rng(41).standard_normal(5 * m.nnodes)
    A = per_element(u, m)
    m.upload_vars(A)
    assert np.array_equal(m.download(), u)
    nm.evolve_myln(m, 1)
    v = m.download()
    out = [[np.zeros_like(a) for a in row] for row in A]
    m.download_vars(out)
    assert all(np.array_equal(o, w) for ro, rw in zip(out, per_element(v, m)) for o, w in zip(ro, rw))


def test_imported_mesh_runs_the_oracle_rhs():
    """the fixture tables of a refined shell handed over with nmx_mesh_import,
    then the RHS of a random state against the oracle on its own mesh"""
    print("import: refined shell fixture", flush=True)
    with open(os.path.join(HERE, "golden", "mesh_tables.json")) as fh:
        case = json.load(fh)["shell_luni1_refined"]
    K = [0.8, -0.5, 0.33166247903554]
    m = nm.Mesh({"scalarwave1_k": "%.17g %.17g %.17g" % tuple(K)}, device=0)
    pats = [{"type": 1, "dom": p, "s0": case["rin"], "s1": case["rout"],
             "bbox": [0.0, 1.0, -1.0, 1.0, -1.0, 1.0]} for p in range(6)]
    elms = [{"p": e["p"], "loc": e["loc"], "n": 4, "bbox": e["bbox"],
             "fnb": [f[0] for f in e["faces"]], "nbface": [f[1] for f in e["faces"]]}
            for e in case["elms"]]
    nm.amr_import_mesh(m, pats, elms)
    om = OMesh("shell", 1, 4, rin=case["rin"], rout=case["rout"])
    om, _ = om.adapted(np.zeros(5 * om.nnodes), hflag=np.array(case["refine"][0], np.int32))
    assert m.nelem == om.nelem and m.nnodes == om.nnodes
    u, _ = om.profile(K, 0.1)
    u = u + 0.1 * np.random.default_rng(42).uniform(-1, 1, u.shape)
    m.upload(u)
    r = nm.evolve_setrhs_mesh(m, t=0.1)
    ro = om.rhs(u, 0.1, K)
    assert float(np.abs(r - ro).max() / np.abs(ro).max()) <= 1e-12
