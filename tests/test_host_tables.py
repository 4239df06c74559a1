"""The product's host mesh code (libnmx, host-only context, no GPU calls)
against the oracle: element order, neighbour lists, ajsurf modes,
orientation codes (integers, bit-exact), coordinates, geometry and dt."""
import numpy as np
import pytest

import nmesh_amd as nm
from oracle.orc import OMesh


def product(kind, luni, n, **kw):
    m = nm.Mesh(device=-1)
    pars = {"amr_mesh_type": "BoxMesh" if kind == "box" else "Shell",
            "amr_luni": luni, "amr_n0": n, "amr_n1": n, "amr_n2": n}
    if kind == "box":
        pars["amr_BoxMesh_dout"] = kw.get("dout", 8.0)
    else:
        pars["amr_Shell_rin"] = kw.get("rin", 2.0)
        pars["amr_Shell_rout"] = kw.get("rout", 20.0)
        pars["amr_hrefine_sphere_radius"] = kw.get("radius", 10)
        pars["amr_hrefine_sphere_levels"] = kw.get("levels", 0)
    for k, v in pars.items():
        m.set_par(k, v)
    nm.amr_setup_mesh(m)
    return m


# the 1056-element refined shell is built over element ranges on several
# host threads (64 elements per thread at least, up to 16 threads); the
# smaller meshes on one
CASES = [("box", 2, 5, {}), ("box", 1, 3, {"dout": 1.0}), ("shell", 1, 6, {}),
         ("shell", 2, 4, {}), ("shell", 2, 4, {"radius": 3, "levels": 2}),
         ("shell", 1, 5, {"radius": 9, "levels": 1})]


@pytest.mark.parametrize("kind,luni,n,kw", CASES)
def test_tables_bit_exact(kind, luni, n, kw):
    p = product(kind, luni, n, **kw)
    o = OMesh(kind, luni, n, dout=kw.get("dout", 8.0), rin=kw.get("rin", 2.0),
              rout=kw.get("rout", 20.0), radius=kw.get("radius", 10), levels=kw.get("levels", 0))
    assert p.nelem == o.nelem and p.nnodes == o.nnodes
    for e in range(o.nelem):
        assert np.array_equal(p.elm_info(e), o.elm_info(e)), e
        assert np.array_equal(p.elm_bbox(e), o.elm_bbox(e))
        for f in range(6):
            pb, pf = p.elm_fnb(e, f)
            ob, of = o.elm_fnb(e, f)
            assert np.array_equal(pb, ob) and np.array_equal(pf, of)
    X, x = o.coords()
    assert np.array_equal(p.coords(), x)
    g, sg, _ = o.geometry()
    pg, psg = p.geometry()
    assert np.array_equal(pg, g) and np.array_equal(psg, sg)
    assert p.dt == o.dt()


def test_amr_mesh_has_interpolated_faces():
    p = product("shell", 2, 4, radius=3, levels=2)
    modes = np.array([p.elm_info(e)[10:16] for e in range(p.nelem)])
    assert (modes == 3).sum() > 0
    assert p.nelem > 384


def test_mixed_order_tables_match():
    rng = np.random.default_rng(4)
    p = product("shell", 1, 5)
    o = OMesh("shell", 1, 5)
    n = rng.integers(3, 8, p.nelem).astype(np.int32)
    p.set_elm_n(n)
    o.set_n(n)
    for e in range(p.nelem):
        assert np.array_equal(p.elm_info(e), o.elm_info(e))
    _, x = o.coords()
    assert np.array_equal(p.coords(), x)


def test_explicit_dt_survives_mesh_changes():
    """nmx_set_dt records an explicit dt: a p-change, an h-refinement or a
    new dtfac keep it; dt = auto brings back the rule (storage.c:1305-1386)."""
    p = product("shell", 1, 4)
    auto = p.dt
    p.dt = 1.25e-3
    p.padapt(np.full(p.nelem, 5, np.int32))
    assert p.dt == 1.25e-3
    flags = np.zeros(p.nelem, np.int32)
    flags[3] = 1
    p.hadapt(flags)
    p.set_par("dtfac", 0.1)
    assert p.dt == 1.25e-3 and p.par("dt") != "auto"
    p.set_par("dt", "auto")
    o = OMesh("shell", 1, 4)
    o, _ = o.adapted(np.zeros(5 * o.nnodes), n=np.full(o.nelem, 5, np.int32))
    o, _ = o.adapted(np.zeros(5 * o.nnodes), hflag=flags)
    assert p.dt == o.dt(dtfac=0.1) and p.dt != auto
