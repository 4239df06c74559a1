"""Integer mesh tables of the oracle and of the product host code against
literal fixtures derived from nmesh's own traversal rules
(tests/golden/make_mesh_tables.py states the rules with their reference
lines): element order and boxes, face-neighbour lists in nmesh's order,
the neighbours' faces, ajsurf modes and orientation codes.  The fixtures do
not use the oracle's geometric neighbour search."""
import json
import os

import numpy as np
import pytest

import nmesh_amd as nm
from oracle.orc import OMesh

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "mesh_tables.json")) as fh:
    FIX = json.load(fh)
N = 4


def oracle_mesh(case):
    if case["kind"] == "box":
        m = OMesh("box", case["luni"], N, dout=case["dout"])
    else:
        m = OMesh("shell", case["luni"], N, rin=case["rin"], rout=case["rout"])
    for flags in case["refine"]:
        m, _ = m.adapted(np.zeros(5 * m.nnodes), hflag=np.array(flags, np.int32))
    return m


def product_mesh(case):
    m = nm.Mesh(device=-1)
    pars = {"amr_luni": case["luni"], "amr_n0": N, "amr_n1": N, "amr_n2": N}
    if case["kind"] == "box":
        pars.update(amr_mesh_type="BoxMesh", amr_BoxMesh_dout=case["dout"])
    else:
        pars.update(amr_mesh_type="Shell", amr_Shell_rin=case["rin"], amr_Shell_rout=case["rout"])
    for k, v in pars.items():
        m.set_par(k, v)
    nm.amr_setup_mesh(m)
    for flags in case["refine"]:
        m.hadapt(np.array(flags, np.int32))
    return m


def check(m, case):
    E = case["elms"]
    assert m.nelem == len(E)
    for e, ref in enumerate(E):
        info = m.elm_info(e)
        assert info[0] == ref["p"] and info[1] == len(ref["loc"]), e
        assert np.allclose(m.elm_bbox(e), ref["bbox"], rtol=0, atol=1e-14), e
        for f in range(6):
            nb, nbf = m.elm_fnb(e, f)
            assert list(nb) == ref["faces"][f][0], (e, f, list(nb), ref["faces"][f][0])
            assert list(nbf) == ref["faces"][f][1], (e, f)
            assert info[4 + f] == len(ref["faces"][f][0])
            assert info[10 + f] == ref["ajmode"][f], (e, f)
            if ref["ajmode"][f] == 2:
                assert info[16 + f] == ref["orient"][f], (e, f)


@pytest.mark.parametrize("name", sorted(FIX))
def test_oracle_tables_match_fixture(name):
    check(oracle_mesh(FIX[name]), FIX[name])


@pytest.mark.parametrize("name", sorted(FIX))
def test_product_tables_match_fixture(name):
    check(product_mesh(FIX[name]), FIX[name])


def test_fixture_covers_the_cases():
    """the fixtures hold coarse faces against 4 finer neighbours, both
    inside a patch and across patches, and oriented cross-patch copies"""
    E = FIX["shell_luni1_refined"]["elms"]
    cross = [len(e["faces"][f][0]) for e in E for f in range(6)
             if e["faces"][f][0] and E[e["faces"][f][0][0]]["p"] != e["p"]]
    assert max(cross) == 4
    assert any(m == 2 for e in FIX["shell_luni1"]["elms"] for m in e["ajmode"])
    assert max(len(e["faces"][f][0]) for e in FIX["box_two_levels"]["elms"] for f in range(6)) >= 4


def fixture_import(case, n=N, edit=None):
    """nmx_mesh_import of a fixture mesh into a host-only context"""
    if case["kind"] == "box":
        d = case["dout"]
        pats = [{"type": 0, "bbox": [-d, d, -d, d, -d, d]}]
    else:
        pats = [{"type": 1, "dom": p, "s0": case["rin"], "s1": case["rout"],
                 "bbox": [0.0, 1.0, -1.0, 1.0, -1.0, 1.0]} for p in range(6)]
    elms = [{"p": e["p"], "loc": e["loc"], "n": n, "bbox": list(e["bbox"]),
             "fnb": [list(f[0]) for f in e["faces"]], "nbface": [list(f[1]) for f in e["faces"]]}
            for e in case["elms"]]
    if edit:
        edit(elms)
    m = nm.Mesh(device=-1)
    nm.amr_import_mesh(m, pats, elms)
    return m


@pytest.mark.parametrize("name", sorted(FIX))
def test_import_of_nmesh_tables_equals_build(name):
    """nmx_mesh_import takes nmesh's tables (here the fixtures) and the
    result equals the built mesh: boxes, neighbours, modes, coordinates, dt"""
    case = FIX[name]
    m = fixture_import(case)
    check(m, case)
    b = product_mesh(case)
    assert m.nelem == b.nelem and m.nnodes == b.nnodes
    for e in range(m.nelem):
        assert np.array_equal(m.elm_info(e), b.elm_info(e))
    assert np.array_equal(m.coords(), b.coords())
    assert m.dt == b.dt


def test_import_refuses_tables_that_differ():
    case = FIX["shell_luni1_refined"]
    big = [i for i, e in enumerate(case["elms"]) if any(len(f[0]) > 1 for f in e["faces"])][0]
    f = [k for k in range(6) if len(case["elms"][big]["faces"][k][0]) > 1][0]

    def swap(elms):        # two neighbours of one face in the other order
        elms[big]["fnb"][f] = elms[big]["fnb"][f][::-1]
        elms[big]["nbface"][f] = elms[big]["nbface"][f][::-1]

    def face(elms):        # a wrong neighbour face
        elms[0]["nbface"][1] = [(elms[0]["nbface"][1][0] + 1) % 6]

    def box(elms):         # a box one ulp off
        elms[5]["bbox"][1] = np.nextafter(elms[5]["bbox"][1], 9.0)

    def order(elms):       # two elements swapped in the list
        elms[2], elms[3] = elms[3], elms[2]

    for edit, words in ((swap, "neighbour list differs"), (face, "neighbour list differs"),
                        (box, "box differs"), (order, "depth-first")):
        with pytest.raises(nm.NmxError, match=words):
            fixture_import(case, edit=edit)
