"""Pins of the oracle's mesh adaptation with state interpolation
(orc_mesh_hadapt, orc_adapt_state after refine.c:156-537 and
storage.c:493-1000).

- Exactness: the Lagrange interpolation of the element's own polynomial
  space is exact, so a field that is a polynomial of degree <= n-1 in each
  coordinate on a Cartesian box is reproduced at the new points by
  p-refinement (n -> n' >= n, and n -> n' < n when the degree allows),
  h-refinement (children), and coarsening (parents).
- Refining and then coarsening gives back any state (the children hold the
  parent's polynomial exactly).
- The coarsening average: with a different constant on each child, a parent
  point inside one child takes that child's constant, and a point on a face,
  an edge or the centre shared by 2, 4 or 8 children takes the mean of their
  constants (make_parent_elm, storage.c:706-745).
- The new mesh's tables: siblings in Morton order replace their parent in
  place; coarsening takes the largest sibling n.
"""
import numpy as np
import pytest

from oracle.orc import OMesh


def poly(x, deg):
    """sum of monomials x^a y^b z^c, a, b, c <= deg, fixed coefficients"""
    rng = np.random.default_rng(deg)
    out = np.zeros(x.shape[1])
    for a in range(deg + 1):
        for b in range(deg + 1):
            for c in range(deg + 1):
                out += rng.uniform(-1, 1) * x[0] ** a * x[1] ** b * x[2] ** c
    return out


def state(om, deg):
    _, x = om.coords()
    x = x.reshape(3, -1)
    u = np.stack([poly(x, deg) * (v + 1) for v in range(5)])
    return u.ravel()


def check(new, unew, deg):
    ref = state(new, deg)
    assert np.abs(unew - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("n,n2,deg", [(4, 6, 3), (6, 4, 3), (5, 5, 4), (3, 8, 2)])
def test_p_change_is_exact_on_polynomials(n, n2, deg):
    om = OMesh("box", 1, n, dout=2.0)
    u = state(om, deg)
    new, unew = om.adapted(u, n=np.full(om.nelem, n2, dtype=np.int32))
    check(new, unew, deg)


def test_h_refine_and_coarsen_are_exact_on_polynomials():
    om = OMesh("box", 1, 5, dout=2.0)
    u = state(om, 4)
    flag = np.zeros(om.nelem, np.int32)
    flag[[0, 5]] = 1
    new, unew = om.adapted(u, hflag=flag)
    assert new.nelem == 8 + 14
    check(new, unew, 4)
    back, ub = new.adapted(unew, hflag=-np.ones(new.nelem, np.int32))
    # one pass: the two runs of 8 new children coarsen back; the level-1
    # elements are not 8 siblings in a row at that moment and stay
    assert back.nelem == 8
    check(back, ub, 4)


def test_refine_then_coarsen_is_the_identity():
    om = OMesh("shell", 1, 5, rin=2.0, rout=6.0)
    u = np.random.default_rng(1).standard_normal(5 * om.nnodes)
    flag = np.zeros(om.nelem, np.int32)
    flag[[3, 17, 40]] = 1
    ref, ur = om.adapted(u, hflag=flag)
    cflag = np.zeros(ref.nelem, np.int32)
    # the children of the refined elements: level 2
    for e in range(ref.nelem):
        if ref.elm_info(e)[1] == 2:
            cflag[e] = -1
    back, ub = ref.adapted(ur, hflag=cflag)
    assert back.nelem == om.nelem
    for e in range(om.nelem):
        assert np.array_equal(back.elm_info(e), om.elm_info(e))
    assert np.abs(ub - u).max() <= 1e-12 * np.abs(u).max()


@pytest.mark.parametrize("n", [4, 5])
def test_coarsening_averages_shared_points(n):
    om = OMesh("box", 1, n, dout=2.0)
    consts = np.arange(1.0, 9.0) ** 2
    u = np.zeros((5, om.nnodes))
    for e in range(8):
        off = int(om.elm_info(e)[3])
        u[:, off:off + n ** 3] = consts[e]
    par, up = om.adapted(u.ravel(), hflag=-np.ones(8, np.int32))
    assert par.nelem == 1
    up = up.reshape(5, n, n, n)[0]       # [k, j, i]
    X, _ = par.coords()
    X = X.reshape(3, n, n, n)
    for k in range(n):
        for j in range(n):
            for i in range(n):
                xs = [X[0, k, j, i], X[1, k, j, i], X[2, k, j, i]]
                sides = [[0, 1] if abs(v) < 1e-12 else ([1] if v > 0 else [0]) for v in xs]
                kids = [a + 2 * b + 4 * c for a in sides[0] for b in sides[1] for c in sides[2]]
                assert abs(up[k, j, i] - consts[kids].mean()) < 1e-12, (i, j, k)


def test_coarsening_takes_the_largest_sibling_order():
    om = OMesh("box", 1, 4, dout=2.0)
    nv = np.array([4, 6, 3, 5, 4, 4, 7, 2], np.int32)
    om.set_n(nv)
    u = state(om, 1)
    par, up = om.adapted(u, hflag=-np.ones(8, np.int32))
    assert par.nelem == 1 and par.elm_info(0)[2] == 7
    check(par, up, 1)


def test_product_host_tables_after_adaptation_match_the_oracle():
    """nmx_mesh_hadapt / nmx_mesh_padapt on a host-only context: element
    order, neighbour tables, coordinates and dt equal the oracle's"""
    import nmesh_amd as nm
    pars = {"amr_mesh_type": "Shell", "amr_luni": 1, "amr_n0": 5, "amr_n1": 5, "amr_n2": 5,
            "amr_Shell_rin": 2.0, "amr_Shell_rout": 6.0}
    m = nm.Mesh(pars, device=-1)
    nm.amr_setup_mesh(m)
    om = OMesh("shell", 1, 5, rin=2.0, rout=6.0)
    flag = np.zeros(om.nelem, np.int32)
    flag[[3, 17, 40]] = 1
    m.hadapt(flag)
    om2, _ = om.adapted(np.zeros(5 * om.nnodes), hflag=flag)
    rng = np.random.default_rng(6)
    nv = rng.integers(3, 9, om2.nelem).astype(np.int32)
    m.padapt(nv)
    om3, _ = om2.adapted(np.zeros(5 * om2.nnodes), n=nv)
    cflag = np.array([-1 if om3.elm_info(e)[1] == 2 else 0 for e in range(om3.nelem)], np.int32)
    cflag[:8] = -1                      # the first 8 level-1 elements are siblings too
    m.hadapt(cflag)
    om4, _ = om3.adapted(np.zeros(5 * om3.nnodes), hflag=cflag)
    assert m.nelem == om4.nelem
    for e in range(om4.nelem):
        assert np.array_equal(m.elm_info(e), om4.elm_info(e)), e
    assert np.array_equal(m.coords(), om4.coords()[1])
    assert m.dt == om4.dt()


def test_switch_to_fv_and_back():
    """DG -> FV on uniform points of the same n (Lagrange from LGL) and back
    (Lagrange from the uniform points, the inverse of the first matrix):
    exact on polynomials of degree <= n-1, and the round trip returns any
    state; the host tables after a switch equal the oracle's"""
    import nmesh_amd as nm
    om = OMesh("box", 1, 5, dout=2.0)
    u = state(om, 4)
    to_fv = np.array([1, 0, 1, 1, 0, 0, 1, 0], np.int32)
    fvm, ufv = om.adapted(u, fv=to_fv)
    check(fvm, ufv, 4)                  # the polynomial at the uniform points
    v = np.random.default_rng(8).standard_normal(5 * om.nnodes)
    fvm, vf = om.adapted(v, fv=to_fv)
    back, vb = fvm.adapted(vf, fv=np.zeros(8, np.int32))
    assert np.abs(vb - v).max() <= 1e-12 * np.abs(v).max()
    m = nm.Mesh({"amr_mesh_type": "BoxMesh", "amr_luni": 1, "amr_n0": 5, "amr_n1": 5, "amr_n2": 5,
                 "amr_BoxMesh_dout": 2.0}, device=-1)
    nm.amr_setup_mesh(m)
    m.switch_fv(to_fv)
    for e in range(8):
        assert np.array_equal(m.elm_info(e), fvm.elm_info(e)), e
    assert np.array_equal(m.coords(), fvm.coords()[1])
    assert m.dt == fvm.dt()
