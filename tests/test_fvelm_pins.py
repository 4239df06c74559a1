"""Pins of the oracle's FV-element path (fv_divf for scalarwave1, the
reference's default FV parameters) and of the product host tables for FV
elements.

- WENOm3 (rec1d_p/m_WENO3 with the uniform WENOm3 weights, rec1d.c:115-232,
  875-905): exact on linear data; on smooth data it tends to the combination
  with the optimal weights (1, 2) / 3 (a swapped weight fails); at a step it
  keeps the smooth side.
- The whole FV-element RHS: a state linear in x, y, z gives the exact RHS
  (r_pi = div c, r_c = grad pi, r_phi = pi) on every element not touching
  the outer boundary, for WENOm3_2 and WENOm5_2.  Every reconstruction,
  closure, face flux, metric factor and the dnfn_extrap1 weights (4/3, -1/3)
  are exact on linear data, so an index, sign, weight or factor slip fails.
- Uniform points as uniform_x places them (Legendre.c:626-663); coordinates,
  tables and dt of the product host bit-exact to the oracle, Cartesian and
  cubed-sphere shell, all FV or FV in some patches only.
- basis_pw_linear (interpolate.c:53-102), the interpolation from an FV
  neighbour's face: reproduces linear data exactly, is a partition of
  unity, and equals numpy's interp inside the points.
- Curvilinear FV elements (cubed-sphere shell), the plane-wave RHS against
  its exact time derivative:
  - coordinates_midpoint_data = yes (metric and normals at the midpoints):
    deep inside the elements the error falls at third order or better, and
    everywhere at first order, as on the Cartesian box (the closures at the
    element faces);
  - coordinates_midpoint_data = no (the reference's default: a midpoint
    takes the metric of the node on its left, coordinates.c:467-483) is
    first order deep inside (the metric sits half a spacing off) and does
    not converge at the nodes next to element faces, where the half-width
    end cell takes the metric of its neighbour node (fv.c:386-391).
  A midpoint metric evaluated at the wrong place, along the wrong direction
  or without det(dXb/dx) fails the first; the second fixes which node the
  default mode takes.
  On a Cartesian box both modes give the identical RHS.
"""
import numpy as np
import pytest

import nmesh_amd as nm
from oracle import orc
from oracle.orc import OMesh, lib


def _p(a):
    import ctypes as C
    return a.ctypes.data_as(C.c_void_p)


def rec_m3_p(u, im):
    u = np.ascontiguousarray(u, dtype=np.float64)
    return lib().orc_rec1d_p_WENOm3(_p(u), int(im))


def rec_m3_m(u, im):
    u = np.ascontiguousarray(u, dtype=np.float64)
    return lib().orc_rec1d_m_WENOm3(_p(u), int(im))


def test_wenom3_exact_on_linear_data_and_optimal_weights_on_smooth_data():
    rng = np.random.default_rng(1)
    for _ in range(100):
        a, b = rng.uniform(-2, 2, 2)
        x = np.arange(5.0) + rng.uniform(-3, 3)
        u = a + b * x
        mid = a + b * (x[1] + 0.5)
        assert abs(rec_m3_p(u, 1) - mid) < 1e-13 * (1 + abs(mid))
        assert abs(rec_m3_m(u, 1) - mid) < 1e-13 * (1 + abs(mid))
    # smooth data (beta << eps = 1e-6): the optimal weights (1, 2) / 3
    h = 1e-4
    x = np.arange(5.0) * h
    u = np.sin(x + 0.3)
    lin_p = (1 / 3) * (-0.5 * u[0] + 1.5 * u[1]) + (2 / 3) * (0.5 * u[1] + 0.5 * u[2])
    lin_m = (2 / 3) * (0.5 * u[1] + 0.5 * u[2]) + (1 / 3) * (1.5 * u[2] - 0.5 * u[3])
    assert abs(rec_m3_p(u, 1) - lin_p) < 1e-12
    assert abs(rec_m3_m(u, 1) - lin_m) < 1e-12


def test_wenom3_keeps_the_smooth_side_at_a_step():
    u = np.array([0.0, 0.0, 1.0, 1.0])
    # midpoint between u[1] = 0 and u[2] = 1 from the left: the smooth
    # left stencil (0, 0) dominates
    assert abs(rec_m3_p(u, 1)) < 1e-10
    assert abs(rec_m3_m(u, 1) - 1.0) < 1e-10


def test_uniform_points_as_uniform_x():
    for n in range(2, 12):
        om = OMesh("box", 0, n, dout=1.0, fv=True)
        X, x = om.coords()
        xs = np.unique(X[0])
        N = n - 1
        h = 2.0 / N
        ref = np.empty(n)
        ref[0], ref[N] = -1.0, 1.0
        for j in range(1, N // 2 + 1):
            y = 0.0 if j == N / 2.0 else -1.0 + j * h
            ref[j], ref[N - j] = y, -y
        assert np.array_equal(np.sort(xs), np.sort(ref))


def test_dnfn_extrap1_is_linear_extrapolation_from_the_fv_difference_locations():
    """The FV differences of a line live at h/4 (face to first midpoint),
    then h, 2h, ... (midpoint to midpoint), and L - h/4: on samples of any
    linear function at those places the extrapolation returns its exact
    values at the faces 0 and L."""
    rng = np.random.default_rng(4)
    for n in range(3, 11):
        h = 1.0
        L = (n - 1) * h
        xs = np.array([h / 4] + [i * h for i in range(1, n - 1)] + [L - h / 4])
        a, b = rng.uniform(-2, 2, 2)
        df = np.ascontiguousarray(a + b * xs)
        lib().orc_dnfn_extrap1(n, _p(df))
        assert abs(df[0] - a) < 1e-13 and abs(df[-1] - (a + b * L)) < 1e-13


def _extrap(df, s1=0.0, s2=None, opt=0):
    df = np.array(df, dtype=np.float64)   # a copy: the oracle works in place
    orc.set_fv_divf_extrap(s1, s2, opt)
    try:
        lib().orc_dnfn_extrap1(len(df), _p(df))
    finally:
        orc.set_fv_divf_extrap()
    return df


def test_dnfn_extrap1_s1_s2_opt_limits():
    """fv_divf_extrap_s1 / _s2 / _opt (rec1d_compute_1s1_u, rec1d.c:542-609):
    on linear data at the FV difference locations the ratio s =
    1.3333333333 |d1| / |d2| is 1 - 2.5e-11 (the reference's truncated 4/3),
    so s1 = 1 already switches to the end value while s1 = 0.999 does not,
    and s2 = 1 keeps the extrapolation while s2 = 0.9 does not; opt 1 / 2
    switch off on opposite / equal signs of d1 and d2, opt -1 always, but
    only while s1 > 0 or s2 < DBL_MAX turns the limit on; with 3 points the
    right end sees the already extrapolated left end."""
    xs = np.array([0.25, 1.0, 2.0, 3.0, 3.75])
    lin = 0.5 + 2.0 * xs
    exact = [0.5, 0.5 + 2.0 * 4.0]
    for kw, extrap in [({}, True), ({"s1": 0.999}, True), ({"s1": 1.0}, False),
                       ({"s2": 1.0}, True), ({"s2": 0.9}, False),
                       ({"s1": 0.5, "opt": 1}, True), ({"s1": 0.5, "opt": 2}, False),
                       ({"s1": 0.5, "opt": -1}, False), ({"opt": -1}, True)]:
        out = _extrap(lin, **kw)
        want = exact if extrap else [lin[0], lin[-1]]
        assert abs(out[0] - want[0]) < 1e-13 and abs(out[-1] - want[1]) < 1e-13, kw
        assert np.array_equal(out[1:-1], lin[1:-1])
    # a zigzag: d1 and d2 of opposite signs at both ends
    zz = np.array([0.0, 1.0, 0.5, 1.0, 0.0])
    out = _extrap(zz, s1=1e-3, opt=1)
    assert out[0] == 0.0 and out[-1] == 0.0
    out = _extrap(zz, s1=1e-3, opt=2)
    assert abs(out[0] - (-1.0 / 3.0)) < 1e-15 and abs(out[-1] - (-1.0 / 3.0)) < 1e-15
    # 3 points, left first: [0, 1, 2] -> left extrapolates to -1/3; the
    # right end then has d2 = -4/3, d1 = -1, s = 0.99999999997 < 1.2 and
    # keeps 2 (with the unmodified left value s would be 1.333 and give 7/3)
    out = _extrap([0.0, 1.0, 2.0], s1=1.2)
    assert abs(out[0] + 1.0 / 3.0) < 1e-15 and out[1] == 1.0 and out[2] == 2.0


def _ends(u, mode, s1=0.0, s2=None, opt=0):
    u = np.array(u, dtype=np.float64)
    out = np.zeros(2)
    orc.set_fv_rec(mode, s1, s2, opt)
    try:
        lib().orc_fv_rec_ends(len(u), _p(u), _p(out))
    finally:
        orc.set_fv_rec("WENOm3_2")
    return out


def test_fv_rec_end_midpoints_by_mode():
    """fv_rec's end midpoints (rec1d.c:1032-1186): the _2 modes average the
    two end points, the _1 modes take the end point, WENOm3/5/Z average
    unless fv_rec_WENOm_s1/_s2/_opt take the end point; s = |d1| / |d2| has
    no factor here, so on linear data s = 1 exactly: s1 = 1 keeps the
    average, s1 = 1 + 1e-9 takes the end point, s2 = 1 takes it,
    s2 = 1 + 1e-9 does not; opt 1 takes it on a zigzag"""
    lin = 0.5 + 2.0 * np.arange(6.0)
    avg = [0.5 * lin[0] + 0.5 * lin[1], 0.5 * lin[-2] + 0.5 * lin[-1]]
    end = [lin[0], lin[-1]]
    for base in ("WENOm3", "WENOm5", "WENOmZ"):
        assert np.array_equal(_ends(lin, base + "_2"), avg)
        assert np.array_equal(_ends(lin, base + "_1"), end)
        assert np.array_equal(_ends(lin, base), avg)
        for kw, want in [({"s1": 1.0}, avg), ({"s1": 1.0 + 1e-9}, end), ({"s2": 1.0}, end),
                         ({"s2": 1.0 + 1e-9}, avg), ({"s1": 0.5, "opt": 2}, end),
                         ({"opt": -1}, avg), ({"s1": 0.5, "opt": -1}, end)]:
            assert np.array_equal(_ends(lin, base, **kw), want), (base, kw)
    zz = np.array([0.0, 1.0, 0.5, 0.7, 1.0, 0.0])
    assert np.array_equal(_ends(zz, "WENOm3", s1=1e-3, opt=1), [0.0, 0.0])


def _pm(u, im, mode):
    u = np.array(u, dtype=np.float64)
    out = np.zeros(2)
    orc.set_fv_rec(mode)
    try:
        lib().orc_fv_rec_pm(len(u), _p(u), im, _p(out))
    finally:
        orc.set_fv_rec("WENOm3_2")
    return out


def test_wenot_closed_form_and_symmetry():
    """WENOmT_2's interior (rec1d_p/m_WENOT, rec1d.c:1412-1578): on the
    zigzag 0, 1, 0, 1, 0 the smoothness indicators are 25/3, 13/3, 25/3 and
    the stencils give -7/6, 1/6, 5/6, so the weights d_k / b_k with
    d = (1, 6, 3) give 0.2096 (WENO5's squared indicators give 0.1922, and
    the test tells them apart); linear point values are reproduced at the
    midpoints (every stencil is exact on them); rec_m is rec_p of the
    reversed line"""
    from fractions import Fraction as Fr
    b = [Fr(25, 3), Fr(13, 3), Fr(25, 3)]
    q = [Fr(-7, 6), Fr(1, 6), Fr(5, 6)]
    d = [1, 6, 3]
    a = [d[k] / b[k] for k in range(3)]
    want = float(sum(a[k] * q[k] for k in range(3)) / sum(a))
    u = [0.0, 1.0, 0.0, 1.0, 0.0, 1.0, 0.0]
    got = _pm(u, 2, "WENOmT_2")[0]
    assert abs(got - want) < 1e-14
    a5 = [d[k] / b[k] ** 2 for k in range(3)]
    w5 = float(sum(a5[k] * q[k] for k in range(3)) / sum(a5))
    assert abs(_pm(u, 2, "WENOm5_2")[0] - w5) < 1e-14 and abs(want - w5) > 1e-2
    rng = np.random.default_rng(13)
    for _ in range(20):
        x = np.arange(8.0)
        c = rng.standard_normal(2)
        lin = c[0] + c[1] * x
        for im in range(2, 5):
            p, m = _pm(lin, im, "WENOmT_2")
            exact = c[0] + c[1] * (im + 0.5)
            assert abs(p - exact) < 1e-12 * (1 + abs(exact)) and abs(m - exact) < 1e-12 * (1 + abs(exact))
        v = rng.standard_normal(8)
        for im in range(1, 5):
            m = _pm(v, im, "WENOmT_2")[1]
            p = _pm(v[::-1].copy(), 8 - 2 - im, "WENOmT_2")[0]
            assert abs(m - p) < 1e-14 * (1 + abs(p))


def test_fv_rec_1_takes_the_neighbouring_points():
    """fv_rec = 1 (rec1d_p_1 / rec1d_m_1, rec1d.c:16-27): the values left
    and right of each midpoint are the grid points themselves"""
    v = np.random.default_rng(14).standard_normal(7)
    for im in range(6):
        assert np.array_equal(_pm(v, im, "1"), [v[im], v[im + 1]])


def test_fv_rec_end_point_kept_skips_the_extrapolation():
    """where WENOm keeps the end point (fv_stat_WENOm_1or2, fv.c:45-79),
    dnfn_extrap1 leaves that end's df alone (fv.c:558-573): with
    s1 = 1e300 every end keeps its point, so the RHS equals the one of the
    _1 mode (the same end points, extrapolated) at all nodes off the element
    faces, and differs on them"""
    n = 6
    om = OMesh("box", 1, n, dout=2.0, fv=True)
    u, _ = om.profile(K, 0.1)
    u = u + 0.1 * np.random.default_rng(9).uniform(-1, 1, u.shape)
    try:
        orc.set_fv_rec("WENOm5", s1=1e300)
        r_kept = om.rhs(u, 0.1, K, "upwind")
        orc.set_fv_rec("WENOm5_1")
        r_one = om.rhs(u, 0.1, K, "upwind")
    finally:
        orc.set_fv_rec("WENOm3_2")
    I = np.indices((n, n, n)).reshape(3, -1)
    inner = np.all((I > 0) & (I < n - 1), axis=0)
    mask = np.tile(np.tile(inner, om.nelem), 5)
    assert np.array_equal(r_kept[mask], r_one[mask])
    assert np.abs(r_kept[~mask] - r_one[~mask]).max() > 1e-6


@pytest.mark.parametrize("rec", ["WENOm3_2", "WENOm5_2", "WENOmZ_2"])
@pytest.mark.parametrize("n", [5, 7])
def test_fv_element_rhs_exact_for_linear_states(rec, n):
    orc.set_fv_rec(rec)
    om = OMesh("box", 2, n, dout=8.0, fv=True)
    X, x = om.coords()
    nn = om.nnodes
    rng = np.random.default_rng(n)
    a = rng.uniform(-1, 1, 5)
    B = rng.uniform(-0.5, 0.5, (5, 3))
    B[4] = 0.0
    u = a[:, None] + B @ x
    r = om.rhs(u.ravel(), 0.0, [0, 0, 0], "upwind").reshape(5, -1)
    ex = np.zeros_like(u)
    ex[0] = B[1, 0] + B[2, 1] + B[3, 2]
    ex[1:4] = B[0][:, None]
    ex[4] = u[0]
    inner = np.zeros(nn, bool)
    for e in range(om.nelem):
        info = om.elm_info(e)
        bb = om.elm_bbox(e)
        if bb.min() > -8 + 1e-9 and bb.max() < 8 - 1e-9:
            inner[info[3]:info[3] + info[2] ** 3] = True
    assert inner.sum() > 0
    assert np.abs(r[:, inner] - ex[:, inner]).max() < 1e-12
    orc.set_fv_rec("WENOm3_2")


def test_fv_element_rhs_converges_for_a_plane_wave():
    orc.set_fv_rec("WENOm5_2")
    K = [0.8, -0.5, 0.33166247903554]
    errs = []
    for luni in (1, 2, 3):
        om = OMesh("box", luni, 6, dout=2.0, fv=True)
        u, du = om.profile(K, 0.3)
        errs.append(np.abs(om.rhs(u, 0.3, K) - du).max())
    # first order in the max norm: the closures next to element faces are
    # second-order midpoint values, differenced over one spacing
    assert errs[1] < errs[0] and errs[2] < errs[1]
    assert errs[1] / errs[2] > 1.8
    orc.set_fv_rec("WENOm3_2")


K = [0.8, -0.5, 0.33166247903554]


def shell_errors(n, mid):
    """max |RHS - du/dt| of the plane wave on FV shell elements: over all
    nodes, and over the nodes at least n/4 points from every element face"""
    om = OMesh("shell", 1, n, rin=2.0, rout=6.0, fv=True)
    u, du = om.profile(K, 0.2)
    orc.set_fv_midpoint(mid)
    try:
        r = om.rhs(u, 0.2, K, "upwind")
    finally:
        orc.set_fv_midpoint(False)
    e = np.abs(r - du).reshape(5, om.nelem, n, n, n)[:4]
    q = n // 4
    return e.max(), e[:, :, q:-q, q:-q, q:-q].max()


def order(e0, e1, n0, n1):
    return np.log(e0 / e1) / np.log((n1 - 1) / (n0 - 1))


def test_shell_fv_with_midpoint_metric_converges():
    orc.set_fv_rec("WENOm5_2")
    ns = (12, 16, 20)
    errs = [shell_errors(n, True) for n in ns]
    for (a0, d0), (a1, d1), n0, n1 in zip(errs[:-1], errs[1:], ns[:-1], ns[1:]):
        assert order(d0, d1, n0, n1) > 2.4, (n0, n1, d0, d1)
        assert 0.8 < order(a0, a1, n0, n1) < 1.6, (n0, n1, a0, a1)
    orc.set_fv_rec("WENOm3_2")


def test_shell_fv_default_metric_first_order_inside_inconsistent_at_faces():
    orc.set_fv_rec("WENOm5_2")
    ns = (12, 16, 20)
    errs = [shell_errors(n, False) for n in ns]
    for (a0, d0), (a1, d1), n0, n1 in zip(errs[:-1], errs[1:], ns[:-1], ns[1:]):
        assert 0.7 < order(d0, d1, n0, n1) < 1.4, (n0, n1, d0, d1)
        assert a1 > 1.0 and a1 > 0.95 * a0, (n0, n1, a0, a1)
    orc.set_fv_rec("WENOm3_2")


def test_midpoint_metric_on_cartesian_elements_is_the_node_metric():
    om = OMesh("box", 1, 7, dout=2.0, fv=True)
    u, _ = om.profile(K, 0.1)
    u = u + 0.1 * np.random.default_rng(3).uniform(-1, 1, u.shape)
    r0 = om.rhs(u, 0.1, K, "upwind")
    orc.set_fv_midpoint(True)
    try:
        r1 = om.rhs(u, 0.1, K, "upwind")
    finally:
        orc.set_fv_midpoint(False)
    assert np.array_equal(r0, r1)


@pytest.mark.parametrize("kind,luni,n,pats", [("box", 1, 5, [0]), ("box", 2, 6, [0]),
                                              ("shell", 1, 6, [5, 4, 3, 2, 1, 0]),
                                              ("shell", 1, 5, [0, 2, 4])])
def test_host_tables_of_fv_meshes_match_the_oracle(kind, luni, n, pats):
    ps = " ".join(str(p) for p in pats)
    m = nm.Mesh({"amr_mesh_type": "BoxMesh" if kind == "box" else "Shell", "amr_luni": luni,
                 "amr_n0": n, "amr_n1": n, "amr_n2": n, "amr_BoxMesh_dout": 8.0,
                 "amr_Shell_rin": 2.0, "amr_Shell_rout": 20.0,
                 "amr_uniform_p": ps, "amr_fv_p": ps}, device=-1)
    nm.amr_setup_mesh(m)
    om = OMesh(kind, luni, n, dout=8.0, rin=2.0, rout=20.0, fv_patches=pats)
    X, x = om.coords()
    assert np.array_equal(m.coords(), x)
    assert m.dt == om.dt()
    for e in range(om.nelem):
        assert np.array_equal(m.elm_info(e), om.elm_info(e))


@pytest.mark.parametrize("n", [4, 6, 9])
def test_fv_dt_uses_uniform_dtfac(n):
    """adapt_node_dt_and_mesh_dt (utilities.c:1264-1304): on FV elements the
    smallest spacing is between two uniform points, so dt = uniform_dtfac *
    hmin: on a Cartesian FV box hmin is the uniform spacing width / (n - 1)
    (closed form), on a DG box dtfac * the first LGL spacing; the host's dt
    equals the oracle's, and follows both parameters"""
    pars = {"amr_mesh_type": "BoxMesh", "amr_luni": 1, "amr_n0": n, "amr_n1": n, "amr_n2": n,
            "amr_BoxMesh_dout": 8.0}
    width = 8.0   # luni 1: two elements of dout per direction
    xl = lgl(n)
    for fv in (False, True):
        extra = {"amr_uniform_p": "0", "amr_fv_p": "0"} if fv else {}
        m = nm.Mesh(dict(pars, **extra), device=-1)
        nm.amr_setup_mesh(m)
        om = OMesh("box", 1, n, dout=8.0, fv=fv)
        h = width / (n - 1) if fv else width * (xl[1] - xl[0]) / 2
        expect = (0.125 if fv else 0.25) * h
        assert abs(om.dt() - expect) <= 1e-15 * expect
        assert m.dt == om.dt()
        m.set_par("uniform_dtfac", "0.1")
        m.set_par("dtfac", "0.3")
        assert m.dt == om.dt(0.3, 0.1)
        assert abs(m.dt - (0.1 if fv else 0.3) * h) <= 1e-15 * h


def lgl(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_lgl(n, _p(x), _p(w))
    return x


def _pw_parab_cells(x, xp, g):
    """the piecewise parabolic interpolant of g on the points xp, stated by
    cells: below the first midpoint the parabola through points 0, 1, 2,
    above the last one through the last three, elsewhere the parabola
    through the nearest point and its two neighbours (numpy's polyfit)"""
    n = len(xp)
    if x < 0.5 * (xp[0] + xp[1]):
        m = 1
    elif x >= 0.5 * (xp[n - 2] + xp[n - 1]):
        m = n - 2
    else:
        m = max(i for i in range(1, n - 1) if x >= 0.5 * (xp[i - 1] + xp[i]))
    pts = xp[m - 1:m + 2]
    return float(np.polyval(np.polyfit(pts, g(pts), 2), x))


@pytest.mark.parametrize("mode,deg", [("linear", 2), ("parabolic", 2), ("parabolic", 4),
                                      ("linear Lagrange_if_dg", 4)])
def test_face_values_across_dg_fv_faces(mode, deg):
    """set_ajsurf_vl on the shell with FV patches 0, 2, 4 (the reference's
    sw7FV-DG layout): a DG element takes its FV neighbour's face values by
    piecewise linear or parabolic interpolation from the uniform points
    (surface.c:1182-1245, fv_surface_interp = linear or parabolic), an FV
    element its DG neighbour's by Lagrange interpolation from the LGL
    points.  The neighbour's face plane is set to g(x1) + g(x2) of its own
    in-plane point coordinates with g(x) = x^deg, even so symmetric under
    the face orientations, so the expected values at my face point (a, b)
    are G(x_a) + G(x_b) of my own point coordinates: G = numpy's piecewise
    linear interp of g on the uniform points, or the parabola by cells
    through the three nearest uniform points (x^2 itself: the parabolas
    reproduce it), for a DG receiver, and g itself (Lagrange on 5 LGL points
    is exact to degree 4) for an FV receiver.  With Lagrange_if_dg the DG
    receiver takes Lagrange rows on the uniform points: x^4 itself."""
    n = 5
    om = OMesh("shell", 1, n, rin=2.0, rout=6.0, fv_patches=[0, 2, 4])
    xl = lgl(n)
    xu = np.linspace(-1.0, 1.0, n)

    def g(x):
        return x ** deg

    if mode == "linear Lagrange_if_dg":
        G = g   # Lagrange on the 5 uniform points is exact to degree 4
    elif mode == "linear":
        def G(x):
            return np.interp(x, xu, g(xu))
    else:
        def G(x):
            return np.array([_pw_parab_cells(t, xu, g) for t in np.ravel(x)]).reshape(np.shape(x))
        if deg == 2:
            assert np.abs(G(xl) - xl ** 2).max() < 1e-15
    checked = {0: 0, 1: 0}
    orc.set_fv_surface_interp(mode)
    try:
        for e in range(om.nelem):
            info = om.elm_info(e)
            fv_me = info[0] in (0, 2, 4)
            for f in range(6):
                nb, nbf = om.elm_fnb(e, f)
                if len(nb) != 1:
                    continue
                ninfo = om.elm_info(nb[0])
                fv_nb = ninfo[0] in (0, 2, 4)
                if fv_nb == fv_me:
                    continue
                assert info[10 + f] == 3           # interpolated, never copied
                # neighbour's nodes: g of its coordinates along its face's plane
                od = nbf[0] // 2
                od1, od2 = (1, 2) if od == 0 else ((0, 2) if od == 1 else (0, 1))
                xn = xu if fv_nb else xl
                u = np.zeros((5, om.nnodes))
                I = np.indices((n, n, n))[::-1]                 # I[0] = i, I[1] = j, I[2] = k
                blk = g(xn[I[od1]]) + g(xn[I[od2]])             # [k, j, i]
                off = ninfo[3]
                u[1, off:off + n ** 3] = blk.ravel()
                out = np.zeros(n * n)
                lib().orc_face_values(om.h, _p(np.ascontiguousarray(u.ravel())), e, f, 1, _p(out))
                xm = xu if fv_me else xl
                Gm = g if fv_me else G
                a = np.arange(n)
                expect = Gm(xm[a][None, :]) + Gm(xm[a][:, None])    # [b, a]
                assert np.abs(out.reshape(n, n) - expect).max() < 1e-13, (e, f, fv_me)
                checked[int(fv_me)] += 1
    finally:
        orc.set_fv_surface_interp("linear")
    assert checked[0] > 0 and checked[1] > 0


def test_pw_parab_basis_by_cells():
    """basis_pw_parab (interpolate.c:104-224) against the cell statement of
    the same interpolant: for every point count 3..9 and random x in
    [-1, 1] (and the points and midpoints themselves) the weights of the
    oracle's face rows applied to random data equal the parabola through the
    three points nearest x; they sum to one and reproduce quadratics"""
    rng = np.random.default_rng(7)
    for n in range(3, 10):
        xp = np.linspace(-1.0, 1.0, n)
        mids = 0.5 * (xp[1:] + xp[:-1])
        for x in np.concatenate([rng.uniform(-1.0, 1.0, 30), xp, mids]):
            w = np.array([orc.pw_parab(k, x, xp) for k in range(n)])
            assert abs(w.sum() - 1.0) < 1e-13
            assert abs(w @ (0.3 - xp + 2.0 * xp ** 2) - (0.3 - x + 2.0 * x * x)) < 1e-13
            vals = rng.standard_normal(n)
            ref = _pw_parab_cells(x, xp, lambda t: np.interp(t, xp, vals))
            assert abs(w @ vals - ref) < 1e-12, (n, x)
    # fewer than 3 points: the piecewise linear basis
    xp = np.array([-1.0, 1.0])
    assert orc.pw_parab(0, 0.5, xp) == _pwl(0, 0.5, xp)


def test_pw_linear_face_interpolation_from_fv_neighbours():
    """The DG side of a DG/FV face on the shell takes its neighbour's face
    values by piecewise linear interpolation from the uniform points: a
    field linear in the patch coordinates (lam, A, B) of the FV patch is
    reproduced exactly at the DG face points, so the DG elements' RHS next
    to FV elements equals the one of the all-DG mesh with the same field.
    Checked through the oracle RHS of a state that is linear in the FV
    patches' coordinates and zero-curvature across the face."""
    # 1D: linear data exact, partition of unity, numpy interp inside
    from oracle.orc import lib as L
    rng = np.random.default_rng(2)
    for n in range(2, 9):
        xp = np.linspace(-1.0, 1.0, n)
        for x in rng.uniform(-1.0, 1.0, 20):
            w = np.array([_pwl(k, x, xp) for k in range(n)])
            assert abs(w.sum() - 1.0) < 1e-14
            assert abs(w @ (0.3 + 2.0 * xp) - (0.3 + 2.0 * x)) < 1e-13
            vals = rng.standard_normal(n)
            assert abs(w @ vals - np.interp(x, xp, vals)) < 1e-13


def _pwl(k, x, xp):
    """basis_pw_linear restated (interpolate.c:53-102)"""
    np_ = len(xp)
    if np_ <= 1:
        return 1.0
    if x < xp[1]:
        return [(x - xp[1]) / (xp[0] - xp[1]), (x - xp[0]) / (xp[1] - xp[0])][k] if k < 2 else 0.0
    if x >= xp[np_ - 2]:
        n0, n1 = np_ - 2, np_ - 1
        if k == n0:
            return (x - xp[n1]) / (xp[n0] - xp[n1])
        if k == n1:
            return (x - xp[n0]) / (xp[n1] - xp[n0])
        return 0.0
    if k < np_ - 1 and xp[k] <= x < xp[k + 1]:
        return (x - xp[k + 1]) / (xp[k] - xp[k + 1])
    if k > 0 and xp[k - 1] <= x < xp[k]:
        return (x - xp[k - 1]) / (xp[k] - xp[k - 1])
    return 0.0


def test_amr_fv_if_uniform_and_nmax():
    """amr_fv_if_uniform = yes puts FV exactly in the uniform patches
    (amr_use_fv_if_P_UNIFORM, mesh.c:86-91), whatever amr_fv_p says;
    amr_n0 above amr_nmax and two radial shells are refused"""
    base = {"amr_mesh_type": "Shell", "amr_luni": 1, "amr_n0": 5, "amr_n1": 5, "amr_n2": 5,
            "amr_Shell_rin": 2.0, "amr_Shell_rout": 6.0, "amr_uniform_p": "0 2 4"}
    m = nm.Mesh(dict(base, amr_fv_if_uniform="yes"), device=-1)
    nm.amr_setup_mesh(m)
    m2 = nm.Mesh(dict(base, amr_fv_p="0 2 4"), device=-1)
    nm.amr_setup_mesh(m2)
    for e in range(m.nelem):
        assert np.array_equal(m.elm_info(e), m2.elm_info(e))
    assert m.dt == m2.dt
    with pytest.raises(nm.NmxError):   # uniform patches without FV are not built
        nm.amr_setup_mesh(nm.Mesh(dict(base), device=-1))
    with pytest.raises(nm.NmxError):
        nm.amr_setup_mesh(nm.Mesh(dict(base, amr_nmax=4), device=-1))
    with pytest.raises(nm.NmxError):
        nm.amr_setup_mesh(nm.Mesh(dict(base, amr_uniform_p="", amr_Shell_r1=3.0), device=-1))


def test_unbuilt_fv_options_are_refused():
    m = nm.Mesh({"amr_mesh_type": "BoxMesh"}, device=-1)
    m.set_par("fv_divf_extrap_s1", "0.5")
    m.set_par("fv_divf_extrap_s2", "DBL_MAX")
    m.set_par("fv_divf_extrap_opt", "1")
    for mode in ("WENOm3", "WENOm5_1", "WENOmZ", "1"):
        m.set_par("fv_rec", mode)
    m.set_par("fv_rec_WENOm_s1", "0.5")
    m.set_par("fv_surface_interp", "parabolic Lagrange_if_dg")
    for k, v in [("fv_surface_interp", "WENO"), ("fv_surface_interp", "Lagrange_if_dg"),
                 ("fv2dg_interp_use_extrap1", "yes"), ("fv_rec", "WENO3_2"),
                 ("coordinates_midpoint_data", "sometimes")]:
        with pytest.raises(nm.NmxError):
            m.set_par(k, v)
    for k, v in [("amr_dir_active1", "no"), ("coordinates_surface_metric", "sqrtgdiagdet"),
                 ("amr_Shell_nr", "2"), ("dg_outerBC_flux_fac", "1 0"), ("evolve_u_surfaces", "maybe")]:
        with pytest.raises(nm.NmxError):
            m.set_par(k, v)
    for k, v in [("dg_outerBC_flux_fac", "1 0 0"), ("evolve_redo_troubled", "step"),
                 ("evolve_u_surfaces", "yes"), ("amr_fv_if_uniform", "yes"), ("amr_nmax", "20")]:
        m.set_par(k, v)
    m.set_par("fv_surface_interp", "linear")
    m.set_par("fv_surface_interp", "parabolic")
    m.set_par("coordinates_midpoint_data", "yes all")
