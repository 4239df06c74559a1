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
  cubed-sphere shell.
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


@pytest.mark.parametrize("kind,luni,n", [("box", 1, 5), ("box", 2, 6), ("shell", 1, 6)])
def test_host_tables_of_fv_meshes_match_the_oracle(kind, luni, n):
    m = nm.Mesh({"amr_mesh_type": "BoxMesh" if kind == "box" else "Shell", "amr_luni": luni,
                 "amr_n0": n, "amr_n1": n, "amr_n2": n, "amr_BoxMesh_dout": 8.0,
                 "amr_Shell_rin": 2.0, "amr_Shell_rout": 20.0,
                 "amr_uniform_p": "0" if kind == "box" else "0 1 2 3 4 5",
                 "amr_fv_p": "0" if kind == "box" else "5 4 3 2 1 0"}, device=-1)
    nm.amr_setup_mesh(m)
    om = OMesh(kind, luni, n, dout=8.0, rin=2.0, rout=20.0, fv=True)
    X, x = om.coords()
    assert np.array_equal(m.coords(), x)
    assert m.dt == om.dt()
    for e in range(om.nelem):
        assert np.array_equal(m.elm_info(e), om.elm_info(e))
