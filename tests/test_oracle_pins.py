"""Pins of the oracle (the CPU restatement of nmesh's DG path) against
closed forms, invariants and textbook routines.  Each check is chosen so that
a plausible slip (wrong weight, transposed D, dropped metric term, wrong
normal sign, wrong RK coefficient) fails it."""
import numpy as np
import numpy.polynomial.legendre as npleg
import pytest

from oracle.orc import OMesh, lib
import ctypes as C

K = np.array([0.8, -0.5, 0.33166247903554])


def _lgl(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_lgl(n, x.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p))
    return x, w


def _D(n):
    DT = np.zeros(n * n)
    lib().orc_dt_matrix(n, DT.ctypes.data_as(C.c_void_p))
    return DT.reshape(n, n)      # row i holds D_i. (DT[j + i n] = D_ij)


@pytest.mark.parametrize("n", range(2, 14))
def test_lgl_nodes_are_roots_of_dPN_and_ends(n):
    """LGL nodes: -1, +1 and the roots of P_N' (Legendre.c:149-213)."""
    x, w = _lgl(n)
    N = n - 1
    assert x[0] == -1.0 and x[-1] == 1.0
    if N >= 2:
        dP = npleg.legder([0] * N + [1])
        roots = np.sort(npleg.legroots(dP))
        assert np.allclose(x[1:-1], roots, rtol=0, atol=1e-14)
    assert np.allclose(x, -x[::-1], atol=1e-15)          # symmetric
    assert abs(w.sum() - 2.0) < 1e-14
    assert abs(w[0] - 2.0 / (N * (N + 1))) < 1e-15        # end weight


@pytest.mark.parametrize("n", range(2, 12))
def test_lgl_quadrature_exact_to_degree_2N_minus_1(n):
    x, w = _lgl(n)
    for deg in range(0, 2 * (n - 1)):
        exact = 0.0 if deg % 2 else 2.0 / (deg + 1)
        assert abs((w * x ** deg).sum() - exact) < 1e-13


@pytest.mark.parametrize("n", range(2, 12))
def test_differentiation_matrix_exact_on_polynomials(n):
    """D x^k = k x^(k-1) for k <= N; rows sum to zero; D_00 = -N(N+1)/4."""
    x, _ = _lgl(n)
    D = _D(n)
    for kk in range(n):
        assert np.allclose(D @ x ** kk, kk * x ** max(kk - 1, 0) * (kk > 0), atol=1e-11)
    assert np.allclose(D.sum(1), 0, atol=1e-12)
    N = n - 1
    if N >= 1:
        assert abs(D[0, 0] + N * (N + 1) / 4.0) < 1e-11


def test_box_mesh_config1_dt_and_counts():
    """BoxMesh [-8,8]^3, luni 2, n=5: 64 elements, 8000 nodes, dt 0.172673."""
    m = OMesh("box", 2, 5, dout=8.0)
    assert m.nelem == 64 and m.nnodes == 8000
    x, _ = _lgl(5)
    hmin = 2.0 * (1.0 - x[3])      # element width 4: spacing 2*(1 - x_{N-1})
    assert abs(m.dt() - 0.25 * hmin) < 1e-15
    assert abs(m.dt() - 0.172673) < 1e-6


def test_box_elements_in_morton_order():
    """Children replace their parent in place; child c takes the upper half
    in direction d when bit d of c is set (connections.c:1175-1206)."""
    m = OMesh("box", 2, 3, dout=8.0)
    for e in range(64):
        c1, c2 = e // 8, e % 8
        lo = np.array([-8.0] * 3)
        for c, L in ((c1, 8.0), (c2, 4.0)):
            lo = lo + L * np.array([(c >> d) & 1 for d in range(3)])
        bb = m.elm_bbox(e)
        assert np.allclose(bb[0::2], lo) and np.allclose(bb[1::2], lo + 4.0)


@pytest.mark.parametrize("kind,luni,n", [("box", 1, 5), ("shell", 1, 5), ("shell", 2, 4)])
def test_face_neighbours_share_their_points(kind, luni, n):
    """Every conforming face copy (alias or oriented) pairs points that
    coincide physically, and every face without neighbour lies on a patch
    face marked as outer boundary."""
    m = OMesh(kind, luni, n, dout=8.0, rin=2.0, rout=20.0)
    _, x = m.coords()
    for e in range(m.nelem):
        info = m.elm_info(e)
        off = info[3]
        for f in range(6):
            nb, nbf = m.elm_fnb(e, f)
            if len(nb) == 0:
                assert m.pat_bound(info[0], f) == 1
                continue
            assert info[10 + f] in (1, 2)
            d = f // 2
            # my face points: physical positions must appear on the neighbour's face
            ninfo = m.elm_info(nb[0])
            noff = ninfo[3]
            nd = nbf[0] // 2
            idx = np.indices((n, n, n)).reshape(3, -1)[::-1]   # i, j, k  (ijk = i + n(j + n k))
            mine = np.where(idx[d] == (n - 1) * (f % 2))[0]
            theirs = np.where(idx[nd] == (n - 1) * (nbf[0] % 2))[0]
            P = x[:, off + mine].T
            Q = x[:, noff + theirs].T
            dist = np.sqrt(((P[:, None, :] - Q[None, :, :]) ** 2).sum(-1)).min(1)
            assert dist.max() < 1e-12 * max(1.0, np.abs(P).max())


def test_shell_geometry_against_finite_differences():
    """dXdx of the cubed-shell map against central differences of the
    inverse map X(x) (coordtrans_CubedSphere.c:259-509)."""
    m = OMesh("shell", 1, 5, rin=2.0, rout=20.0)
    X, x = m.coords()
    g, sg, det = m.geometry()
    rng = np.random.default_rng(1)
    for node in rng.choice(m.nnodes, 40, replace=False):
        e = np.searchsorted([m.elm_info(q)[3] for q in range(m.nelem)], node, side="right") - 1
        p = m.elm_info(e)[0]
        dom = p
        dirn = dom // 2
        ia = 1 if dirn in (0, 2) else 0
        ib = 0 if dirn == 2 else 2
        xx = x[:, node]
        pm = 1.0 if dom % 2 else -1.0

        def Xof(v):
            r = v[dirn]
            A, B = v[ia] / r, v[ib] / r
            s = np.sqrt(1 + A * A + B * B)
            a0, a1 = pm * 2.0 / s, pm * 20.0 / s
            return np.array([(r - a0) / (a1 - a0), A, B])
        assert np.allclose(Xof(xx), X[:, node], atol=1e-12)
        h = 1e-6
        J = np.zeros((3, 3))
        for b in range(3):
            dv = np.zeros(3)
            dv[b] = h
            J[:, b] = (Xof(xx + dv) - Xof(xx - dv)) / (2 * h)
        assert np.allclose(g[:, node].reshape(3, 3), J, rtol=1e-6, atol=1e-8)


@pytest.mark.parametrize("numflux", ["upwind", "LLF"])
def test_rhs_plane_wave_spectral_convergence(numflux):
    """scalarwave1's own known answer (scalarwave1_analyze, scalarwave1.c:710-751):
    the RHS of the exact plane wave tends to its time derivative, spectrally
    in n on a box mesh."""
    errs = []
    for n in (4, 6, 8, 10, 12):
        m = OMesh("box", 1, n, dout=3.0)
        u, du = m.profile(K, 0.3)
        r = m.rhs(u, 0.3, K, numflux)
        errs.append(np.abs(r - du).max())
    errs = np.array(errs)
    assert np.all(errs[1:] < errs[:-1] * 0.05)
    assert errs[-1] < 1e-8


def test_rhs_plane_wave_shell_converges():
    errs = []
    for n in (4, 6, 8, 10):
        m = OMesh("shell", 1, n, rin=2.0, rout=6.0)
        u, du = m.profile(K, 0.3)
        errs.append(np.abs(m.rhs(u, 0.3, K) - du).max())
    assert np.all(np.diff(np.log(errs)) < -1.5)


def test_rhs_is_linear_and_zero_for_zero_state():
    m = OMesh("shell", 1, 4, rin=2.0, rout=20.0)
    rng = np.random.default_rng(2)
    k0 = [0.0, 0.0, 0.0]
    u = rng.standard_normal(5 * m.nnodes)
    v = rng.standard_normal(5 * m.nnodes)
    assert np.abs(m.rhs(np.zeros_like(u), 0.0, k0)).max() == 0.0
    lhs = m.rhs(2.0 * u - 3.0 * v, 0.0, k0)
    rhs = 2.0 * m.rhs(u, 0.0, k0) - 3.0 * m.rhs(v, 0.0, k0)
    assert np.abs(lhs - rhs).max() < 1e-11 * np.abs(lhs).max()


def _energy(m, u):
    """discrete energy sum w J (pi^2 + |c|^2) over all nodes (box, J const)"""
    _, x = m.coords()
    u = u.reshape(5, -1)
    return float((u[0] ** 2 + u[1] ** 2 + u[2] ** 2 + u[3] ** 2).sum())


def test_upwind_flux_closed_form_equals_eigen_form():
    """The R-matrix upwind flux of scalarwave1 equals A+ ui + A- ua with
    A+- = (A +- |A|)/2, A = [[0,-n^T],[-n,0]]: checked through the RHS of a
    box mesh (normals +-e_d) and a shell mesh (generic normals) by comparing
    upwind and LLF, which differ only in the tangential c jump."""
    rng = np.random.default_rng(3)
    for kind in ("box", "shell"):
        m = OMesh(kind, 1, 4, dout=2.0, rin=2.0, rout=6.0)
        u = rng.standard_normal(5 * m.nnodes)
        k0 = [0.0, 0.0, 0.0]
        ru = m.rhs(u, 0.0, k0, "upwind")
        rl = m.rhs(u, 0.0, k0, "LLF")
        d = (ru - rl).reshape(5, -1)
        # pi and phi rows agree (F_pi identical for both fluxes)
        assert np.abs(d[0]).max() < 1e-10 * np.abs(ru).max()
        assert np.abs(d[4]).max() == 0.0
        assert np.abs(d[1:4]).max() > 1e-3


def _time_order(method):
    """Self-convergence in dt on a fixed mesh: the spatial error is common to
    all runs, so differences between successive dt isolate the time error;
    their ratio is 2^p for a method of order p (evolve_test, evolve_test.c:100-136)."""
    m = OMesh("box", 1, 8, dout=3.0)
    dt0 = m.dt()
    T = 8 * dt0
    sols = []
    for nsteps in (32, 64, 128):
        u0, _ = m.profile(K, 0.0)
        u, t = m.evolve(u0, T / nsteps, nsteps, 0.0, K, method)
        sols.append(u)
    d1 = np.abs(sols[0] - sols[1]).max()
    d2 = np.abs(sols[1] - sols[2]).max()
    return d1 / d2


def test_rk4_fourth_order_in_time():
    assert 14.0 < _time_order("RK4") < 19.0


def test_ssprk3_third_order_in_time():
    assert 7.0 < _time_order("sspRK3") < 10.0


def test_euler_first_order_in_time():
    assert 1.8 < _time_order("Euler") < 2.3


def test_upwind_energy_decays_with_outgoing_boundary():
    """With a zero exterior state the upwind scheme dissipates energy."""
    m = OMesh("box", 2, 5, dout=8.0)
    _, x = m.coords()
    r2 = (x ** 2).sum(0)
    u = np.zeros((5, m.nnodes))
    u[4] = np.exp(-r2)
    u[1:4] = -2 * x * np.exp(-r2)
    u = u.ravel()
    e0 = _energy(m, u)
    u1, _ = m.evolve(u, m.dt(), 40, 0.0, [0.0, 0.0, 0.0], "RK4")
    assert _energy(m, u1) < e0
