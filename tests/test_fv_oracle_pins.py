"""Pins of the FV oracle (oracle/fv_oracle.c): WENO5 reconstruction as
rec1d.c:1190-1352 defines it, the LLF flux of numflux1d.c:85-112, and the
special-relativistic Euler physics (absent from the reference, pinned by
physics here).

Each check is chosen so a plausible slip fails it: a wrong stencil
coefficient or index breaks exactness on quadratics and the 5th order; a
swapped ideal weight breaks the match with the linear 5th-order combination;
a wrong eigenvalue breaks the Jacobian comparison; a wrong flux sign, face
index or direction breaks the RHS convergence to the exact flux difference.
"""
import numpy as np
import pytest

from nmesh_amd.fv import blob_initial_data
from oracle import orc

GAM = 2.0


def cell_avg_quadratic(c, x0):
    """cell averages of c0 + c1 x + c2 x^2 over [x0, x0 + 1]"""
    a, b = x0, x0 + 1.0
    return c[0] + c[1] * (a + b) / 2 + c[2] * (a * a + a * b + b * b) / 3


def test_weno5_exact_on_cell_averages_of_quadratics():
    rng = np.random.default_rng(1)
    for _ in range(200):
        c = rng.uniform(-2, 2, 3)
        x = np.arange(-3, 4, dtype=float) + rng.uniform(-5, 5)
        u = cell_avg_quadratic(c, x)
        # face between cells 3 and 4 of u (x[3] + 1)
        xf = x[3] + 1.0
        exact = c[0] + c[1] * xf + c[2] * xf * xf
        assert abs(orc.rec_p(u, 3) - exact) <= 1e-12 * (1 + abs(exact))
        assert abs(orc.rec_m(u, 3) - exact) <= 1e-12 * (1 + abs(exact))


def _sine_errors(N, rec):
    k = 2 * np.pi / N
    x0 = np.arange(-3, N + 3)
    avg = (np.cos(k * x0) - np.cos(k * (x0 + 1))) / k
    return max(abs(rec(avg, i + 3) - np.sin(k * (i + 1))) for i in range(N))


@pytest.mark.parametrize("rec", [orc.rec_p, orc.rec_m])
def test_weno5_fifth_order_and_ideal_weights_on_smooth_data(rec):
    e64, e128 = _sine_errors(64, rec), _sine_errors(128, rec)
    assert e64 / e128 > 25.0          # 2^5 = 32
    # smooth data: the nonlinear weights tend to the ideal (1, 6, 3) / 10, so
    # the result tends to the linear 5th-order combination
    N = 256
    k = 2 * np.pi / N
    x0 = np.arange(-3, N + 3)
    avg = (np.cos(k * x0) - np.cos(k * (x0 + 1))) / k
    for i in range(0, N, 17):
        s = avg[i + 1:i + 7]
        lin_p = (2 * s[0] - 13 * s[1] + 47 * s[2] + 27 * s[3] - 3 * s[4]) / 60
        lin_m = (-3 * s[1] + 27 * s[2] + 47 * s[3] - 13 * s[4] + 2 * s[5]) / 60
        lin = lin_p if rec is orc.rec_p else lin_m
        assert abs(rec(avg, i + 3) - lin) < 5e-8


def test_weno5_is_essentially_non_oscillatory_at_a_step():
    # the smooth stencil alone is used: beta = 0 gets weight ~ 1/eps^2
    assert abs(orc.rec_p(np.array([0., 0., 0., 1., 1., 1.]), 2)) < 1e-60
    assert abs(orc.rec_m(np.array([0., 0., 0., 1., 1., 1.]), 2) - 1.0) < 1e-15
    # constant data: exactly the constant (all weights ideal)
    assert orc.rec_p(np.full(6, 0.7), 2) == pytest.approx(0.7, abs=1e-16)


@pytest.mark.parametrize("rec", [orc.rec_pz, orc.rec_mz])
def test_wenoz_exact_on_quadratics_fifth_order_and_eno(rec):
    """WENO-Z (rec1d.c:1627-1781): the same stencils as WENO5 with weights
    g_k (1 + tau5 / (beta_k + eps)), tau5 = |beta0 - beta2|."""
    rng = np.random.default_rng(6)
    for _ in range(100):
        c = rng.uniform(-2, 2, 3)
        x = np.arange(-3, 4, dtype=float) + rng.uniform(-5, 5)
        u = cell_avg_quadratic(c, x)
        xf = x[3] + 1.0
        exact = c[0] + c[1] * xf + c[2] * xf * xf
        assert abs(rec(u, 3) - exact) <= 1e-12 * (1 + abs(exact))
    e64, e128 = _sine_errors(64, rec), _sine_errors(128, rec)
    assert e64 / e128 > 25.0
    step = np.array([0., 0., 0., 1., 1., 1.])
    if rec is orc.rec_pz:
        assert abs(rec(step, 2)) < 1e-30
    else:
        assert abs(rec(step, 2) - 1.0) < 1e-15


def test_wenoz_differs_from_weno5_only_in_the_weights():
    # with beta0 = beta2 (tau5 = 0) the Z weights are the ideal ones exactly
    u = np.array([1.0, 2.0, 2.5, 2.0, 1.0, 0.3])     # symmetric about index 2
    lin = (2 * u[0] - 13 * u[1] + 47 * u[2] + 27 * u[3] - 3 * u[4]) / 60
    assert abs(orc.rec_pz(u, 2) - lin) < 1e-14
    # while WENO5-JS on the same data is not at the ideal weights
    assert abs(orc.rec_p(u, 2) - lin) > 1e-6


def test_weno5_m_mirrors_p():
    rng = np.random.default_rng(2)
    for _ in range(100):
        u = rng.standard_normal(6)
        assert abs(orc.rec_m(u, 1) - orc.rec_p(u[::-1].copy(), 3)) <= 1e-14 * (1 + np.abs(u).max())


def random_prims(rng, n, vmax=0.98):
    out = []
    for _ in range(n):
        rho = 10 ** rng.uniform(-1, 1)
        p = 10 ** rng.uniform(-2, 1)
        d = rng.standard_normal(3)
        v = d / np.linalg.norm(d) * rng.uniform(0, vmax)
        out.append(np.array([rho, v[0], v[1], v[2], p]))
    return out


def test_c2p_inverts_p2c():
    rng = np.random.default_rng(3)
    for P in random_prims(rng, 300):
        U = orc.sr_p2c(GAM, P)
        for guess in (P[4], 10 * P[4], 0.01 * P[4]):
            Q, it = orc.sr_c2p(GAM, U, guess)
            assert it > 0
            assert np.abs(Q - P).max() <= 1e-9 * np.abs(P).max(), (P, Q)


def test_eigenvalues_are_those_of_the_flux_jacobian():
    rng = np.random.default_rng(4)
    for P in random_prims(rng, 40, vmax=0.9):
        U0 = orc.sr_p2c(GAM, P)
        for d in range(3):
            F0, _, lam = orc.sr_flux(GAM, d, P)
            J = np.zeros((5, 5))
            for j in range(5):
                h = 1e-6 * max(abs(U0[j]), 1e-3)
                Fs = []
                for sgn in (1, -1):
                    U = U0.copy()
                    U[j] += sgn * h
                    Q, _ = orc.sr_c2p(GAM, U, P[4])
                    Fs.append(orc.sr_flux(GAM, d, Q)[0])
                J[:, j] = (Fs[0] - Fs[1]) / (2 * h)
            ev = np.sort(np.linalg.eigvals(J).real)
            assert abs(ev[0] - lam[0]) < 1e-5 and abs(ev[-1] - lam[1]) < 1e-5, (ev, lam)
            # the three middle eigenvalues are v_d (entropy and shear waves)
            assert np.abs(ev[1:4] - P[1 + d]).max() < 1e-5


def test_llf_is_consistent_and_fluxes_follow_the_definitions():
    rng = np.random.default_rng(5)
    for P in random_prims(rng, 50):
        for d in range(3):
            F, U, lam = orc.sr_flux(GAM, d, P)
            assert np.array_equal(orc.sr_llf(GAM, d, P, P), F)
            vd = P[1 + d]
            assert F[0] == pytest.approx(U[0] * vd)
            assert F[4] == pytest.approx(U[1 + d] - U[0] * vd, rel=1e-12, abs=1e-12)
    # at rest the acoustic speeds are +-c_s
    P = np.array([1.0, 0, 0, 0, 1.0])
    cs = np.sqrt(GAM * 1.0 / (1.0 * (1 + GAM / (GAM - 1))))
    assert np.allclose(orc.sr_flux(GAM, 0, P)[2], [-cs, cs], rtol=0, atol=1e-15)


def test_constant_state_has_zero_rhs():
    N = 16
    fv = orc.OFV(N)
    P = np.empty((5, N ** 3))
    P[:] = np.array([1.3, 0.2, -0.1, 0.05, 0.7])[:, None]
    assert np.abs(fv.rhs_prim(P)).max() == 0.0


def _advected_rhs_error(N, d, amp=0.2, v=0.3):
    """RHS of an exactly averaged profile varying along d (uniform v_d, p)
    against the exact flux difference -(F(x_{i+1}) - F(x_i))."""
    k = 2 * np.pi / N
    x = np.arange(N)
    avg = 1 + amp * (np.cos(k * x) - np.cos(k * (x + 1))) / k
    shape = [1, 1, 1]
    shape[2 - d] = N
    P = np.zeros((5, N, N, N))
    P[0] = avg.reshape(shape)
    P[1 + d] = v
    P[4] = 1.0
    fv = orc.OFV(N)
    r = fv.rhs_prim(P.reshape(5, -1)).reshape(5, N, N, N)

    def Fx(xf):
        Q = np.array([1 + amp * np.sin(k * xf), 0, 0, 0, 1.0])
        Q[1 + d] = v
        return orc.sr_flux(GAM, d, Q)[0]

    exact = np.array([-(Fx(i + 1.0) - Fx(float(i))) for i in range(N)])   # [i][var]
    idx = [0, 0, 0]
    err = 0.0
    for i in range(N):
        idx[2 - d] = i
        err = max(err, np.abs(r[(slice(None),) + tuple(idx)] - exact[i]).max())
    return err


@pytest.mark.parametrize("d", [0, 1, 2])
def test_rhs_converges_at_fifth_order_to_the_exact_flux_difference(d):
    e32, e64 = _advected_rhs_error(32, d), _advected_rhs_error(64, d)
    assert e64 < 1e-5
    assert e32 / e64 > 20.0


def test_rhs_conserves_every_component():
    N = 16
    P = blob_initial_data(N, seed=7, nblobs=4)
    r = orc.OFV(N).rhs_prim(P).reshape(5, -1)
    assert np.abs(r.sum(1)).max() < 1e-11 * np.abs(r).max() * N ** 3


def test_blob_initial_data_has_the_stated_ranges():
    P = blob_initial_data(32, seed=2026)
    assert P[0].min() >= 1.0 and P[0].max() < 1.0 + 32
    assert np.abs(P[1:4]).max() <= 0.3 and np.all(P[4] == 1.0)
