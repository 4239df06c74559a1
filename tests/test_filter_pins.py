"""Pins of the oracle's exponential filter (orc_filter, after filter.c:20-92).

- Legendre products are eigenfunctions: on a box (constant Jacobian) the
  filter multiplies u = P_a(Xb) P_b(Yb) P_c(Zb) by f(a) f(b) f(c) with
  f(i) = exp(-alp (i/N)^s), N = n-1-dn.  A wrong mode index, a transposed
  analysis or a missing discrete norm of the top mode (c_N = 2/N on LGL
  nodes) fails this.
- alp = 0 is the identity; a constant field is kept for any alp.
- On the curvilinear shell, where J varies per node, the filter equals
  J^-jpow (F x F x F)(J^jpow u) with F = V diag(f) V^-1 built from the
  Legendre-Vandermonde matrix and an exact solve (numpy), not from the
  quadrature-based analysis the oracle follows; jpow = 1, 0 and -1.
"""
import numpy as np
import pytest
from numpy.polynomial import legendre as L

from oracle import orc
from oracle.orc import OMesh


def lgl(n):
    x = np.zeros(n)
    w = np.zeros(n)
    orc.lib().orc_lgl(n, x.ctypes.data, w.ctypes.data)
    return x


def fac(n, alp, s, dn):
    N = float(n - 1 - dn)
    if N <= 0:
        N = 1.0
    return np.exp(-alp * (np.arange(n) / N) ** s)


def elm_view(om, u, n):
    """(5, nelem, k, j, i) view of a state of a mesh with one order n"""
    return u.reshape(5, om.nelem, n, n, n)


@pytest.mark.parametrize("n,alp,s,dn", [(5, 36.0, 32.0, 0), (6, 4.0, 2.0, 0), (7, 2.0, 4.0, 1),
                                        (4, 1.0, 1.0, 2)])
def test_legendre_products_are_eigenfunctions(n, alp, s, dn):
    om = OMesh("box", 1, n, dout=2.0)
    x = lgl(n)
    f = fac(n, alp, s, dn)
    rng = np.random.default_rng(n)
    for _ in range(4):
        a, b, c = rng.integers(0, n, 3)
        Pa = L.legval(x, np.eye(n)[a])
        Pb = L.legval(x, np.eye(n)[b])
        Pc = L.legval(x, np.eye(n)[c])
        blk = Pc[:, None, None] * Pb[None, :, None] * Pa[None, None, :]
        u = np.zeros((5, om.nelem, n, n, n))
        u[:] = blk
        u[2] *= -2.5
        out = elm_view(om, om.filter(u.ravel(), 31, alp, s, dn, 1), n)
        expect = u * (f[a] * f[b] * f[c])
        assert np.abs(out - expect).max() <= 1e-13 * max(1.0, np.abs(u).max()), (a, b, c)


def test_zero_strength_is_identity_and_constants_are_kept():
    om = OMesh("box", 1, 6, dout=2.0)
    rng = np.random.default_rng(3)
    u = rng.standard_normal(5 * om.nnodes)
    assert np.abs(om.filter(u, 31, 0.0, 32.0, 0, 1) - u).max() < 1e-13 * np.abs(u).max()
    c = np.full(5 * om.nnodes, 1.7)
    assert np.abs(om.filter(c, 31, 36.0, 2.0, 0, 1) - c).max() < 1e-13


def test_varmask_selects_variables():
    om = OMesh("box", 1, 5, dout=2.0)
    rng = np.random.default_rng(4)
    u = rng.standard_normal(5 * om.nnodes)
    full = om.filter(u, 31, 3.0, 2.0, 0, 1).reshape(5, -1)
    part = om.filter(u, 0b10010, 3.0, 2.0, 0, 1).reshape(5, -1)
    u5 = u.reshape(5, -1)
    for v in range(5):
        ref = full[v] if v in (1, 4) else u5[v]
        assert np.array_equal(part[v], ref)


@pytest.mark.parametrize("jpow", [1, 0, -1])
def test_shell_filter_equals_dense_vandermonde_form(jpow):
    n, alp, s, dn = 5, 3.0, 4.0, 0
    om = OMesh("shell", 1, n, rin=2.0, rout=6.0)
    x = lgl(n)
    V = L.legvander(x, n - 1)                       # V[j, i] = P_i(x_j)
    F = V @ np.diag(fac(n, alp, s, dn)) @ np.linalg.inv(V)
    _, _, det = om.geometry()                       # det(dXb/dx) = 1/J
    J = (1.0 / det).reshape(om.nelem, n, n, n)
    rng = np.random.default_rng(5)
    u = rng.standard_normal(5 * om.nnodes)
    out = elm_view(om, om.filter(u, 31, alp, s, dn, jpow), n)
    uu = elm_view(om, u, n)
    w = uu * J[None] ** jpow
    w = np.einsum("il,vekjl->vekji", F, w)                           # along i (x)
    w = np.einsum("jl,vekli->vekji", F, w)                           # along j (y)
    w = np.einsum("kl,velji->vekji", F, w)                           # along k (z)
    expect = w * J[None] ** (-jpow)
    assert np.abs(out - expect).max() <= 1e-12 * np.abs(expect).max()
