"""Pins of the oracle's MRS limiter (orc_limit_MRS after limiter.c:56-196)
and of its place in the RK steps (RungeKutta.c:19-188, evolve.c:143-144).

- A closed form: on the 8-element box, element 0 carries pi = x + 1 (in
  [-1, 1], average 0) and everything else is 0.  With alpha h^1.5 = 0.5 the
  neighbours' range is widened to [-0.5, 0.5], so theta = phi(0.5 / 1) =
  0.5 / 1.1 and element 0 becomes 0 + theta (x + 1); with alpha = 0 it
  collapses to its average 0; the zero elements, whose range is a point
  (denominator 0), keep theta = 1.  A wrong 1.1, a wrong power of h or a
  missing widening fails this.  (A constant that the quadrature does not
  reproduce to the last bit gives a denominator of one rounding error, as
  in the reference, and theta 0 or 1 at random: the pin avoids it.)
- Every element keeps its LGL-quadrature average, for any data.
- After limiting, an element's values stay below the widened maximum of
  its neighbourhood and above the widened minimum.
- A smooth field with a generous alpha is left unchanged (to rounding).
- In the RK steps the limiter acts on every later stage's input and on the
  final u: one Euler step with the limiter equals the unlimited step
  followed by one limiter application.

minmodB (orc_limit_minmodB after limiter.c:200-380):
- a closed form: u = x has normalized-Legendre slope c100 = sqrt(2/3) s
  and neighbour mean differences sqrt(2) 2 s, so beta / sqrt(3) times them
  equals c100 for beta = 0.5 (unchanged) and halves it for beta = 0.25 (the
  element becomes its mean plus half its slope); a wrong sqrt(3), mode
  index or normalization fails this;
- element means are kept for any data; a generous bound changes nothing;
- its place in the steps, as for MRS.
"""
import numpy as np
import pytest

from oracle import orc
from oracle.orc import OMesh, lib


def lgl_w(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_lgl(n, x.ctypes.data, w.ctypes.data)
    return w


def averages(om, u, n):
    w = lgl_w(n)
    W = w[:, None, None] * w[None, :, None] * w[None, None, :]
    return (u.reshape(5, om.nelem, n, n, n) * W).sum(axis=(2, 3, 4)) / 8.0


def test_closed_form_two_values():
    n = 5
    om = OMesh("box", 1, n, dout=2.0)
    _, x = om.coords()
    x = x.reshape(3, -1)
    u = np.zeros((5, om.nnodes))
    off = int(om.elm_info(0)[3])
    sl = slice(off, off + n ** 3)
    u[0, sl] = x[0, sl] + 1.0
    alpha = 0.5 / 2.0 ** 1.5                       # element width 2: alpha h^1.5 = 0.5
    v = om.limit_MRS(u.ravel(), alpha).reshape(5, -1)
    theta = 0.5 / 1.1
    assert np.abs(v[0, sl] - theta * (x[0, sl] + 1.0)).max() < 1e-14
    rest = np.ones(om.nnodes, bool)
    rest[sl] = False
    assert np.array_equal(v[0, rest], u[0, rest]) and np.array_equal(v[1:], u[1:])
    v0 = om.limit_MRS(u.ravel(), 0.0).reshape(5, -1)
    assert np.abs(v0[0, sl]).max() < 1e-14


@pytest.mark.parametrize("alpha", [0.0, 0.3, 5.0])
def test_averages_kept_and_values_bounded(alpha):
    n = 5
    om = OMesh("shell", 1, n, rin=2.0, rout=6.0)
    u = np.random.default_rng(1).standard_normal(5 * om.nnodes)
    v = om.limit_MRS(u, alpha)
    a0, a1 = averages(om, u, n), averages(om, v, n)
    assert np.abs(a1 - a0).max() < 1e-13
    U = u.reshape(5, om.nelem, -1)
    V = v.reshape(5, om.nelem, -1)
    for e in range(om.nelem):
        nbs = set()
        for f in range(6):
            nbs.update(om.elm_fnb(e, f)[0].tolist())
        bb = om.elm_bbox(e)
        ah = alpha * max(bb[1] - bb[0], bb[3] - bb[2], bb[5] - bb[4]) ** 1.5
        for q in range(5):
            hi = max([U[q, j].max() for j in nbs] + [a0[q, e] + ah])
            lo = min([U[q, j].min() for j in nbs] + [a0[q, e] - ah])
            assert V[q, e].max() <= hi + 1e-12 and V[q, e].min() >= lo - 1e-12


def test_smooth_field_with_generous_alpha_is_unchanged():
    om = OMesh("box", 2, 6, dout=2.0)
    u, _ = om.profile([0.8, -0.5, 0.33166247903554], 0.3)
    v = om.limit_MRS(u, 50.0)
    # theta = 1: qbar + (q - qbar) gives q back up to one rounding
    assert np.abs(v - u).max() <= 4e-16 * np.abs(u).max()


def test_limiter_in_the_steps():
    om = OMesh("box", 1, 5, dout=2.0)
    rng = np.random.default_rng(2)
    u = rng.standard_normal(5 * om.nnodes)
    K = [0.8, -0.5, 0.33166247903554]
    dt = om.dt()
    plain, _ = om.evolve(u, dt, 1, 0.0, K, "Euler")
    orc.set_limiter_MRS(0.2)
    try:
        lim, _ = om.evolve(u, dt, 1, 0.0, K, "Euler")
    finally:
        orc.set_limiter_MRS(None)
    assert np.array_equal(lim, om.limit_MRS(plain, 0.2))
    assert not np.array_equal(lim, plain)


# ---- minmodB (limiter.c:200-380, arXiv:1506.06140) ----

def test_minmodB_linear_field_closed_form():
    """u = x on the 4x4x4 box: an element with neighbours on both x faces
    has c100 = sqrt(2/3) s and neighbour c000 differences sqrt(2) 2 s (s the
    slope per unit Xb), so beta / sqrt(3) times them is c100 exactly for
    beta = 0.5 (the TVD value: unchanged) and c100 / 2 for beta = 0.25 (the
    element becomes its mean plus half its slope); alpha = 0 makes the
    bound 0, so the minmod always decides"""
    n = 4
    om = OMesh("box", 2, n, dout=2.0)
    _, x = om.coords()
    x = x.reshape(3, -1)
    u = np.zeros((5, om.nnodes))
    u[0] = x[0]
    inner = []
    for e in range(om.nelem):
        bb = om.elm_bbox(e)
        if bb[0] > -2 + 1e-9 and bb[1] < 2 - 1e-9:
            inner.append(e)
    assert len(inner) == 32
    v = om.limit_minmodB(u.ravel(), 0.0, 0.5).reshape(5, -1)
    h = om.limit_minmodB(u.ravel(), 0.0, 0.25).reshape(5, -1)
    for e in inner:
        off = int(om.elm_info(e)[3])
        sl = slice(off, off + n ** 3)
        assert np.abs(v[0, sl] - u[0, sl]).max() < 1e-14
        mean = 0.5 * (om.elm_bbox(e)[0] + om.elm_bbox(e)[1])
        assert np.abs(h[0, sl] - (mean + 0.5 * (u[0, sl] - mean))).max() < 1e-14
    assert np.array_equal(v[1:], u[1:])


@pytest.mark.parametrize("alpha,beta,scale", [(0.0, 0.5, False), (0.5, 1.0, True), (0.2, 0.5, False)])
def test_minmodB_keeps_averages(alpha, beta, scale):
    n = 5
    om = OMesh("shell", 1, n, rin=2.0, rout=6.0)
    u = np.random.default_rng(4).standard_normal(5 * om.nnodes)
    v = om.limit_minmodB(u, alpha, beta, scale)
    assert np.abs(averages(om, v, n) - averages(om, u, n)).max() < 1e-13
    assert not np.array_equal(v, u)


def test_minmodB_generous_bound_leaves_the_state():
    om = OMesh("box", 2, 5, dout=2.0)
    u = np.random.default_rng(5).standard_normal(5 * om.nnodes)
    assert np.array_equal(om.limit_minmodB(u, 1e6, 0.5), u)


def test_minmodB_in_the_steps():
    om = OMesh("box", 1, 5, dout=2.0)
    u = np.random.default_rng(6).standard_normal(5 * om.nnodes)
    K = [0.8, -0.5, 0.33166247903554]
    dt = om.dt()
    plain, _ = om.evolve(u, dt, 1, 0.0, K, "Euler")
    orc.set_limiter_minmodB(0.1, 0.5, False)
    try:
        lim, _ = om.evolve(u, dt, 1, 0.0, K, "Euler")
    finally:
        orc.set_limiter_minmodB(None)
    assert np.array_equal(lim, om.limit_minmodB(plain, 0.1, 0.5))
    assert not np.array_equal(lim, plain)
