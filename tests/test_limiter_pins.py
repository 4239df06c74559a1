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
