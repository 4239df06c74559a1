"""Pins of the oracle's numerical flux, lift and dt against closed forms
that do not go through the oracle's own formulas.

- The upwind and LLF fluxes, all five rows, against A+ ui + A- ua and
  (f(ui) + f(ua) - (ua - ui)) / 2 with A = d(f.n)/du built in numpy
  (scalarwave1.c:39-48,324-499; numflux1d.c:85-112).
- Discrete conservation on box meshes with non-zero jumps: LGL quadrature
  is summation by parts, so sum_e sum_ijk W J r equals minus the outer
  boundary quadrature of fnum.  A lift of 2/w_end, a wrong sqrtgdiag or a
  wrong normal sign breaks it.
- The discrete energy rate u^T (W J) RHS(u) on the box against the face sum
  of pi c.n - pi F_pi - c.F_c with F = fnum - f.n taken from numpy's closed
  form, both sides of every face found by coordinates.
- A single-element impulse: every other element's RHS is exactly the face
  term -(fnum - f.n) sqrtgdiag / w_end (dg.c:142-143; coordinates.c:796-798)
  at the nodes it shares with the impulse and zero elsewhere, on a box and on
  the cubed-sphere shell (cross-patch orientations included).  The pairing is
  found from coordinates, the normal and sqrtgdiag from the oracle's dX/dx
  (itself pinned by finite differences of the map).
- dt = dtfac * min hmin with hmin the smallest axial distance between
  neighbouring points at the 8 element corners, or the bounding-box diagonal
  over sqrt(3) (find_hmin, utilities.c:1028-1128), by brute force over
  corners with coordinates from a numpy statement of xyz_of_lamAB_CubSph
  (coordtrans_CubedSphere.c:38-120).
"""
import ctypes as C

import numpy as np
import numpy.polynomial.legendre as npleg
import pytest

from oracle.orc import OMesh, lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def lgl(n):
    """LGL nodes and weights from numpy's Legendre module"""
    N = n - 1
    x = np.concatenate(([-1.0], np.sort(npleg.legroots(npleg.legder([0] * N + [1]))), [1.0]))
    PN = npleg.legval(x, [0] * N + [1])
    return x, 2.0 / (N * (N + 1) * PN ** 2)


def flux_jacobian(nrm):
    """A = d(f.n)/du for u = (pi, cx, cy, cz, phi), f.n = (-c.n, -pi n, 0)"""
    A = np.zeros((5, 5))
    A[0, 1:4] = -nrm
    A[1:4, 0] = -nrm
    return A


def fnum_closed(numflux, nrm, ui, ua):
    A = flux_jacobian(nrm)
    if numflux == "upwind":
        w, V = np.linalg.eigh(A)
        Ap = V @ np.diag(np.maximum(w, 0)) @ V.T
        Am = V @ np.diag(np.minimum(w, 0)) @ V.T
        f = Ap @ ui + Am @ ua
        f[4] = 0.0
        return f
    return 0.5 * (A @ ui + A @ ua - (ua - ui))


def orc_fnum(numflux, nrm, ui, ua):
    out = np.zeros(5)
    lib().orc_numflux({"upwind": 0, "LLF": 1}[numflux], _p(np.ascontiguousarray(nrm, float)),
                      _p(np.ascontiguousarray(ui, float)), _p(np.ascontiguousarray(ua, float)), _p(out))
    return out


def _normals(rng):
    """unit normals reaching each of the three R branches of the upwind flux
    (nz != 0; nz = 0, nx != 0; nx = nz = 0) and the axis normals"""
    out = []
    for _ in range(20):
        out.append(rng.standard_normal(3))
        v = rng.standard_normal(3)
        v[2] = 0.0
        out.append(v)
    out += [np.array(e, float) for e in ([1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1],
                                          [0, 0, -1], [0, 2, 0], [0, -3, 0])]
    return [v / np.linalg.norm(v) for v in out]


@pytest.mark.parametrize("numflux", ["upwind", "LLF"])
def test_numflux_all_rows_closed_form(numflux):
    rng = np.random.default_rng(11)
    for nrm in _normals(rng):
        ui, ua = rng.standard_normal(5), rng.standard_normal(5)
        f = orc_fnum(numflux, nrm, ui, ua)
        g = fnum_closed(numflux, nrm, ui, ua)
        assert np.abs(f - g).max() < 1e-14 * max(1.0, np.abs(g).max()), (nrm, f, g)
        # consistency: f*(u, u) = f(u).n on the four flux rows
        fu = orc_fnum(numflux, nrm, ui, ui)
        assert np.abs(fu[:4] - (flux_jacobian(nrm) @ ui)[:4]).max() < 1e-14


def _elements(m):
    """per element: (offset, n, bbox)"""
    out = []
    for e in range(m.nelem):
        info = m.elm_info(e)
        out.append((int(info[3]), int(info[2]), m.elm_bbox(e)))
    return out


def _face_nodes(n, f):
    """ijk of the nodes of face f, in (a, b) order"""
    d = f // 2
    idx = np.indices((n, n, n)).reshape(3, -1)[::-1]      # i, j, k of ijk
    return np.where(idx[d] == (n - 1) * (f % 2))[0]


@pytest.mark.parametrize("numflux", ["upwind", "LLF"])
@pytest.mark.parametrize("luni,n", [(1, 5), (2, 4)])
def test_box_discrete_conservation(numflux, luni, n):
    dout = 3.0
    m = OMesh("box", luni, n, dout=dout)
    rng = np.random.default_rng(5 + luni)
    u = rng.standard_normal(5 * m.nnodes)
    r = m.rhs(u, 0.0, [0.0, 0.0, 0.0], numflux).reshape(5, -1)
    U = u.reshape(5, -1)
    x, w = lgl(n)
    W3 = (w[None, None, :] * w[None, :, None] * w[:, None, None]).ravel()   # ijk = i + n(j + n k)
    vol = np.zeros(4)
    bnd = np.zeros(4)
    for off, nn, bb in _elements(m):
        h = bb[1::2] - bb[0::2]
        J = np.prod(h / 2)
        vol += (r[:4, off:off + n ** 3] * (W3 * J)).sum(1)
        for f in range(6):
            d = f // 2
            if abs(abs(bb[f]) - dout) > 1e-12:
                continue
            nrm = np.zeros(3)
            nrm[d] = 1.0 if f % 2 else -1.0
            nodes = _face_nodes(n, f)
            wa = np.outer(w, w).ravel()        # (a, b) quadrature on the face
            for q, node in enumerate(nodes):
                fnum = fnum_closed(numflux, nrm, U[:, off + node], np.zeros(5))
                bnd += -fnum[:4] * wa[q] * J * 2.0 / h[d]
    assert np.abs(vol - bnd).max() < 1e-12 * np.abs(bnd).max(), (vol, bnd)


def _pairing(x, elems, e, f, tol=1e-9):
    """for each node of face f of element e, the global index of the
    coinciding node of the one other element that holds all of the face's
    points (its face neighbour on a conforming mesh), or -1 without one"""
    off, n, _ = elems[e]
    mine = x[:, off + _face_nodes(n, f)].T
    scale = tol * max(1.0, np.abs(mine).max())
    for q, (o, nn, _) in enumerate(elems):
        if q == e:
            continue
        P = x[:, o:o + nn ** 3].T
        d = np.abs(mine[:, None, :] - P[None, :, :]).max(2)
        k = d.argmin(1)
        if (d[np.arange(len(mine)), k] < scale).all():
            return o + k
    return -np.ones(len(mine), int)


@pytest.mark.parametrize("numflux", ["upwind", "LLF"])
def test_box_energy_rate_equals_face_sum(numflux):
    n, dout = 4, 3.0
    m = OMesh("box", 1, n, dout=dout)
    _, x = m.coords()
    elems = _elements(m)
    rng = np.random.default_rng(21)
    u = rng.standard_normal(5 * m.nnodes)
    U = u.reshape(5, -1)
    r = m.rhs(u, 0.0, [0.0, 0.0, 0.0], numflux).reshape(5, -1)
    _, w = lgl(n)
    W3 = (w[None, None, :] * w[None, :, None] * w[:, None, None]).ravel()
    rate = 0.0
    face = 0.0
    wa = np.outer(w, w).ravel()
    for e, (off, nn, bb) in enumerate(elems):
        h = bb[1::2] - bb[0::2]
        J = np.prod(h / 2)
        sl = slice(off, off + n ** 3)
        rate += ((U[:4, sl] * r[:4, sl]).sum(0) * W3 * J).sum()
        for f in range(6):
            d = f // 2
            nrm = np.zeros(3)
            nrm[d] = 1.0 if f % 2 else -1.0
            pair = _pairing(x, elems, e, f)
            for q, node in enumerate(_face_nodes(n, f)):
                ui = U[:, off + node]
                ua = U[:, pair[q]] if pair[q] >= 0 else np.zeros(5)
                F = fnum_closed(numflux, nrm, ui, ua) - flux_jacobian(nrm) @ ui
                face += wa[q] * J * 2.0 / h[d] * (ui[0] * (ui[1:4] @ nrm) - ui[0] * F[0] - ui[1:4] @ F[1:4])
    assert abs(rate - face) < 1e-12 * abs(face), (rate, face)
    if numflux == "upwind":
        assert rate < 0.0


@pytest.mark.parametrize("numflux", ["upwind", "LLF"])
@pytest.mark.parametrize("kind", ["box", "shell"])
def test_single_element_impulse_face_terms(kind, numflux):
    n = 5
    if kind == "box":
        m = OMesh("box", 2, n, dout=3.0)
        e0, nface = 21, 6             # interior element: six face neighbours
    else:
        m = OMesh("shell", 1, n, rin=2.0, rout=6.0)
        e0, nface = 9, 5              # patch 1, outer layer: three siblings, two other patches
    _, x = m.coords()
    g, _, _ = m.geometry()
    elems = _elements(m)
    _, w = lgl(n)
    oo_wend = 1.0 / w[0]
    rng = np.random.default_rng(31)
    u = np.zeros((5, m.nnodes))
    o0 = elems[e0][0]
    u[:, o0:o0 + n ** 3] = rng.standard_normal((5, n ** 3))
    r = m.rhs(u.ravel(), 0.0, [0.0, 0.0, 0.0], numflux).reshape(5, -1)
    expect = np.zeros_like(r)
    touched = 0
    for e, (off, nn, bb) in enumerate(elems):
        if e == e0:
            continue
        h = bb[1::2] - bb[0::2]
        for f in range(6):
            d = f // 2
            pair = _pairing(x, elems, e, f)
            nodes = _face_nodes(n, f)
            src = [p for p in pair if o0 <= p < o0 + n ** 3]
            if len(src) != len(nodes):
                continue            # not the face shared with the impulse
            touched += 1
            for q, node in enumerate(nodes):
                gi = off + node
                row = g[3 * d:3 * d + 3, gi]
                s = np.linalg.norm(row)
                nrm = (1.0 if f % 2 else -1.0) * row / s
                sqrtgdiag = (2.0 / h[d]) * s
                ua = u[:, pair[q]]
                F = fnum_closed(numflux, nrm, np.zeros(5), ua)       # f(0).n = 0
                expect[:4, gi] += -F[:4] * sqrtgdiag * oo_wend
    assert touched == nface
    mask = np.ones(m.nnodes, bool)
    mask[o0:o0 + n ** 3] = False
    err = np.abs(r[:, mask] - expect[:, mask]).max()
    assert err < 1e-13 * np.abs(expect).max(), err
    assert np.abs(expect).max() > 0.1


def shell_xyz(dom, s0, s1, lam, A, B):
    """xyz_of_lamAB_CubSph for a CubedShell centred at 0 with constant radii
    (coordtrans_CubedSphere.c:38-120): dir = dom/2, pm = 2(dom%2) - 1"""
    d, pm = dom // 2, 2 * (dom % 2) - 1
    q = np.sqrt(1.0 + A * A + B * B)
    a0, a1 = pm * s0 / q, pm * s1 / q
    r = (a1 - a0) * lam + a0
    if d == 0:
        return np.array([r, A * r, B * r])
    if d == 1:
        return np.array([A * r, r, B * r])
    return np.array([B * r, A * r, r])


@pytest.mark.parametrize("luni,n,rin,rout", [(1, 5, 2.0, 20.0), (1, 4, 2.0, 6.0), (2, 6, 2.0, 20.0),
                                              (3, 8, 2.0, 20.0)])
def test_shell_dt_brute_force(luni, n, rin, rout):
    m = OMesh("shell", luni, n, rin=rin, rout=rout)
    x, _ = lgl(n)
    hmins = []
    for e in range(m.nelem):
        p = int(m.elm_info(e)[0])
        bb = m.elm_bbox(e)
        Xb = [0.5 * ((bb[2 * d + 1] - bb[2 * d]) * x + (bb[2 * d + 1] + bb[2 * d])) for d in range(3)]
        pt = lambda i, j, k: shell_xyz(p, rin, rout, Xb[0][i], Xb[1][j], Xb[2][k])
        diag = np.linalg.norm(shell_xyz(p, rin, rout, bb[1], bb[3], bb[5]) -
                              shell_xyz(p, rin, rout, bb[0], bb[2], bb[4]))
        h = diag / np.sqrt(3.0)
        for c in range(8):
            I = [(n - 1) * ((c >> d) & 1) for d in range(3)]
            for d in range(3):
                J = list(I)
                J[d] += -1 if I[d] else 1
                h = min(h, np.linalg.norm(pt(*J) - pt(*I)))
        hmins.append(h)
        assert abs(lib().orc_mesh_elm_hmin(m.h, e) - h) < 1e-13 * h, e
    dt = 0.25 * min(hmins)
    assert abs(m.dt() - dt) < 1e-13 * dt, (m.dt(), dt)
