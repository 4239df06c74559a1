"""Pins of the oracle's troubled-element machinery (oracle/trouble.py and
orc_trouble_score / orc_trouble_rdmp):

- the score state machine against sequences written out from the rules of
  evolve_set_trouble_score (trouble.c:27-110) and the mesh score of
  evolve_read_trouble_score_mesh (trouble.c:129-161), clamp included;
- the RDMP indicator against brute force on planted values: a value just
  inside or just outside m - d, M + d, the neighbour widening, NaN, FV;
- the orchestration against closed forms: no trouble is the plain RK step;
  one troubled element in step mode is the RK step on the mesh with that
  element switched; one troubled element in Euler substep mode has its
  update from the switched mesh and every other element its DG update;
  the switch back happens after exactly 11 * substeps good calls.

Planted mistakes (in the oracle, on the CPU) must fail these pins.
"""
import numpy as np
import pytest

from oracle.orc import OMesh, trouble_score
from oracle.trouble import OTroubled

K = [0.8, -0.5, 0.33166247903554]


# -- the score state machine -------------------------------------------------

def run_scores(seq, start=0):
    s = np.array([start], dtype=np.int32)
    out = []
    for t in seq:
        trouble_score(s, [t])
        out.append(int(s[0]))
    return out


def test_score_sequences_follow_the_rules():
    # 1: start at 1 or raise; 0: reset; -1: start at -1 or lower
    assert run_scores([1, 1, 1, 0, -1, -1, -1, 1, -1]) == [1, 2, 3, 0, -1, -2, -3, 1, -1]
    assert run_scores([-1, -1, 1, 1, 0, 0]) == [-1, -2, 1, 2, 0, 0]
    assert run_scores([1], start=-5) == [1]
    assert run_scores([-1], start=7) == [-1]


def test_score_clamp():
    big = 1 << 30
    assert run_scores([1, 1], start=big - 1) == [big, big]
    assert run_scores([-1, -1], start=-big + 1) == [-big, -big]


def test_mesh_score_is_max_if_positive_else_min():
    s = np.array([0, -3, 2, -7], dtype=np.int32)
    assert trouble_score(s) == 2
    s = np.array([0, -3, -1, -7], dtype=np.int32)
    assert trouble_score(s) == -7
    s = np.array([0, 0, 0], dtype=np.int32)
    assert trouble_score(s) == 0
    s = np.array([-2, -4], dtype=np.int32)
    tr = np.array([-1, 1], dtype=np.int32)
    assert trouble_score(s, tr) == 1 and list(s) == [-3, 1]


# -- the RDMP indicator ---------------------------------------------------------

def neighbours(m, e):
    out = set()
    for f in range(6):
        nb, _ = m.elm_fnb(e, f)
        out.update(int(x) for x in nb)
    return out


def brute_rdmp(m, u, up, delta0, eps):
    info = np.array([m.elm_info(e) for e in range(m.nelem)])
    off, np_ = info[:, 3], info[:, 2] ** 3
    nn = m.nnodes
    fv = m.fv_flags()
    tr = np.zeros(m.nelem, dtype=np.int32)
    for e in range(m.nelem):
        els = [e] + sorted(neighbours(m, e))
        out = False
        for v in range(5):
            vals = np.concatenate([up[v * nn + off[j]: v * nn + off[j] + np_[j]] for j in els])
            lo, hi = vals.min(), vals.max()
            d = max(delta0, eps * (hi - lo))
            q = u[v * nn + off[e]: v * nn + off[e] + np_[e]]
            if not np.all((q >= lo - d) & (q <= hi + d)):
                out = True
        tr[e] = (0 if out else -1) if fv[e] else (1 if out else 0)
    return tr


def test_rdmp_planted_values():
    m = OMesh("box", 2, 4)
    rng = np.random.default_rng(3)
    up = rng.uniform(-1, 1, 5 * m.nnodes)
    u = up.copy()
    tr = m.trouble_rdmp(u, up, 1e-4, 1e-3)
    assert (tr == 0).all()                        # u = u_p: all in bounds, DG keeps
    info = np.array([m.elm_info(e) for e in range(m.nelem)])
    nn = m.nnodes
    e = 21
    els = [e] + sorted(neighbours(m, e))
    v = 3
    vals = np.concatenate([up[v * nn + info[j, 3]: v * nn + info[j, 3] + 64] for j in els])
    hi, lo = vals.max(), vals.min()
    d = max(1e-4, 1e-3 * (hi - lo))
    g = v * nn + info[e, 3] + 17
    u[g] = hi + d * (1 - 1e-9)
    assert m.trouble_rdmp(u, up, 1e-4, 1e-3)[e] == 0
    u[g] = hi + d * (1 + 1e-9)
    tr = m.trouble_rdmp(u, up, 1e-4, 1e-3)
    assert tr[e] == 1 and (tr == brute_rdmp(m, u, up, 1e-4, 1e-3)).all()
    u[g] = lo - d * (1 + 1e-9)
    assert m.trouble_rdmp(u, up, 1e-4, 1e-3)[e] == 1
    u[g] = np.nan
    assert m.trouble_rdmp(u, up, 1e-4, 1e-3)[e] == 1
    # delta0 dominates a flat state
    upc = np.zeros_like(up)
    uc = upc.copy()
    uc[g] = 0.9e-4
    assert m.trouble_rdmp(uc, upc, 1e-4, 1e-3)[e] == 0
    uc[g] = 1.1e-4
    assert m.trouble_rdmp(uc, upc, 1e-4, 1e-3)[e] == 1


def test_rdmp_neighbour_widening_and_fv():
    """a value above the element's own u_p maximum but inside a face
    neighbour's is in bounds; FV elements in bounds say -1"""
    m = OMesh("box", 2, 4)
    info = np.array([m.elm_info(e) for e in range(m.nelem)])
    nn = m.nnodes
    up = np.zeros(5 * nn)
    e = 21
    nb = sorted(neighbours(m, e))[0]
    up[0 * nn + info[nb, 3] + 5] = 1.0            # neighbour's maximum 1
    u = up.copy()
    u[0 * nn + info[e, 3] + 9] = 0.99
    assert m.trouble_rdmp(u, up)[e] == 0
    far = [j for j in range(m.nelem) if j != e and j not in neighbours(m, e) and j != nb][0]
    u2 = up.copy()
    u2[0 * nn + info[far, 3] + 9] = 0.5           # far from the bump: out of bounds
    tr = m.trouble_rdmp(u2, up)
    assert tr[far] == (1 if nb not in neighbours(m, far) else 0)
    fv = np.zeros(m.nelem, dtype=np.int32)
    fv[e] = 1
    mf, upf = m.adapted(up, fv=fv)
    assert mf.trouble_rdmp(upf, upf)[e] == -1
    assert (mf.trouble_rdmp(upf, upf) == brute_rdmp(mf, upf, upf, 1e-4, 1e-3)).all()


# -- the orchestration ---------------------------------------------------------

def smooth_state(m, seed=0):
    u, _ = m.profile(K, 0.0)
    return u


def schedule(table):
    """indicator from a table {call: {element: value}}, 0 elsewhere"""
    def ind(call, mesh, u, up):
        tr = np.zeros(mesh.nelem, dtype=np.int32)
        for e, v in table.get(call, {}).items():
            tr[e] = v
        return tr
    return ind


@pytest.mark.parametrize("method", ["RK4", "sspRK3", "Euler"])
def test_no_trouble_is_the_plain_step(method):
    m = OMesh("box", 2, 4)
    u = smooth_state(m)
    dt = m.dt()
    ref, _ = m.evolve(u, dt, 1, 0.0, K, method)
    ot = OTroubled(m, K, method, schedule({}))
    got = ot.step(u, 0.0, dt)
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()
    assert ot.calls == {"RK4": 4, "sspRK3": 3, "Euler": 1}[method]
    assert ot.redos == 0 and ot.switches == 0


@pytest.mark.parametrize("method", ["RK4", "sspRK3"])
def test_step_redo_is_the_step_on_the_switched_mesh(method):
    m = OMesh("box", 2, 4)
    u = smooth_state(m)
    dt = m.dt()
    e = 21
    ot = OTroubled(m, K, method, schedule({0: {e: 1}}), redo="step")
    got = ot.step(u, 0.0, dt)
    fv = np.zeros(m.nelem, dtype=np.int32)
    fv[e] = 1
    ms, us = m.adapted(u, fv=fv)
    ref, _ = ms.evolve(us, dt, 1, 0.0, K, method)
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()
    assert (ot.m.fv_flags() == fv).all()
    assert ot.calls == 1 and ot.redos == 1 and ot.switches == 1


def test_euler_substep_redo_of_one_element():
    """Euler, substep mode, one troubled element: its update is the Euler
    update on the mesh with it switched (from u_p interpolated), every other
    element keeps its DG update"""
    m = OMesh("shell", 1, 4)
    u = smooth_state(m)
    dt = m.dt()
    e = 9
    ot = OTroubled(m, K, "Euler", schedule({0: {e: 1}}))
    got = ot.step(u, 0.0, dt)
    fv = np.zeros(m.nelem, dtype=np.int32)
    fv[e] = 1
    ms, us = m.adapted(u, fv=fv)
    sw, _ = ms.evolve(us, dt, 1, 0.0, K, "Euler")
    dg, _ = m.evolve(u, dt, 1, 0.0, K, "Euler")
    info = np.array([m.elm_info(j) for j in range(m.nelem)])
    nn = m.nnodes
    ix = (np.arange(5)[:, None] * nn + info[e, 3] + np.arange(64)[None, :]).ravel()
    mask = np.zeros(5 * nn, bool)
    mask[ix] = True
    scale = np.abs(dg).max()
    assert np.abs(got[mask] - sw[mask]).max() <= 1e-13 * scale
    assert np.abs(got[~mask] - dg[~mask]).max() <= 1e-15 * scale
    assert (ot.m.fv_flags() == fv).all() and ot.score[e] == 1


def test_switch_back_after_notroubles_good_calls():
    """an FV element scored -1 at every call goes back to DG when its score
    reaches -11 * substeps: after 11 steps in substep mode, 11 in step mode"""
    m = OMesh("box", 2, 4)
    e = 21
    fv = np.zeros(m.nelem, dtype=np.int32)
    fv[e] = 1
    m1, u = m.adapted(smooth_state(m), fv=fv)
    dt = m.dt()

    def good(call, mesh, u, up):
        return np.where(mesh.fv_flags() != 0, -1, 0).astype(np.int32)

    for method, redo in [("RK4", "substep"), ("Euler", "substep"), ("RK4", "step")]:
        ot = OTroubled(m1, K, method, good, redo=redo)
        x, t = u.copy(), 0.0
        back_at = None
        for step in range(1, 14):
            x = ot.step(x, t, dt)
            t += dt
            if back_at is None and ot.m.fv_flags()[e] == 0:
                back_at = step
        assert back_at == 11, (method, redo, back_at)
        assert ot.switches == 1
