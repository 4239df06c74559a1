"""GPU parity of the troubled-element machinery (nmx_trouble = hook or
RDMP inside nmx_evolve) against the oracle's restatement
(oracle/trouble.py: evolve_myln, Evolve_mesh with redone substeps,
evolve_set_trouble_score, trouble.c:27-400, RungeKutta.c:19-344).

The test's TROUBLE function is a table of values per scoring call (so both
sides decide the same, whatever the rounding), handed to the device through
nmx_set_trouble_hook; the hook also records the u and u_p it is given.
Every call's u and u_p, every step's state, scores, FV flags, redo and
switch counts are compared, 1e-12 relative to the field maximum:

- RK4, substep mode, on a box: two adjacent elements troubled in one
  substep, another later, all FV elements good from some call on, so they
  switch back after 11 * 4 calls;
- sspRK3, substep mode, on an h-refined shell: two troubled elements that
  share a non-conforming face (each redone against the other as it was
  before the switch), Euler on the same mesh;
- RK4 and sspRK3 in step mode (the step redone from u_p on the switched
  mesh);
- the library's RDMP indicator on a state with a planted bump.
"""
import ctypes as C

import numpy as np
import pytest

import nmesh_amd as nm
from oracle.orc import OMesh
from oracle.trouble import OTroubled

pytestmark = pytest.mark.gpu
K = [0.8, -0.5, 0.33166247903554]
TOL = 1e-12

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
        _hip.hipMemcpyAsync.restype = C.c_int
        _hip.hipStreamSynchronize.argtypes = [C.c_void_p]
        _hip.hipStreamSynchronize.restype = C.c_int
    return _hip


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


class Hook:
    """the device TROUBLE function: reads u and u_p (element blocks) into
    the flat layout, asks the indicator, writes tr"""

    def __init__(self, m, indicator):
        info = np.array([m.elm_info(e) for e in range(m.nelem)])
        self.off, self.np = info[:, 3], info[:, 2] ** 3
        self.nn = m.nnodes
        self.m = m
        self.indicator = indicator
        self.calls = 0
        self.log = []

    def flat(self, ptr, stream):
        blk = np.zeros(5 * self.nn)
        assert hip().hipMemcpyAsync(blk.ctypes.data, ptr, blk.nbytes, 2, stream) == 0
        assert hip().hipStreamSynchronize(stream) == 0
        out = np.empty_like(blk)
        for e in range(len(self.off)):
            o, k = self.off[e], self.np[e]
            out.reshape(5, -1)[:, o:o + k] = blk[5 * o:5 * o + 5 * k].reshape(5, k)
        return out

    def __call__(self, u, up, tr, n, stream):
        try:
            U, UP = self.flat(u, stream), self.flat(up, stream)
            t = np.ascontiguousarray(self.indicator(self.calls, self.m, U, UP), dtype=np.int32)
            assert t.size == n
            # in stream order; t stays alive until the copy is done
            assert hip().hipMemcpyAsync(tr, t.ctypes.data, 4 * n, 1, stream) == 0
            assert hip().hipStreamSynchronize(stream) == 0
            self.log.append((U, UP))
            self.calls += 1
            return 0
        except Exception as exc:      # an exception cannot cross the C frame
            print("hook failed:", exc, flush=True)
            return 1


def table_indicator(table, good_from=None):
    """{call: {element: value}}; from call good_from on, FV elements get -1"""
    def ind(call, mesh, u, up):
        tr = np.zeros(mesh.nelem, dtype=np.int32)
        if good_from is not None and call >= good_from:
            tr[mesh.fv_flags() != 0] = -1
        for e, v in table.get(call, {}).items():
            tr[e] = v
        return tr
    return ind


def box(method, redo, trouble="hook", extra=None):
    pars = {"amr_mesh_type": "BoxMesh", "amr_luni": 2, "amr_n0": 4, "amr_n1": 4, "amr_n2": 4,
            "scalarwave1_k": "%.17g %.17g %.17g" % tuple(K), "evolve_method": method,
            "nmx_trouble": trouble, "evolve_redo_troubled": redo}
    pars.update(extra or {})
    m = nm.Mesh(pars, device=0)
    nm.amr_setup_mesh(m)
    return m, OMesh("box", 2, 4)


def refined_shell(method, redo, trouble="hook"):
    pars = {"amr_mesh_type": "Shell", "amr_luni": 1, "amr_n0": 4, "amr_n1": 4, "amr_n2": 4,
            "amr_Shell_rin": 2.0, "amr_Shell_rout": 6.0,
            "scalarwave1_k": "%.17g %.17g %.17g" % tuple(K), "evolve_method": method,
            "nmx_trouble": trouble, "evolve_redo_troubled": redo}
    m = nm.Mesh(pars, device=0)
    nm.amr_setup_mesh(m)
    om = OMesh("shell", 1, 4, rin=2.0, rout=6.0)
    flag = np.zeros(om.nelem, np.int32)
    flag[[3, 17, 40]] = 1
    m.hadapt(flag)
    om, _ = om.adapted(np.zeros(5 * om.nnodes), hflag=flag)
    return m, om


def nonconforming_pair(om):
    """a level-1 element and a level-2 face neighbour"""
    for e in range(om.nelem):
        if om.elm_info(e)[1] != 1:
            continue
        for f in range(6):
            nb, _ = om.elm_fnb(e, f)
            if len(nb) > 1:
                return e, int(nb[0])
    raise AssertionError("no non-conforming face")


def run_both(m, om, method, redo, indicator, nsteps, u0, dt):
    m.time = 0.0
    m.dt = dt
    m.upload(u0)
    hook = Hook(m, indicator)
    m.set_trouble_hook(hook)
    ot = OTroubled(om, K, method, indicator, redo=redo)
    u, t = u0.copy(), 0.0
    try:
        for step in range(nsteps):
            print("trouble step", step, flush=True)
            u = ot.step(u, t, dt)
            t += dt
            nm.evolve_myln(m, 1)
            assert hook.calls == ot.calls, (step, hook.calls, ot.calls)
            got = m.download()
            assert rel(got, u) <= TOL, (step, rel(got, u))
            sc, redos, switches = m.trouble_scores()
            assert np.array_equal(sc, ot.score), step
            assert np.array_equal(m.fv_flags(), ot.m.fv_flags()), step
            assert (redos, switches) == (ot.redos, ot.switches), (step, redos, switches, ot.redos, ot.switches)
        for k, ((U, UP), (Uo, UPo, _)) in enumerate(zip(hook.log, ot.log)):
            assert rel(U, Uo) <= TOL and rel(UP, UPo) <= TOL, k
    finally:
        m.set_trouble_hook(None)
    return ot, hook


@pytest.mark.parametrize("graph,path", [("yes", "plain"), ("no", "plain"), ("yes", "batched")])
def test_rk4_substep_redo_and_switch_back_on_box(graph, path):
    """the stages from captured graphs (one per RK stage, replayed every
    step and recaptured after each switch) and eager"""
    m, om = box("RK4", "substep", extra={"nmx_graph": graph, "nmx_path": path})
    u0, _ = om.profile(K, 0.0)
    dt = om.dt()
    table = {1: {21: 1, 22: 1}, 6: {42: 1}, 7: {21: 1}}
    ot, hook = run_both(m, om, "RK4", "substep", table_indicator(table, good_from=9), 14, u0, dt)
    assert ot.redos == 3 and hook.calls == 56
    # all three went FV; 42 and 22 scored -1 from call 9, 21 from call 9 too
    # (its 1 at call 7 was reset): back to DG after 44 good calls
    assert not ot.m.fv_flags().any()


@pytest.mark.parametrize("method", ["sspRK3", "Euler"])
def test_substep_redo_across_a_nonconforming_face(method):
    m, om = refined_shell(method, "substep")
    a, b = nonconforming_pair(om)
    u0, _ = om.profile(K, 0.0)
    u0 = u0 + 0.05 * np.random.default_rng(7).uniform(-1, 1, u0.shape)
    dt = om.dt()
    table = {1: {a: 1, b: 1}} if method == "sspRK3" else {0: {a: 1, b: 1}, 2: {5: 1}}
    ot, _ = run_both(m, om, method, "substep", table_indicator(table), 3, u0, dt)
    fv = ot.m.fv_flags()
    assert fv[a] and fv[b]


@pytest.mark.parametrize("method", ["RK4", "sspRK3"])
def test_step_redo_on_box(method):
    m, om = box(method, "step")
    u0, _ = om.profile(K, 0.0)
    dt = om.dt()
    table = {0: {21: 1, 22: 1}, 2: {5: 1}}
    ot, _ = run_both(m, om, method, "step", table_indicator(table), 3, u0, dt)
    assert ot.redos == 2


def test_rdmp_indicator_on_a_bump():
    """the library's RDMP indicator (nmx_trouble = RDMP) against the
    oracle's on a smooth state with a planted bump"""
    m, om = box("RK4", "substep", trouble="RDMP")
    u0, _ = om.profile(K, 0.0)
    info = np.array([om.elm_info(e) for e in range(om.nelem)])
    nn = om.nnodes
    u0[0 * nn + info[21, 3] + 21] += 3.0
    u0[4 * nn + info[42, 3] + 10] -= 2.0
    dt = om.dt()
    m.time = 0.0
    m.dt = dt
    m.upload(u0)

    def rdmp(call, mesh, u, up):
        return mesh.trouble_rdmp(u, up, 1e-4, 1e-3)

    ot = OTroubled(om, K, "RK4", rdmp)
    u, t = u0.copy(), 0.0
    for step in range(3):
        print("rdmp step", step, flush=True)
        u = ot.step(u, t, dt)
        t += dt
        nm.evolve_myln(m, 1)
        assert rel(m.download(), u) <= TOL, step
        sc, redos, switches = m.trouble_scores()
        assert np.array_equal(sc, ot.score), step
        assert np.array_equal(m.fv_flags(), ot.m.fv_flags()), step
        assert (redos, switches) == (ot.redos, ot.switches), step
    assert ot.redos >= 1 and ot.m.fv_flags().any()


def test_trouble_refusals():
    m, om = box("RK4", "substep")
    m.upload(om.profile(K, 0.0)[0])
    with pytest.raises(nm.NmxError):
        nm.evolve_myln(m, 1)                 # hook mode without a hook
    m.set_par("nmx_trouble", "none")
    nm.evolve_myln(m, 1)
    with pytest.raises(nm.NmxError):
        m.set_par("nmx_trouble", "sometimes")
