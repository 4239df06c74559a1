#!/usr/bin/env python3
"""Benchmark of the DG RHS + RK stage hot path on one MI355X.

Headline workload (the mixed-order AMR configuration the north-star target
is stated on): scalar wave (scalarwave1) on a cubed-sphere shell r in [2, 20]
refined to level 3 and, inside nested spheres of radius 6 and 3, to levels 4
and 5 (19,200 elements); per-element n drawn uniformly from {4..10} points
per direction (N in {3..9}, PCG64 seed 2026; 8,213,159 nodes); FP64, upwind
flux, outgoing boundaries, RK4 at dtfac 0.25, batched order-bucket launches
with hipGraph-captured steps and the autotuned launch shape.  Initial data:
a Gaussian shell psi = exp(-(r-8)^2) (1 + sum_m a_m Y_2m) with a_m ~
U[-0.1, 0.1] (PCG64 seed 2026), Phi_i = d_i psi, Pi = 0.

One "step" is one full RK4 step (4 stages).  The unit of work is a
node-update: one RHS evaluation at one node for all 5 fields, fused with its
stage update.  Secondary legs: the plain per-order kernel on the uniform
N=7 shell (6x8^3 elements), per-element launches, the adjoint, FV, filter
and CRC.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# algorithmic bytes per node-update of a fused DG RHS + RK4 stage on
# curvilinear elements: 8 * (25 state doubles + 9 dXdx)
BYTES_PER_NU_CURV = 272.0
# the least the batched path moves per node-update: 20 state doubles on
# average over the 4 stages (dXdx recomputed)
MIN_BYTES_PER_NU = 160.0


def flops_per_nu(n):
    """FP64 operations of one DG node-update of scalarwave1 at n points per
    direction: 12 derivatives along the element's reference directions (of
    pi and of the three c_i) at 2n each, and 36 for their combination with
    the metric into grad pi and div c and the RHS.  Face terms, the metric's
    own recomputation and the stage combination are not counted."""
    return 24.0 * n + 36.0


def shell_initial_data(x, seed=2026):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-0.1, 0.1, 5)
    c1 = 0.5 * np.sqrt(15.0 / np.pi)
    c0 = 0.25 * np.sqrt(5.0 / np.pi)
    # S = x^T A x / r^2 = sum_m a_m Y_2m (real harmonics)
    A = np.zeros((3, 3))
    A[0, 1] = A[1, 0] = 0.5 * c1 * a[0]            # Y_2,-2 ~ xy
    A[1, 2] = A[2, 1] = 0.5 * c1 * a[1]            # Y_2,-1 ~ yz
    A += c0 * a[2] * np.diag([-1.0, -1.0, 2.0])    # Y_2,0 ~ 3z^2 - r^2
    A[0, 2] = A[2, 0] = 0.5 * c1 * a[3]            # Y_2,1 ~ xz
    A[0, 0] += 0.5 * c1 * a[4]                     # Y_2,2 ~ x^2 - y^2
    A[1, 1] -= 0.5 * c1 * a[4]
    r2 = (x ** 2).sum(0)
    r = np.sqrt(r2)
    Ax = A @ x
    Q = (x * Ax).sum(0)
    S = Q / r2
    G = np.exp(-(r - 8.0) ** 2)
    dG = -2.0 * (r - 8.0) * G
    psi = G * (1.0 + S)
    dS = (2.0 * Ax * r2 - 2.0 * x * Q) / r2 ** 2
    dpsi = dG * x / r * (1.0 + S) + G * dS
    u = np.zeros((5, x.shape[1]))
    u[1:4] = dpsi
    u[4] = psi
    return u.ravel()


def c2_params():
    return {
        "amr_mesh_type": "Shell", "amr_Shell_rin": 2.0, "amr_Shell_rout": 20.0,
        "amr_luni": 3, "amr_n0": 8, "amr_n1": 8, "amr_n2": 8,
        "scalarwave1_k": "0 0 0", "scalarwave1_numflux": "upwind",
        "evolve_method": "RK4", "dtfac": 0.25, "dt": "auto",
    }


def load_traffic(name):
    p = os.path.join(ROOT, "profiles", name)
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return None


def cpu_baseline(x_unused, budget_s=12.0):
    """Oracle (OpenMP C restatement of the reference algorithm) on the same
    mesh and operator, timed on the host cores for a bounded number of RK4
    steps."""
    from oracle.orc import OMesh, lib
    om = OMesh("shell", 3, 8, rin=2.0, rout=20.0)
    X, x = om.coords()
    u = shell_initial_data(x)
    dt = om.dt()
    k = [0.0, 0.0, 0.0]
    # one step first, then as many as fit the budget
    t0 = time.perf_counter()
    u, t = om.evolve(u, dt, 1, 0.0, k, "RK4")
    one = time.perf_counter() - t0
    nmore = max(0, min(8, int((budget_s - one) / max(one, 1e-3))))
    t0 = time.perf_counter()
    if nmore:
        u, t = om.evolve(u, dt, nmore, t, k, "RK4")
    el = time.perf_counter() - t0
    steps = nmore if nmore else 1
    secs = el if nmore else one
    rate = steps * 4 * om.nnodes / secs
    return {"value": rate, "unit": "node-updates/s", "cores": int(lib().orc_num_threads()),
            "kind": "port",
            "sample": "%d RK4 step(s) of the same 3072-element N=7 shell (%d node-updates) "
                      "with the oracle's OpenMP C restatement" % (steps, steps * 4 * om.nnodes)}


# extra library parameters of the C4 meshes (--amr-par name=value) and of
# the FV grid (--fv-par name=value)
C4_EXTRA = {}
FV_EXTRA = {}


def c4_mesh(nm, path, extra=None):
    """Shell r in [2,20] refined to level 3, then in nested spheres of radius
    6 and 3 (levels 4 and 5): 19,200 elements; per-element N drawn uniformly
    from {3..9} (PCG64, seed 2026); scalar wave, upwind, outgoing, RK4."""
    p = c2_params()
    p.update({"amr_n0": 5, "amr_n1": 5, "amr_n2": 5, "amr_hrefine_sphere_radius": 3,
              "amr_hrefine_sphere_levels": 2, "nmx_path": path})
    p.update(C4_EXTRA)
    p.update(extra or {})
    m = nm.Mesh(p, device=0)
    nm.amr_setup_mesh(m)
    rng = np.random.default_rng(2026)
    m.set_elm_n(rng.integers(4, 11, m.nelem).astype(np.int32))
    return m


def c4_headline(nm, steps, warmup, have_torch):
    """The headline: the C4 mesh on the batched path, W untimed steps, then
    exactly K timed steps bracketed by device synchronisations."""
    import torch
    m = c4_mesh(nm, "batched")
    x = m.coords()
    m.upload(shell_initial_data(x))
    nm.evolve_myln(m, 1)                   # autotune (launch shape, tiles) and graph capture
    if warmup:
        nm.evolve_myln(m, warmup)
    m.synchronize()
    if have_torch:
        torch.cuda.synchronize()
    t0 = time.perf_counter()
    nm.evolve_myln(m, steps)
    m.synchronize()
    if have_torch:
        torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    dev_s = m.last_evolve_ms() * 1e-3      # HIP events on the library's stream around the steps
    info = np.array([m.elm_info(e) for e in range(m.nelem)])
    n = info[:, 2].astype(np.float64)
    nn = m.nnodes
    stages = 4 * steps
    stage_s = dev_s / stages
    flops_stage = float((n ** 3 * flops_per_nu(n)).sum())
    buckets = {int(k): int((info[:, 2] == k).sum()) for k in np.unique(info[:, 2])}
    res = {"m": m, "wall": wall, "dev_s": dev_s, "nodes": nn, "elements": m.nelem,
           "stage_us": stage_s * 1e6, "flops_per_stage": flops_stage,
           "achieved": BYTES_PER_NU_CURV * nn / stage_s / 1e9,
           "achieved_min": MIN_BYTES_PER_NU * nn / stage_s / 1e9,
           "tflops": flops_stage / stage_s / 1e12, "buckets": buckets,
           "shape": m.par("nmx_batch"), "dt": m.dt}
    return res


def c4_cpu_baseline(budget_s=12.0):
    """The oracle (OpenMP C restatement of the reference algorithm) on the
    same C4 mesh, orders and operator: as many RK4 steps as fit the budget
    (at least one), timed on the host cores."""
    from oracle.orc import OMesh, lib
    om = OMesh("shell", 3, 5, rin=2.0, rout=20.0, radius=3, levels=2)
    om.set_n(np.random.default_rng(2026).integers(4, 11, om.nelem).astype(np.int32))
    X, x = om.coords()
    u = shell_initial_data(x)
    dt = om.dt()
    k = [0.0, 0.0, 0.0]
    t0 = time.perf_counter()
    u, t = om.evolve(u, dt, 1, 0.0, k, "RK4")
    one = time.perf_counter() - t0
    steps, secs = 1, one
    nmore = max(0, min(4, int((budget_s - one) / max(one, 1e-3))))
    if nmore:
        t0 = time.perf_counter()
        u, t = om.evolve(u, dt, nmore, t, k, "RK4")
        steps, secs = nmore, time.perf_counter() - t0
    rate = steps * 4 * om.nnodes / secs
    return {"value": rate, "unit": "node-updates/s", "cores": int(lib().orc_num_threads()),
            "kind": "port",
            "sample": "%d RK4 step(s) of the same 19,200-element mixed-order AMR shell (%d "
                      "node-updates) with the oracle's OpenMP C restatement" % (steps, steps * 4 * om.nnodes)}


def batched_leg(nm, steps, plain_steps):
    """Config-4 AMR mesh: order-bucketed launches (hipGraph-captured RK4 steps)
    against one launch per element per stage."""
    out = {}
    for path, nsteps in (("batched", steps), ("perelement", plain_steps)):
        if nsteps == 0:
            continue
        m = c4_mesh(nm, path)
        x = m.coords()
        m.upload(shell_initial_data(x))
        nm.evolve_myln(m, 1)            # warm-up (and autotune, which picks the launch shape)
        m.synchronize()
        shapes = [m.par("nmx_batch")] if path == "batched" else [None]
        if path == "batched":
            shapes += [q for q in ("persistent", "buckets", "legacy") if q != shapes[0]]
        for shape in shapes:
            if shape is not None and shape != m.par("nmx_batch"):
                m.set_par("nmx_batch", shape)
                nm.evolve_myln(m, 1)
                m.synchronize()
            t0 = time.perf_counter()
            nm.evolve_myln(m, nsteps)
            m.synchronize()
            wall = time.perf_counter() - t0
            rate = 4.0 * nsteps * m.nnodes / wall
            dev_s = m.last_evolve_ms() * 1e-3
            res = {"value": rate, "steps": nsteps, "ms_per_step": 1e3 * wall / nsteps,
                   "device_ms_per_step": 1e3 * dev_s / nsteps,
                   "achieved_GBps": BYTES_PER_NU_CURV * 4 * nsteps * m.nnodes / dev_s / 1e9,
                   "achieved_min_GBps": MIN_BYTES_PER_NU * 4 * nsteps * m.nnodes / dev_s / 1e9}
            if shape is not None:
                res["launch_shape"] = shape
                out.setdefault("shapes", {})[shape] = res["ms_per_step"]
            if path not in out or res["value"] > out[path]["value"]:
                out[path] = res
        out["nodes"] = m.nnodes
        out["elements"] = m.nelem
        del m
    if "perelement" in out:
        out["speedup_batched_vs_perelement"] = out["batched"]["value"] / out["perelement"]["value"]
    out["roofline_frac_batched"] = out["batched"]["achieved_GBps"] / HBM_PEAK_GBS
    out["roofline_frac_min_batched"] = out["batched"]["achieved_min_GBps"] / HBM_PEAK_GBS
    out["workload"] = ("scalarwave1 DG, cubed-sphere shell r in [2,20], levels 3-5 (19,200 elements), "
                       "N uniform in {3..9} (%d nodes), upwind, outgoing, RK4" % out["nodes"])
    return out


# FV (config 3): algorithmic figures per cell-update (one RHS evaluation at
# one cell, all 5 fields; sspRK3 does 3 per cell per step).  Bytes: the c2p
# pass reads the conserved state and the pressure guess and writes the
# primitives (11 doubles); the stage pass reads the primitives once and the
# stage registers (win, u_p, u_acc: 5, 15 and 10 doubles in stages 1, 2, 3)
# and writes u_acc, w (10, 10, 5): (11 + 5 + (15 + 25 + 15) / 3) * 8 = 275 B.
# FLOPs (hand count of the kernel's formulas, a division or a square root
# counted as one): 10 WENO5 reconstructions of 68 FLOPs, 2 SR fluxes with
# eigenvalues of 60, LLF and accumulation 32, physicality checks 10: 852
# per face, 3 faces per cell; c2p about 115 (5 Newton iterations of 20 and
# the primitives) and the stage combination 15: about 2690 per cell-update.
FV_BYTES_PER_CU = 275.0
FV_FLOPS_PER_CU = 2690.0


def fv_leg(steps=50, warmup=8, N=256, cpu=True):
    """Config 3: FV WENO5 + LLF, special-relativistic Euler (Gamma = 2) on a
    periodic 256^3 grid of 16^3-cell blocks, 32 seeded Gaussian density
    blobs, per-cell velocities U[-0.3, 0.3], p = 1; sspRK3 at dt = 0.25."""
    from nmesh_amd.fv import FvGrid, blob_initial_data
    P = blob_initial_data(N)
    pars = {"fv_n": N, "evolve_method": "sspRK3"}
    pars.update(FV_EXTRA)
    g = FvGrid(pars, device=0)
    g.setup()
    g.upload_prim(P)
    g.evolve(warmup)                 # includes the graph capture
    t0 = time.perf_counter()
    g.evolve(steps)
    wall = time.perf_counter() - t0
    dev_s = g.last_evolve_ms() * 1e-3
    cu = 3.0 * steps * N ** 3
    out = {"workload": "SR-Euler FV WENO5 + LLF, %d^3 periodic cells in 16^3 blocks, sspRK3, %d steps" % (N, steps),
           "value": cu / wall, "unit": "cell-updates/s", "ms_per_step": 1e3 * wall / steps,
           "device_ms_per_step": 1e3 * dev_s / steps, "c2p_failed_blocks": g.c2p_failures(),
           "bytes_per_cell_update": FV_BYTES_PER_CU, "flops_per_cell_update": FV_FLOPS_PER_CU,
           "achieved_GBps": FV_BYTES_PER_CU * cu / dev_s / 1e9,
           "achieved_TFLOPs": FV_FLOPS_PER_CU * cu / dev_s / 1e12}
    out["hbm_frac"] = out["achieved_GBps"] / HBM_PEAK_GBS
    out["fp64_frac"] = out["achieved_TFLOPs"] / FP64_PEAK_TFLOPS
    del g
    if cpu:
        from oracle.orc import OFV, lib
        n = 64
        Pc = blob_initial_data(n)
        ofv = OFV(n)
        u = ofv.p2c(Pc)
        t0 = time.perf_counter()
        ofv.evolve(u, Pc, 0.25, 2)
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": 3 * 2 * n ** 3 / el, "unit": "cell-updates/s",
                               "cores": int(lib().orc_num_threads()), "kind": "port",
                               "sample": "2 sspRK3 steps of the same scheme on a 64^3 grid (oracle, OpenMP C)"}
    return out


def c1_params():
    return {"amr_mesh_type": "BoxMesh", "amr_BoxMesh_dout": 8.0, "amr_mesh_xc": "0 0 0",
            "amr_luni": 2, "amr_n0": 5, "amr_n1": 5, "amr_n2": 5,
            "scalarwave1_k": "0 0 0", "scalarwave1_numflux": "upwind",
            "evolve_method": "RK4", "dtfac": 0.25, "dt": "auto"}


def fvelm_leg(nm, steps=20):
    """scalarwave1 on FV elements inside the mesh (fv_divf): a BoxMesh
    [-8,8]^3 of 16^3 elements, all FV, n = 8 uniform points per direction
    (2,097,152 cells), WENOm5_2, upwind fnum, RK4.  Bytes per cell-update as
    the DG stage on Cartesian elements: 160 B (stage average of the 15-25
    state doubles read and written)."""
    p = {"amr_mesh_type": "BoxMesh", "amr_BoxMesh_dout": 8.0, "amr_luni": 4,
         "amr_n0": 8, "amr_n1": 8, "amr_n2": 8, "amr_uniform_p": "0", "amr_fv_p": "0",
         "fv_rec": "WENOm5_2", "scalarwave1_k": "0 0 0", "evolve_method": "RK4"}
    m = nm.Mesh(p, device=0)
    nm.amr_setup_mesh(m)
    x = m.coords()
    r2 = (x ** 2).sum(0)
    u = np.zeros((5, x.shape[1]))
    u[4] = np.exp(-r2)
    u[1:4] = -2.0 * x * u[4]
    m.upload(u.ravel())
    nm.evolve_myln(m, 2)
    m.synchronize()
    t0 = time.perf_counter()
    nm.evolve_myln(m, steps)
    m.synchronize()
    wall = time.perf_counter() - t0
    dev_s = m.last_evolve_ms() * 1e-3
    nu = 4.0 * steps * m.nnodes
    gbs = MIN_BYTES_PER_NU * nu / dev_s / 1e9
    out = {"workload": "scalarwave1 on FV elements (fv_divf, WENOm5_2, upwind), BoxMesh 16^3 "
                       "elements of 8^3 uniform points (%d cells), RK4" % m.nnodes,
           "value": nu / wall, "unit": "cell-updates/s", "steps": steps,
           "ms_per_step": 1e3 * wall / steps, "bytes_per_cell_update": MIN_BYTES_PER_NU,
           "achieved_GBps": gbs, "hbm_frac": gbs / HBM_PEAK_GBS,
           "kernel": "fvelm_stage_kernel<8,5,*>"}
    del m
    out["shell"] = fvelm_shell_leg(nm, steps)
    return out


def fvelm_shell_leg(nm, steps=20):
    """The same on cubed-sphere elements: the shell r in [2,20] at level 3
    (6x8^3 elements), all FV, n = 8 uniform points per direction (1,572,864
    cells), WENOm5_2 with coordinates_midpoint_data = yes (the metric at the
    midpoints, as nmesh's curved-mesh FV parfiles set it), upwind, RK4."""
    p = {"amr_mesh_type": "Shell", "amr_Shell_rin": 2.0, "amr_Shell_rout": 20.0, "amr_luni": 3,
         "amr_n0": 8, "amr_n1": 8, "amr_n2": 8, "amr_uniform_p": "0 1 2 3 4 5",
         "amr_fv_p": "0 1 2 3 4 5", "fv_rec": "WENOm5_2", "coordinates_midpoint_data": "yes",
         "scalarwave1_k": "0 0 0", "evolve_method": "RK4"}
    m = nm.Mesh(p, device=0)
    nm.amr_setup_mesh(m)
    m.upload(shell_initial_data(m.coords()))
    nm.evolve_myln(m, 2)
    m.synchronize()
    t0 = time.perf_counter()
    nm.evolve_myln(m, steps)
    m.synchronize()
    wall = time.perf_counter() - t0
    dev_s = m.last_evolve_ms() * 1e-3
    nu = 4.0 * steps * m.nnodes
    gbs = MIN_BYTES_PER_NU * nu / dev_s / 1e9
    out = {"workload": "scalarwave1 on FV elements of the cubed-sphere shell (fv_divf, WENOm5_2, "
                       "midpoint metric, upwind), 6x8^3 elements of 8^3 uniform points (%d cells), RK4" % m.nnodes,
           "value": nu / wall, "unit": "cell-updates/s", "steps": steps,
           "ms_per_step": 1e3 * wall / steps, "device_us_per_stage": dev_s / (4 * steps) * 1e6,
           "bytes_per_cell_update": MIN_BYTES_PER_NU, "achieved_GBps": gbs, "hbm_frac": gbs / HBM_PEAK_GBS,
           "kernel": "fvelm_stage_kernel<8,5,*,2,*> (shell)"}
    del m
    return out


def trouble_leg(nm, steps=20, redo_steps=3):
    """Troubled elements on the C4 mesh, batched path, RK4 with
    evolve_redo_troubled = substep (trouble.c:27-400, RungeKutta.c:221-344).
    1. Scoring only: the RDMP indicator after every substep with
       nmx_trouble_delta0 = 1e30, so no element is troubled; ms per step
       against the headline's graph-captured steps.
    2. Redo: a TROUBLE hook marks 4 elements far apart at the second scoring
       call (step 0, substep 2) and none after. Those elements take their
       substep back, become FV and redo it; the step is timed, switch
       included, then steps on the mixed DG/FV mesh."""
    import ctypes as C
    m = c4_mesh(nm, "batched", {"nmx_trouble": "RDMP", "evolve_redo_troubled": "substep",
                                "nmx_trouble_delta0": "1e30"})
    u0 = shell_initial_data(m.coords())
    m.upload(u0)
    dt = m.dt
    nm.evolve_myln(m, 1)
    m.synchronize()
    t0 = time.perf_counter()
    nm.evolve_myln(m, steps)
    m.synchronize()
    score_ms = 1e3 * (time.perf_counter() - t0) / steps
    _, redos0, switches0 = m.trouble_scores()
    print("trouble leg: scoring only %.3f ms per step" % score_ms, file=sys.stderr, flush=True)

    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    hip.hipMemcpyAsync.restype = C.c_int
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    hip.hipStreamSynchronize.restype = C.c_int
    ne = m.nelem
    marked = [ne // 8, 3 * ne // 8, 5 * ne // 8, 7 * ne // 8]
    zero = np.zeros(ne, np.int32)
    mark = zero.copy()
    mark[marked] = 1
    calls = [0]

    def hook(u, up, tr, n, stream):
        t = mark if calls[0] == 1 else zero
        calls[0] += 1
        if n != ne or hip.hipMemcpyAsync(tr, t.ctypes.data, 4 * n, 1, stream) != 0:
            return 1
        return int(hip.hipStreamSynchronize(stream) != 0)

    m.set_par("nmx_trouble", "hook")
    m.set_trouble_hook(hook)
    try:
        m.time = 0.0
        m.dt = dt
        m.upload(u0)
        m.synchronize()
        per_step = []
        for k in range(redo_steps):
            t0 = time.perf_counter()
            nm.evolve_myln(m, 1)
            m.synchronize()
            per_step.append(1e3 * (time.perf_counter() - t0))
            print("trouble leg: redo step %d %.1f ms" % (k, per_step[-1]), file=sys.stderr, flush=True)
        _, redos, switches = m.trouble_scores()
        fv = int(m.fv_flags().sum())
    finally:
        m.set_trouble_hook(None)
    del m
    return {"workload": "C4 mesh (19,200 elements, n in {4..10}), batched path, RK4, "
                        "evolve_redo_troubled = substep",
            "scoring_only": {"indicator": "RDMP, delta0 1e30 (no element troubled)",
                             "ms_per_step": score_ms, "steps": steps, "redos": redos0,
                             "switches": switches0},
            "redo": {"troubled_elements": marked, "at_call": 1, "ms_per_step": per_step,
                     "redos": redos, "switches": switches, "fv_elements_after": fv}}


def adapt_leg(nm):
    """Mesh adaptation between steps on the C4 mesh (refine.c:174-506,
    storage.c:501-760): every 50th element of the coarsest level (3)
    h-refined into 8 children of its n, those children coarsened back into their parent, then every
    element's n raised by one (at most 10). Each call rebuilds the host
    tables and the device mirror and interpolates the state on the device;
    wall time per call. Refining and coarsening back is exact for the
    parent's polynomial, so the state must come back to rounding."""
    m = c4_mesh(nm, "batched")
    u0 = shell_initial_data(m.coords())
    m.upload(u0)
    m.synchronize()
    ne0 = m.nelem
    lev = np.array([m.elm_info(e)[1] for e in range(ne0)])
    flag = np.zeros(ne0, np.int32)
    flag[np.where(lev == lev.min())[0][::50]] = 1     # children stay 2:1 with their neighbours
    t0 = time.perf_counter()
    m.hadapt(flag)
    m.synchronize()
    t_ref = time.perf_counter() - t0
    ne1 = m.nelem
    back = np.zeros(ne1, np.int32)
    pos = 0
    for i in range(ne0):
        if flag[i]:
            back[pos:pos + 8] = -1
            pos += 8
        else:
            pos += 1
    t0 = time.perf_counter()
    m.hadapt(back)
    m.synchronize()
    t_coarse = time.perf_counter() - t0
    err = float(np.abs(m.download() - u0).max() / np.abs(u0).max())
    n = np.array([m.elm_info(e)[2] for e in range(m.nelem)], np.int32)
    t0 = time.perf_counter()
    m.padapt(np.minimum(n + 1, 10).astype(np.int32))
    m.synchronize()
    t_p = time.perf_counter() - t0
    nodes_p = m.nnodes
    del m
    return {"workload": "C4 mesh (19,200 elements, n in {4..10}): every 50th level-3 element h-refined, "
                        "coarsened back, then n + 1 (at most 10) everywhere",
            "refine_ms": 1e3 * t_ref, "elements_after_refine": ne1,
            "coarsen_ms": 1e3 * t_coarse, "roundtrip_rel_err": err,
            "p_change_ms": 1e3 * t_p, "nodes_after_p_change": nodes_p}


def filter_leg(nm, reps=100):
    """The FILTER bin's exponential filter (alp 36, s 32, dn 0, Jacobian
    power 1) of all 5 evolved fields on the 3072-element N=7 shell, in place,
    reps back-to-back launches.  Algorithmic bytes per node: 5 doubles read
    and 5 written (the Jacobian is recomputed): 80 B."""
    m = nm.Mesh(dict(c2_params(), evolve_filter_all="yes"), device=0)
    nm.amr_setup_mesh(m)
    x = m.coords()
    m.upload(shell_initial_data(x))
    nm.evolve_filter_evosys_mesh(m)
    m.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        nm.evolve_filter_evosys_mesh(m)
    m.synchronize()
    wall = time.perf_counter() - t0
    per = wall / reps
    gbs = 80.0 * m.nnodes / per / 1e9
    out = {"workload": "exponential filter of pi, cx, cy, cz, phi on the 3072-element N=7 shell "
                       "(1,572,864 nodes), alp 36, s 32, J-weighted",
           "value": m.nnodes / per, "unit": "node-filterings/s", "us_per_launch": per * 1e6,
           "bytes_per_node": 80.0, "achieved_GBps": gbs, "hbm_frac": gbs / HBM_PEAK_GBS,
           "kernel": "filter_kernel<8,2,512>"}
    del m
    return out


def crc_leg(nm, reps=50):
    """The checkpoint's "vars" CRC-64 of the state register on the device
    (crc_elm_kernel + crc_xor_kernel), on the config-4 AMR shell (19,200
    elements, N in 4..10, 8,213,159 nodes: 328.5 MB of doubles) and the
    3072-element N=7 shell (62.9 MB).  Algorithmic bytes: the state once,
    40 B per node.  Wall time per call includes the launch, the stream
    synchronisation and the 8-byte read-back; the kernel's own duration is in
    the rocprofv3 summary.  The host CRC of the same bytes (table-driven,
    one core) is timed beside it on a 64 MB sample."""
    from nmesh_amd import checkpoint as ck
    out = {}
    for name, m in (("amr_shell_19200", c4_mesh(nm, "batched")),
                    ("shell_3072_n7", nm.Mesh(c2_params(), device=0))):
        if name == "shell_3072_n7":
            nm.amr_setup_mesh(m)
        m.upload(np.random.default_rng(9).standard_normal(5 * m.nnodes))
        ck.vars_crc(m)
        t0 = time.perf_counter()
        for _ in range(reps):
            c = ck.vars_crc(m)
        per = (time.perf_counter() - t0) / reps
        nbytes = 40.0 * m.nnodes
        out[name] = {"nodes": m.nnodes, "bytes": nbytes, "us_per_call": per * 1e6,
                     "GBps_wall": nbytes / per / 1e9, "crc": "%016x" % c}
        del m
    buf = np.random.default_rng(10).integers(0, 256, 64 << 20, dtype=np.uint8).tobytes()
    t0 = time.perf_counter()
    ck.crc64(buf)
    host = time.perf_counter() - t0
    out["host_crc64_GBps_1core"] = len(buf) / host / 1e9
    out["workload"] = ("vars CRC-64 (reflected ECMA-182) of a state register on the device, "
                       "nmesh's checkpoint order (per element phi, pi, cx, cy, cz)")
    out["kernel"] = "crc_elm_kernel"
    return out


def adjoint_leg(nm, steps=256, every=16, reps=3):
    """Discrete adjoint cost against the forward step.  C5: the config-1 box
    (64 elements, N=4, 8,000 nodes), 256 RK4 steps, psi = exp(-r^2) with
    Phi_i = d_i psi; device time of nmx_evolve(256) and nmx_adjoint_evolve(256)
    (HIP events), and wall time of the torch.autograd forward (checkpoints
    every 16 steps) and its backward.  Also the adjoint/forward device-time
    ratio per step on the 3072-element N=7 shell."""
    import torch
    from nmesh_amd import adjoint as adj
    out = {}
    m = nm.Mesh(c1_params(), device=0)
    nm.amr_setup_mesh(m)
    x = m.coords()
    r2 = (x ** 2).sum(0)
    psi = np.exp(-r2)
    u0 = np.zeros((5, m.nnodes))
    u0[4] = psi
    u0[1:4] = -2 * x * psi
    u0 = u0.ravel()
    rng = np.random.default_rng(2026)
    lam0 = torch.from_numpy(rng.standard_normal(u0.shape)).cuda()
    u0_t = torch.from_numpy(u0).cuda()
    m.upload(u0)
    nm.evolve_myln(m, every)                    # autotune, graph capture
    adj.adjoint_evolve(m, lam0, every)
    fwd, bwd = [], []
    for _ in range(reps):
        m.upload(u0)
        nm.evolve_myln(m, steps)
        fwd.append(m.last_evolve_ms())
        adj.adjoint_evolve(m, lam0, steps)
        bwd.append(m.last_adjoint_ms())
    out["forward_ms"] = min(fwd)
    out["adjoint_ms"] = min(bwd)
    out["adjoint_vs_forward"] = min(bwd) / min(fwd)
    bnd = adj.outer_boundary_nodes(m)
    target = torch.zeros(bnd.numel(), dtype=torch.float64, device="cuda")
    tf, tb = [], []
    for _ in range(reps):
        u = u0_t.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        traj = adj.evolve(m, u, steps, every)
        loss = ((traj[-1][4 * m.nnodes:][bnd] - target) ** 2).sum()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        (g,) = torch.autograd.grad(loss, u)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        tf.append(t1 - t0)
        tb.append(t2 - t1)
    out["autograd_forward_ms"] = 1e3 * min(tf)
    out["autograd_backward_ms"] = 1e3 * min(tb)
    out["autograd_backward_vs_forward"] = min(tb) / min(tf)
    del m
    # the same ratio on the 3072-element N=7 shell (HBM-bound kernels)
    m = nm.Mesh(c2_params(), device=0)
    nm.amr_setup_mesh(m)
    xs = m.coords()
    us = shell_initial_data(xs)
    m.upload(us)
    lam = torch.from_numpy(rng.standard_normal(us.shape)).cuda()
    nm.evolve_myln(m, 8)
    adj.adjoint_evolve(m, lam, 8)
    nm.evolve_myln(m, 32)
    f2 = m.last_evolve_ms()
    adj.adjoint_evolve(m, lam, 32)
    b2 = m.last_adjoint_ms()
    out["shell_forward_ms_per_step"] = f2 / 32
    out["shell_adjoint_ms_per_step"] = b2 / 32
    out["shell_adjoint_vs_forward"] = b2 / f2
    del m
    # and on the C4 mixed-order AMR shell (the batched path's adjoint)
    m = c4_mesh(nm, "batched")
    uc = shell_initial_data(m.coords())
    m.upload(uc)
    lam = torch.from_numpy(rng.standard_normal(uc.shape)).cuda()
    nm.evolve_myln(m, 8)
    adj.adjoint_evolve(m, lam, 8)
    nm.evolve_myln(m, 16)
    f3 = m.last_evolve_ms()
    adj.adjoint_evolve(m, lam, 16)
    b3 = m.last_adjoint_ms()
    out["c4_forward_ms_per_step"] = f3 / 16
    out["c4_adjoint_ms_per_step"] = b3 / 16
    out["c4_adjoint_vs_forward"] = b3 / f3
    del m
    out["workload"] = ("C5: config-1 box (64 elements, N=4, 8000 nodes), upwind, zero exterior, "
                       "%d RK4 steps, checkpoints every %d steps; plus 32 RK4 steps of the "
                       "3072-element N=7 shell and 16 of the C4 mixed-order AMR shell" % (steps, every))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200, help="timed RK4 steps of the headline (C4 batched)")
    ap.add_argument("--warmup", type=int, default=20, help="untimed RK4 steps before them")
    ap.add_argument("--shape-steps", type=int, default=40,
                    help="RK4 steps timed for each other batched launch shape (0: skip)")
    ap.add_argument("--perelement-steps", type=int, default=1,
                    help="RK4 steps of the per-element launch leg (0: skip)")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain uniform-shell leg")
    ap.add_argument("--plain-steps", type=int, default=400)
    ap.add_argument("--plain-warmup", type=int, default=40)
    ap.add_argument("--plain-n", type=int, default=8, help="points per direction of the plain leg")
    ap.add_argument("--only-plain", action="store_true", help="run only the plain uniform-shell leg")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--amr-steps", type=int, default=40, help="RK4 steps of the batched AMR leg")
    ap.add_argument("--only-amr", action="store_true", help="run only the batched AMR leg (profiling)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint-cost leg")
    ap.add_argument("--no-fv", action="store_true", help="skip the FV (SR-Euler WENO5) leg")
    ap.add_argument("--only-fv", action="store_true", help="run only the FV leg")
    ap.add_argument("--fv-steps", type=int, default=50, help="sspRK3 steps of the FV leg")
    ap.add_argument("--only-adjoint", action="store_true", help="run only the adjoint-cost leg")
    ap.add_argument("--only-filter", action="store_true", help="run only the exponential-filter leg")
    ap.add_argument("--only-fvelm", action="store_true", help="run only the FV-element leg")
    ap.add_argument("--only-crc", action="store_true", help="run only the checkpoint CRC leg")
    ap.add_argument("--only-trouble", action="store_true", help="run only the troubled-element leg")
    ap.add_argument("--no-trouble", action="store_true", help="skip the troubled-element leg")
    ap.add_argument("--only-adapt", action="store_true", help="run only the mesh-adaptation leg")
    ap.add_argument("--no-adapt", action="store_true", help="skip the mesh-adaptation leg")
    ap.add_argument("--no-crc", action="store_true", help="skip the checkpoint CRC leg")
    ap.add_argument("--no-fvelm", action="store_true", help="skip the FV-element leg")
    ap.add_argument("--no-filter", action="store_true", help="skip the exponential-filter leg")
    ap.add_argument("--path", default="plain")
    ap.add_argument("--graph", default="yes", help="capture RK steps in hipGraphs (yes|no)")
    ap.add_argument("--geometry", default="analytic", help="dXdx recomputed (analytic) or streamed (stored)")
    ap.add_argument("--xcd", default="auto", help="contiguous element ranges per XCD (yes|no|auto)")
    ap.add_argument("--block", default="auto",
                    help="nodes per thread: full (1), half (2), quarter (4) or auto (autotuned)")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="instead of the workload, run 10 device copies of MIB MiB with "
                         "torch (counter calibration for rocprofv3 --pmc) and exit")
    ap.add_argument("--amr-par", action="append", default=[], metavar="NAME=VALUE",
                    help="a library parameter of the C4 meshes (repeatable)")
    ap.add_argument("--fv-par", action="append", default=[], metavar="NAME=VALUE",
                    help="a library parameter of the FV grid (repeatable)")
    args = ap.parse_args()
    for kv in args.amr_par:
        k, v = kv.split("=", 1)
        C4_EXTRA[k] = v
    for kv in args.fv_par:
        k, v = kv.split("=", 1)
        FV_EXTRA[k] = v
    if args.copy_calib:
        import torch
        n = args.copy_calib * 1024 * 1024 // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calib_mib": args.copy_calib, "copies": 10}))
        return
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: the hot path is measured on one "
              "GPU's share of the mesh; multi-GPU runs are not part of this build.",
              file=sys.stderr)
        sys.exit(2)

    import nmesh_amd as nm
    try:
        import torch
        have_torch = torch.cuda.is_available()
    except Exception:
        have_torch = False

    if args.only_amr:
        out = batched_leg(nm, args.amr_steps, 0) if args.amr_steps else {}
        print(json.dumps(out), flush=True)
        return
    if args.only_plain:
        print(json.dumps(plain_leg(nm, args, have_torch)), flush=True)
        return
    if args.only_adjoint:
        print(json.dumps(adjoint_leg(nm)), flush=True)
        return
    if args.only_filter:
        print(json.dumps(filter_leg(nm)), flush=True)
        return
    if args.only_fvelm:
        print(json.dumps(fvelm_leg(nm)), flush=True)
        return
    if args.only_crc:
        print(json.dumps(crc_leg(nm)), flush=True)
        return
    if args.only_trouble:
        print(json.dumps(trouble_leg(nm)), flush=True)
        return
    if args.only_adapt:
        print(json.dumps(adapt_leg(nm)), flush=True)
        return
    if args.only_fv:
        print(json.dumps(fv_leg(args.fv_steps, cpu=not args.no_cpu)), flush=True)
        return
    # the headline: C4 on the batched path
    h = c4_headline(nm, args.steps, args.warmup, have_torch)
    mh = h.pop("m")
    shapes = {h["shape"]: 1e3 * h["wall"] / args.steps}
    if args.shape_steps:
        for shape in ("persistent", "buckets", "legacy"):
            if shape in shapes:
                continue
            mh.set_par("nmx_batch", shape)
            nm.evolve_myln(mh, 1)
            mh.synchronize()
            t0 = time.perf_counter()
            nm.evolve_myln(mh, args.shape_steps)
            mh.synchronize()
            shapes[shape] = 1e3 * (time.perf_counter() - t0) / args.shape_steps
    del mh
    nn = h["nodes"]
    value = 4.0 * args.steps * nn / h["wall"]
    traffic = load_traffic("traffic_dg_batched_c4.json")
    out = {
        "metric": "DG RHS node-updates/s (FP64)",
        "value": value,
        "unit": "node-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * h["wall"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "scalarwave1 DG, mixed-order AMR cubed-sphere shell r in [2,20], "
                               "levels 3-5 (19,200 elements), n uniform in {4..10} points per "
                               "direction (%d nodes), upwind, outgoing, RK4, batched order "
                               "buckets, hipGraph-captured steps" % nn,
                   "elements": h["elements"], "nodes": nn, "dt": h["dt"], "path": "batched",
                   "launch_shape": h["shape"], "order_buckets": h["buckets"]},
        "roofline": {"bound": "hbm", "achieved": h["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": h["achieved"] / HBM_PEAK_GBS,
                     "traffic": traffic["bytes_per_launch"] if traffic else None,
                     "traffic_source": ("profile figure: profiles/%s, rocprofv3 --pmc FETCH_SIZE and "
                                        "WRITE_SIZE summed over one stage's launches, calibrated"
                                        % traffic["summary"]) if traffic else None,
                     "kernel": "one RK4 stage of the batched path (%s launch shape: every order "
                               "bucket's launches of the stage)" % h["shape"],
                     "avg_launch_us": h["stage_us"],
                     "bytes_per_node_update": BYTES_PER_NU_CURV,
                     "min_bytes_per_node_update": MIN_BYTES_PER_NU,
                     "achieved_min": h["achieved_min"], "frac_min": h["achieved_min"] / HBM_PEAK_GBS,
                     "fp64": {"flops_per_node_update": "24 n + 36", "flops_per_stage": h["flops_per_stage"],
                              "achieved_TFLOPs": h["tflops"], "peak_TFLOPs": FP64_PEAK_TFLOPS,
                              "frac": h["tflops"] / FP64_PEAK_TFLOPS}},
        "batched_shapes_ms_per_step": shapes,
    }
    if args.perelement_steps:
        mp = c4_mesh(nm, "perelement")
        mp.upload(shell_initial_data(mp.coords()))
        nm.evolve_myln(mp, 1)
        mp.synchronize()
        t0 = time.perf_counter()
        nm.evolve_myln(mp, args.perelement_steps)
        mp.synchronize()
        wall_pe = time.perf_counter() - t0
        del mp
        pe = 4.0 * args.perelement_steps * nn / wall_pe
        out["perelement"] = {"value": pe, "unit": "node-updates/s", "steps": args.perelement_steps,
                             "ms_per_step": 1e3 * wall_pe / args.perelement_steps,
                             "speedup_batched_vs_perelement": value / pe}
    if not args.no_plain:
        out["plain_shell"] = plain_leg(nm, args, have_torch)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_leg(nm)
    if not args.no_fv:
        out["fv_sr_euler"] = fv_leg(args.fv_steps, cpu=not args.no_cpu)
    if not args.no_filter:
        out["filter"] = filter_leg(nm)
    if not args.no_fvelm:
        out["fv_elements"] = fvelm_leg(nm)
    if not args.no_crc:
        out["checkpoint_crc"] = crc_leg(nm)
    if not args.no_adapt:
        out["adaptation"] = adapt_leg(nm)
    if not args.no_trouble:
        try:
            out["troubled_elements"] = trouble_leg(nm)
        except nm.NmxError as exc:    # recorded in the line; the headline stands
            out["troubled_elements"] = {"error": str(exc)}
    if not args.no_cpu:
        out["cpu_baseline"] = c4_cpu_baseline()
    print(json.dumps(out), flush=True)


def plain_leg(nm, args, have_torch):
    """The plain per-order kernel on the uniform N=7 shell (6x8^3 elements,
    n = args.plain_n points per direction), graph-captured RK4 steps."""
    import torch
    p = c2_params()
    if args.plain_n != 8:
        p.update({"amr_n0": args.plain_n, "amr_n1": args.plain_n, "amr_n2": args.plain_n})
    m = nm.Mesh(p, device=0)
    m.set_par("nmx_path", args.path)
    m.set_par("nmx_graph", args.graph)
    m.set_par("nmx_geometry", args.geometry)
    if args.xcd != "auto":
        m.set_par("nmx_xcd", args.xcd)
    if args.block != "auto":
        m.set_par("nmx_block", args.block)
    nm.amr_setup_mesh(m)
    x = m.coords()
    m.upload(shell_initial_data(x))
    nn = m.nnodes
    if args.block == "auto" and args.xcd == "auto":
        nm.autotune(m)            # times tile size / XCD mapping on this mesh
    if args.plain_warmup:
        nm.evolve_myln(m, args.plain_warmup)
    m.synchronize()
    if have_torch:
        torch.cuda.synchronize()
    t0 = time.perf_counter()
    nm.evolve_myln(m, args.plain_steps)
    m.synchronize()
    if have_torch:
        torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    dev_ms = m.last_evolve_ms()    # HIP events on the library's stream around the steps
    launches = 4 * args.plain_steps
    avg_launch_s = dev_ms * 1e-3 / launches
    # algorithmic bytes per node-update of a fused RHS + RK4 stage: read w,
    # u, u_p and write u, w (25 doubles) plus 9 doubles of dXdx per node on
    # curvilinear elements, face-neighbour re-reads counted 0: 272 B
    achieved = BYTES_PER_NU_CURV * nn / avg_launch_s / 1e9
    # the least this kernel must move: stage 1 reads u and writes u_acc, w
    # (15 doubles), stages 2-3 read w, u_acc, u_p and write u_acc, w (25),
    # stage 4 reads w, u_acc and writes u_acc (15); dXdx recomputed (0) or
    # streamed (9)
    g = 9.0 if args.geometry == "stored" else 0.0
    min_bytes_per_nu = 8.0 * ((15 + 25 + 25 + 15) / 4.0 + g)
    achieved_min = min_bytes_per_nu * nn / avg_launch_s / 1e9
    tflops = flops_per_nu(args.plain_n) * nn / avg_launch_s / 1e12
    traffic = load_traffic("traffic_dg_stage_shell_n8.json") if args.plain_n == 8 else None
    out = {"workload": "scalarwave1 DG, cubed-sphere shell r in [2,20], 6x8^3 elements, "
                       "n=%d (%d nodes), upwind, outgoing, RK4, fused RHS+stage" % (args.plain_n, nn),
           "value": 4.0 * args.plain_steps * nn / wall, "unit": "node-updates/s",
           "steps": args.plain_steps, "warmup": args.plain_warmup,
           "ms_per_step": 1e3 * wall / args.plain_steps, "elements": m.nelem, "nodes": nn,
           "dt": m.dt, "path": args.path,
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS,
                        "traffic": traffic["bytes_per_launch"] if traffic else None,
                        "traffic_source": "profile figure: profiles/dg_stage_kernel_shell_n8_rk4.md" if traffic else None,
                        "kernel": "dg_stage_kernel, n=%d, %s tile, %s geometry" % (args.plain_n, m.par("nmx_block"),
                                                                                    args.geometry),
                        "tile": m.par("nmx_block"), "xcd_map": m.par("nmx_xcd"),
                        "avg_launch_us": avg_launch_s * 1e6,
                        "bytes_per_node_update": BYTES_PER_NU_CURV,
                        "min_bytes_per_node_update": min_bytes_per_nu,
                        "achieved_min": achieved_min, "frac_min": achieved_min / HBM_PEAK_GBS,
                        "fp64": {"flops_per_node_update": flops_per_nu(args.plain_n),
                                 "achieved_TFLOPs": tflops, "frac": tflops / FP64_PEAK_TFLOPS}}}
    del m
    return out

if __name__ == "__main__":
    main()
