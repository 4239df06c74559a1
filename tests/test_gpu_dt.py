"""The time step's device reductions:

- nmx_mesh_dt_device (per-element find_hmin from the device's node
  coordinates, minimised with wavefront shuffles) against the mesh's dt set
  up on the host and the oracle's, on the C1 box, the C2 shell, a refined
  mixed-order shell, a mesh mixing DG and FV patches and elements of two
  points per direction; within one rounding (both sides evaluate the same
  distances without fused multiply-adds);
- nmx_fv_max_speed (the CFL reduction over the SR-Euler characteristic
  speeds) against numpy on the c2p primitives, and against the relativistic
  velocity addition (v + c_s) / (1 + v c_s) on a uniform flow.
"""
import numpy as np
import pytest

import nmesh_amd as nm
from oracle.orc import OMesh

pytestmark = pytest.mark.gpu


def dg_case(which):
    if which == "c1":
        pars = {"amr_mesh_type": "BoxMesh", "amr_BoxMesh_dout": 8.0, "amr_luni": 2,
                "amr_n0": 5, "amr_n1": 5, "amr_n2": 5}
        om = OMesh("box", 2, 5, dout=8.0)
    elif which == "c2":
        pars = {"amr_mesh_type": "Shell", "amr_Shell_rin": 2.0, "amr_Shell_rout": 20.0, "amr_luni": 3,
                "amr_n0": 8, "amr_n1": 8, "amr_n2": 8}
        om = OMesh("shell", 3, 8, rin=2.0, rout=20.0)
    elif which == "refined_mixed":
        pars = {"amr_mesh_type": "Shell", "amr_Shell_rin": 2.0, "amr_Shell_rout": 20.0, "amr_luni": 1,
                "amr_n0": 4, "amr_n1": 4, "amr_n2": 4, "amr_hrefine_sphere_radius": 7,
                "amr_hrefine_sphere_levels": 1}
        om = OMesh("shell", 1, 4, rin=2.0, rout=20.0, radius=7, levels=1)
    elif which == "n2":
        pars = {"amr_mesh_type": "BoxMesh", "amr_BoxMesh_dout": 1.0, "amr_luni": 1,
                "amr_n0": 2, "amr_n1": 2, "amr_n2": 2}
        om = OMesh("box", 1, 2, dout=1.0)
    else:   # DG and FV patches on the shell
        pars = {"amr_mesh_type": "Shell", "amr_Shell_rin": 2.0, "amr_Shell_rout": 6.0, "amr_luni": 1,
                "amr_n0": 5, "amr_n1": 5, "amr_n2": 5, "amr_uniform_p": "0 2 4", "amr_fv_p": "0 2 4"}
        om = OMesh("shell", 1, 5, rin=2.0, rout=6.0, fv_patches=[0, 2, 4])
    m = nm.Mesh(pars, device=0)
    nm.amr_setup_mesh(m)
    if which == "refined_mixed":
        nvec = np.random.default_rng(3).integers(3, 10, om.nelem).astype(np.int32)
        m.set_elm_n(nvec)
        om.set_n(nvec)
    return m, om


@pytest.mark.parametrize("which", ["c1", "c2", "refined_mixed", "n2", "dg_fv"])
def test_device_dt_matches_host_and_oracle(which):
    print("device dt", which, flush=True)
    m, om = dg_case(which)
    host = m.dt
    dev = m.dt_device()
    ref = om.dt()
    assert host == ref
    assert abs(dev - host) <= 2.3e-16 * host, (dev, host)
    if which == "c1":
        assert abs(dev - 0.172673) < 5e-7   # the reference's stated C1 dt


def sr_speeds(P, gam=2.0):
    rho, v, p = P[0], P[1:4], P[4]
    v2 = (v ** 2).sum(0)
    h = 1 + gam / (gam - 1) * p / rho
    cs2 = gam * p / (rho * h)
    cs = np.sqrt(cs2)
    out = 0.0
    for d in range(3):
        vd = v[d]
        root = cs * np.sqrt((1 - v2) * (1 - v2 * cs2 - vd * vd * (1 - cs2)))
        lm = (vd * (1 - cs2) - root) / (1 - v2 * cs2)
        lp = (vd * (1 - cs2) + root) / (1 - v2 * cs2)
        out = max(out, np.abs(lm).max(), np.abs(lp).max())
    return out


def test_fv_max_speed_matches_numpy():
    print("fv max speed", flush=True)
    from nmesh_amd.fv import FvGrid, blob_initial_data
    N = 32
    g = FvGrid({"fv_n": N}, device=0)
    g.setup()
    g.upload_prim(blob_initial_data(N, seed=21, nblobs=6))
    a = g.max_speed()
    P = g.download(1).reshape(5, -1)
    ref = sr_speeds(P)
    assert abs(a - ref) <= 1e-14 * ref
    assert 0 < a < 1


def test_fv_max_speed_of_a_uniform_flow_is_the_velocity_addition():
    print("fv max speed uniform", flush=True)
    from nmesh_amd.fv import FvGrid
    N = 16
    g = FvGrid({"fv_n": N}, device=0)
    g.setup()
    rho, vx, p, gam = 1.3, 0.4, 0.7, 2.0
    P = np.zeros((5, N ** 3))
    P[0] = rho
    P[1] = vx
    P[4] = p
    g.upload_prim(P.ravel())
    h = 1 + gam / (gam - 1) * p / rho
    cs = np.sqrt(gam * p / (rho * h))
    expect = (vx + cs) / (1 + vx * cs)     # along x, the largest
    assert abs(g.max_speed() - expect) <= 1e-14
