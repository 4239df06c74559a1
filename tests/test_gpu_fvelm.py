"""GPU parity of scalarwave1 on FV elements (fv_divf with the reference's
default FV parameters; libnmx fvelm_stage_kernel) against the oracle.

The kernel and the oracle evaluate the same reconstructions and metric
factors; the numerical flux is the closed form on the device and the
R-matrix / LLF form in the oracle, so the RHS agrees to ~1e-14 relative.
The bar is 1e-12 relative to the field maximum, per RHS and per step.
Cartesian boxes and cubed-sphere shells (metric recomputed on the device at
nodes and midpoints), coordinates_midpoint_data no and yes; meshes with FV
elements in some patches and DG in the others (order buckets of both kinds,
DG/FV faces interpolated), h-refined ones among them.
"""
import numpy as np
import pytest

import nmesh_amd as nm
from oracle import orc
from oracle.orc import OMesh

pytestmark = pytest.mark.gpu
K = [0.8, -0.5, 0.33166247903554]
TOL = 1e-12


def make(luni, n, rec, numflux="upwind", k=K, graph="yes", kind="box", mid="no", extra=None):
    pars = {"amr_mesh_type": "BoxMesh" if kind == "box" else "Shell", "amr_luni": luni,
            "amr_n0": n, "amr_n1": n, "amr_n2": n, "amr_BoxMesh_dout": 2.0,
            "amr_Shell_rin": 2.0, "amr_Shell_rout": 6.0,
            "amr_uniform_p": "0" if kind == "box" else "0 1 2 3 4 5",
            "amr_fv_p": "0" if kind == "box" else "0 1 2 3 4 5", "fv_rec": rec,
            "coordinates_midpoint_data": mid,
            "scalarwave1_numflux": numflux, "scalarwave1_k": "%.17g %.17g %.17g" % tuple(k),
            "nmx_graph": graph, "nmx_graph_steps": 2}
    pars.update(extra or {})
    m = nm.Mesh(pars, device=0)
    nm.amr_setup_mesh(m)
    orc.set_fv_rec(rec)
    orc.set_fv_midpoint(mid != "no")
    om = OMesh(kind, luni, n, dout=2.0, rin=2.0, rout=6.0, fv=True)
    return m, om


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("rec", ["WENOm3_2", "WENOm5_2", "WENOmZ_2", "WENOmT_2"])
@pytest.mark.parametrize("numflux", ["upwind", "LLF"])
@pytest.mark.parametrize("luni,n", [(1, 5), (2, 8), (1, 10)])
def test_fv_element_rhs_matches_oracle(rec, numflux, luni, n):
    print("fvelm rhs", rec, numflux, luni, n, flush=True)
    m, om = make(luni, n, rec, numflux)
    u, _ = om.profile(K, 0.2)
    rng = np.random.default_rng(n)
    u = u + 0.1 * rng.uniform(-1, 1, u.shape)
    m.upload(u)
    r = nm.evolve_setrhs_mesh(m, t=0.2)
    r_o = om.rhs(u, 0.2, K, numflux)
    assert rel(r, r_o) <= TOL


@pytest.mark.parametrize("rec", ["WENOm3_2", "WENOmZ_2"])
def test_fv_elements_of_3_points_match_oracle(rec):
    """the smallest FV element (3 points per direction: one interior point,
    both ends extrapolated from it), alone and next to DG elements"""
    print("fvelm n=3", rec, flush=True)
    m, om = make(2, 3, rec)
    u, _ = om.profile(K, 0.2)
    u = u + 0.1 * np.random.default_rng(3).uniform(-1, 1, u.shape)
    m.upload(u)
    assert rel(nm.evolve_setrhs_mesh(m, t=0.2), om.rhs(u, 0.2, K, "upwind")) <= TOL
    m, om = make_mixed(3, rec, levels=1)
    u, _ = om.profile(K, 0.2)
    u = u + 0.1 * np.random.default_rng(4).uniform(-1, 1, u.shape)
    m.upload(u)
    r = nm.evolve_setrhs_mesh(m, t=0.2)
    r_o = om.rhs(u, 0.2, K, "upwind")
    orc.set_fv_midpoint(False)
    assert rel(r, r_o) <= TOL


@pytest.mark.parametrize("s1,s2,opt", [(0.5, None, 0), (0.0, 1.2, 1), (0.9, 3.0, 2), (1.0, None, -1)])
@pytest.mark.parametrize("where", ["box", "mixed_refined"])
def test_fv_divf_extrap_limits_match_oracle(s1, s2, opt, where):
    """fv_divf_extrap_s1 / _s2 / _opt (the kernels' limited dnfn_extrap1,
    fvelm_ex_kernels.hip) against the oracle's, on an FV box and on an
    h-refined mixed shell; the limit changes the RHS"""
    print("fv extrap limits", s1, s2, opt, where, flush=True)
    extra = {"fv_divf_extrap_s1": s1, "fv_divf_extrap_s2": "DBL_MAX" if s2 is None else s2,
             "fv_divf_extrap_opt": opt}
    if where == "box":
        m, om = make(1, 6, "WENOmZ_2")
        for k, v in extra.items():
            m.set_par(k, v)
    else:
        m, om = make_mixed(5, "WENOm5_2", levels=1, extra=extra)
    u, _ = om.profile(K, 0.2)
    u = u + 0.2 * np.random.default_rng(11).uniform(-1, 1, u.shape)
    m.upload(u)
    r = nm.evolve_setrhs_mesh(m, t=0.2)
    r_plain = om.rhs(u, 0.2, K, "upwind")
    orc.set_fv_divf_extrap(s1, s2, opt)
    try:
        r_o = om.rhs(u, 0.2, K, "upwind")
    finally:
        orc.set_fv_divf_extrap()
        orc.set_fv_midpoint(False)
    assert rel(r, r_o) <= TOL
    assert rel(r_plain, r_o) > 1e-6


@pytest.mark.parametrize("rec,s1,s2,opt", [("WENOm3_1", 0.0, None, 0), ("WENOm5_1", 0.0, None, 0),
                                           ("WENOm5", 0.5, None, 0), ("WENOmZ", 0.3, 3.0, 1),
                                           ("WENOm3", 1e300, None, 0), ("WENOmT_2", 0.0, None, 0),
                                           ("1", 0.0, None, 0)])
@pytest.mark.parametrize("where", ["box", "mixed_refined"])
def test_fv_rec_end_modes_match_oracle(rec, s1, s2, opt, where):
    """fv_rec's _1 modes and WENOm3/5/Z with fv_rec_WENOm_s1/_s2/_opt (the
    end midpoints, and dnfn_extrap1 skipped where the end point was kept)
    against the oracle; plus an RK4 step on the box"""
    print("fv rec ends", rec, s1, s2, opt, where, flush=True)
    extra = {"fv_rec_WENOm_s1": s1, "fv_rec_WENOm_s2": "DBL_MAX" if s2 is None else s2,
             "fv_rec_WENOm_opt": opt}
    if where == "box":
        m, om = make(1, 6, rec, extra=extra)
    else:
        m, om = make_mixed(5, rec, levels=1, extra=extra)
    orc.set_fv_rec(rec, s1, s2, opt)
    try:
        u, _ = om.profile(K, 0.2)
        u = u + 0.2 * np.random.default_rng(12).uniform(-1, 1, u.shape)
        m.upload(u)
        r = nm.evolve_setrhs_mesh(m, t=0.2)
        assert rel(r, om.rhs(u, 0.2, K, "upwind")) <= TOL
        if where == "box":
            m.time = 0.2
            v, _ = om.evolve(u, m.dt, 1, 0.2, K, "RK4")
            nm.evolve_myln(m, 1)
            assert rel(m.download(), v) <= TOL
    finally:
        orc.set_fv_rec("WENOm3_2")
        orc.set_fv_midpoint(False)


@pytest.mark.parametrize("method", ["RK4", "sspRK3", "Euler"])
@pytest.mark.parametrize("rec", ["WENOm3_2", "WENOm5_2", "WENOmZ_2"])
def test_fv_element_steps_match_oracle(method, rec):
    print("fvelm steps", method, rec, flush=True)
    m, om = make(1, 6, rec)
    m.set_par("evolve_method", method)
    u, _ = om.profile(K, 0.0)
    m.upload(u)
    t = 0.0
    for s in range(3):
        u, t = om.evolve(u, m.dt, 1, t, K, method)
        nm.evolve_myln(m, 1)
        assert rel(m.download(), u) <= TOL, "step %d" % s
    # captured graphs give the same states
    m2, _ = make(1, 6, rec)
    m2.set_par("evolve_method", method)
    m2.upload(om.profile(K, 0.0)[0])
    nm.evolve_myln(m2, 3)
    assert np.array_equal(m2.download(), m.download())


@pytest.mark.parametrize("mid", ["no", "yes"])
@pytest.mark.parametrize("rec", ["WENOm3_2", "WENOm5_2", "WENOmZ_2"])
@pytest.mark.parametrize("numflux,n,geometry", [("upwind", 6, "analytic"), ("LLF", 9, "stored")])
def test_fv_shell_rhs_matches_oracle(mid, rec, numflux, n, geometry):
    print("fvelm shell rhs", mid, rec, numflux, n, geometry, flush=True)
    m, om = make(1, n, rec, numflux, kind="shell", mid=mid, extra={"nmx_geometry": geometry})
    u, _ = om.profile(K, 0.2)
    u = u + 0.1 * np.random.default_rng(n).uniform(-1, 1, u.shape)
    m.upload(u)
    r = nm.evolve_setrhs_mesh(m, t=0.2)
    r_o = om.rhs(u, 0.2, K, numflux)
    orc.set_fv_midpoint(False)
    assert rel(r, r_o) <= TOL


@pytest.mark.parametrize("mid", ["no", "yes"])
def test_fv_shell_steps_match_oracle(mid):
    """RK4 on FV shell elements, checked every step; the midpoint mode is
    read at each launch, so switching it on an existing mesh takes effect"""
    print("fvelm shell steps", mid, flush=True)
    m, om = make(1, 7, "WENOm5_2", kind="shell", mid="no" if mid == "yes" else "yes")
    m.set_par("coordinates_midpoint_data", mid)
    orc.set_fv_midpoint(mid == "yes")
    u, _ = om.profile(K, 0.0)
    m.upload(u)
    t = 0.0
    try:
        for s in range(3):
            u, t = om.evolve(u, m.dt, 1, t, K, "RK4")
            nm.evolve_myln(m, 1)
            assert rel(m.download(), u) <= TOL, "step %d" % s
    finally:
        orc.set_fv_midpoint(False)


def make_mixed(n, rec, mid="yes", numflux="upwind", levels=0, pats=(0, 2, 4), extra=None):
    """the shell with FV elements in patches pats and DG elsewhere (the
    reference's sw7FV-DG layout for 0 2 4), optionally h-refined"""
    ps = " ".join(str(p) for p in pats)
    pars = {"amr_mesh_type": "Shell", "amr_luni": 1, "amr_n0": n, "amr_n1": n, "amr_n2": n,
            "amr_Shell_rin": 2.0, "amr_Shell_rout": 6.0, "amr_hrefine_sphere_radius": 4,
            "amr_hrefine_sphere_levels": levels, "amr_uniform_p": ps, "amr_fv_p": ps,
            "fv_rec": rec, "coordinates_midpoint_data": mid, "scalarwave1_numflux": numflux,
            "scalarwave1_k": "%.17g %.17g %.17g" % tuple(K), "nmx_graph_steps": 2}
    pars.update(extra or {})
    m = nm.Mesh(pars, device=0)
    nm.amr_setup_mesh(m)
    orc.set_fv_rec(rec)
    orc.set_fv_midpoint(mid != "no")
    om = OMesh("shell", 1, n, rin=2.0, rout=6.0, radius=4, levels=levels, fv_patches=list(pats))
    return m, om


@pytest.mark.parametrize("mid", ["no", "yes"])
@pytest.mark.parametrize("rec,numflux,n,geometry", [("WENOm3_2", "upwind", 6, "analytic"),
                                                    ("WENOm5_2", "upwind", 7, "analytic"),
                                                    ("WENOmZ_2", "LLF", 9, "stored")])
def test_mixed_dg_fv_rhs_matches_oracle(mid, rec, numflux, n, geometry):
    """DG and FV elements side by side: FV order buckets next to DG ones,
    DG/FV faces interpolated each stage (Lagrange into FV elements,
    piecewise linear into DG elements)"""
    print("mixed dg/fv rhs", mid, rec, numflux, n, geometry, flush=True)
    m, om = make_mixed(n, rec, mid, numflux, extra={"nmx_geometry": geometry})
    u, _ = om.profile(K, 0.2)
    u = u + 0.1 * np.random.default_rng(n).uniform(-1, 1, u.shape)
    m.upload(u)
    r = nm.evolve_setrhs_mesh(m, t=0.2)
    r_o = om.rhs(u, 0.2, K, numflux)
    orc.set_fv_midpoint(False)
    assert rel(r, r_o) <= TOL


@pytest.mark.parametrize("case", ["refined", "all_fv_batched", "fv_in_1_3_5"])
def test_fv_buckets_rhs_matches_oracle(case):
    """FV order buckets elsewhere: an h-refined mixed shell (FV/FV and DG/FV
    faces with several neighbours), every element FV on the batched path,
    and the other half of the patches FV"""
    print("fv buckets rhs", case, flush=True)
    if case == "refined":
        m, om = make_mixed(5, "WENOm5_2", levels=1)
    elif case == "all_fv_batched":
        m, om = make_mixed(6, "WENOm5_2", pats=(0, 1, 2, 3, 4, 5), extra={"nmx_path": "batched"})
    else:
        m, om = make_mixed(6, "WENOm3_2", mid="no", pats=(1, 3, 5))
    u, _ = om.profile(K, 0.2)
    u = u + 0.1 * np.random.default_rng(5).uniform(-1, 1, u.shape)
    m.upload(u)
    r = nm.evolve_setrhs_mesh(m, t=0.2)
    r_o = om.rhs(u, 0.2, K, "upwind")
    orc.set_fv_midpoint(False)
    assert rel(r, r_o) <= TOL


@pytest.mark.parametrize("mode", ["linear Lagrange_if_dg", "parabolic Lagrange_if_dg"])
def test_mixed_lagrange_if_dg_matches_oracle(mode):
    """fv_surface_interp with Lagrange_if_dg: DG elements take their FV
    neighbours' faces by Lagrange rows on the uniform points, FV elements
    keep the piecewise rows, on an h-refined mixed shell"""
    print("mixed", mode, flush=True)
    m, om = make_mixed(5, "WENOm5_2", levels=1, extra={"fv_surface_interp": mode})
    u, _ = om.profile(K, 0.2)
    u = u + 0.1 * np.random.default_rng(15).uniform(-1, 1, u.shape)
    m.upload(u)
    orc.set_fv_surface_interp(mode)
    try:
        assert rel(nm.evolve_setrhs_mesh(m, t=0.2), om.rhs(u, 0.2, K, "upwind")) <= TOL
    finally:
        orc.set_fv_surface_interp("linear")
        orc.set_fv_midpoint(False)


@pytest.mark.parametrize("levels", [0, 1])
def test_mixed_parabolic_surface_interp_matches_oracle(levels):
    """fv_surface_interp = parabolic: DG elements take their FV neighbours'
    face values by the piecewise parabolic rows (basis_pw_parab); the RHS
    and RK4 steps against the oracle's with the same mode; switching the
    parameter on a built mesh rebuilds the face tables (back to linear)
    and keeps the state"""
    print("mixed parabolic", levels, flush=True)
    m, om = make_mixed(5, "WENOm5_2", levels=levels, extra={"fv_surface_interp": "parabolic"})
    u, _ = om.profile(K, 0.2)
    u = u + 0.1 * np.random.default_rng(8).uniform(-1, 1, u.shape)
    m.upload(u)
    orc.set_fv_surface_interp("parabolic")
    try:
        r = nm.evolve_setrhs_mesh(m, t=0.2)
        assert rel(r, om.rhs(u, 0.2, K, "upwind")) <= TOL
        m.time = 0.0
        v, t = om.evolve(u, m.dt, 2, 0.0, K, "RK4")
        nm.evolve_myln(m, 2)
        assert rel(m.download(), v) <= TOL
    finally:
        orc.set_fv_surface_interp("linear")
    m.upload(u)
    m.set_par("fv_surface_interp", "linear")
    assert np.array_equal(m.download(), u)
    r_lin = nm.evolve_setrhs_mesh(m, t=0.2)
    assert rel(r_lin, om.rhs(u, 0.2, K, "upwind")) <= TOL
    assert rel(r_lin, r) > 1e-8
    orc.set_fv_midpoint(False)


def test_mixed_dg_fv_steps_match_oracle():
    """RK4 on the mixed shell, every step against the oracle; a captured
    multi-step graph (order buckets on stream lanes) gives the same states
    bitwise; the filter with basis_filter_fv = no smooths the DG elements
    only"""
    print("mixed dg/fv steps", flush=True)
    m, om = make_mixed(6, "WENOm5_2")
    u, _ = om.profile(K, 0.0)
    m.upload(u)
    t = 0.0
    try:
        for s in range(3):
            u, t = om.evolve(u, m.dt, 1, t, K, "RK4")
            nm.evolve_myln(m, 1)
            assert rel(m.download(), u) <= TOL, "step %d" % s
        m2, _ = make_mixed(6, "WENOm5_2")
        m2.upload(om.profile(K, 0.0)[0])
        nm.evolve_myln(m2, 3)
        assert np.array_equal(m2.download(), m.download())
        # filter: refused on FV elements unless basis_filter_fv = no
        m.set_par("evolve_filter_all", "yes")
        m.set_par("evolve_filter_alp", "4")
        m.set_par("evolve_filter_s", "6")
        with pytest.raises(nm.NmxError):
            nm.evolve_filter_evosys_mesh(m)
        m.set_par("basis_filter_fv", "no")
        v = u + 0.05 * np.random.default_rng(6).uniform(-1, 1, u.shape)
        m.upload(v)
        nm.evolve_filter_evosys_mesh(m)
        assert rel(m.download(), om.filter(v, 31, 4.0, 6.0, 0, 1)) <= TOL
    finally:
        orc.set_fv_midpoint(False)


def test_fv_elements_refuse_unbuilt_configurations():
    pars = {"amr_mesh_type": "Shell", "amr_luni": 1, "amr_n0": 5, "amr_n1": 5, "amr_n2": 5,
            "amr_uniform_p": "0 2 4", "amr_fv_p": "0 2 4", "nmx_path": "perelement"}
    m = nm.Mesh(pars, device=0)
    with pytest.raises(nm.NmxError):
        nm.amr_setup_mesh(m)          # FV elements on the per-element path
    with pytest.raises(nm.NmxError):
        m.set_par("coordinates_midpoint_data", "maybe")
    m = nm.Mesh({"amr_mesh_type": "BoxMesh", "amr_uniform_p": "0"}, device=0)
    with pytest.raises(nm.NmxError):
        nm.amr_setup_mesh(m)          # DG on uniform points
