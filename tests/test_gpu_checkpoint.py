"""GPU tests of the checkpoint variables file and the device "vars" CRC
(libnmx crc_elm_kernel / crc_xor_kernel, nmx_checkpoint_*_vars) against
the oracle (orc_vars_crc, orc_write_variables_bin), bit-exact:

- the device CRC of a register equals the oracle's CRC of the same state,
  on Cartesian, shell, mixed-order refined and FV meshes and on every
  register;
- the file written from the device equals the oracle's file byte for byte;
- reading it back restores the state bitwise, and a run restarted from it
  continues exactly as the uninterrupted run;
- damaged or foreign files are refused and leave the register unchanged;
  variables of other names in a file are skipped.
"""
import os

import numpy as np
import pytest

import nmesh_amd as nm
from nmesh_amd import checkpoint as ck
from nmesh_amd.evolve import REG_P, REG_R, REG_U
from oracle.orc import OMesh

pytestmark = pytest.mark.gpu
K = [0.8, -0.5, 0.33166247903554]


def make(case):
    if case == "box":
        pars = {"amr_mesh_type": "BoxMesh", "amr_luni": 2, "amr_n0": 5, "amr_n1": 5, "amr_n2": 5,
                "amr_BoxMesh_dout": 2.0}
        om = OMesh("box", 2, 5, dout=2.0)
        return nm.Mesh(pars, device=0), om, None
    if case == "shell":
        pars = {"amr_mesh_type": "Shell", "amr_luni": 1, "amr_n0": 7, "amr_n1": 7, "amr_n2": 7,
                "amr_Shell_rin": 2.0, "amr_Shell_rout": 20.0}
        return nm.Mesh(pars, device=0), OMesh("shell", 1, 7, rin=2.0, rout=20.0), None
    if case == "amr":
        pars = {"amr_mesh_type": "Shell", "amr_luni": 1, "amr_n0": 4, "amr_n1": 4, "amr_n2": 4,
                "amr_Shell_rin": 2.0, "amr_Shell_rout": 20.0, "amr_hrefine_sphere_radius": 7,
                "amr_hrefine_sphere_levels": 1}
        om = OMesh("shell", 1, 4, rin=2.0, rout=20.0, radius=7, levels=1)
        nvec = np.random.default_rng(5).integers(2, 11, om.nelem).astype(np.int32)
        return nm.Mesh(pars, device=0), om, nvec
    if case == "fv":
        pars = {"amr_mesh_type": "Shell", "amr_luni": 1, "amr_n0": 6, "amr_n1": 6, "amr_n2": 6,
                "amr_Shell_rin": 2.0, "amr_Shell_rout": 6.0, "amr_uniform_p": "0 2 4",
                "amr_fv_p": "0 2 4"}
        return nm.Mesh(pars, device=0), OMesh("shell", 1, 6, rin=2.0, rout=6.0, fv_patches=[0, 2, 4]), None
    raise ValueError(case)


def setup(case):
    m, om, nvec = make(case)
    nm.amr_setup_mesh(m)
    if nvec is not None:
        m.set_elm_n(nvec)
        om.set_n(nvec)
    return m, om


@pytest.mark.parametrize("case", ["box", "shell", "amr", "fv"])
def test_device_vars_crc_equals_oracle(case):
    print("vars crc", case, flush=True)
    m, om = setup(case)
    rng = np.random.default_rng(11)
    for reg in (REG_U, REG_R, REG_P):
        u = rng.standard_normal(5 * om.nnodes)
        m.upload(u, reg)
        assert ck.vars_crc(m, reg) == om.vars_crc(u), reg
    # one flipped bit anywhere changes it
    u = rng.standard_normal(5 * om.nnodes)
    m.upload(u)
    c0 = ck.vars_crc(m)
    v = u.copy()
    b = v.view(np.uint64)
    b[int(rng.integers(0, b.size))] ^= np.uint64(1) << np.uint64(int(rng.integers(0, 64)))
    m.upload(v)
    assert ck.vars_crc(m) != c0 and ck.vars_crc(m) == om.vars_crc(v)


@pytest.mark.parametrize("case", ["shell", "amr", "fv"])
def test_save_equals_oracle_file_and_loads_back(case, tmp_path):
    print("variables.bin", case, flush=True)
    m, om = setup(case)
    u = np.random.default_rng(12).standard_normal(5 * om.nnodes)
    m.upload(u)
    mine = os.path.join(tmp_path, "mine.bin")
    ref = os.path.join(tmp_path, "oracle.bin")
    ck.checkpoint_save_EvoVars(m, mine)
    om.write_variables_bin(u, ref)
    with open(mine, "rb") as f1, open(ref, "rb") as f2:
        assert f1.read() == f2.read()
    m.upload(np.zeros_like(u), REG_P)
    ck.checkpoint_load_Vars(m, mine, REG_P)
    assert np.array_equal(m.download(REG_P), u)


def test_restart_continues_exactly(tmp_path):
    """5 RK4 steps straight = 3 steps, save, load into a new mesh, 2 steps"""
    print("restart", flush=True)
    pars = {"amr_mesh_type": "Shell", "amr_luni": 1, "amr_n0": 6, "amr_n1": 6, "amr_n2": 6,
            "amr_Shell_rin": 2.0, "amr_Shell_rout": 20.0,
            "scalarwave1_k": "%.17g %.17g %.17g" % tuple(K)}
    a = nm.Mesh(pars, device=0)
    nm.amr_setup_mesh(a)
    nm.scalarwave1_set_profile(a, t=0.0)
    nm.evolve_myln(a, 3)
    path = os.path.join(tmp_path, "variables.bin")
    ck.checkpoint_save_EvoVars(a, path)
    t3, crc3 = a.time, ck.vars_crc(a)
    nm.evolve_myln(a, 2)
    b = nm.Mesh(pars, device=0)
    nm.amr_setup_mesh(b)
    ck.checkpoint_load_Vars(b, path)
    assert ck.vars_crc(b) == crc3
    b.time = t3
    nm.evolve_myln(b, 2)
    assert np.array_equal(b.download(), a.download())


def test_damaged_and_foreign_files(tmp_path):
    print("damaged files", flush=True)
    m, om = setup("box")
    u = np.random.default_rng(13).standard_normal(5 * om.nnodes)
    m.upload(u)
    good = os.path.join(tmp_path, "good.bin")
    ck.checkpoint_save_EvoVars(m, good)
    data = open(good, "rb").read()
    v = np.random.default_rng(14).standard_normal(5 * om.nnodes)
    m.upload(v)
    bad = os.path.join(tmp_path, "bad.bin")
    cases = {
        "truncated": data[: len(data) // 2],
        "no end marker": data.replace(b"EndOfFile.\n", b""),
        "unknown element": data.replace(b"\n0_00\n", b"\n3_00\n", 1),
        "wrong point count": data.replace(b"\n125\n", b"\n64\n", 1),
        "missing element": data[: data.index(b"{\n0_01\n")] + b"EndOfFile.\n",
    }
    for what, d in cases.items():
        assert d != data, what
        with open(bad, "wb") as f:
            f.write(d)
        with pytest.raises(nm.NmxError):
            ck.checkpoint_load_Vars(m, bad)
        assert np.array_equal(m.download(), v), what     # register unchanged
    with pytest.raises(nm.NmxError):
        ck.checkpoint_load_Vars(m, os.path.join(tmp_path, "absent.bin"))
    # a mesh of another order refuses the file
    m2 = nm.Mesh({"amr_mesh_type": "BoxMesh", "amr_luni": 2, "amr_n0": 4, "amr_n1": 4, "amr_n2": 4,
                  "amr_BoxMesh_dout": 2.0}, device=0)
    nm.amr_setup_mesh(m2)
    with pytest.raises(nm.NmxError):
        ck.checkpoint_load_Vars(m2, good)
    # an extra variable of another name (with array dims) is skipped
    names, elms = read_blocks(data)
    out = [b"number of variables and variable list in this file:\n%d\n" % (len(names) + 1)]
    out += [n.encode() + b"\n" for n in names + ["other_var"]]
    out.append(b"\n\nvariable-list-data:\n\n")
    for name, np_, arrays in elms:
        out.append(b"{\n%s\n%d\n" % (name.encode(), np_))
        for vli, a in enumerate(arrays):
            out.append(b"%d\n" % vli + a + b"\n")
        out.append(b"5 2 1 1\n" + np.array([1.5, -2.5], dtype="<f8").tobytes() + b"\n}\n\n")
    out.append(b"EndOfFile.\n")
    with open(bad, "wb") as f:
        f.write(b"".join(out))
    ck.checkpoint_load_Vars(m, bad)
    assert np.array_equal(m.download(), u)


def read_blocks(data):
    """(variable names, [(element name, np, [raw arrays in list order])])"""
    pos = 0

    def line():
        nonlocal pos
        end = data.index(b"\n", pos)
        s = data[pos:end].decode()
        pos = end + 1
        return s

    line()
    nv = int(line())
    names = [line() for _ in range(nv)]
    while line() != "variable-list-data:":
        pass
    elms = []
    while True:
        s = line()
        if s == "":
            continue
        if s == "EndOfFile.":
            return names, elms
        name, np_ = line(), int(line())
        arrays = []
        for _ in range(nv):
            line()
            arrays.append(data[pos:pos + 8 * np_])
            pos += 8 * np_ + 1
        assert line() == "}"
        elms.append((name, np_, arrays))
