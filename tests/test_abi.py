"""The C ABI library loads on any host and exports every entry point
include/nmx.h declares (no GPU call is made)."""
import ctypes as C
import os
import re

import nmesh_amd as nm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "nmx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(nmx_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(nm.LIB_PATH)
    names = declared()
    assert len(names) > 20
    for n in names:
        assert hasattr(lib, n), n


def test_binding_covers_header():
    nm.lib()
    from nmesh_amd import _lib
    assert set(declared()) <= set(_lib._proto_names)


def test_host_only_context_rejects_device_calls():
    m = nm.Mesh(device=-1)
    nm.amr_setup_mesh(m)
    try:
        nm.evolve_myln(m)
    except nm.NmxError as e:
        assert "host-only" in str(e)
    else:
        raise AssertionError("expected NmxError")


def test_bad_parameter_values_raise():
    m = nm.Mesh(device=-1)
    for name, val in [("evolve_method", "RK5"), ("scalarwave1_numflux", "HLLC"),
                      ("no_such_par", "1"), ("nmx_path", "warp")]:
        try:
            m.set_par(name, val)
        except nm.NmxError:
            pass
        else:
            raise AssertionError(name)
