"""nmx_vars_crc_noalloc (the device "vars" CRC with its scratch in a static
device array, no allocation) through the C ABI against nmx_vars_crc, on two
calls in a row with the second on a smaller mesh, so stale element
contributions of the first call must not leak into the second; and the
element limit refused with a clear error."""
import ctypes as C

import numpy as np
import pytest

import nmesh_amd as nm
from nmesh_amd._lib import lib
from nmesh_amd.evolve import REG_U

pytestmark = pytest.mark.gpu


def crc_pair(m, reg=REG_U):
    a = C.c_uint64()
    b = C.c_uint64()
    assert lib().nmx_vars_crc(m._ctx, int(reg), C.byref(a)) == 0, lib().nmx_last_error()
    assert lib().nmx_vars_crc_noalloc(m._ctx, int(reg), C.byref(b)) == 0, lib().nmx_last_error()
    return a.value, b.value


def mesh(luni, n, seed, nvec_seed=None):
    pars = {"amr_mesh_type": "Shell", "amr_luni": luni, "amr_n0": n, "amr_n1": n, "amr_n2": n,
            "amr_Shell_rin": 2.0, "amr_Shell_rout": 20.0}
    m = nm.Mesh(pars, device=0)
    nm.amr_setup_mesh(m)
    if nvec_seed is not None:
        m.set_elm_n(np.random.default_rng(nvec_seed).integers(2, 11, m.nelem).astype(np.int32))
    m.upload(np.random.default_rng(seed).standard_normal(5 * m.nnodes))
    return m


def test_noalloc_crc_equals_allocating_crc_twice_second_smaller():
    print("noalloc crc: 384-element mixed-order shell, then 48-element shell", flush=True)
    big = mesh(2, 5, 1, nvec_seed=3)
    small = mesh(1, 6, 2)
    assert small.nelem < big.nelem
    a1, b1 = crc_pair(big)
    a2, b2 = crc_pair(small)
    assert a1 == b1 and a2 == b2
    assert a1 != a2
    # the small mesh again after the big one, and a changed state
    a3, b3 = crc_pair(big)
    u = small.download()
    u[7] += 1.0
    small.upload(u)
    a4, b4 = crc_pair(small)
    assert (a3, a4) == (b3, b4) and a3 == a1 and a4 != a2


def test_noalloc_crc_refuses_too_many_elements():
    print("noalloc crc: element limit", flush=True)
    pars = {"amr_mesh_type": "BoxMesh", "amr_luni": 5, "amr_n0": 2, "amr_n1": 2, "amr_n2": 2,
            "amr_BoxMesh_dout": 1.0}
    m = nm.Mesh(pars, device=0)
    nm.amr_setup_mesh(m)
    a = C.c_uint64()
    b = C.c_uint64()
    assert m.nelem == 1 << 15          # at the limit: taken
    assert lib().nmx_vars_crc(m._ctx, REG_U, C.byref(a)) == 0
    assert lib().nmx_vars_crc_noalloc(m._ctx, REG_U, C.byref(b)) == 0 and a.value == b.value
    flags = np.zeros(m.nelem, np.int32)
    flags[0] = 1
    m.hadapt(flags)                    # 7 more elements: refused
    assert m.nelem == (1 << 15) + 7
    assert lib().nmx_vars_crc_noalloc(m._ctx, REG_U, C.byref(b)) != 0
    assert b"at most 32768 elements" in lib().nmx_last_error()
