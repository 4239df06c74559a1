"""nmesh's checkpoint variables file and CRCs over libnmx.

Mirrors the reference's checkpoint calls for the evolved variables:

- ``checkpoint_save_EvoVars(mesh, fname)`` (checkpoint_save.c:337-356):
  writes variables.bin, scalarwave1's phi, pi, cx, cy, cz per element in
  nmesh's format (little-endian doubles under nmesh's element names);
- ``checkpoint_load_Vars(mesh, fname)`` (checkpoint_load.c:396-480): reads
  one back; every element of the mesh must be in it with its five variables,
  other variables are skipped, and on error the state is left as it was;
- ``vars_crc(mesh)``: the "vars" entry of nmesh_CRCs (norms.c:480-520),
  computed on the device;
- ``crc64`` / ``crc64_combine``: crc64_continue and crc64_0start_combine
  (crc64.c:153-256);
- ``read_CRCs(fname)``: the CRC file checkpoint_save_CRCs writes
  (checkpoint_save.c:482-507), as a dict.
"""
import ctypes as C

from ._lib import check, lib
from .evolve import REG_U


def checkpoint_save_EvoVars(mesh, fname, reg=REG_U):
    check(lib().nmx_checkpoint_save_vars(mesh._ctx, int(reg), str(fname).encode()))


def checkpoint_load_Vars(mesh, fname, reg=REG_U):
    check(lib().nmx_checkpoint_load_vars(mesh._ctx, int(reg), str(fname).encode()))


def vars_crc(mesh, reg=REG_U):
    out = C.c_uint64(0)
    check(lib().nmx_vars_crc(mesh._ctx, int(reg), C.byref(out)))
    return int(out.value)


def crc64(data, crc=0):
    b = bytes(memoryview(data).cast("B"))
    return int(lib().nmx_crc64(C.c_uint64(crc), b, len(b)))


def crc64_combine(crc1, crc2, nbytes2):
    return int(lib().nmx_crc64_combine(C.c_uint64(crc1), C.c_uint64(crc2), C.c_uint64(nbytes2)))


def read_CRCs(fname):
    """{"pars": .., "pats": .., "elms": .., "nbinfo": .., "vars": ..}"""
    out = {}
    with open(fname) as f:
        for line in f:
            if ":" in line:
                k, v = line.split(":", 1)
                out[k.strip()] = int(v.strip())
    return out
