"""nmesh_amd: MI355X-native DG/FV evolution hot path for nmesh meshes.

The compute path is libnmx.so (hand-written HIP kernels for gfx950 behind a
C ABI, include/nmx.h); this package holds its ctypes binding and a host-side
mirror of nmesh's evolution interface.
"""
from ._lib import LIB_PATH, NmxError, lib  # noqa: F401
from .evolve import (REG_P, REG_R, REG_U, Mesh, amr_import_mesh, amr_setup_mesh, autotune,  # noqa: F401
                     evolve_Euler_mesh, evolve_filter_evosys_mesh, evolve_mesh, evolve_myln,
                     evolve_RK4_mesh, evolve_setrhs_mesh, evolve_sspRK3_mesh,
                     scalarwave1_set_profile)
