"""ctypes binding of libnmx.so (the C ABI declared in include/nmx.h).

The library is the product: it must be built (``python -c "import
__graft_entry__ as g; g.build()"``) before use.  There is no fallback; a
missing or unloadable library raises immediately.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnmx.so")

_lib = None


class NmxError(RuntimeError):
    """An nmx_* call returned a non-zero status."""


def _proto(lib):
    vp, i, l, d = C.c_void_p, C.c_int, C.c_long, C.c_double
    P = C.POINTER
    sig = {
        "nmx_last_error": (C.c_char_p, []),
        "nmx_ctx_create": (i, [P(vp), i]),
        "nmx_ctx_destroy": (i, [vp]),
        "nmx_set_stream": (i, [vp, vp]),
        "nmx_synchronize": (i, [vp]),
        "nmx_set_param": (i, [vp, C.c_char_p, C.c_char_p]),
        "nmx_get_param": (i, [vp, C.c_char_p, C.c_char_p, i]),
        "nmx_mesh_build": (i, [vp]),
        "nmx_mesh_import": (i, [vp, i, vp, vp, i, vp, vp, vp, vp, vp, vp, vp, vp]),
        "nmx_mesh_set_elm_n": (i, [vp, vp, i]),
        "nmx_mesh_hadapt": (i, [vp, vp, i]),
        "nmx_mesh_padapt": (i, [vp, vp, i]),
        "nmx_mesh_switch_fv": (i, [vp, vp, i]),
        "nmx_mesh_fv_flags": (i, [vp, vp, i]),
        "nmx_mesh_dt_device": (i, [vp, P(d)]),
        "nmx_mesh_nelem": (i, [vp, P(i)]),
        "nmx_mesh_nnodes": (i, [vp, P(l)]),
        "nmx_mesh_elm_info": (i, [vp, i, vp]),
        "nmx_mesh_elm_bbox": (i, [vp, i, vp]),
        "nmx_mesh_elm_fnb": (i, [vp, i, i, vp, vp, P(i)]),
        "nmx_mesh_coords": (i, [vp, vp]),
        "nmx_mesh_geometry": (i, [vp, vp, vp]),
        "nmx_get_dt": (i, [vp, P(d)]),
        "nmx_set_dt": (i, [vp, d]),
        "nmx_get_time": (i, [vp, P(d)]),
        "nmx_set_time": (i, [vp, d]),
        "nmx_evolve_mesh": (i, [vp]),
        "nmx_get_iteration": (i, [vp, P(l)]),
        "nmx_state_upload": (i, [vp, i, vp]),
        "nmx_state_download": (i, [vp, i, vp]),
        "nmx_state_upload_elm": (i, [vp, i, i, i, vp, i]),
        "nmx_state_download_elm": (i, [vp, i, i, i, vp, i]),
        "nmx_state_upload_ptrs": (i, [vp, i, vp, i, i]),
        "nmx_state_download_ptrs": (i, [vp, i, vp, i, i]),
        "nmx_state_set_profile": (i, [vp, i, d]),
        "nmx_setrhs": (i, [vp, i, i, d]),
        "nmx_autotune": (i, [vp]),
        "nmx_set_trouble_hook": (i, [vp, vp, vp]),
        "nmx_get_trouble_scores": (i, [vp, vp, i, P(l), P(l)]),
        "nmx_evolve": (i, [vp, i]),
        "nmx_filter": (i, [vp, i]),
        "nmx_state_device_ptr": (i, [vp, i, P(vp), P(l)]),
        "nmx_adjoint_setrhs": (i, [vp, i, i]),
        "nmx_adjoint_evolve": (i, [vp, i]),
        "nmx_last_adjoint_ms": (i, [vp, P(d)]),
        "nmx_checkpoint_alloc": (i, [vp, i]),
        "nmx_checkpoint_store": (i, [vp, i, i]),
        "nmx_checkpoint_load": (i, [vp, i, i]),
        "nmx_state_copy_in": (i, [vp, i, vp]),
        "nmx_state_copy_out": (i, [vp, i, vp]),
        "nmx_fv_create": (i, [P(vp), i]),
        "nmx_fv_destroy": (i, [vp]),
        "nmx_fv_set_stream": (i, [vp, vp]),
        "nmx_fv_synchronize": (i, [vp]),
        "nmx_fv_set_param": (i, [vp, C.c_char_p, C.c_char_p]),
        "nmx_fv_get_param": (i, [vp, C.c_char_p, C.c_char_p, i]),
        "nmx_fv_setup": (i, [vp]),
        "nmx_fv_ncells": (i, [vp, P(l)]),
        "nmx_fv_get_dt": (i, [vp, P(d)]),
        "nmx_fv_set_dt": (i, [vp, d]),
        "nmx_fv_get_time": (i, [vp, P(d)]),
        "nmx_fv_set_time": (i, [vp, d]),
        "nmx_fv_upload_prim": (i, [vp, vp]),
        "nmx_fv_upload": (i, [vp, i, vp]),
        "nmx_fv_download": (i, [vp, i, vp]),
        "nmx_fv_setrhs": (i, [vp]),
        "nmx_fv_max_speed": (i, [vp, P(d)]),
        "nmx_fv_evolve": (i, [vp, i]),
        "nmx_fv_last_evolve_ms": (i, [vp, P(d)]),
        "nmx_fv_status": (i, [vp, P(l)]),
        "nmx_last_evolve_ms": (i, [vp, P(d)]),
        "nmx_checkpoint_save_vars": (i, [vp, i, C.c_char_p]),
        "nmx_checkpoint_load_vars": (i, [vp, i, C.c_char_p]),
        "nmx_vars_crc": (i, [vp, i, P(C.c_uint64)]),
        "nmx_vars_crc_noalloc": (i, [vp, i, P(C.c_uint64)]),
        "nmx_crc64": (C.c_uint64, [C.c_uint64, vp, C.c_size_t]),
        "nmx_crc64_combine": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint64]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return sig


def exported_symbols():
    """Names declared in include/nmx.h that the binding expects."""
    return sorted(_proto_names)


_proto_names = None


def lib():
    """Load libnmx.so once; raise if it has not been built."""
    global _lib, _proto_names
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "libnmx.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        # A process must hold one HIP runtime.  torch, when installed, ships
        # its own libamdhip64; loading it first lets libnmx bind to the same
        # runtime (the other order leaves torch without devices).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        _proto_names = list(_proto(L).keys())
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().nmx_last_error()
        raise NmxError("nmx error %d: %s" % (rc, msg.decode() if msg else ""))
    return rc
