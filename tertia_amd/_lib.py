"""ctypes binding of libqss.so (the C ABI declared in include/qss.h).

The library is built in-tree by ``tertia_amd.build``; this module never falls
back to anything else: if ``libqss.so`` is missing or cannot be loaded the
import of the product path raises.
"""
from __future__ import annotations

import ctypes
import os

from .params import QssDriver, QssParams

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqss.so")

_dp = ctypes.POINTER(ctypes.c_double)
_dpp = ctypes.POINTER(ctypes.c_void_p)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p


class QssH5Set(ctypes.Structure):
    """qss_h5_set: one dataset of qss_h5_write."""
    _fields_ = [("name", ctypes.c_char_p), ("kind", ctypes.c_int32), ("rank", ctypes.c_int32),
                ("dims", ctypes.c_int64 * 3), ("data", ctypes.c_void_p)]

# every symbol declared in include/qss.h: (restype, argtypes)
SIGNATURES = {
    "qss_create": (ctypes.c_int, [ctypes.POINTER(QssParams), ctypes.c_int, ctypes.c_uint32,
                                  ctypes.POINTER(_vp)]),
    "qss_destroy": (ctypes.c_int, [_vp]),
    "qss_last_error": (ctypes.c_char_p, []),
    "qss_info": (ctypes.c_int, [_vp, _i64p, _i64p]),
    "qss_stream": (_vp, [_vp]),
    "qss_sweep": (ctypes.c_int, [_vp, _dp]),
    "qss_sweep_range": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp]),
    "qss_reset": (ctypes.c_int, [_vp]),
    "qss_layer_step": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp]),
    "qss_beam_layer": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "qss_guess": (ctypes.c_int, [_vp, ctypes.c_int]),
    "qss_push_deposit": (ctypes.c_int, [_vp, ctypes.c_int]),
    "qss_iterate": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "qss_copy_c_to_p": (ctypes.c_int, [_vp]),
    "qss_transform2d": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp,
                                       ctypes.c_double]),
    "qss_transform2d_batched": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _ip, _dpp, _dpp,
                                               ctypes.c_double]),
    "qss_get_layer_field": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp]),
    "qss_set_layer_field": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp]),
    "qss_get_history": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp]),
    "qss_set_history": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp]),
    "qss_layer_state_bytes": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_size_t)]),
    "qss_pack_layer": (ctypes.c_int, [_vp, _vp]),
    "qss_unpack_layer": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
    "qss_set_transform_method": (ctypes.c_int, [_vp, ctypes.c_int]),
    "qss_transform_method": (ctypes.c_int, [_vp]),
    "qss_fact_choice": (ctypes.c_int, [_vp, _ip]),
    "qss_set_fact_choice": (ctypes.c_int, [_vp, _ip]),
    "qss_transform2d_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_int, _vp, _vp, ctypes.c_double, _vp]),
    "qss_layer_record_doubles": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_size_t)]),
    "qss_pack_layer_record": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "qss_unpack_layer_record": (ctypes.c_int, [_vp, _dp, ctypes.c_size_t]),
    "qss_driver_seed": (ctypes.c_int, [_vp, ctypes.POINTER(QssDriver), _i64p]),
    "qss_driver_push": (ctypes.c_int, [_vp]),
    "qss_time_step": (ctypes.c_int, [_vp, _dp]),
    "qss_driver_get": (ctypes.c_int, [_vp, _i32p, _dp, _dp, _dp, _i32p]),
    "qss_driver_set": (ctypes.c_int, [_vp, ctypes.c_int64, _i32p, _dp, _dp, _dp, _i32p]),
    "qss_driver_density": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp]),
    "qss_driver_timing": (ctypes.c_int, [_vp, _dp, _dp]),
    "qss_history_ptr": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp)]),
    "qss_layer_ptr": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
    "qss_get_particles": (ctypes.c_int, [_vp, _i32p, _i32p, _dp, _dp, _dp, _dp, _dp, _i64p]),
    "qss_set_particles": (ctypes.c_int, [_vp, ctypes.c_int64, _i32p, _i32p, _dp, _dp, _dp,
                                         _dp, _dp, _i64p]),
    "qss_get_spectral": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "qss_set_spectral": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "qss_enable_timing": (ctypes.c_int, [_vp, ctypes.c_int]),
    "qss_autotune": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]),
    "qss_set_gemm_variant": (ctypes.c_int, [_vp, ctypes.c_int]),
    "qss_gemm_variants": (ctypes.c_int, [ctypes.c_int]),
    "qss_gemm_choice": (ctypes.c_int, [_vp, _ip]),
    "qss_set_bins": (ctypes.c_int, [_vp, _dp, ctypes.c_int]),
    "qss_sweep_checkpointed": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp]),
    "qss_sweep_vjp": (ctypes.c_int, [_vp, ctypes.c_void_p, ctypes.c_int, _dp]),
    "qss_copy_history": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_void_p]),
    "qss_device_memory": (ctypes.c_int, [ctypes.c_int, _i64p, _i64p]),
    "qss_select_wake": (ctypes.c_int, [_vp, ctypes.c_int]),
    "qss_set_gemm_choice": (ctypes.c_int, [_vp, _ip]),
    "qss_set_particle_split": (ctypes.c_int, [_vp, ctypes.c_int]),
    "qss_particle_split": (ctypes.c_int, [_vp]),
    "qss_enable_graph": (ctypes.c_int, [_vp, ctypes.c_int]),
    "qss_timing": (ctypes.c_int, [_vp, _dp, _dp, _i64p, _dp]),
    "qss_h5_write": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(QssH5Set)]),
    "qss_h5_info": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, _i32p, _i32p, _i64p]),
    "qss_h5_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int32,
                                   ctypes.c_int64, ctypes.c_int64, _vp]),
    "qss_h5_list": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t]),
    "qss_save": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_int, ctypes.c_double,
                                ctypes.c_int]),
    "qss_load": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_int, _dp]),
}

_lib = None


class QssError(RuntimeError):
    """A negative status from the C ABI (the reference exit()s instead)."""

    def __init__(self, code, where, msg):
        super().__init__("%s failed with %d: %s" % (where, code, msg))
        self.code = code


def load():
    """Load libqss.so and declare every exported signature."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libqss.so is not built (run python -m tertia_amd.build): %s"
                              % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(code, where):
    if code != 0:
        msg = load().qss_last_error()
        raise QssError(code, where, msg.decode() if msg else "")
    return code
