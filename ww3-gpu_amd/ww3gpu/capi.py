"""ctypes declarations of every entry point in include/ww3gpu_c.h."""
from __future__ import annotations

import ctypes
import enum
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_PKG), "libww3gpu.so")

c_int, c_long, c_float, c_void_p, c_char_p = (ctypes.c_int, ctypes.c_long, ctypes.c_float,
                                               ctypes.c_void_p, ctypes.c_char_p)
f32p = ctypes.POINTER(ctypes.c_float)
i32p = ctypes.POINTER(ctypes.c_int)


class Status(enum.IntEnum):
    """ww_kokkos_c.hpp:42-50 of the reference."""
    OK = 0
    NOT_INITIALISED = 1
    BAD_SHAPE = 2
    KERNEL = 3
    RUNTIME = 4
    UNSUPPORTED = 5
    TABLE = 6


class WW3GpuError(RuntimeError):
    def __init__(self, where: str, code: int):
        super().__init__(f"{where}: status {code} ({Status(code).name if code in Status._value2member_map_ else '?'})")
        self.code = code


# ww_xyp3_batch and ww_ktp3_batch (prop3_c.hpp:25-32, 80-85)
_XYP3 = ([c_void_p, c_int] * 4 + [c_float] + [c_void_p] * 21 + [c_int, c_int] + [c_void_p] * 4 +
         [c_int, c_void_p])
_KTP3 = ([c_void_p, c_int, c_void_p, c_int] + [c_void_p] * 10 + [c_int, c_int, c_float, c_float, c_int] +
         [c_void_p] * 2 + [c_int] + [c_void_p] * 3 + [c_int] + [c_void_p] * 3)

# name -> (restype, argtypes).  Pointer arguments are declared as c_void_p so callers
# can pass either host pointers (numpy .ctypes.data) or device pointers (tensor.data_ptr()).
SIGNATURES = {
    "ww_kokkos_init": (c_int, [c_int]),
    "ww_kokkos_finalize": (None, []),
    "ww_kokkos_flag": (c_int, [c_char_p]),
    "ww_kokkos_preset_status": (c_int, []),
    "ww_kokkos_strict": (c_int, []),
    "ww_kokkos_fallback_slot": (c_int, [c_char_p]),
    "ww_kokkos_fallback_add": (None, [c_int]),
    "ww_snl1_init": (c_int, [c_int, c_int, c_int] + [c_float] * 10 + [c_void_p]),
    "ww_snl1": (None, [c_int, c_int] + [c_void_p] * 5),
    "ww_snl1_enabled": (c_int, []),
    "ww_snl1_last_error": (c_int, []),
    "ww3gpu_snl1_dev": (c_int, [c_int, c_int] + [c_void_p] * 6),
    "ww3gpu_snl1_tables": (c_int, [c_int] + [c_void_p] * 5),
    "ww3gpu_snl1_set_af11": (c_int, [c_int, c_void_p, c_int]),
    "ww3gpu_insnl1": (c_int, [c_int, c_int, c_float, c_float, c_float] + [c_void_p] * 7),
    "ww3gpu_snl1_set_variant": (c_int, [c_int, c_int]),
    "ww3gpu_snl1_kernel_variant": (c_int, [c_int]),
    "ww3gpu_snl1_lds_bytes": (c_long, [c_int]),
    "ww3gpu_devlibm_eval": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_long, c_void_p]),
    "ww3gpu_version": (c_char_p, []),
    "ww_st4_init": (c_int, [c_int, c_void_p, c_int, c_void_p, c_int]),
    "ww_st4_set_real": (c_int, [c_int, c_int, c_void_p, c_int, c_void_p, c_void_p]),
    "ww_st4_set_int": (c_int, [c_int, c_int, c_void_p, c_int, c_void_p, c_void_p]),
    "ww_st4_commit": (c_int, [c_int]),
    "ww_st4_enabled": (c_int, [c_int]),
    "ww_st4_last_error": (c_int, []),
    "ww_srce_init": (c_int, [c_int, c_void_p, c_int, c_void_p, c_int] + [c_void_p] * 5),
    "ww_srce": (None, [c_int] * 5 + [c_void_p] * 3 + [c_float] * 8 + [c_void_p] * 26),
    "ww_srce_batch": (None, [c_int] * 5 + [c_float, c_int] + [c_void_p] * 9),
    "ww_srce_enabled": (c_int, []),
    "ww_srce_last_error": (c_int, []),
    "ww3gpu_srce_batch_dev": (c_int, [c_int] * 5 + [c_float, c_int] + [c_void_p] * 10),
    "ww3gpu_srce_substeps": (c_int, [c_int]),
    "ww3gpu_srce_point_substeps": (c_long, [c_int]),
    "ww3gpu_srce_timing": (c_int, [c_int, c_int]),
    "ww3gpu_srce_group_lanes": (c_int, [c_int]),
    "ww3gpu_srce_width_thresholds": (c_int, [c_int, c_int]),
    "ww3gpu_srce_timers": (c_int, [c_int, c_void_p, c_void_p, c_int]),
    "ww3gpu_insin4_tables": (c_int, [c_void_p, c_int, c_void_p, c_int] + [c_void_p] * 7),
    "ww_sds4": (None, [c_int, c_int] + [c_void_p] * 8 + [c_int, c_int] + [c_void_p] * 4),
    "ww_sin4": (None, [c_int, c_int] + [c_void_p] * 17),
    "ww_spr4": (None, [c_int, c_int] + [c_void_p] * 20),
    "ww_va_host_read": (c_int, [c_int, c_void_p, c_int, c_int]),
    "ww_va_host_write": (c_int, [c_int, c_void_p, c_int, c_int]),
    "ww_va_host_read_cols": (c_int, [c_int, c_void_p, c_int, c_int, c_int, c_void_p]),
    "ww_va_host_wrote_cols": (c_int, [c_int, c_void_p, c_int, c_int, c_int, c_void_p]),
    "ww_va_host_update": (c_int, [c_int, c_void_p, c_int, c_int]),
    "ww_va_device_zero_cols": (c_int, [c_int, c_void_p, c_int, c_int, c_int, c_void_p]),
    "ww_va_device": (c_int, [c_int, c_int]),
    "ww_va_report": (None, []),
    "ww_srce_batch_va": (None, [c_int] * 5 + [c_float, c_void_p, c_int, c_int, c_int] + [c_void_p] * 9),
    "ww3gpu_va_state": (c_int, [c_int]),
    "ww_xyp3_batch": (c_int, _XYP3),
    "ww3gpu_xyp3_batch_dev": (c_int, _XYP3 + [c_void_p]),
    "ww_xyp3_batch_va": (c_int, [c_int] + _XYP3[:-2] + [c_void_p, c_int, c_int]),
    "ww_xyp3_batch_mpi": (c_int, [c_int] + _XYP3[:-2] + [c_void_p, c_int, c_int, c_int, c_int] +
                          [c_void_p] * 3 + [c_int]),
    "ww_xyp3_mpi_pack": (c_int, [c_int, c_void_p, c_int, c_int, c_void_p, c_int, c_void_p]),
    "ww_xyp3_mpi_unpack": (c_int, [c_int, c_void_p, c_int, c_int, c_void_p, c_int, c_void_p]),
    "ww_ktp3_batch": (c_int, _KTP3),
    "ww3gpu_ktp3_batch_dev": (c_int, _KTP3 + [c_void_p]),
    "ww_ktp3_batch_va": (c_int, [c_int] + _KTP3[:-4] + [c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "ww3gpu_prop3_last_error": (c_int, []),
    "ww3gpu_prop3_stats": (c_int, [c_void_p, c_int]),
    "ww3gpu_snlsur_init": (c_int, [c_int, ctypes.c_ulonglong]),
    "ww3gpu_snlsur_nparams": (ctypes.c_longlong, [c_int, c_void_p]),
    "ww3gpu_snlsur_params": (c_int, [c_int, c_int, c_void_p, ctypes.c_longlong]),
    "ww3gpu_snlsur_synth_dev": (c_int, [c_int, c_int, ctypes.c_ulonglong, c_float] + [c_void_p] * 4),
    "ww3gpu_snlsur_train_dev": (c_int, [c_int, c_int] + [c_void_p] * 4 + [c_float, c_void_p, c_void_p]),
    "ww3gpu_snlsur_fit": (c_int, [c_int, c_int, c_int, c_int, c_float, ctypes.c_ulonglong, c_float,
                                  c_void_p, c_void_p]),
    "ww3gpu_snlsur_dev": (c_int, [c_int, c_int] + [c_void_p] * 7),
    "ww3gpu_snlsur_grads": (c_int, [c_int, c_int] + [c_void_p] * 4),
    "ww3gpu_snlsur_forward": (c_int, [c_int, c_int, c_void_p, c_void_p]),
    "ww3gpu_sur_gemm_dev": (c_int, [c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_int, c_void_p, c_int,
                                    c_void_p]),
    "ww3gpu_snlsur_last_error": (c_int, []),
}

_lib = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libww3gpu.so and declare its entry points.  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libww3gpu.so not built: {path} (run __graft_entry__.build())")
    L = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def lib() -> ctypes.CDLL:
    return load()


def ptr(a) -> int:
    """Address of a numpy array or a torch tensor."""
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data


def check(where: str, code: int) -> None:
    if code != Status.OK:
        raise WW3GpuError(where, code)
