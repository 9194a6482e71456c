"""ctypes binding of libscrimp_hip.so (declarations mirror include/scrimp_hip.h)."""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libscrimp_hip.so"

MP_OK = 0
STATUS = {
    -1: "MP_ERR_BAD_WINDOW", -2: "MP_ERR_SHORT_SERIES", -3: "MP_ERR_NULL", -4: "MP_ERR_OOM",
    -5: "MP_ERR_DEVICE", -6: "MP_ERR_TOO_LONG", -7: "MP_ERR_BAD_ARG", -8: "MP_ERR_NO_DEVICE",
    -9: "MP_ERR_ADMISSION", -10: "MP_ERR_NOT_FOUND",
}


class ScrimpError(RuntimeError):
    """A negative status returned by the C ABI."""

    def __init__(self, status: int, where: str):
        self.status = status
        super().__init__(f"{where}: {STATUS.get(status, status)} ({status})")


class Opts(C.Structure):
    _fields_ = [
        ("exclusion", C.c_int32), ("band_k", C.c_int32), ("tile_len", C.c_int32),
        ("refine", C.c_int32), ("index_init", C.c_int32), ("accum64", C.c_int32), ("walk_form", C.c_int32),
        ("reserved", C.c_int32),
    ]


class Timing(C.Structure):
    _fields_ = [
        ("total_ms", C.c_double), ("stats_ms", C.c_double), ("seeds_ms", C.c_double),
        ("kernel_ms", C.c_double), ("finalize_ms", C.c_double), ("cells", C.c_double),
        ("work_items", C.c_int64), ("band_k", C.c_int32), ("tile_len", C.c_int32),
        ("walk_form", C.c_int32), ("reserved", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_L = C.POINTER(C.c_int64)

# name -> (restype, argtypes)
SIGNATURES = {
    "mp_status_string": (C.c_char_p, [C.c_int]),
    "mp_opts_default": (None, [C.POINTER(Opts)]),
    "mp_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "mp_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "mp_ctx_destroy": (C.c_int, [_P]),
    "mp_ctx_sync": (C.c_int, [_P]),
    "mp_dev_alloc": (C.c_int, [_P, C.c_uint64, C.POINTER(_P)]),
    "mp_dev_free": (C.c_int, [_P, _P]),
    "mp_memcpy_h2d": (C.c_int, [_P, _P, _P, C.c_uint64]),
    "mp_memcpy_d2h": (C.c_int, [_P, _P, _P, C.c_uint64]),
    "mp_selfjoin_f64": (C.c_int, [_D, C.c_int64, C.c_int32, _D, _L, C.POINTER(Opts), C.POINTER(Timing)]),
    "mp_selfjoin_f64_dev": (C.c_int, [_P, _P, C.c_int64, C.c_int32, _P, _P, C.POINTER(Opts), C.POINTER(Timing)]),
    "mp_selfjoin_f64_dev_noalloc": (C.c_int, [_P, _P, C.c_int64, C.c_int32, _P, _P, C.POINTER(Opts),
                                              C.POINTER(Timing)]),
    "mp_prescrimp_f64": (C.c_int, [_D, C.c_int64, C.c_int32, C.c_double, _D, _L, C.POINTER(Timing)]),
    "mp_prescrimp_f64_dev": (C.c_int, [_P, _P, C.c_int64, C.c_int32, C.c_double, _P, _P, C.POINTER(Timing)]),
    "mp_selfjoin_f32r": (C.c_int, [_D, C.c_int64, C.c_int32, _D, _L, C.POINTER(Opts), C.POINTER(Timing)]),
    "mp_selfjoin_f32r_dev": (C.c_int, [_P, _P, C.c_int64, C.c_int32, _P, _P, C.POINTER(Opts), C.POINTER(Timing)]),
    "mp_eval_diag_block_triangle_f64": (C.c_int, [_D, _L, _D, _L, _D, C.c_int64, C.c_int64, _D, _D,
                                                  C.c_int64, _D, _D, _D, _D, C.c_int64, C.c_int64]),
    "mp_batch_create": (C.c_int, [_P, C.c_int32, _L, C.POINTER(C.c_int32), C.POINTER(Opts), C.POINTER(_P)]),
    "mp_batch_run": (C.c_int, [_P, _P, _P, _P, C.c_int32, C.POINTER(Timing)]),
    "mp_batch_tune": (C.c_int, [_P, _P, _P, _P, C.c_int32, C.POINTER(Timing)]),
    "mp_batch_info": (C.c_int, [_P, _L, _L, _L, C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "mp_batch_destroy": (C.c_int, [_P]),
    "mp_server_create": (C.c_int, [_P, C.c_uint64, C.c_int64, C.POINTER(Opts), C.POINTER(_P)]),
    "mp_server_submit": (C.c_int, [_P, _D, C.c_int64, C.c_int32, _D, _L, C.POINTER(C.c_uint64)]),
    "mp_server_poll": (C.c_int, [_P, C.c_uint64, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "mp_server_wait": (C.c_int, [_P, C.c_uint64, C.POINTER(C.c_double)]),
    "mp_server_completed": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_uint64), C.POINTER(C.c_int32),
                                      C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int32)]),
    "mp_server_stats": (C.c_int, [_P, _L, _L, C.POINTER(C.c_double), _L, _L, _L]),
    "mp_server_destroy": (C.c_int, [_P]),
    "mp_gen_philox4x32_10": (None, [C.POINTER(C.c_uint32), C.c_uint64, C.POINTER(C.c_uint32)]),
    "mp_gen_normals": (None, [C.c_uint64, C.c_uint32, C.c_int64, C.c_int64, _D]),
    "mp_gen_uniforms": (None, [C.c_uint64, C.c_uint32, C.c_int64, C.c_int64, _D]),
    "mp_gen_random_walk": (None, [C.c_uint64, C.c_uint32, C.c_int64, _D]),
    "mp_gen_precision_series": (C.c_int, [C.c_uint64, C.c_int64, C.c_int32, _D, _D]),
}

_lib = None


def lib() -> C.CDLL:
    """Load libscrimp_hip.so; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m thesiscode_amd.build` "
                "(the HIP library is required; there is no CPU fallback)")
        handle = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | getattr(os, "RTLD_GLOBAL", 0))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status: int, where: str) -> None:
    if status != MP_OK:
        raise ScrimpError(status, where)
