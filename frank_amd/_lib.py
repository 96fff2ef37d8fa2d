"""ctypes binding of ``libfrank_amd.so`` (declarations: include/frank_amd.h).

The library is built in tree by ``frank_amd/build.py``.  There is no CPU
fallback: if the shared library is missing, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libfrank_amd.so"
HEADER = HERE.parent / "include" / "frank_amd.h"

FRK_OK = 0
FRK_ERR_ARG = -1
FRK_ERR_HIP = -2
FRK_ERR_CAPACITY = -3
FRK_ERR_UNSUPPORTED = -4

TILE_DENSE, TILE_LOWRANK, TILE_EMPTY = 0, 1, 2
LRA_NAIVE, LRA_ROUNDED, LRA_FAST = 0, 1, 2
KERNEL_LAPLACE, KERNEL_GAUSSIAN, KERNEL_IDENTITY, KERNEL_GAUSSIAN_DELL = 0, 1, 2, 3
ADMIS_POSITION, ADMIS_GEOMETRY = 0, 1
PREC_F64, PREC_F32 = 0, 1


class FrankError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"frank_amd error {code}: {msg}")
        self.code = code


_lib = None

vp = C.c_void_p
i32 = C.c_int32
i64 = C.c_int64
dbl = C.c_double
pdbl = C.POINTER(C.c_double)
pi32 = C.POINTER(C.c_int32)
pi64 = C.POINTER(C.c_int64)

_SIGS = {
    "frk_last_error": ([], C.c_char_p),
    "frk_version": ([], C.c_int),
    "frk_ctx_create": ([C.POINTER(vp)], C.c_int),
    "frk_ctx_destroy": ([vp], C.c_int),
    "frk_ctx_synchronize": ([vp], C.c_int),
    "frk_blr_create_kernel": ([vp, C.c_int, C.c_int, i64, vp, dbl, dbl, C.c_int, dbl, C.c_int, dbl, C.c_int,
                               C.c_int, C.c_int, C.POINTER(vp)], C.c_int),
    "frk_blr_create": ([vp, C.c_int, C.c_int, vp, vp, vp, C.c_int, dbl, C.c_int, C.c_int, C.POINTER(vp)], C.c_int),
    "frk_blr_destroy": ([vp], C.c_int),
    "frk_blr_create_kernel_rect": ([vp, C.c_int, C.c_int, i64, vp, i64, i64, C.c_int, C.c_int, dbl, dbl, dbl, C.c_int,
                                    dbl, C.c_int, C.c_int, C.c_int, C.POINTER(vp)], C.c_int),
    "frk_blr_qr": ([vp, C.POINTER(vp)], C.c_int),
    "frk_kernel_apply": ([vp, C.c_int, C.c_int, i64, vp, dbl, dbl, vp, vp, C.c_int], C.c_int),
    "frk_qrf_destroy": ([vp], C.c_int),
    "frk_qrf_get_t": ([vp, C.c_int, C.c_int, vp, vp], C.c_int),
    "frk_blr_apply_q": ([vp, vp, vp, C.c_int], C.c_int),
    "frk_ctx_stream": ([vp, C.POINTER(vp)], C.c_int),
    "frk_ctx_counters": ([vp, pi64, C.c_int, C.c_int], C.c_int),
    "frk_ctx_time_class": ([vp, C.c_char_p], C.c_int),
    "frk_ctx_set_rank_cap": ([vp, C.c_int], C.c_int),
    "frk_ctx_op_classes": ([vp, C.c_char_p, C.c_int], C.c_int),
    "frk_ctx_op_roofline": ([vp, C.c_char_p, vp], C.c_int),
    "frk_ctx_op_stats": ([vp, C.c_char_p, pi64, pdbl, pdbl, pdbl, pi64], C.c_int),
    "frk_ctx_reset_op_stats": ([vp], C.c_int),
    "frk_blr_clone": ([vp, C.POINTER(vp)], C.c_int),
    "frk_blr_copy": ([vp, vp], C.c_int),
    "frk_blr_convert": ([vp, C.c_int, C.POINTER(vp)], C.c_int),
    "frk_blr_refine": ([vp, vp, vp, C.c_int, dbl, C.c_int, C.POINTER(C.c_int), C.POINTER(dbl)], C.c_int),
    "frk_blr_set_lra": ([vp, C.c_int], C.c_int),
    "frk_blr_shape": ([vp, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
    "frk_blr_tile_info": ([vp, C.c_int, C.c_int, pi32, pi32, pi64, pi64], C.c_int),
    "frk_blr_get_dense": ([vp, C.c_int, C.c_int, vp], C.c_int),
    "frk_blr_set_dense": ([vp, C.c_int, C.c_int, vp], C.c_int),
    "frk_blr_get_lowrank": ([vp, C.c_int, C.c_int, vp, vp, vp], C.c_int),
    "frk_blr_set_lowrank": ([vp, C.c_int, C.c_int, C.c_int, vp, vp, vp], C.c_int),
    "frk_blr_getrf": ([vp], C.c_int),
    "frk_blr_getrf_steps": ([vp, C.c_int], C.c_int),
    "frk_blr_solve": ([vp, vp, C.c_int, C.c_int], C.c_int),
    "frk_blr_trsm": ([vp, vp, C.c_int, C.c_int, C.c_int], C.c_int),
    "frk_blr_matvec": ([vp, vp, vp, C.c_int], C.c_int),
    "frk_blr_kernel_residual": ([vp, vp, vp, vp, C.c_int], C.c_int),
    "frk_dgetrf_nopiv_batched": ([vp, vp, C.c_int, C.c_int, i64, vp], C.c_int),
    "frk_dgetrf_nopiv_batched_noalloc": ([vp, vp, C.c_int, C.c_int, i64, vp], C.c_int),
    "frk_dtrsm": ([vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp, i64, vp, i64], C.c_int),
    "frk_dgemm": ([vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dbl, vp, i64, vp, i64, dbl, vp, i64], C.c_int),
    "frk_dlr_add_batched": ([vp, C.c_int, C.c_int, vp, vp, dbl, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp],
                            C.c_int),
    "frk_slr_add_batched": ([vp, C.c_int, C.c_int, vp, vp, dbl, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp],
                            C.c_int),
    "frk_dgemm_tiles_batched": ([vp, C.c_int, vp, vp, vp, dbl, dbl, C.c_int, C.c_int, C.c_int], C.c_int),
    "frk_dgeqrt_batched": ([vp, C.c_int, C.c_int, vp, i64, vp, i64], C.c_int),
    "frk_dlarfb_batched": ([vp, C.c_int, C.c_int, vp, i64, vp, i64, vp, C.c_int], C.c_int),
    "frk_dtpqrt_batched": ([vp, C.c_int, C.c_int, vp, i64, vp, vp, i64], C.c_int),
    "frk_dtpmqrt_batched": ([vp, C.c_int, C.c_int, vp, vp, i64, vp, vp, C.c_int, C.c_int], C.c_int),
    "frk_dqr_batched": ([vp, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp], C.c_int),
    "frk_drq_batched": ([vp, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp], C.c_int),
    "frk_dsvd_batched": ([vp, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp], C.c_int),
    "frk_dtruncated_geqp3_batched": ([vp, C.c_int, vp, vp, vp, vp, dbl, C.c_int, vp, vp, vp, vp, vp], C.c_int),
    "frk_drsvd": ([vp, C.c_int, C.c_int, vp, i64, C.c_int, vp, i64, vp, i64, vp, i64], C.c_int),
    "frk_did": ([vp, C.c_int, C.c_int, vp, i64, C.c_int, vp, i64, vp], C.c_int),
    "frk_dcompress": ([vp, C.c_int, C.c_int, vp, i64, dbl, C.c_int, C.c_int, vp, vp, vp, vp], C.c_int),
}


def header_symbols() -> list:
    """Every ``frk_*`` function declared in include/frank_amd.h."""
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(frk_\w+)\s*\(", text, flags=re.M)))


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: run frank_amd/build.py (no CPU fallback exists)")
        L = C.CDLL(str(LIB_PATH))
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(rc: int) -> int:
    if rc != FRK_OK:
        msg = lib().frk_last_error()
        raise FrankError(rc, msg.decode() if msg else "")
    return rc
