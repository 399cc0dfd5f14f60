"""ctypes binding of libnalgebra_rocm.so (the C ABI in include/nalgebra_rocm.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback anywhere in the package.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnalgebra_rocm.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "nalgebra_rocm.h")

OK = 0
ERR_HIP = 1
ERR_ALLOC = 2
ERR_UNSUPPORTED = 3
ERR_NOT_CONVERGED = 4
EXACT = 0
FAST = 1
PREC_FP64 = 0
PREC_FP32 = 1
PREC_BF16 = 2
PREC_BF16X3 = 3
PREC_GMRES = 32  # OR-ed: FGMRES-based refinement preconditioned by the factors
PREC_WINDOW = 64  # OR-ed: pivots searched in the next 128 rows of each 64-column block

_vp = C.c_void_p
_i = C.c_int
_d = C.c_double
_f = C.c_float
_sz = C.c_size_t
_pd = C.c_ssize_t
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)

_PROTOS = {
    "nrocm_create": (_i, [C.POINTER(_vp)]),
    "nrocm_destroy": (_i, [_vp]),
    "nrocm_set_stream": (_i, [_vp, _vp]),
    "nrocm_get_stream": (_i, [_vp, C.POINTER(_vp)]),
    "nrocm_status_string": (C.c_char_p, [_i]),
    "nrocm_last_error": (C.c_char_p, [_vp]),
    "nrocm_version": (_i, []),
    "nrocm_dgemm": (_i, [_vp, _sz, _sz, _sz, _d, _vp, _pd, _pd, _vp, _pd, _pd, _d, _vp, _pd, _pd]),
    "nrocm_sgemm": (_i, [_vp, _sz, _sz, _sz, _f, _vp, _pd, _pd, _vp, _pd, _pd, _f, _vp, _pd, _pd]),
    "nrocm_dgetrf": (_i, [_vp, _i, _i, _vp, _i, _vp, _vp, _i]),
    "nrocm_dgetrs": (_i, [_vp, _i, _i, _vp, _i, _vp, _vp, _i, _vp, _i]),
    "nrocm_dgetrs_noalloc": (_i, [_vp, _i, _i, _vp, _i, _vp, _vp, _i, _vp, _i]),
    "nrocm_dpotrf": (_i, [_vp, _i, _vp, _i, _vp, _i]),
    "nrocm_dpotrs": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _i]),
    "nrocm_dpotrf_substitute": (_i, [_vp, _i, _vp, _i, _d, _vp, _i]),
    "nrocm_ddiag_det": (_i, [_vp, _i, _vp, _i, _vp]),
    "nrocm_batch_create": (_i, [_vp, _i, _ip, _i, C.POINTER(_vp)]),
    "nrocm_batch_destroy": (_i, [_vp]),
    "nrocm_batch_set_graph": (_i, [_vp, _i]),
    "nrocm_batch_autotune": (_i, [_vp, _vp, _i, _vp]),
    "nrocm_batch_schedule": (_i, [_vp, _i, _vp]),
    "nrocm_batch_set_schedule": (_i, [_vp, _i, _vp]),
    "nrocm_dpotrf_potrs_batched": (_i, [_vp, _vp, _vp, _vp, _vp, _i]),
    "nrocm_dpotrs_backward_batched": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "nrocm_dgetrf_getrs_batched": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _i]),
    "nrocm_dgetrs_backward_batched": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "nrocm_dgeqrf": (_i, [_vp, _i, _i, _vp, _i, _vp]),
    "nrocm_sgeqrf": (_i, [_vp, _i, _i, _vp, _i, _vp]),
    "nrocm_dqr_qtmul": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _i]),
    "nrocm_dqr_rsolve": (_i, [_vp, _i, _vp, _i, _vp, _i, _vp, _i, _vp]),
    "nrocm_dqr_qmul": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _i]),
    "nrocm_dqr_rmul": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _i]),
    "nrocm_sgels_batched": (_i, [_vp, _vp, _i, _vp, _vp, _vp, _i, _vp]),
    "nrocm_dgesv_ir": (_i, [_vp, _i, _vp, _i, _vp, _vp, _i, _d, _i, _ip, _dp]),
    "nrocm_ir_fallback": (_i, [_vp, _ip]),
    "nrocm_gen_uniform": (None, [C.c_uint64, _sz, _d, _d, _vp]),
    "nrocm_gen_normal": (None, [C.c_uint64, _sz, _vp]),
    "nrocm_gen_spd_batch": (None, [C.c_uint64, _i, _ip, _i, _vp, _vp, _vp, _vp]),
}

_lib = None


class NrocmError(RuntimeError):
    def __init__(self, fn: str, status: int, detail: str = ""):
        self.status = status
        super().__init__(f"{fn} failed with status {status}: {detail}")


def header_functions():
    """Names of every function the public header declares."""
    with open(HEADER) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(nrocm_\w+)\s*\(", text)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m nalgebra_amd.build` "
                "(this package has no CPU fallback)")
        _lib = C.CDLL(LIB_PATH)
        for name, (res, args) in _PROTOS.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args
    return _lib


def check(fn: str, status: int, handle=None):
    if status != OK:
        detail = ""
        if handle is not None and status > 0:
            detail = (lib().nrocm_last_error(handle) or b"").decode()
        elif status < 0:
            detail = f"invalid argument #{-status}"
        else:
            detail = (lib().nrocm_status_string(status) or b"").decode()
        raise NrocmError(fn, status, detail)
    return status
