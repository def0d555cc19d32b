"""ctypes binding of libbss.so (the C ABI declared in include/bss.h).

The library is built in-tree next to this file (``__graft_entry__.build()`` or
``make -C blockstructuredsolvers_amd/csrc``).  There is no fallback: if the shared
library is missing, importing the solver API raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbss.so")

BSS_F32, BSS_F64 = 0, 1
BSS_MODE_PLAIN, BSS_MODE_MIXED = 0, 1
BSS_SCHED_AUTO, BSS_SCHED_REFERENCE = 0, 1
BSS_OK, BSS_NOT_CONVERGED = 0, 1
BSS_E_ARG, BSS_E_HIP, BSS_E_NOMEM, BSS_E_STATE, BSS_E_UNSUPPORTED = -1, -2, -3, -4, -5

BSS_Q_N_USE = 0
BSS_Q_LEVELS = 1
BSS_Q_WORKSPACE_BYTES = 2
BSS_Q_LEVEL_N = 16
BSS_Q_LEVEL_CHAINS = 48
BSS_Q_LEVEL_MAXLEN = 80
BSS_Q_LAST_ITERS = 112
BSS_Q_REFACTORED = 113
BSS_Q_FALLBACK = 114

BSS_OPT_TOL = 0
BSS_OPT_MAX_ITERS = 1
BSS_OPT_PIVOT_TOL = 2
BSS_OPT_CHAIN_LEN = 3
BSS_OPT_BASE_MAX = 4
BSS_OPT_PRESERVE_INPUT = 5
BSS_OPT_FACTOR_WAVES = 6
BSS_OPT_GRAPH = 7
BSS_OPT_PROFILE_KINDS = 8
BSS_OPT_PROFILE_MAX_LEVEL = 9
BSS_OPT_FALLBACK = 10
BSS_OPT_FACTOR_SWEEP = 11
BSS_OPT_CHAIN_LEN_LEVEL = 32

KERNEL_KINDS = {"factor": 0, "assemble_diag": 1, "forward": 2, "assemble_rhs": 3,
                "backward": 4, "scatter": 5, "residual": 6, "residual_init": 7,
                "berr": 8}

# every symbol include/bss.h declares, with its ctypes signature
_VP, _I, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
SIGNATURES = {
    "bss_plan_create": ([_I, _I, _I, _I, _I, _I, _I, ctypes.POINTER(_VP)], _I),
    "bss_plan_create_group": ([_I, ctypes.POINTER(_I), _I, _I, _I, _I, _I, ctypes.POINTER(_VP)], _I),
    "bss_plan_system_sizes": ([_VP, ctypes.POINTER(_I)], _I),
    "bss_set_option": ([_VP, _I, _D], _I),
    "bss_plan_get": ([_VP, _I, ctypes.POINTER(ctypes.c_int64)], _I),
    "bss_plan_set_workspace": ([_VP, _VP, ctypes.c_size_t], _I),
    "bss_plan_bind": ([_VP, _VP, _VP, _VP], _I),
    "bss_build_smoother": ([_I, _I, _I, _VP, _VP, _VP, _VP, _VP, _VP, _I, _VP], _I),
    "bss_factorize": ([_VP, _VP], _I),
    "bss_solve": ([_VP, _VP], _I),
    "bss_factorize_solve": ([_VP, _VP], _I),
    "bss_factorize_solve_noalloc": ([_I, _I, _I, _I, _VP, _VP, _VP, _VP], _I),
    "bss_info": ([_VP, ctypes.POINTER(_I)], _I),
    "bss_last_backward_error": ([_VP, ctypes.POINTER(_D)], _I),
    "bss_sync": ([_VP], _I),
    "bss_plan_destroy": ([_VP], _I),
    "bss_last_error": ([ctypes.c_char_p, ctypes.c_size_t], _I),
    "bss_reference_schedule": ([_I, ctypes.POINTER(_I), ctypes.POINTER(_I), _I], _I),
    "bss_plan_level_chains": ([_VP, _I, ctypes.POINTER(_I), _I], _I),
    "bss_batch_create": ([_I, ctypes.POINTER(_I), ctypes.POINTER(_I), _I, _I, _I, _I,
                          ctypes.POINTER(_VP)], _I),
    "bss_batch_set_option": ([_VP, _I, _D], _I),
    "bss_batch_layout": ([_VP, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                          ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)], _I),
    "bss_batch_get": ([_VP, _I, ctypes.POINTER(ctypes.c_int64)], _I),
    "bss_batch_bind": ([_VP, _VP, _VP, _VP], _I),
    "bss_batch_factorize": ([_VP, _VP], _I),
    "bss_batch_solve": ([_VP, _VP], _I),
    "bss_batch_info": ([_VP, ctypes.POINTER(_I)], _I),
    "bss_batch_last_backward_error": ([_VP, ctypes.POINTER(_D)], _I),
    "bss_batch_sync": ([_VP], _I),
    "bss_batch_destroy": ([_VP], _I),
    "bss_profile_enable": ([_VP, _I], _I),
    "bss_profile_read": ([_VP, _I, _I, ctypes.POINTER(_I), ctypes.POINTER(_D)], _I),
}

_lib = None


class BssError(RuntimeError):
    """A libbss call returned a negative status."""

    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed ({code}): {msg}")
        self.code = code


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"libbss.so is not built ({LIB_PATH}); run __graft_entry__.build() or "
                "`make -C blockstructuredsolvers_amd/csrc`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    lib().bss_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(fn: str, rc: int) -> int:
    if rc < 0:
        raise BssError(fn, rc, last_error())
    return rc


def call(fn: str, *args) -> int:
    return check(fn, getattr(lib(), fn)(*args))
