"""ctypes binding of libmossq_hip.so (the C ABI declared in include/mossq_hip.h).

The library is loaded from this package directory.  There is no CPU fallback:
if the library is missing or no GPU is present, every device entry point
raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libmossq_hip.so"

MOSSQ_OK = 0
MOSSQ_EINVAL = 1
MOSSQ_EOOM = 2
MOSSQ_EDEVICE = 3
MOSSQ_EUNSUPPORTED = 4

OP_GATE = 0
OP_CHANNEL = 1


class MossqError(RuntimeError):
    """A nonzero status from libmossq_hip (message from mossq_last_error)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[mossq status {code}] {msg}")
        self.code = code


class MossqOOM(MossqError):
    pass


class MossqUnsupported(MossqError):
    pass


class OpRecord(C.Structure):
    """Mirror of struct mossq_op."""

    _fields_ = [
        ("kind", C.c_int32),
        ("k", C.c_int32),
        ("pos", C.c_int32 * 5),
        ("mat", C.c_int32),
        ("dmat", C.c_int32),
        ("reserved", C.c_int32),
    ]


# (name, restype, argtypes)
_P = C.c_void_p
_I = C.c_int
_U64 = C.c_uint64
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)
SIGNATURES = [
    ("mossq_version", _I, []),
    ("mossq_last_error", C.c_char_p, []),
    ("mossq_device_count", _I, [C.POINTER(C.c_int)]),
    ("mossq_ctx_create", _I, [_I, _U64, C.POINTER(_P)]),
    ("mossq_ctx_destroy", _I, [_P]),
    ("mossq_sync", _I, [_P]),
    ("mossq_mem_in_use", _I, [_P, C.POINTER(_U64)]),
    ("mossq_alloc", _I, [_P, _U64, C.POINTER(_P)]),
    ("mossq_free", _I, [_P, _P]),
    ("mossq_memset_zero", _I, [_P, _P, _U64]),
    ("mossq_copy_h2d", _I, [_P, _P, _P, _U64]),
    ("mossq_copy_d2h", _I, [_P, _P, _P, _U64]),
    ("mossq_copy_d2d", _I, [_P, _P, _P, _U64]),
    ("mossq_sv_init_zero", _I, [_P, _P, _I]),
    ("mossq_dm_init_zero", _I, [_P, _P, _I]),
    ("mossq_dm_from_sv", _I, [_P, _P, _P, _I]),
    ("mossq_sv_apply_gate", _I, [_P, _P, _I, _I, _IP, _DP]),
    ("mossq_dm_apply_gate", _I, [_P, _P, _I, _I, _IP, _DP]),
    ("mossq_dm_apply_channel", _I, [_P, _P, _I, _I, _IP, _DP, _I]),
    ("mossq_obs_create", _I, [_P, _I, _I, _IP, _IP, _DP, C.POINTER(_P)]),
    ("mossq_obs_destroy", _I, [_P]),
    ("mossq_sv_expectation", _I, [_P, _P, _P, _DP]),
    ("mossq_sv_cross_expectation", _I, [_P, _P, _P, _P, _DP]),
    ("mossq_sv_apply_operator", _I, [_P, _P, _P, _P]),
    ("mossq_dm_expectation", _I, [_P, _P, _P, _DP]),
    ("mossq_dm_apply_operator", _I, [_P, _P, _P, _P]),
    ("mossq_dot", _I, [_P, _P, _P, _U64, _DP]),
    ("mossq_dm_trace", _I, [_P, _P, _I, _DP]),
    ("mossq_dm_diag_bounds", _I, [_P, _P, _I, _DP]),
    ("mossq_probabilities", _I, [_P, _P, _I, _I, _I, _IP, _DP]),
    ("mossq_program_create", _I, [_P, _I, _I, _I, C.POINTER(OpRecord), _DP, _U64, _I, C.POINTER(_P)]),
    ("mossq_program_destroy", _I, [_P]),
    ("mossq_program_sweeps", _I, [_P, C.POINTER(C.c_int)]),
    ("mossq_program_run", _I, [_P, _P, _P]),
    ("mossq_plan_emulate", _I, [_I, _I, _I, C.POINTER(OpRecord), _DP, _U64, _DP, C.POINTER(C.c_int)]),
    ("mossq_plan_emulate_adjoint", _I, [_I, _I, C.POINTER(OpRecord), _DP, _U64, _DP, _DP, _DP, C.POINTER(C.c_int)]),
    ("mossq_plan_emulate_density_adjoint", _I,
     [_I, _I, C.POINTER(OpRecord), _DP, _U64, _I, _DP, _DP, _DP, C.POINTER(C.c_int)]),
    ("mossq_program_update", _I, [_P, _P, _I, C.POINTER(OpRecord), _DP, _U64]),
    ("mossq_program_update_many", _I, [_P, _I, _P, _P, _P, _P, _P, _I, C.POINTER(C.c_int)]),
    ("mossq_program_autotune", _I, [_P, _P, _I, C.POINTER(C.c_int), _DP]),
    ("mossq_program_capture", _I, [_P, _P, _P, C.POINTER(_P)]),
    ("mossq_graph_launch", _I, [_P, _P]),
    ("mossq_graph_destroy", _I, [_P]),
    ("mossq_adjoint", _I, [_P, _P, _P, _P, _I, _P, _DP, _DP]),
    ("mossq_adjoint_noalloc", _I, [_P, _P, _P, _P, _P, _DP, _DP]),
    ("mossq_batch_expectation", _I, [_P, _I, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P), _DP]),
    ("mossq_batch_adjoint", _I, [_P, _I, C.POINTER(_P), C.POINTER(_P), C.POINTER(C.c_int64), _DP, _DP]),
    ("mossq_batch_adjoint_init", _I,
     [_P, _I, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P), C.POINTER(C.c_int64), _DP, _DP]),
    ("mossq_batch_capture", _I, [_P, _I, C.POINTER(_P), C.POINTER(_P), C.POINTER(C.c_int64), C.POINTER(_P)]),
    ("mossq_batch_graph_run", _I, [_P, _P, _DP, _DP]),
    ("mossq_sv_expectation_batch", _I, [_P, _P, _P, _I, _DP]),
    ("mossq_dm_expectation_batch", _I, [_P, _P, _P, _I, _DP]),
    ("mossq_server_create", _I, [_P, _U64, C.POINTER(_P)]),
    ("mossq_server_destroy", _I, [_P]),
    ("mossq_server_submit", _I, [_P, _I, _I, _I, C.POINTER(OpRecord), _DP, _U64, _P, C.POINTER(C.c_int64)]),
    ("mossq_server_step", _I, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("mossq_server_completed", _I, [_P, C.POINTER(C.c_int64), _I, C.POINTER(C.c_int)]),
    ("mossq_server_poll", _I, [_P, C.c_int64, C.POINTER(C.c_int), _DP, _DP, _I]),
    ("mossq_server_memory", _I, [_P, C.POINTER(_U64), C.POINTER(_U64)]),
    ("mossq_batch_launch_tables", _I,
     [_I, _IP, _I, _IP, _U64, C.POINTER(C.c_uint32), C.POINTER(C.c_int), C.POINTER(_U64)]),
    ("mossq_event_record", _I, [_P, _I]),
    ("mossq_event_elapsed_ms", _I, [_P, _I, _I, C.POINTER(C.c_float)]),
]

_lib = None


def load() -> C.CDLL:
    """Load libmossq_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m mossquantum_amd.build` "
            "(there is no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH), mode=C.RTLD_GLOBAL)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int) -> None:
    if status == MOSSQ_OK:
        return
    msg = load().mossq_last_error().decode(errors="replace")
    if status == MOSSQ_EOOM:
        raise MossqOOM(status, msg)
    if status == MOSSQ_EUNSUPPORTED:
        raise MossqUnsupported(status, msg)
    raise MossqError(status, msg)


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_DP)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(_IP)


def device_count() -> int:
    n = C.c_int(0)
    check(load().mossq_device_count(C.byref(n)))
    return n.value


class Context:
    """A device context: one HIP stream, an optional device-memory budget."""

    def __init__(self, device: int = 0, budget_bytes: int = 0):
        self.lib = load()
        h = _P()
        check(self.lib.mossq_ctx_create(device, budget_bytes, C.byref(h)))
        self.h = h
        self.device = device

    def close(self) -> None:
        if self.h:
            check(self.lib.mossq_ctx_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.lib.mossq_ctx_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def sync(self) -> None:
        check(self.lib.mossq_sync(self.h))

    def mem_in_use(self) -> int:
        v = C.c_uint64(0)
        check(self.lib.mossq_mem_in_use(self.h, C.byref(v)))
        return v.value

    def record(self, slot: int) -> None:
        check(self.lib.mossq_event_record(self.h, slot))

    def elapsed_ms(self, s0: int, s1: int) -> float:
        v = C.c_float(0)
        check(self.lib.mossq_event_elapsed_ms(self.h, s0, s1, C.byref(v)))
        return float(v.value)


_default_ctx: Context | None = None


def default_context() -> Context:
    """The process-wide context on device 0 (budget from MOSSQ_MEM_BUDGET bytes)."""
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(int(os.environ.get("MOSSQ_DEVICE", "0")),
                               int(os.environ.get("MOSSQ_MEM_BUDGET", "0")))
    return _default_ctx
