"""C-ABI checks that need no GPU: the library loads, exports every entry point
declared in include/mossq_hip.h, and refuses to run without a device (no CPU
fallback)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import mossquantum_amd as mq
from mossquantum_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "mossq_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mossq_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_abi():
    syms = declared_symbols()
    assert "mossq_sv_apply_gate" in syms and "mossq_adjoint" in syms and len(syms) >= 30


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for s in declared_symbols():
        assert hasattr(lib, s), s


def test_bindings_cover_the_header():
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert set(declared_symbols()) == bound


def test_op_record_layout():
    assert C.sizeof(_lib.OpRecord) == 40


def test_no_cpu_fallback_without_device():
    if mq.device_count() > 0:
        pytest.skip("a device is visible")
    h = C.c_void_p()
    st = _lib.load().mossq_ctx_create(0, 0, C.byref(h))
    assert st == _lib.MOSSQ_EDEVICE
    assert "no HIP device" in _lib.load().mossq_last_error().decode()
    with pytest.raises(_lib.MossqError):
        mq.Context(0)
