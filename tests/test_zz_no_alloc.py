"""mossq_adjoint_noalloc: the pure-state adjoint with its scratch in the
library's file-scope device array, against mossq_adjoint (GPU tests, through
the C ABI).  Two calls in a row, the second with a smaller register, give the
existing entry point's values and gradients; the call adds nothing to the
context's device memory; registers whose scratch does not fit are refused."""
import ctypes as C

import numpy as np
import pytest

import mossquantum_amd as mq
from mossquantum_amd import _lib
from mossquantum_amd.workloads import hea_circuit, tfim, xxz

pytestmark = pytest.mark.gpu
EXP, GRAD = 1e-10, 1e-9


def _adjoint(ctx, fn, prog, dob, inp=None, out=None):
    v = C.c_double(0)
    g = np.zeros(max(prog.lowered.nderiv, 1))
    args = [ctx.h, prog.h, dob.h, None if inp is None else inp.storage.ptr]
    if fn == "mossq_adjoint":
        args.append(0)
    args += [None if out is None else out.storage.ptr, C.byref(v), _lib.dptr(g)]
    st = getattr(ctx.lib, fn)(*args)
    return st, v.value, g[:prog.lowered.nderiv]


def test_noalloc_adjoint_matches_adjoint_twice_in_a_row():
    ctx = _lib.Context(0)  # a fresh context: any allocation would show in its memory count
    try:
        rng = np.random.default_rng(70)
        cases = [(18, 4, True, tfim), (11, 3, False, xxz)]  # second call: a smaller register, unfused
        for n, layers, fuse, H in cases:
            tpl = hea_circuit(n, layers, rng.uniform(0, 2 * np.pi, 2 * n * layers))
            prog = mq.Program(tpl, n, False, ctx, fuse=fuse)
            dob = mq.DeviceObservable(H(n), n, ctx)
            held = ctx.mem_in_use()
            st, v, g = _adjoint(ctx, "mossq_adjoint_noalloc", prog, dob)
            assert st == _lib.MOSSQ_OK, ctx.lib.mossq_last_error().decode()
            assert ctx.mem_in_use() == held  # nothing allocated
            st, v0, g0 = _adjoint(ctx, "mossq_adjoint", prog, dob)
            assert st == _lib.MOSSQ_OK
            assert abs(v - v0) < EXP and np.max(np.abs(g - g0)) < GRAD, (n, v - v0)
            assert v == v0 and np.array_equal(g, g0)  # same kernels, same reduction order
    finally:
        ctx.close()


def test_noalloc_adjoint_with_input_and_output_state(ctx):
    n = 14
    rng = np.random.default_rng(71)
    psi = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    psi /= np.linalg.norm(psi)
    tpl = hea_circuit(n, 2, rng.uniform(0, 2 * np.pi, 4 * n))
    prog = mq.Program(tpl, n, False, ctx, fuse=True)
    dob = mq.DeviceObservable(xxz(n), n, ctx)
    inp = mq.StateVector.from_host(psi, ctx)
    outs = [mq.StateVector.zeros(n, ctx) for _ in range(2)]
    st, v, g = _adjoint(ctx, "mossq_adjoint_noalloc", prog, dob, inp, outs[0])
    assert st == _lib.MOSSQ_OK
    st, v0, g0 = _adjoint(ctx, "mossq_adjoint", prog, dob, inp, outs[1])
    assert st == _lib.MOSSQ_OK
    assert abs(v - v0) < EXP and np.max(np.abs(g - g0)) < GRAD
    assert np.array_equal(outs[0].to_host(), outs[1].to_host())
    assert np.array_equal(inp.to_host(), psi)  # the input is not modified


def test_noalloc_adjoint_refuses_what_does_not_fit(ctx):
    n = 21  # two 32 MiB vectors alone pass the 64 MiB array
    prog = mq.Program(hea_circuit(n, 1), n, False, ctx, fuse=False)
    dob = mq.DeviceObservable(tfim(n), n, ctx)
    st, _, _ = _adjoint(ctx, "mossq_adjoint_noalloc", prog, dob)
    assert st == _lib.MOSSQ_EINVAL
    assert "scratch" in ctx.lib.mossq_last_error().decode()
    dprog = mq.Program(hea_circuit(4, 1, noise=(0.01, 0.02)), 4, True, ctx)
    st, _, _ = _adjoint(ctx, "mossq_adjoint_noalloc", dprog, mq.DeviceObservable(tfim(4), 4, ctx))
    assert st == _lib.MOSSQ_EUNSUPPORTED
