"""Edge cases and full-size properties on the device: empty circuits and
parameterless objectives, one-qubit registers, the headline sizes checked
through size-independent properties (fused against unfused, norm, trace,
adjoint against parameter shift), batches of one."""
import numpy as np
import pytest

import oracle
import mossquantum_amd as mq
from mossquantum_amd import lower
from mossquantum_amd.gates import haar_unitary
from mossquantum_amd.workloads import haar_layers, hea_angles, hea_circuit, random_circuit, tfim, xxz

pytestmark = pytest.mark.gpu
AMP, EXP, GRAD = 1e-12, 1e-10, 1e-9


def test_empty_circuits_and_parameterless_objectives(ctx):
    for n in (1, 5, 13):
        v = np.random.default_rng(n).standard_normal(1 << n).astype(np.complex128)
        v /= np.linalg.norm(v)
        for fuse in (True, False):
            sv = mq.apply_(mq.QCircuit(), mq.StateVector.from_host(v, ctx), mq.HIPBackend(ctx, fuse=fuse))
            assert np.array_equal(sv.to_host(), v)
        obj = mq.ExpectationObjective(mq.QCircuit([mq.HGate(1)]), tfim(n), mq.StateVector.from_host(v, ctx))
        val, g = mq.value_and_gradient(mq.AdjointGradient(), obj, [])
        want = oracle.adjoint(lower(mq.QCircuit([mq.HGate(1)]), n, False), tfim(n).groups(), state_in=v)[0]
        assert g.shape == (0,) and abs(val - want) < EXP
    dobj = mq.ExpectationObjective(mq.QCircuit([mq.AmplitudeDamping(1, 0.3)]), tfim(6), mq.DensityMatrix.zeros(6, ctx))
    val, g = mq.value_and_gradient(mq.AdjointGradient(), dobj, [])
    assert g.shape == (0,) and np.isfinite(val)


def test_one_qubit_registers(ctx):
    th = 0.7
    for is_dm in (False, True):
        st = mq.DensityMatrix.zeros(1, ctx) if is_dm else mq.StateVector.zeros(1, ctx)
        c = mq.QCircuit([mq.RyGate(1, th), mq.AmplitudeDamping(1, 0.25)] if is_dm else [mq.RyGate(1, th)])
        out = mq.simulate(c, st).state
        p1 = np.sin(th / 2) ** 2 * (0.75 if is_dm else 1.0)
        assert abs(mq.probabilities(out)[1] - p1) < 1e-15


def test_headline_size_properties(ctx):
    """The 26-qubit depth-20 circuit (1 GiB state): the fused plan equals the
    unfused sweeps amplitude for amplitude, and the norm stays one."""
    n, d = 26, 20
    c = random_circuit(n, d, haar_layers(n, d, 2026))
    a = mq.StateVector.zeros(n, ctx)
    mq.Program(c, n, False, ctx, fuse=True).run(a.storage)
    ha = a.to_host()
    del a
    b = mq.StateVector.zeros(n, ctx)
    mq.Program(c, n, False, ctx, fuse=False).run(b.storage)
    hb = b.to_host()
    del b
    assert np.max(np.abs(ha - hb)) < AMP
    assert abs(np.vdot(ha, ha).real - 1) < 1e-12


def test_density_headline_size_properties(ctx):
    """The 13-qubit noisy density matrix (1 GiB): fused against unfused, unit
    trace, Hermitian diagonal, energy agreement."""
    n, L = 13, 8
    c = hea_circuit(n, L, hea_angles(n, L, 7), noise=(0.01, 0.02))
    res = {}
    for fuse in (True, False):
        dm = mq.DensityMatrix.zeros(n, ctx)
        mq.Program(c, n, True, ctx, fuse=fuse).run(dm.storage)
        res[fuse] = (mq.expectation(tfim(n), dm), mq.trace(dm), mq.purity(dm), mq.probabilities(dm, [1, 13]))
        del dm
    (e1, t1, p1, q1), (e2, t2, p2, q2) = res[True], res[False]
    assert abs(e1 - e2) < EXP and abs(t1 - 1) < 1e-12 and abs(t2 - 1) < 1e-12
    assert abs(p1 - p2) < 1e-12 and np.max(np.abs(q1 - q2)) < 1e-14


def test_adjoint_at_large_size_matches_parameter_shift(ctx):
    """A 20-qubit fused adjoint against the parameter-shift gradient of the same
    objective (different algorithm, same device state path)."""
    n = 20
    tpl = hea_circuit(n, 1)
    th = np.random.default_rng(20).uniform(0, 2 * np.pi, tpl.nparameters())
    obj = mq.ExpectationObjective(tpl, xxz(n), mq.StateVector.zeros(n, ctx))
    v, g = mq.value_and_gradient(mq.AdjointGradient(), obj, th)
    est = mq.estimate_gradient(mq.ParameterShift(), obj, th)
    assert abs(est.value - v) < EXP and np.max(np.abs(est.gradient - g)) < GRAD


def test_batches_of_one(ctx):
    n = 10
    c = hea_circuit(n, 2, np.random.default_rng(1).uniform(0, 6, 4 * n))
    vals, grads = mq.batched_adjoint([c], [xxz(n)], [n], ctx)
    v, g = oracle.adjoint(lower(c, n, False), xxz(n).groups())
    assert abs(vals[0] - v) < EXP and np.max(np.abs(grads[0] - g)) < GRAD
    nop = mq.QCircuit([mq.QuantumGate((1, 2), haar_unitary(np.random.default_rng(3), 4))])
    vals, grads = mq.batched_adjoint([nop], [xxz(n)], [n], ctx)
    assert grads[0].shape == (0,)
    assert abs(vals[0] - oracle.adjoint(lower(nop, n, False), xxz(n).groups())[0]) < EXP


def test_block_cache_under_a_budget():
    """The library keeps released scratch blocks for reuse; on a context with
    a memory budget they count as in use, and an allocation that only fits
    without them releases them first.  Repeated calls reuse the blocks and
    give identical results."""
    import ctypes as C
    n = 18
    budget = 48 << 20
    small = mq._lib.Context(0, budget)
    try:
        c = hea_circuit(n, 1, np.random.default_rng(3).uniform(0, 6.3, 2 * n))
        want = oracle.adjoint(lower(c, n, False), xxz(n).groups())
        vals = []
        for _ in range(2):
            v, g = mq.batched_adjoint([c], [xxz(n)], [n], small, fuse=False)
            vals.append((v[0], g[0]))
            assert abs(v[0] - want[0]) < EXP
        assert vals[0][0] == vals[1][0] and np.array_equal(vals[0][1], vals[1][1])
        held = small.mem_in_use()
        assert held > 0  # the scratch blocks stay with the context
        ptr = C.c_void_p()
        # a buffer that only fits once the cached blocks are released
        size = budget - held + (1 << 20)
        mq._lib.check(small.lib.mossq_alloc(small.h, size, C.byref(ptr)))
        assert small.mem_in_use() <= budget
        mq._lib.check(small.lib.mossq_free(small.h, ptr))
        with pytest.raises(mq._lib.MossqError):
            mq._lib.check(small.lib.mossq_alloc(small.h, budget + 1, C.byref(ptr)))
    finally:
        small.close()
