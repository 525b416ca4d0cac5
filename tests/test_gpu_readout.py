"""Readout on the device: marginals from the HIP library, seeded draws and
counts on the host, following the reference's execution_api tests
(test/integration/execution_api.jl:30-110)."""
import numpy as np
import pytest

import oracle
from oracle import sampling as ref
import mossquantum_amd as mq

pytestmark = pytest.mark.gpu


def bell_state():
    v = np.zeros(4, dtype=np.complex128)
    v[0] = v[3] = 1 / np.sqrt(2)
    return v


def test_marginals_and_order(ctx):
    a = np.sqrt(np.array([0.1, 0.2, 0.3, 0.4], dtype=np.complex128))
    sv = mq.StateVector.from_host(a)
    assert np.allclose(mq.probabilities(sv), [0.1, 0.2, 0.3, 0.4], atol=1e-15)
    assert np.allclose(mq.probabilities(sv, [1]), [0.4, 0.6], atol=1e-15)
    assert np.allclose(mq.probabilities(sv, [2]), [0.3, 0.7], atol=1e-15)
    assert np.allclose(mq.probabilities(sv, [2, 1]), [0.1, 0.3, 0.2, 0.4], atol=1e-15)
    dm = mq.DensityMatrix.from_host(np.outer(a, a.conj()))
    assert np.allclose(mq.probabilities(dm, [2, 1]), [0.1, 0.3, 0.2, 0.4], atol=1e-15)
    assert np.array_equal(mq.probabilities(mq.StateVector.from_host(bell_state()), []), [1.0])


def test_seeded_sampling_and_counts(ctx):
    bell = mq.StateVector.from_host(bell_state())
    rho = mq.DensityMatrix.from_host(np.outer(bell_state(), bell_state().conj()))
    draws = mq.sample(bell, 64, np.random.default_rng(91))
    assert draws.shape == (64, 2) and np.array_equal(draws[:, 0], draws[:, 1])
    assert np.array_equal(draws, mq.sample(bell, 64, np.random.default_rng(91)))
    assert np.array_equal(draws, mq.sample(rho, 64, np.random.default_rng(91)))
    # bit-exact against the oracle's restatement on the same marginal and uniforms
    p = mq.probabilities(bell)
    assert draws.tolist() == ref.draw_samples(list(np.random.default_rng(91).random(64)), p, 2)
    ones = int(draws[:, 0].sum())
    want = {k: v for k, v in (("00", 64 - ones), ("11", ones)) if v}
    assert mq.counts(bell, 64, np.random.default_rng(91)) == want
    assert mq.counts(rho, 64, np.random.default_rng(91)) == want
    basis = np.zeros(4, dtype=np.complex128)
    basis[1] = 1  # qubit 1 set, qubit 2 clear
    b = mq.StateVector.from_host(basis)
    assert mq.sample(b, 3, np.random.default_rng(4)).tolist() == [[1, 0]] * 3
    assert mq.sample(b, 3, np.random.default_rng(4), qubits=[2, 1]).tolist() == [[0, 1]] * 3
    assert mq.counts(b, 3, np.random.default_rng(4)) == {"10": 3}
    assert mq.counts(b, 3, np.random.default_rng(4), qubits=[2, 1]) == {"01": 3}
    assert mq.sample(bell, 0, np.random.default_rng(4)).shape == (0, 2)
    assert mq.counts(bell, 0, np.random.default_rng(4)) == {}
    assert mq.sample(bell, 3, np.random.default_rng(4), qubits=[]).shape == (3, 0)
    assert mq.counts(bell, 3, np.random.default_rng(4), qubits=[]) == {"": 3}
    assert np.array_equal(bell.to_host(), bell_state())


def test_sampling_larger_register_matches_oracle(ctx):
    """A 12-qubit random state: device marginal against the oracle's, draws
    against the loop restatement on the device marginal."""
    n = 12
    rng = np.random.default_rng(12)
    v = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    v /= np.linalg.norm(v)
    sv = mq.StateVector.from_host(v)
    qs = [3, 11, 1, 7]
    p = mq.probabilities(sv, qs)
    want = oracle.probabilities(v, n, False, qs)
    assert np.max(np.abs(p - want / want.sum())) < 1e-14
    u = list(np.random.default_rng(5).random(2000))
    assert mq.sample(sv, 2000, np.random.default_rng(5), qubits=qs).tolist() == ref.draw_samples(u, p, 4)


def test_batch_dimensions(ctx):
    basis = np.zeros(4, dtype=np.complex128)
    basis[1] = 1
    batch = mq.BatchStateVector([mq.StateVector.from_host(bell_state()), mq.StateVector.from_host(basis)])
    P = mq.probabilities(batch)
    assert P.shape == (4, 2) and np.allclose(P[:, 1], [0, 1, 0, 0])
    assert np.allclose(mq.probabilities(batch, [2, 1])[:, 1], [0, 0, 1, 0])
    draws = mq.sample(batch, 16, np.random.default_rng(11))
    assert draws.shape == (16, 2, 2)
    assert np.array_equal(draws[:, 0, 0], draws[:, 1, 0])
    assert np.all(draws[:, 0, 1] == 1) and np.all(draws[:, 1, 1] == 0)
    h = mq.counts(batch, 16, np.random.default_rng(11))
    assert len(h) == 2 and sum(h[0].values()) == 16 and h[1] == {"10": 16}
    assert mq.sample(batch, 0, np.random.default_rng(11)).shape == (0, 2, 2)


def test_validation(ctx):
    bell = mq.StateVector.from_host(bell_state())
    for qs in ([0], [3], [1, 1]):
        with pytest.raises(ValueError):
            mq.probabilities(bell, qs)
        with pytest.raises(ValueError):
            mq.sample(bell, 2, np.random.default_rng(1), qubits=qs)
    with pytest.raises(ValueError):
        mq.sample(bell, -1, np.random.default_rng(1))
    with pytest.raises(ValueError):
        mq.counts(bell, -1, np.random.default_rng(1))
    with pytest.raises(ValueError):
        mq.probabilities(mq.StateVector.from_host(np.array([2.0, 0.0])))
    with pytest.raises(ValueError):
        mq.sample(mq.StateVector.from_host(np.array([2.0, 0.0])), 0, np.random.default_rng(1))
    with pytest.raises(ValueError, match="sum to one"):
        mq.probabilities(mq.DensityMatrix.from_host(np.diag([0.8, 0.8])))
    with pytest.raises(ValueError, match="real"):
        mq.probabilities(mq.DensityMatrix.from_host(np.diag([0.5j, 1.0])))
    with pytest.raises(ValueError, match="nonnegative"):
        mq.probabilities(mq.DensityMatrix.from_host(np.diag([1.1, -0.1])))


def test_measure_collapses_like_the_reference(ctx):
    """measure_ (reference measure!(backend, state, position),
    execution/simulation.jl:42-51, measurement.jl:4-17): outcome drawn from
    [p0, 1 - p0] with the given uniform, probability returned, and the state
    collapsed by |0><o|/sqrt(p) (auto_reset) or |o><o|/sqrt(p); the same
    uniforms through the oracle's sampler and a host collapse give the same
    outcome and state.  Density matrices, batches and a bad position too."""
    rng = np.random.default_rng(81)
    n, pos = 7, 3
    for auto_reset in (True, False):
        for seed in (1, 2, 3):
            psi = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
            psi /= np.linalg.norm(psi)
            sv = mq.StateVector.from_host(psi, ctx)
            o, p = mq.measure_(sv, pos, auto_reset=auto_reset, rng=np.random.default_rng(seed))
            u = np.random.default_rng(seed).random()
            bit = (np.arange(1 << n) >> (pos - 1)) & 1
            p0 = float(np.sum(np.abs(psi[bit == 0]) ** 2))
            i = ref.discrete_sample(u, [p0, 1 - p0])
            assert o == i - 1 and abs(p - [p0, 1 - p0][i - 1]) < 1e-13
            m = np.zeros((2, 2), dtype=np.complex128)
            m[0 if auto_reset else o, o] = 1 / np.sqrt(p)
            want = psi.copy()
            oracle.apply_gate(want, n, [pos - 1], m)
            assert np.max(np.abs(sv.to_host() - want)) < 1e-12
            got = mq.probabilities(sv, [pos])
            assert abs(got[0 if auto_reset else o] - 1) < 1e-12
    # density matrix
    m_ = 5
    A = rng.standard_normal((1 << m_, 1 << m_)) + 1j * rng.standard_normal((1 << m_, 1 << m_))
    rho = A @ A.conj().T
    rho /= np.trace(rho)
    dm = mq.DensityMatrix.from_host(rho, ctx)
    o, p = mq.measure_(dm, 2, auto_reset=False, rng=np.random.default_rng(5))
    d = np.real(np.diag(rho))
    p0 = float(np.sum(d[((np.arange(1 << m_) >> 1) & 1) == 0]))
    i = ref.discrete_sample(np.random.default_rng(5).random(), [p0, 1 - p0])
    assert o == i - 1 and abs(p - [p0, 1 - p0][i - 1]) < 1e-13
    mm = np.zeros((2, 2), dtype=np.complex128)
    mm[o, o] = 1 / np.sqrt(p)
    want = rho.reshape(-1, order="F").copy()
    oracle.dm_apply_gate(want, m_, [1], mm)
    assert np.max(np.abs(dm.to_host().reshape(-1, order="F") - want)) < 1e-12
    assert abs(mq.trace(dm) - 1) < 1e-12
    # batch: one draw per column, in column order
    cols = [mq.apply_(mq.RyGate(1, t), mq.StateVector.zeros(2, ctx)) for t in (0.3, 2.9, np.pi)]
    batch = mq.BatchStateVector(cols)  # one buffer, the columns copied in
    cols = batch.columns
    outs, probs = mq.measure_(batch, 1, rng=np.random.default_rng(9))
    us = np.random.default_rng(9).random(3)
    for j, (t, u) in enumerate(zip((0.3, 2.9, np.pi), us)):
        p0 = np.cos(t / 2) ** 2
        i = ref.discrete_sample(u, [p0, 1 - p0])
        assert outs[j] == i - 1 and abs(probs[j] - [p0, 1 - p0][i - 1]) < 1e-12
        assert abs(mq.probabilities(cols[j], [1])[0] - 1) < 1e-12  # reset to 0
    with pytest.raises(ValueError):
        mq.measure_(mq.StateVector.zeros(2, ctx), 3)
