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
