"""Host readout sampling (mossquantum_amd.readout) against the oracle's loop
restatement of the reference's _draw_samples! and discrete_sample, with the
same uniforms.  No device needed."""
import numpy as np
import pytest

from mossquantum_amd import readout
from oracle import sampling as ref


class FixedUniforms:
    """An rng handing out a fixed list of uniforms through rand()."""

    def __init__(self, u):
        self.u = list(u)

    def rand(self):
        return self.u.pop(0)


# ---------------------------------------------------------- oracle pins ----
def test_oracle_point_mass():
    p = [0.0, 0.0, 1.0, 0.0]
    for u in (0.0, 0.3, 0.999999):
        assert ref.draw_samples([u], p, 2) == [[0, 1]]
        assert ref.discrete_sample(u, p) == 3


def test_oracle_hand_worked_draws():
    # cumulative (sequential) = 0.1, 0.30000000000000004, 0.6000000000000001, 1.0
    p = [0.1, 0.2, 0.3, 0.4]
    u = [0.05, 0.1, 0.2, 0.45, 0.7, 0.99]
    # 0.1 lands on the first boundary: entries <= 0.1 number 1, so bin 2 (basis 1)
    want = [0, 1, 1, 2, 3, 3]
    got = [s[0] + 2 * s[1] for s in ref.draw_samples(u, p, 2)]
    assert got == want
    assert [ref.discrete_sample(x, p) - 1 for x in u] == want


def test_oracle_trailing_zeros_never_drawn():
    p = [0.5, 0.5, 0.0, 0.0]
    u = list(np.linspace(0, 1, 257, endpoint=False))
    assert all(s[1] == 0 for s in ref.draw_samples(u, p, 2))


def test_oracle_draws_lie_in_their_interval():
    rng = np.random.default_rng(3)
    for _ in range(20):
        p = rng.random(16) * (rng.random(16) > 0.3)
        p[rng.integers(16)] += 0.1
        c = ref.cumsum(p)
        u = list(rng.random(50))
        for x, s in zip(u, ref.draw_samples(u, p, 4)):
            b = sum(bit << j for j, bit in enumerate(s))
            lo = c[b - 1] if b else 0.0
            assert lo <= x * c[-1] < c[b] and p[b] > 0


# ----------------------------------------------------- product vs oracle ----
@pytest.mark.parametrize("k", [1, 3, 6])
def test_draw_samples_matches_oracle(k):
    rng = np.random.default_rng(10 + k)
    for trial in range(5):
        p = rng.random(1 << k) * (rng.random(1 << k) > 0.4)
        p[0] += 1e-3
        if trial % 2:
            p /= p.sum()
        u = rng.random(400)
        u[:3] = [0.0, np.nextafter(1.0, 0.0), ref.cumsum(p)[0] / ref.cumsum(p)[-1]]
        got = readout.draw_samples(FixedUniforms(u), p, 400, k)
        assert got.shape == (400, k)
        assert got.tolist() == ref.draw_samples(list(u), p, k)


def test_draw_samples_numpy_generator_stream():
    """A numpy Generator supplies its random() stream, one uniform per shot."""
    p = [0.25, 0.0, 0.5, 0.25]
    got = readout.draw_samples(np.random.default_rng(91), p, 64, 2)
    u = list(np.random.default_rng(91).random(64))
    assert got.tolist() == ref.draw_samples(u, p, 2)
    assert np.array_equal(got, readout.draw_samples(np.random.default_rng(91), p, 64, 2))


def test_discrete_sample_matches_oracle():
    rng = np.random.default_rng(4)
    for _ in range(200):
        w = rng.random(4) * (rng.random(4) > 0.3) + np.array([0, 0, 0, 1e-3])
        u = rng.random()
        assert readout.discrete_sample(FixedUniforms([u]), w) == ref.discrete_sample(u, w)


def test_histogram_keys_leftmost_first_qubit():
    draws = np.array([[1, 0], [1, 0], [0, 1]])
    assert readout._histogram(draws) == {"10": 2, "01": 1}
    assert readout._histogram(np.zeros((3, 0), dtype=int)) == {"": 3}


def test_validation():
    for bad in ([-0.1, 1.1], [np.nan, 1.0], [0.0, 0.0], [np.inf, 0.0]):
        with pytest.raises(ValueError):
            readout.draw_samples(FixedUniforms([0.5]), bad, 1, 1)
    with pytest.raises(ValueError):
        readout.discrete_sample(FixedUniforms([0.5]), [])
    with pytest.raises(ValueError):
        readout.discrete_sample(FixedUniforms([0.5]), [0.0, 0.0])
