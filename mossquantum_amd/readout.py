"""Terminal readout: seeded sampling and counts on the host.

The marginal distribution comes from the device (`probabilities`); the draws
follow the reference's host algorithm step for step, so equal probability
vectors and equal uniforms give equal samples:
  cumulative = cumsum(p) (sequential, left to right), total = its last entry,
  per shot: selected = searchsortedlast(cumulative, u * total) + 1, falling
  back to the last positive entry past the end
(reference _draw_samples!, src/execution/probabilities.jl:107-121).

`rng` is a numpy Generator (its `random()` stream) or any object with a
`rand()` method returning a float in [0, 1); one uniform is taken per shot,
in shot order, as the reference does with `rand(rng)`.
"""
from __future__ import annotations

import numpy as np

from .backend import probabilities
from .states import BatchDensityMatrix, BatchStateVector


def _uniforms(rng, count: int) -> np.ndarray:
    if rng is None:
        rng = np.random.default_rng()
    if hasattr(rng, "random"):
        return np.asarray(rng.random(count), dtype=np.float64).reshape(count)
    return np.array([float(rng.rand()) for _ in range(count)], dtype=np.float64)


def draw_samples(rng, distribution, shots: int, nbits: int) -> np.ndarray:
    """(shots, nbits) int array of readout bits, column j = bit j of the drawn
    basis index (reference _draw_samples!, probabilities.jl:107-121)."""
    p = np.asarray(distribution, dtype=np.float64)
    if not np.all(np.isfinite(p) & (p >= 0)):
        raise ValueError("invalid sampling distribution")
    cumulative = np.cumsum(p)
    total = cumulative[-1] if cumulative.size else 0.0
    if not (np.isfinite(total) and total > 0):
        raise ValueError("sampling probabilities must have positive finite sum")
    last_positive = int(np.nonzero(p > 0)[0][-1])
    u = _uniforms(rng, shots)
    # searchsortedlast(c, x) = number of entries <= x, so the 0-based basis
    # index is that count, or the last positive entry when it runs off the end
    basis = np.searchsorted(cumulative, u * total, side="right")
    basis = np.where(basis >= p.size, last_positive, basis)
    bits = (basis[:, None] >> np.arange(nbits)[None, :]) & 1
    return bits.astype(np.int64).reshape(shots, nbits)


def discrete_sample(rng, weights) -> int:
    """1-based index drawn from finite nonnegative weights with one uniform
    (reference discrete_sample, src/execution/sampling.jl:3-18)."""
    w = np.asarray(weights, dtype=np.float64)
    if w.size == 0:
        raise ValueError("cannot sample an empty distribution")
    if not np.all(np.isfinite(w) & (w >= 0)):
        raise ValueError("sampling weights must be finite and nonnegative")
    total = 0.0
    for x in w:
        total += float(x)
    if not (np.isfinite(total) and total > 0):
        raise ValueError("sampling weights must have positive finite sum")
    threshold = float(_uniforms(rng, 1)[0]) * total
    cumulative = 0.0
    last_positive = 1
    for i, x in enumerate(w, start=1):
        if x > 0:
            last_positive = i
        cumulative += float(x)
        if threshold < cumulative:
            return i
    return last_positive


def _qubit_list(state, qubits):
    qs = list(range(1, state.n + 1)) if qubits is None else [int(q) for q in qubits]
    if len(set(qs)) != len(qs) or not all(1 <= q <= state.n for q in qs):
        raise ValueError("readout qubits must be distinct and in range")
    return qs


def sample(state, shots: int, rng=None, qubits=None) -> np.ndarray:
    """Terminal readout samples without collapsing `state`: rows are shots,
    columns follow `qubits` in the requested order; batches add a third axis
    (reference sample, src/execution/probabilities.jl:130-168)."""
    if int(shots) != shots or shots < 0:
        raise ValueError("shots must be nonnegative")
    shots = int(shots)
    qs = _qubit_list(state, qubits)
    if isinstance(state, (BatchStateVector, BatchDensityMatrix)):
        dists = probabilities(state, qs)
        out = np.zeros((shots, len(qs), state.batch_size()), dtype=np.int64)
        for b in range(state.batch_size()):
            out[:, :, b] = draw_samples(rng, dists[:, b], shots, len(qs))
        return out
    return draw_samples(rng, probabilities(state, qs), shots, len(qs))


def _histogram(draws: np.ndarray) -> dict[str, int]:
    hist: dict[str, int] = {}
    for row in draws:
        key = "".join(str(int(b)) for b in row)
        hist[key] = hist.get(key, 0) + 1
    return hist


def counts(state, shots: int, rng=None, qubits=None):
    """Readout histogram; in each key the leftmost character is the first
    requested qubit; one histogram per batch element for batches (reference
    counts, src/execution/probabilities.jl:170-190)."""
    draws = sample(state, shots, rng, qubits)
    if draws.ndim == 2:
        return _histogram(draws)
    return [_histogram(draws[:, :, b]) for b in range(draws.shape[2])]


def measure_(state, position: int, auto_reset: bool = True, rng=None, b=None):
    """Mid-circuit measurement of one qubit, in place (reference
    measure!(backend, state, position; auto_reset, rng),
    src/execution/simulation.jl:42-51, _cpu_apply!(::QMeasure),
    src/execution/measurement.jl:4-17,171-201).

    The device sums the qubit's zero-outcome probability p0; the host draws
    the outcome from [p0, 1 - p0] (normalised as the reference does, one
    uniform from `rng`) and the device applies the collapse as a one-qubit
    gate: |0><o| / sqrt(p) with auto_reset (the qubit returns to 0), else
    |o><o| / sqrt(p).  A density matrix collapses as m rho m^dagger.
    Returns (outcome, probability); a batch returns (outcomes, probabilities)
    in column order, the draws taken column by column after every column's
    distribution is known (measurement.jl:69-91)."""
    from .backend import _normalize, _probability_tolerance, apply_, backend
    from .gates import QuantumGate
    from .states import DensityMatrix, StateVector
    if isinstance(state, (BatchStateVector, BatchDensityMatrix)):
        if not 1 <= int(position) <= state.n:
            raise ValueError(f"qubit position must be between 1 and {state.n}")
        from .backend import _device_p0
        p0s = [_device_p0(c, position) for c in state.columns]
        outs, probs = [], []
        for c, p0 in zip(state.columns, p0s):
            o, p = _collapse(c, position, p0, auto_reset, rng, b, apply_, QuantumGate, _normalize,
                             _probability_tolerance())
            outs.append(o)
            probs.append(p)
        return np.array(outs, dtype=np.int64), np.array(probs)
    if not isinstance(state, (StateVector, DensityMatrix)):
        raise TypeError("measure_ takes state vectors and density matrices")
    if not 1 <= int(position) <= state.n:
        raise ValueError(f"qubit position must be between 1 and {state.n}")
    b = b or backend(state)
    from .backend import _device_p0
    return _collapse(state, position, _device_p0(state, position), auto_reset, rng, b, apply_, QuantumGate,
                     _normalize, _probability_tolerance())


def _collapse(state, position, p0, auto_reset, rng, b, apply_, QuantumGate, normalize, tol):
    dist = normalize(np.array([p0, 1.0 - p0]), tol)
    i = discrete_sample(rng, dist)  # 1-based
    prob = float(dist[i - 1])
    m = np.zeros((2, 2), dtype=np.complex128)
    m[0 if auto_reset else i - 1, i - 1] = 1.0 / np.sqrt(prob)
    apply_(QuantumGate((int(position),), m), state, b)
    return i - 1, prob
