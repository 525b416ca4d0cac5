"""Noise models: rules that expand a circuit into an explicit noisy circuit,
and classical readout errors (reference src/execution/noise_models.jl).

`with_noise` expands once on the host into a plain `QCircuit` of gates and
`QuantumMap`s; each map carries its supermatrix from construction, so running
the expanded circuit never rebuilds one.  On the device the planner merges
each qubit's gate and channel chain into one 4x4 superoperator per LDS-tile
op, so the noise adds no sweeps of its own.
"""
from __future__ import annotations

import copy

import numpy as np

from .channels import QuantumMap
from .circuit import Delay, QCircuit

_TOL = 100 * np.finfo(np.float64).eps  # reference _channel_tolerance(Float64, nothing)


class ReadoutError:
    """Classical confusion matrix: row true+1, column reported+1; the first
    qubit is the least significant local bit; rows sum to one.  scope is
    'terminal', 'measurement' or 'all' (the latter two need one qubit)
    (reference ReadoutError, noise_models.jl:1-29).  A one-qubit error may be
    given as p0_to_1 / p1_to_0 instead of a matrix."""

    def __init__(self, qubits, probabilities=None, scope: str = "terminal", p0_to_1=None, p1_to_0=None):
        qs = [int(qubits)] if np.isscalar(qubits) else [int(q) for q in qubits]
        if not (1 <= len(qs) <= 2 and all(q > 0 for q in qs) and len(set(qs)) == len(qs)):
            raise ValueError("readout error requires one or two distinct positive qubits")
        if scope not in ("terminal", "measurement", "all"):
            raise ValueError("invalid readout scope")
        if scope != "terminal" and len(qs) != 1:
            raise ValueError("intermediate readout errors require one qubit")
        if probabilities is None:
            if p0_to_1 is None or p1_to_0 is None or len(qs) != 1:
                raise ValueError("give a confusion matrix or p0_to_1 and p1_to_0 for one qubit")
            probabilities = [[1 - p0_to_1, p0_to_1], [p1_to_0, 1 - p1_to_0]]
        d = 1 << len(qs)
        p = np.array(probabilities, dtype=np.float64)
        if p.shape != (d, d):
            raise ValueError("readout matrix dimensions must match qubits")
        if not np.all(np.isfinite(p) & (p >= 0)):
            raise ValueError("readout probabilities must be finite and nonnegative")
        if not np.all(np.abs(p.sum(axis=1) - 1) <= _TOL):
            raise ValueError("readout probability rows must sum to one")
        self.qubits = qs
        self.probabilities = p / p.sum(axis=1, keepdims=True)
        self.scope = scope

    def __repr__(self):
        return f"ReadoutError({self.qubits}, scope={self.scope})"


def _check_readout_errors(errors):
    for scope in ("terminal", "measurement"):
        seen: set[int] = set()
        for e in errors:
            ReadoutError(e.qubits, e.probabilities, scope=e.scope)  # revalidate mutable configuration
            if e.scope not in (scope, "all"):
                continue
            if seen & set(e.qubits):
                raise ValueError(f"overlapping readout errors in {scope} scope")
            seen |= set(e.qubits)


def _physical_maps(maps) -> list[QuantumMap]:
    if isinstance(maps, QuantumMap):
        values = [maps]
    else:
        try:
            values = list(maps)
        except TypeError:
            raise ValueError("noise factories must return QuantumMap values") from None
    if not all(isinstance(x, QuantumMap) for x in values):
        raise ValueError("noise factories must return QuantumMap values")
    out = []
    for x in values:
        ks = x.kraus_matrices()
        acc = sum(k.conj().T @ k for k in ks)
        if not np.allclose(acc, np.eye(acc.shape[0]), atol=max(_TOL, 1e-10), rtol=0):
            raise ValueError("noise maps must be trace preserving")
        out.append(copy.deepcopy(x))
    return out


class NoiseRule:
    """Match an instruction type or a predicate; factory(instruction) returns
    one QuantumMap or a list of them (or factory is a fixed QuantumMap).
    Rules run in registration order on leaf instructions, before or after
    them; `qubits` restricts a rule to that exact ordered register (reference
    NoiseRule, noise_models.jl:46-69)."""

    def __init__(self, matcher, factory, when: str = "after", qubits=None):
        if when not in ("before", "after"):
            raise ValueError("when must be 'before' or 'after'")
        if not (isinstance(matcher, type) or callable(matcher)):
            raise ValueError("matcher must be an instruction type or predicate")
        qs = None if qubits is None else tuple(int(q) for q in qubits)
        if qs is not None and (not all(q > 0 for q in qs) or len(set(qs)) != len(qs)):
            raise ValueError("rule qubits must be positive and distinct")
        self.matcher = matcher
        self.factory = factory
        self.when = when
        self.qubits = qs


class NoiseModel:
    """Preparation maps, instruction rules, terminal maps and readout errors
    (reference NoiseModel, noise_models.jl:79-93)."""

    def __init__(self, rules=(), readout=(), preparation=(), terminal=()):
        self.rules = [copy.deepcopy(r) for r in rules]
        errors = [readout] if isinstance(readout, ReadoutError) else list(readout)
        errors = [copy.deepcopy(e) for e in errors]
        _check_readout_errors(errors)
        self.readout = errors
        self.preparation = _physical_maps(preparation)
        self.terminal = _physical_maps(terminal)


class NoisyCircuit:
    """An expanded noise program: `.circuit` and `.readout` (reference
    NoisyCircuit, noise_models.jl:108-123)."""

    def __init__(self, circuit: QCircuit, readout):
        self.circuit = circuit
        self.readout = list(readout)

    def bind(self, values):
        return NoisyCircuit(self.circuit.bind(values), copy.deepcopy(self.readout))

    def nparameters(self):
        return self.circuit.nparameters()

    def active_parameters(self):
        return self.circuit.active_parameters()


def _append_noise(out, instruction, rules, timing):
    for rule in rules:
        if rule.when != timing:
            continue
        if rule.qubits is not None and rule.qubits != tuple(instruction.positions()):
            continue
        if isinstance(rule.matcher, type):
            matched = isinstance(instruction, rule.matcher)
        else:
            matched = rule.matcher(copy.deepcopy(instruction))
            if not isinstance(matched, (bool, np.bool_)):
                raise ValueError("noise predicate must return a bool")
        if not matched:
            continue
        generated = rule.factory if isinstance(rule.factory, QuantumMap) else rule.factory(copy.deepcopy(instruction))
        out.extend(_physical_maps(generated))


def _expand(out, circuit, rules):
    for instruction in circuit:
        if isinstance(instruction, QCircuit):
            _expand(out, instruction, rules)
            continue
        _append_noise(out, instruction, rules, "before")
        if not isinstance(instruction, Delay):
            out.append(copy.deepcopy(instruction))
        _append_noise(out, instruction, rules, "after")


def with_noise(circuit, model: NoiseModel) -> NoisyCircuit:
    """Expand noise once without changing `circuit`; expanding a NoisyCircuit
    again is rejected (reference with_noise, noise_models.jl:152-164)."""
    if isinstance(circuit, NoisyCircuit):
        raise ValueError("noise is already expanded; use the original circuit")
    _check_readout_errors(model.readout)
    out: list = list(_physical_maps(model.preparation))
    _expand(out, circuit, model.rules)
    out.extend(_physical_maps(model.terminal))
    return NoisyCircuit(QCircuit(out), copy.deepcopy(model.readout))


# ------------------------------------------------------------- readout -----
def readout_register(errors, indices, n):
    """Readout qubits widened by every terminal error block they touch, so a
    partial register still marginalizes the full block (reference
    _readout_register, noise_models.jl:166-180)."""
    _check_readout_errors(errors)
    expanded = list(indices)
    for e in errors:
        if not all(q <= n for q in e.qubits):
            raise ValueError("readout qubit outside register")
        if e.scope not in ("terminal", "all") or not set(e.qubits) & set(indices):
            continue
        for q in e.qubits:
            if q not in expanded:
                expanded.append(q)
    return expanded


def apply_readout_errors(p, expanded, errors, nout):
    """Push an ideal distribution over the `expanded` register through the
    terminal errors, then marginalize onto its first `nout` qubits (reference
    readout_probabilities, noise_models.jl:206-245)."""
    p = np.array(p, dtype=np.float64)
    for e in errors:
        if e.scope not in ("terminal", "all"):
            continue
        if not all(q in expanded for q in e.qubits):
            continue
        cols = [expanded.index(q) for q in e.qubits]
        out = np.zeros(p.size)
        for basis in range(p.size):
            actual = sum(((basis >> c) & 1) << j for j, c in enumerate(cols))
            for reported in range(1 << len(cols)):
                target = basis
                for j, c in enumerate(cols):
                    target = (target & ~(1 << c)) | (((reported >> j) & 1) << c)
                out[target] += p[basis] * e.probabilities[actual, reported]
        p = out
    res = np.zeros(1 << nout)
    mask = res.size - 1
    for basis in range(p.size):
        res[basis & mask] += p[basis]
    return res


def readout_probabilities(state, errors, qubits=None):
    """Classical readout distribution of `state` under readout errors; the
    ideal marginal comes from the device (reference readout_probabilities,
    noise_models.jl:206-260).  Batches give one column per member."""
    from .backend import _normalize, _probability_tolerance, probabilities
    from .states import BatchDensityMatrix, BatchStateVector
    errors = [errors] if isinstance(errors, ReadoutError) else list(errors)
    indices = list(range(1, state.n + 1)) if qubits is None else [int(q) for q in qubits]
    if len(set(indices)) != len(indices) or not all(1 <= q <= state.n for q in indices):
        raise ValueError("readout qubits must be distinct and in range")
    expanded = readout_register(errors, indices, state.n)
    if isinstance(state, (BatchStateVector, BatchDensityMatrix)):
        return np.stack([readout_probabilities(c, errors, indices) for c in state.columns], axis=1)
    p = _normalize(probabilities(state, expanded), _probability_tolerance())
    return apply_readout_errors(p, expanded, errors, len(indices))
