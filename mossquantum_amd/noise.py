"""Noise models expanded into explicit noisy circuits (the reference's
NoiseModel / NoiseRule / with_noise interface, src/execution/noise_models.jl).

A model is compiled once into two rule lists (maps before and after an
instruction); `with_noise` then makes one pass over the circuit's leaf
instructions, emitting for each one the maps of its `before` rules, the
instruction itself (an idle `Delay` emits nothing) and the maps of its
`after` rules, with the model's preparation maps in front and its terminal
maps behind.  Every generated map is checked to be trace preserving on its
supermatrix, the form the device applies: with vec(rho) ordering, a map is
trace preserving exactly when sum_a S[(a, a), (b, e)] = delta_be.

On the device the fused planner folds each qubit's gate and channel chain
into one superoperator per tile op, so expanded noise adds no sweeps.

Classical readout errors (the reference's ReadoutError stage) act on sampled
outcomes after the simulation and are outside this backend; a model that
carries them is refused.
"""
from __future__ import annotations

import copy
from dataclasses import dataclass, field

import numpy as np

from .channels import QuantumMap
from .circuit import Delay, QCircuit

_TP_TOL = 1e-10


def _trace_preserving(m: QuantumMap) -> bool:
    d = 1 << m.nqubits()
    S = np.asarray(m.ordered_supermat(), dtype=np.complex128).reshape((d, d, d, d), order="F")
    return bool(np.max(np.abs(np.einsum("aabe->be", S) - np.eye(d))) <= _TP_TOL)


def _as_maps(generated) -> list[QuantumMap]:
    """A factory's result as independent, trace-preserving QuantumMaps."""
    if isinstance(generated, QuantumMap):
        generated = [generated]
    try:
        maps = list(generated)
    except TypeError:
        raise ValueError("noise factories must return QuantumMap values") from None
    for m in maps:
        if not isinstance(m, QuantumMap):
            raise ValueError("noise factories must return QuantumMap values")
        if not _trace_preserving(m):
            raise ValueError("noise maps must be trace preserving")
    return [copy.deepcopy(m) for m in maps]


@dataclass
class NoiseRule:
    """Maps to insert next to matching instructions (reference NoiseRule,
    noise_models.jl:46-69): `matcher` is an instruction type or a predicate
    returning a bool, `factory(instruction)` returns one QuantumMap or several
    (or `factory` is a fixed QuantumMap), `when` is "before" or "after", and
    `qubits` restricts the rule to that exact ordered register."""

    matcher: object
    factory: object
    when: str = "after"
    qubits: tuple | None = None

    def __post_init__(self):
        if self.when not in ("before", "after"):
            raise ValueError("when must be 'before' or 'after'")
        if not (isinstance(self.matcher, type) or callable(self.matcher)):
            raise ValueError("matcher must be an instruction type or predicate")
        if self.qubits is not None:
            self.qubits = tuple(int(q) for q in self.qubits)
            if min(self.qubits, default=1) < 1 or len(set(self.qubits)) != len(self.qubits):
                raise ValueError("rule qubits must be positive and distinct")

    def maps_for(self, op) -> list[QuantumMap]:
        """The maps this rule inserts for `op` (empty when it does not match);
        predicates and factories see independent copies of the instruction."""
        if self.qubits is not None and self.qubits != tuple(op.positions()):
            return []
        if isinstance(self.matcher, type):
            hit = isinstance(op, self.matcher)
        else:
            hit = self.matcher(copy.deepcopy(op))
            if not isinstance(hit, (bool, np.bool_)):
                raise ValueError("noise predicate must return a bool")
        if not hit:
            return []
        return _as_maps(self.factory if isinstance(self.factory, QuantumMap) else self.factory(copy.deepcopy(op)))


@dataclass
class NoiseModel:
    """Instruction rules plus preparation and terminal maps (reference
    NoiseModel, noise_models.jl:79-93)."""

    rules: list = field(default_factory=list)
    readout: object = ()
    preparation: object = ()
    terminal: object = ()

    def __post_init__(self):
        if not isinstance(self.readout, (list, tuple)) or len(self.readout):
            raise ValueError("classical readout errors act on sampled outcomes and are not part of this backend")
        self.rules = [copy.deepcopy(r) for r in self.rules]
        if not all(isinstance(r, NoiseRule) for r in self.rules):
            raise ValueError("rules must be NoiseRule values")
        self.preparation = _as_maps(self.preparation)
        self.terminal = _as_maps(self.terminal)


@dataclass
class NoisyCircuit:
    """An expanded noise program (reference NoisyCircuit, noise_models.jl:108-123)."""

    circuit: QCircuit

    def bind(self, values):
        return NoisyCircuit(self.circuit.bind(values))

    def nparameters(self):
        return self.circuit.nparameters()

    def active_parameters(self):
        return self.circuit.active_parameters()


def with_noise(circuit, model: NoiseModel) -> NoisyCircuit:
    """The circuit with the model's maps inserted, the input left unchanged;
    expanding an already expanded circuit is refused (reference with_noise,
    noise_models.jl:152-169)."""
    if isinstance(circuit, NoisyCircuit):
        raise ValueError("noise is already expanded; use the original circuit")
    before = [r for r in model.rules if r.when == "before"]
    after = [r for r in model.rules if r.when == "after"]
    out = [copy.deepcopy(m) for m in model.preparation]
    for op in circuit.flat():
        out += [m for r in before for m in r.maps_for(op)]
        if not isinstance(op, Delay):
            out.append(copy.deepcopy(op))
        out += [m for r in after for m in r.maps_for(op)]
    out += [copy.deepcopy(m) for m in model.terminal]
    return NoisyCircuit(QCircuit(out))
