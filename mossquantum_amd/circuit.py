"""Circuits and their lowering to device programs.

`QCircuit` mirrors the reference container (src/quantum_circuits/circuits/
circuit.jl:98-294): an ordered list of gates, channels and nested circuits;
active parameters are enumerated in circuit order (circuit.jl:214-239).

`lower()` turns a circuit into the flat instruction records of the C ABI
(`mossq_op`) plus a complex matrix pool: ordered positions (0-based bits),
`ordered_mat` for gates, `ordered_supermat` for channels and one derivative
gate (dU/dtheta)^dagger U per active parameter.
"""
from __future__ import annotations

import numpy as np

from ._lib import OP_CHANNEL, OP_GATE, OpRecord
from .channels import QuantumMap
from .gates import Gate, QuantumGate, reorder_legs


class Delay:
    """An explicit idle interval on one qubit, in the time unit of the noise rule
    that matches it (reference Delay, src/execution/noise_models.jl:31-44).  It
    leaves the state unchanged; `with_noise` replaces it by what its rules
    generate."""

    def __init__(self, position: int, duration: float):
        if int(position) != position or position <= 0:
            raise ValueError("position must be positive")
        duration = float(duration)
        if not (np.isfinite(duration) and duration >= 0):
            raise ValueError("duration must be finite and nonnegative")
        self.position = int(position)
        self.duration = duration

    def positions(self):
        return (self.position,)

    def ordered_positions(self):
        return (self.position,)

    def nqubits(self):
        return 1

    def nparameters(self):
        return 0

    def adjoint(self):
        return self

    def __repr__(self):
        return f"Delay({self.position}; duration={self.duration})"


class QCircuit:
    def __init__(self, ops=()):
        self.ops = []
        for o in ops:
            self.append(o)

    def append(self, op):
        if not isinstance(op, (Gate, QuantumMap, QCircuit, Delay)):
            raise TypeError(f"unsupported instruction {type(op).__name__}")
        self.ops.append(op)
        return self

    push = append

    def __iter__(self):
        return iter(self.ops)

    def __len__(self):
        return len(self.ops)

    def flat(self):
        for o in self.ops:
            if isinstance(o, QCircuit):
                yield from o.flat()
            else:
                yield o

    def has_channel(self):
        return any(isinstance(o, QuantumMap) for o in self.flat())

    def nparameters(self):
        return sum(o.nparameters() for o in self.flat())

    def active_parameters(self):
        out = []
        for o in self.flat():
            out.extend(o.active_parameters() if isinstance(o, Gate) else [])
        return out

    def bind(self, values):
        """Copy with the active parameters replaced in circuit order
        (reference bind_parameters, src/api/parameter_binding.jl:19-63)."""
        values = list(np.asarray(values, dtype=np.float64))
        if len(values) != self.nparameters():
            raise ValueError(f"expected {self.nparameters()} parameters, received {len(values)}")
        if not all(np.isfinite(values)):
            raise ValueError("parameter values must be finite")
        it = iter(values)

        def rebind(c):
            out = QCircuit()
            for o in c.ops:
                if isinstance(o, QCircuit):
                    out.append(rebind(o))
                elif isinstance(o, Gate) and o.nparameters():
                    out.append(o.with_parameters([next(it) for _ in range(o.nparameters())]))
                else:
                    out.append(o)
            return out

        return rebind(self)

    def adjoint(self):
        return QCircuit([o.adjoint() for o in reversed(list(self.flat()))])

    def max_position(self):
        return max((max(o.positions()) for o in self.flat()), default=0)


def fuse_gates(circuit: QCircuit, max_qubits: int = 5) -> QCircuit:
    """Greedy merge of adjacent non-parameterized gates while the combined
    support stays within max_qubits (reference gate_fusion.jl:10-62)."""
    if not 1 <= max_qubits <= 5:
        raise ValueError("max_qubits must be between 1 and 5")
    result = QCircuit()
    pending = None
    for op in circuit.ops:
        eligible = isinstance(op, Gate) and op.parameters() is None and op.nqubits() <= max_qubits
        if not eligible:
            if pending is not None:
                result.append(pending)
                pending = None
            result.append(fuse_gates(op, max_qubits) if isinstance(op, QCircuit) else op)
        elif pending is None:
            pending = op
        else:
            support = sorted(set(pending.positions()) | set(op.positions()))
            if len(support) <= max_qubits:
                combined = _embed(op, support) @ _embed(pending, support)
                pending = QuantumGate(tuple(support), combined)
            else:
                result.append(pending)
                pending = op
    if pending is not None:
        result.append(pending)
    return result


def _embed(g: Gate, support):
    """g's ordered matrix acting on the sorted qubit list `support` (which
    holds g's positions), identity on the rest, in the ordered basis of
    `support` (its first qubit is the least significant local bit).  Built as
    kron(I, U) over (g's qubits, then the others) and a permutation of the
    tensor legs into support order."""
    pos = list(g.ordered_positions())
    data = np.asarray(g.ordered_mat(), dtype=np.complex128)
    if pos == list(support):
        return data
    m = len(support)
    order = pos + [q for q in support if q not in pos]  # qubit at each bit of kron(I, U), low to high
    T = np.kron(np.eye(1 << (m - len(pos)), dtype=np.complex128), data).reshape([2] * (2 * m))
    # axis a of the row (and column) legs is bit m-1-a
    axes = [m - 1 - order.index(support[m - 1 - a]) for a in range(m)]
    return T.transpose(axes + [m + x for x in axes]).reshape(1 << m, 1 << m)


class Lowered:
    """Instruction records + matrix pool of a circuit, ready for mossq_program_create."""

    def __init__(self, n, is_density):
        self.n = n
        self.is_density = is_density
        self.records: list[OpRecord] = []
        self.pool: list[np.ndarray] = []
        self.pool_len = 0
        self.nslots = 0
        self.grad_slot: list[int] = []
        # (pool entry of the gate matrix, gate, first slot, pool offset) per parameterized gate
        self.param_gates: list[tuple[int, Gate, int, int]] = []
        self._rebind = None

    def _push(self, m):
        m = np.asarray(m, dtype=np.complex128)
        off = self.pool_len
        flat = np.asfortranarray(m).reshape(-1, order="F")
        self.pool.append(flat)
        self.pool_len += flat.size
        return off

    def add(self, op):
        r = OpRecord()
        pos = op.ordered_positions()
        if max(pos) > self.n:
            raise ValueError(f"position {max(pos)} outside a {self.n}-qubit register")
        r.k = len(pos)
        for j, p in enumerate(pos):
            r.pos[j] = p - 1
        r.dmat = -1
        if isinstance(op, QuantumMap):
            r.kind = OP_CHANNEL
            r.mat = self._push(op.ordered_supermat())
            self.records.append(r)
            return
        r.kind = OP_GATE
        r.mat = self._push(op.ordered_mat())
        ders = op.differentiate() or []
        first_slot = self.nslots
        if ders:
            if len(ders) != op.nparameters():
                raise ValueError("derivative gates must match the gate's active parameters")
            self.param_gates.append((len(self.pool) - 1, op, first_slot, r.mat))
        self.nslots += len(ders)
        if len(ders) > 1:
            # extra derivative gates ride on identity records placed before the
            # gate: in reverse order they see the same (U^dag y, U^dag lambda)
            ident = np.eye(1 << r.k, dtype=np.complex128)
            for j, dg in enumerate(ders[1:], start=1):
                e = OpRecord()
                e.kind, e.k = OP_GATE, r.k
                for q in range(r.k):
                    e.pos[q] = r.pos[q]
                e.mat = self._push(ident)
                e.dmat = self._push(reorder_legs(dg.positions(), dg.mat())[1])
                self.records.append(e)
                self.grad_slot.append(first_slot + j)
        if ders:
            r.dmat = self._push(reorder_legs(ders[0].positions(), ders[0].mat())[1])
            self.grad_slot.append(first_slot)
        self.records.append(r)

    @property
    def nderiv(self):
        return len(self.grad_slot)

    def slot_gradients(self, device_grads):
        """Device gradient order (records with a derivative, forward order) ->
        active-slot order of the circuit."""
        out = np.zeros(self.nslots)
        out[np.asarray(self.grad_slot, dtype=np.int64)] = device_grads
        return out

    def _rebind_tables(self):
        """One-qubit rotations and phases grouped by kind (slot, pool offset),
        rebuilt vectorised; every other parameterized gate through its object.
        Derivative matrices (dU)^dagger U of these families do not depend on
        the angle, so only U is rewritten."""
        from .gates import PHASEGate, RxGate, RyGate, RzGate
        kinds = {RxGate: "rx", RyGate: "ry", RzGate: "rz", PHASEGate: "phase"}
        fast = {k: ([], []) for k in kinds.values()}
        slow = []
        for li, op, s0, off in self.param_gates:
            kind = kinds.get(type(op))
            if kind is not None and op.nqubits() == 1 and op.nparameters() == 1:
                fast[kind][0].append(s0)
                fast[kind][1].append(off)
            else:
                slow.append((op, s0, off))
        base = np.ascontiguousarray(np.concatenate(self.pool)) if self.pool else np.zeros(1, dtype=np.complex128)
        fast = {k: (np.array(a, dtype=np.int64), np.array(b, dtype=np.int64)) for k, (a, b) in fast.items() if a}
        self._rebind = (base, fast, slow)
        return self._rebind

    def pool_with_slots(self, slot_values, slots=None):
        """The pool with the parameterized gates' matrices rebuilt at
        `slot_values` (active slots in circuit order); `slots` limits the
        rebuild to the gates holding those slots."""
        base, fast, slow = self._rebind or self._rebind_tables()
        v = np.asarray(slot_values, dtype=np.float64)
        pool = base.copy()
        want = None if slots is None else np.asarray(sorted(set(int(x) for x in slots)), dtype=np.int64)
        for kind, (sl, off) in fast.items():
            if want is not None:
                keep = np.isin(sl, want)
                sl, off = sl[keep], off[keep]
                if not sl.size:
                    continue
            th = v[sl]
            c, s = np.cos(th / 2), np.sin(th / 2)
            if kind == "rx":  # [[c, -is], [-is, c]], column-major
                vals = (c, -1j * s, -1j * s, c)
            elif kind == "ry":  # [[c, -s], [s, c]]
                vals = (c, s, -s, c)
            elif kind == "rz":  # diag(e^{-i th/2}, e^{i th/2})
                vals = (np.exp(-0.5j * th), 0.0, 0.0, np.exp(0.5j * th))
            else:  # diag(1, e^{i th})
                vals = (1.0, 0.0, 0.0, np.exp(1j * th))
            for q, x in enumerate(vals):
                pool[off + q] = x
        for op, s0, off in slow:
            k = op.nparameters()
            if want is not None and not np.any((want >= s0) & (want < s0 + k)):
                continue
            g = op.with_parameters([float(x) for x in v[s0:s0 + k]])
            m = np.asfortranarray(g.ordered_mat()).reshape(-1, order="F")
            pool[off:off + m.size] = m
        return pool

    def arrays(self):
        recs = (OpRecord * max(len(self.records), 1))(*self.records)
        pool = np.concatenate(self.pool) if self.pool else np.zeros(1, dtype=np.complex128)
        return recs, np.ascontiguousarray(pool)


def lower(circuit: QCircuit, n: int, is_density: bool) -> Lowered:
    lw = Lowered(n, is_density)
    for op in circuit.flat():
        if isinstance(op, Delay):
            if op.position > n:
                raise ValueError(f"position {op.position} outside a {n}-qubit register")
            continue  # idle: no instruction
        lw.add(op)
    return lw
