"""Pauli-string observables (reference src/quantum_circuits/operators/).

A `QubitsTerm` is coeff times a tensor product of 2x2 matrices on distinct
sites; a `QubitsOperator` groups terms by their (sorted) site tuple.  The group
matrix is sum coeff * kron(reverse(ops)), the first site being the least
significant local bit (reference _get_mat, src/backends/cpu/operators/utils.jl:3-20).
"""
from __future__ import annotations

from collections import OrderedDict

import numpy as np

from .gates import I2, X, Y, Z

OP_MAPPING = {
    "X": X, "Y": Y, "Z": Z, "I": I2,
    "+": np.array([[0, 1], [0, 0]], dtype=np.complex128),
    "-": np.array([[0, 0], [1, 0]], dtype=np.complex128),
    "u": np.array([[1, 0], [0, 0]], dtype=np.complex128),
    "d": np.array([[0, 0], [0, 1]], dtype=np.complex128),
    "0": np.array([[1, 0], [0, 0]], dtype=np.complex128),
    "1": np.array([[0, 0], [0, 1]], dtype=np.complex128),
}


def _op(x):
    if isinstance(x, str):
        return OP_MAPPING[x]
    m = np.array(x, dtype=np.complex128)
    if m.shape != (2, 2):
        raise ValueError("local operators must be 2x2 matrices")
    return m


class QubitsTerm:
    """QubitsTerm({pos: op, ...}, coeff) or QubitsTerm(pos=>op pairs as (pos, op) tuples)."""

    def __init__(self, ops=None, coeff=1.0):
        if ops is None:
            ops = {}
        items = list(ops.items()) if isinstance(ops, dict) else list(ops)
        pos = [int(p) for p, _ in items]
        if any(p <= 0 for p in pos) or len(set(pos)) != len(pos):
            raise ValueError("operator positions must be positive and distinct")
        order = sorted(range(len(pos)), key=lambda i: pos[i])
        self.positions = tuple(pos[i] for i in order)
        self.ops = [_op(items[i][1]) for i in order]
        self.coeff = complex(coeff)

    def __mul__(self, c):
        t = QubitsTerm()
        t.positions, t.ops, t.coeff = self.positions, list(self.ops), self.coeff * c
        return t

    __rmul__ = __mul__

    def adjoint(self):
        t = QubitsTerm()
        t.positions, t.ops, t.coeff = self.positions, [o.conj().T for o in self.ops], self.coeff.conjugate()
        return t

    def group_matrix(self):
        m = np.ones((1, 1), dtype=np.complex128)
        for o in self.ops:  # kron(reverse(ops)): first site least significant
            m = np.kron(o, m)
        return self.coeff * m


class QubitsOperator:
    """Sum of QubitsTerms grouped by site tuple (reference qubits_operator.jl:1-50)."""

    def __init__(self, terms=()):
        self.data: OrderedDict[tuple, list] = OrderedDict()
        for t in terms:
            self.add(t)

    def add(self, term: QubitsTerm):
        self.data.setdefault(term.positions, []).append((list(term.ops), term.coeff))
        return self

    def __iadd__(self, term):
        return self.add(term)

    def terms(self):
        for k, v in self.data.items():
            for ops, c in v:
                t = QubitsTerm()
                t.positions, t.ops, t.coeff = k, ops, c
                yield t

    def adjoint(self):
        return QubitsOperator(t.adjoint() for t in self.terms())

    def nqubits(self):
        return max((max(k) for k in self.data if k), default=0)

    def largest_nterm(self):
        return max((len(k) for k in self.data), default=0)

    def groups(self):
        """[(sites (1-based, sorted), 2^k x 2^k group matrix)] in insertion order."""
        out = []
        for k, items in self.data.items():
            d = 1 << len(k)
            m = np.zeros((d, d), dtype=np.complex128)
            for ops, c in items:
                t = np.ones((1, 1), dtype=np.complex128)
                for o in ops:
                    t = np.kron(o, t)
                m += c * t
            out.append((k, m))
        return out

    def matrix(self, n):
        """Dense 2^n x 2^n matrix (reference matrix(n, x), qubits_operator.jl:229-241)."""
        dim = 1 << n
        M = np.zeros((dim, dim), dtype=np.complex128)
        for t in self.terms():
            full = np.ones((1, 1), dtype=np.complex128)
            ops = dict(zip(t.positions, t.ops))
            for q in range(1, n + 1):
                full = np.kron(ops.get(q, I2), full)
            M += t.coeff * full
        return M


def tfim_chain(n, h=1.0):
    """Open-chain transverse-field Ising: sum Z_i Z_{i+1} + h sum X_i."""
    H = QubitsOperator()
    for i in range(1, n):
        H.add(QubitsTerm({i: "Z", i + 1: "Z"}, 1.0))
    for i in range(1, n + 1):
        H.add(QubitsTerm({i: "X"}, h))
    return H


def xxz_chain(n, jzz=0.5):
    """Open-chain XXZ: sum (X_i X_{i+1} + Y_i Y_{i+1} + jzz Z_i Z_{i+1}); equals the
    reference heisenberg_chain(n; J=1, Jzz=jzz, hz=0) since 2(s+s- + s-s+) = XX + YY
    (reference src/utilities/models/spin_hamiltonians.jl:7-22)."""
    H = QubitsOperator()
    for i in range(1, n):
        H.add(QubitsTerm({i: "X", i + 1: "X"}, 1.0))
        H.add(QubitsTerm({i: "Y", i + 1: "Y"}, 1.0))
        H.add(QubitsTerm({i: "Z", i + 1: "Z"}, jzz))
    return H
