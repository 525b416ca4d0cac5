"""Kraus channels (reference src/quantum_circuits/channels/).

`QuantumMap` snapshots a Kraus list and its ordered supermatrix
S[(a + d c), (b + d e)] = sum_K K[a,b] conj(K[c,e]) (reference
quantum_map.jl:59-67), applied on bits (pos..., pos+n...) of vec(rho).
"""
from __future__ import annotations

import itertools

import numpy as np

from .gates import I2, X, Y, Z, reorder_legs


def compute_supermat(kraus):
    d = kraus[0].shape[0]
    S = np.zeros((d * d, d * d), dtype=np.complex128)
    for K in kraus:
        # S[a + d c, b + d e] = K[a,b] conj(K[c,e]) -> kron(conj(K), K) with K fastest
        S += np.kron(K.conj(), K)
    return S


class QuantumMap:
    def __init__(self, positions, kraus, check_tp=False, atol=None):
        if isinstance(positions, int):
            positions = (positions,)
        positions = tuple(int(p) for p in positions)
        n = len(positions)
        if n == 0:
            raise ValueError("a channel needs at least one position")
        if any(p <= 0 for p in positions) or len(set(positions)) != n:
            raise ValueError("channel positions must be positive and distinct")
        if not kraus:
            raise ValueError("a channel needs at least one Kraus matrix")
        ks = [np.array(k, dtype=np.complex128) for k in kraus]
        for k in ks:
            if k.shape != (1 << n, 1 << n):
                raise ValueError("Kraus dimensions must match the register")
            if not np.all(np.isfinite(k)):
                raise ValueError("Kraus matrices must be finite")
        tol = 1e-12 if atol is None else atol
        if check_tp:
            acc = sum(k.conj().T @ k for k in ks)
            if not np.allclose(acc, np.eye(1 << n), atol=max(tol, 1e-10), rtol=0):
                raise ValueError("noise must be trace preserving")
        self._positions = positions
        self._kraus = ks
        ordered = [reorder_legs(positions, k) for k in ks]
        self._ordered_positions = ordered[0][0]
        self._ordered_kraus = [o[1] for o in ordered]
        self._supermat = compute_supermat(self._ordered_kraus)

    def positions(self):
        return self._positions

    def nqubits(self):
        return len(self._positions)

    def kraus_matrices(self):
        return [k.copy() for k in self._kraus]

    def ordered_positions(self):
        return self._ordered_positions

    def ordered_supermat(self):
        return self._supermat

    def nparameters(self):
        return 0

    def parameters(self):
        return None

    def __repr__(self):
        return f"QuantumMap{self._positions}"


def _prob(v, name, upper=1.0):
    v = float(v)
    if not (np.isfinite(v) and 0 <= v <= upper):
        raise ValueError(f"{name} must be finite and between 0 and {upper}")
    return v


def AmplitudeDamping(pos, gamma):
    g = _prob(gamma, "gamma")
    return QuantumMap(pos, [np.array([[1, 0], [0, np.sqrt(1 - g)]]), np.array([[0, np.sqrt(g)], [0, 0]])],
                      check_tp=True)


def PhaseDamping(pos, gamma):
    g = _prob(gamma, "gamma")
    return QuantumMap(pos, [np.array([[1, 0], [0, np.sqrt(1 - g)]]), np.array([[0, 0], [0, np.sqrt(g)]])],
                      check_tp=True)


def MixedUnitary(pos, unitaries, probabilities):
    if len(unitaries) != len(probabilities) or not unitaries:
        raise ValueError("unitaries and probabilities must match")
    w = np.array([_prob(p, "probability") for p in probabilities])
    if abs(w.sum() - 1) > 1e-12:
        raise ValueError("probabilities must sum to one")
    w = w / w.sum()
    ks = [np.sqrt(p) * np.asarray(u, dtype=np.complex128) for p, u in zip(w, unitaries)]
    return QuantumMap(pos, ks, check_tp=True)


_PAULI = {"I": I2, "X": X, "Y": Y, "Z": Z}


def PauliChannel(pos, words, probabilities):
    """Mixture of Pauli words; the first character acts on the first position."""
    reg = (pos,) if isinstance(pos, int) else tuple(pos)
    mats = []
    for w in words:
        if len(w) != len(reg):
            raise ValueError("Pauli words must match the register")
        m = np.ones((1, 1), dtype=np.complex128)
        for ch in w:
            m = np.kron(_PAULI[ch], m)
        mats.append(m)
    return MixedUnitary(reg, mats, probabilities)


def Depolarizing(pos, p):
    """(1-p) rho + p tr(rho) I/d for one or two qubits (reference channels/one_qubit.jl:49-59)."""
    reg = (pos,) if isinstance(pos, int) else tuple(pos)
    n = len(reg)
    if not 1 <= n <= 2:
        raise ValueError("Depolarizing supports one or two qubits")
    count = 4 ** n
    p = _prob(p, "p", upper=count / (count - 1))
    # Iterators.product order: the first character varies fastest
    words = ["".join(reversed(t)) for t in itertools.product("IXYZ", repeat=n)]
    probs = [p / count] * count
    probs[0] = max(0.0, 1 - (count - 1) * p / count)
    return PauliChannel(reg, words, probs)


def BitFlip(pos, p):
    return PauliChannel(pos, ["I", "X"], [1 - _prob(p, "p"), p])


def PhaseFlip(pos, p):
    return PauliChannel(pos, ["I", "Z"], [1 - _prob(p, "p"), p])


def BitPhaseFlip(pos, p):
    return PauliChannel(pos, ["I", "Y"], [1 - _prob(p, "p"), p])


def GeneralizedAmplitudeDamping(pos, gamma, excited_population=0.0):
    """Damping toward the equilibrium |1> population `excited_population`
    (reference channels/one_qubit.jl:78-86)."""
    g = _prob(gamma, "gamma")
    e = _prob(excited_population, "excited_population")
    a, b = np.sqrt(1 - e), np.sqrt(e)
    return QuantumMap(pos, [a * np.array([[1, 0], [0, np.sqrt(1 - g)]]), a * np.array([[0, np.sqrt(g)], [0, 0]]),
                            b * np.array([[np.sqrt(1 - g), 0], [0, 1]]), b * np.array([[0, 0], [np.sqrt(g), 0]])],
                      check_tp=True)


def ResetChannel(pos, state=0, p=1.0):
    """With probability p replace the qubit by |state> (reference
    channels/one_qubit.jl:88-96)."""
    if state not in (0, 1):
        raise ValueError("reset state must be 0 or 1")
    p = _prob(p, "p")
    a = np.zeros((2, 2))
    b = np.zeros((2, 2))
    a[state, 0] = np.sqrt(p)
    b[state, 1] = np.sqrt(p)
    return QuantumMap(pos, [np.sqrt(1 - p) * np.eye(2), a, b], check_tp=True)


def ThermalRelaxation(pos, T1, T2, duration, excited_population=0.0):
    """Population relaxation with T1 and dephasing with T2 over `duration`
    (common time unit; T2 <= 2 T1; infinite lifetimes allowed): phase damping
    after generalized amplitude damping, Kraus products b*a (reference
    channels/one_qubit.jl:98-115)."""
    T1, T2, duration = float(T1), float(T2), float(duration)
    if not (T1 > 0 and T2 > 0 and T2 / 2 <= T1):
        raise ValueError("require positive T1/T2 and T2 <= 2T1")
    if not (np.isfinite(duration) and duration >= 0):
        raise ValueError("duration must be finite and nonnegative")
    e = _prob(excited_population, "excited_population")
    gamma = -np.expm1(-duration / T1)
    inverse_phi = max(0.0, 1 / T2 - 1 / T1 / 2)
    phase = PhaseDamping(pos, -np.expm1(-2 * duration * inverse_phi))
    amplitude = GeneralizedAmplitudeDamping(pos, gamma, e)
    ks = [b @ a for a in amplitude.kraus_matrices() for b in phase.kraus_matrices()]
    return QuantumMap(pos, ks, check_tp=True)
