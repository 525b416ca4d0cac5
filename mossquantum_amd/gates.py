"""Gate descriptions (host side, storage-agnostic).

Mirrors MOSSQuantum's gate layer (reference src/quantum_circuits/gates/):
positions are 1-based qubit labels, qubit 1 is the least significant bit, the
first listed position is the least significant bit of the gate's local index,
and matrices are (row, column) = (output, input).  `ordered_mat` permutes the
tensor legs into sorted-position order exactly like `_get_norm_order`
(reference abstract_gates.jl:88-121,162-166); the device kernels consume
`ordered_positions` and `ordered_mat`.
"""
from __future__ import annotations

import numpy as np

# ---- standard matrices (reference src/quantum_circuits/gates/Gates.jl:7-93) ----
I2 = np.eye(2, dtype=np.complex128)
X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)
S = np.array([[1, 0], [0, 1j]], dtype=np.complex128)
H = (X + Z) / np.sqrt(2)
T = np.array([[1, 0], [0, np.exp(1j * np.pi / 4)]], dtype=np.complex128)
UP = np.array([[1, 0], [0, 0]], dtype=np.complex128)
DOWN = np.array([[0, 0], [0, 1]], dtype=np.complex128)
SWAP = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=np.complex128)
ISWAP = np.array([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]], dtype=np.complex128)


def rkron(a, b):
    """Reverse Kronecker product: the first factor is the least significant."""
    return np.kron(b, a)


def Rx(theta):
    c, s = np.cos(theta / 2), np.sin(theta / 2)
    return np.array([[c, -1j * s], [-1j * s, c]], dtype=np.complex128)


def Ry(theta):
    c, s = np.cos(theta / 2), np.sin(theta / 2)
    return np.array([[c, -s], [s, c]], dtype=np.complex128)


def Rz(theta):
    return np.array([[np.exp(-1j * theta / 2), 0], [0, np.exp(1j * theta / 2)]], dtype=np.complex128)


def PHASE(phi):
    return np.array([[1, 0], [0, np.exp(1j * phi)]], dtype=np.complex128)


def CONTROL(u):
    """kron(I, |0><0|) + kron(u, |1><1|): the first position is the control (Gates.jl:43-47)."""
    return rkron(UP, I2) + rkron(DOWN, u)


def CONTROLCONTROL(u):
    return (rkron(rkron(UP, UP), I2) + rkron(rkron(UP, DOWN), I2) + rkron(rkron(DOWN, UP), I2)
            + rkron(rkron(DOWN, DOWN), u))


CZ = CONTROL(Z)
CNOT = CONTROL(X)
TOFFOLI = CONTROLCONTROL(X)
FREDKIN = rkron(UP, np.eye(4)) + rkron(DOWN, SWAP)


def _phase_mat_diff(phi):
    return np.array([[0, 0], [0, 1j * np.exp(1j * phi)]], dtype=np.complex128)


def reorder_legs(positions, m):
    """Permute a k-qubit operator to sorted positions (reference _get_norm_order)."""
    pos = list(positions)
    k = len(pos)
    seq = sorted(range(k), key=lambda i: pos[i])
    if seq == list(range(k)):
        return tuple(pos), np.array(m, dtype=np.complex128)
    dim = 1 << k
    idx = np.arange(dim)
    old = np.zeros(dim, dtype=np.int64)
    for i, s in enumerate(seq):  # new local bit i is old local bit seq[i]
        old |= ((idx >> i) & 1) << s
    m = np.asarray(m, dtype=np.complex128)
    return tuple(pos[s] for s in seq), m[np.ix_(old, old)]


class Gate:
    """A unitary on `positions` (1-based) with matrix `mat()` in positions order."""

    def __init__(self, positions, matrix=None):
        if isinstance(positions, int):
            positions = (positions,)
        positions = tuple(int(p) for p in positions)
        if any(p <= 0 for p in positions) or len(set(positions)) != len(positions):
            raise ValueError("gate positions must be positive and distinct")
        self._positions = positions
        self._matrix = None if matrix is None else np.array(matrix, dtype=np.complex128)
        if self._matrix is not None and self._matrix.shape != (1 << len(positions),) * 2:
            raise ValueError("gate matrix dimension does not match its positions")

    # -- description ------------------------------------------------------
    def positions(self):
        return self._positions

    def nqubits(self):
        return len(self._positions)

    def mat(self):
        return self._matrix

    def ordered_positions(self):
        return tuple(sorted(self._positions))

    def ordered_mat(self):
        return reorder_legs(self._positions, self.mat())[1]

    # parameters: none by default
    def parameters(self):
        return None

    def is_parameters(self):
        return ()

    def nparameters(self):
        return 0

    def active_parameters(self):
        return []

    def with_parameters(self, values):
        if values:
            raise ValueError("gate has no parameters")
        return self

    def differentiate(self):
        """One derivative gate (dU/dtheta)^dagger U per active parameter, or None."""
        return None

    def adjoint(self):
        return QuantumGate(self._positions, self.mat().conj().T)

    def shift(self, j):
        g = QuantumGate(tuple(p + j for p in self._positions), self.mat())
        return g

    def __repr__(self):
        return f"{type(self).__name__}{self._positions}"


class QuantumGate(Gate):
    """General gate from an explicit matrix (reference quantum_gate.jl:5-27)."""


class _Fixed(Gate):
    MATRIX = None

    def __init__(self, *positions):
        if len(positions) == 1 and isinstance(positions[0], (tuple, list)):
            positions = tuple(positions[0])
        super().__init__(positions, self.MATRIX)


class XGate(_Fixed):
    MATRIX = X


class YGate(_Fixed):
    MATRIX = Y


class ZGate(_Fixed):
    MATRIX = Z


class HGate(_Fixed):
    MATRIX = H


class SGate(_Fixed):
    MATRIX = S


class TGate(_Fixed):
    MATRIX = T


class SWAPGate(_Fixed):
    MATRIX = SWAP


class iSWAPGate(_Fixed):
    MATRIX = ISWAP


class TOFFOLIGate(_Fixed):
    MATRIX = TOFFOLI


class FREDKINGate(_Fixed):
    MATRIX = FREDKIN


class CZGate(_Fixed):
    MATRIX = CZ


class CNOTGate(Gate):
    """CNOT(control, target) (reference Gates.jl:43-53; the first position is the control)."""

    def __init__(self, control, target=None):
        if target is None:
            control, target = control
        super().__init__((control, target), CNOT)


class _Param1(Gate):
    """One-parameter gate; `isparas` marks the parameter active (trainable)."""

    NQ = 1

    def __init__(self, positions, theta, isparas=True):
        if isinstance(positions, int):
            positions = (positions,)
        self.theta = float(theta)
        self.isparas = bool(isparas)
        super().__init__(positions, None)
        if len(self._positions) != self.NQ:
            raise ValueError(f"{type(self).__name__} acts on {self.NQ} qubit(s)")

    def _m(self, theta):
        raise NotImplementedError

    def _dm(self, theta):
        raise NotImplementedError

    def mat(self):
        return self._m(self.theta)

    def parameters(self):
        return [self.theta]

    def is_parameters(self):
        return (self.isparas,)

    def nparameters(self):
        return 1 if self.isparas else 0

    def active_parameters(self):
        return [self.theta] if self.isparas else []

    def with_parameters(self, values):
        values = list(values)
        if len(values) != self.nparameters():
            raise ValueError("wrong number of parameters")
        if not values:
            return self
        return type(self)(self._positions, values[0], isparas=self.isparas)

    def differentiate(self):
        # (dU/dtheta)^dagger U (reference gates/differentiation.jl:14-97)
        if not self.isparas:
            return None
        dm = self._dm(self.theta)
        return [QuantumGate(self._positions, dm.conj().T @ self.mat())]


class RxGate(_Param1):
    def _m(self, t):
        return Rx(t)

    def _dm(self, t):
        return 0.5 * Rx(t + np.pi)


class RyGate(_Param1):
    def _m(self, t):
        return Ry(t)

    def _dm(self, t):
        return 0.5 * Ry(t + np.pi)


class RzGate(_Param1):
    def _m(self, t):
        return Rz(t)

    def _dm(self, t):
        return 0.5 * Rz(t + np.pi)


class PHASEGate(_Param1):
    def _m(self, t):
        return PHASE(t)

    def _dm(self, t):
        return _phase_mat_diff(t)


class CRxGate(_Param1):
    NQ = 2

    def _m(self, t):
        return CONTROL(Rx(t))

    def _dm(self, t):
        return rkron(DOWN, 0.5 * Rx(t + np.pi))


class CRyGate(_Param1):
    NQ = 2

    def _m(self, t):
        return CONTROL(Ry(t))

    def _dm(self, t):
        return rkron(DOWN, 0.5 * Ry(t + np.pi))


class CRzGate(_Param1):
    NQ = 2

    def _m(self, t):
        return CONTROL(Rz(t))

    def _dm(self, t):
        return rkron(DOWN, 0.5 * Rz(t + np.pi))


class CPHASEGate(_Param1):
    NQ = 2

    def _m(self, t):
        return CONTROL(PHASE(t))

    def _dm(self, t):
        return rkron(DOWN, _phase_mat_diff(t))


def haar_unitary(rng: np.random.Generator, dim: int = 2) -> np.ndarray:
    """Haar-random unitary: QR of a complex Gaussian (Re, Im ~ N(0, 1/2)) with
    the phase fix U = Q diag(R_ii / |R_ii|)."""
    z = (rng.standard_normal((dim, dim)) + 1j * rng.standard_normal((dim, dim))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r)
    return q * (d / np.abs(d))
