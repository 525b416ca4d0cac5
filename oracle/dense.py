"""Textbook dense linear algebra -- TEST INFRASTRUCTURE ONLY.

Independent of both the product code and cpu_ref.c: gates are embedded as full
2^n x 2^n matrices by explicit index mapping (the construction of the
reference's regression_embed_gate, test/regression/kernels.jl:4-17), or applied
by tensor contraction; density matrices evolve as U rho U^dagger and
sum_K K rho K^dagger; observables are sums of Kronecker products.  Used to pin
cpu_ref on small registers (tests/test_oracle.py).
"""
from __future__ import annotations

import numpy as np

I2 = np.eye(2, dtype=np.complex128)
PAULI = {
    "I": I2,
    "X": np.array([[0, 1], [1, 0]], dtype=np.complex128),
    "Y": np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
    "Z": np.array([[1, 0], [0, -1]], dtype=np.complex128),
}


def embed(n, positions1, U):
    """Full matrix of a k-qubit operator U on `positions1` (1-based; the first
    position is the least significant local bit)."""
    bits = [p - 1 for p in positions1]
    k = len(bits)
    dim = 1 << n
    M = np.zeros((dim, dim), dtype=np.complex128)
    mask = sum(1 << b for b in bits)
    for col in range(dim):
        lc = sum(((col >> b) & 1) << j for j, b in enumerate(bits))
        rem = col & ~mask
        for lr in range(1 << k):
            row = rem
            for j, b in enumerate(bits):
                row |= ((lr >> j) & 1) << b
            M[row, col] = U[lr, lc]
    return M


def apply_tensor(psi, n, positions1, U):
    """U applied by tensor contraction (state axis of qubit q is n - q)."""
    k = len(positions1)
    T = np.asarray(psi, dtype=np.complex128).reshape((2,) * n)
    Ut = np.asarray(U, dtype=np.complex128).reshape((2,) * (2 * k))
    in_axes = [k + (k - 1 - j) for j in range(k)]
    st_axes = [n - positions1[j] for j in range(k)]
    R = np.tensordot(Ut, T, axes=(in_axes, st_axes))
    remaining = [a for a in range(n) if a not in st_axes]
    cur = [n - positions1[k - 1 - i] for i in range(k)] + remaining
    return np.moveaxis(R, list(range(n)), cur).reshape(-1)


def operator_matrix(n, terms):
    """sum_t c_t prod_q op_q as a dense matrix; terms = [({pos1: 'X'|2x2}, coeff)]."""
    dim = 1 << n
    M = np.zeros((dim, dim), dtype=np.complex128)
    for ops, c in terms:
        full = np.ones((1, 1), dtype=np.complex128)
        for q in range(1, n + 1):
            o = ops.get(q, "I")
            o = PAULI[o] if isinstance(o, str) else np.asarray(o, dtype=np.complex128)
            full = np.kron(o, full)
        M += c * full
    return M


def tfim_terms(n, h=1.0):
    return [({i: "Z", i + 1: "Z"}, 1.0) for i in range(1, n)] + [({i: "X"}, h) for i in range(1, n + 1)]


def xxz_terms(n, jzz=0.5):
    out = []
    for i in range(1, n):
        out += [({i: "X", i + 1: "X"}, 1.0), ({i: "Y", i + 1: "Y"}, 1.0), ({i: "Z", i + 1: "Z"}, jzz)]
    return out


def kraus_apply(rho, n, positions1, kraus):
    out = np.zeros_like(rho)
    for K in kraus:
        Kf = embed(n, positions1, K)
        out += Kf @ rho @ Kf.conj().T
    return out


def ry(t):
    c, s = np.cos(t / 2), np.sin(t / 2)
    return np.array([[c, -s], [s, c]], dtype=np.complex128)


def rz(t):
    return np.diag([np.exp(-0.5j * t), np.exp(0.5j * t)])


CNOT_CTRL_FIRST = np.array([[1, 0, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0], [0, 1, 0, 0]], dtype=np.complex128)


def hea_dense_ops(n, layers, angles):
    """The hardware-efficient ansatz as (positions1, matrix) pairs, built
    independently of the product's gate classes."""
    ops = []
    k = 0
    for _ in range(layers):
        for q in range(1, n + 1):
            ops.append(((q,), ry(angles[k])))
            ops.append(((q,), rz(angles[k + 1])))
            k += 2
        for q in range(1, n):
            ops.append(((q, q + 1), CNOT_CTRL_FIRST))
    return ops


def hea_energy(n, layers, angles, H):
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1
    for pos, U in hea_dense_ops(n, layers, angles):
        psi = apply_tensor(psi, n, list(pos), U)
    return float(np.real(np.vdot(psi, H @ psi))), psi
