"""Seeded circuit and observable builders for the benchmark configurations.

All randomness comes from numpy's PCG64 (np.random.default_rng(seed)), so the
inputs are reproducible across machines; committed fixtures hold the drawn
angles and unitaries.
"""
from __future__ import annotations

import numpy as np

from .channels import AmplitudeDamping, Depolarizing
from .circuit import QCircuit
from .gates import CNOTGate, CZGate, QuantumGate, RyGate, RzGate, haar_unitary
from .operators import tfim_chain, xxz_chain


def hea_angles(n: int, layers: int, seed: int) -> np.ndarray:
    """Angles in slot order (layer, qubit, theta then phi), uniform in [0, 2pi)."""
    return np.random.default_rng(seed).uniform(0.0, 2 * np.pi, size=2 * n * layers)


def hea_circuit(n: int, layers: int, angles=None, noise=None) -> QCircuit:
    """Hardware-efficient ansatz: per layer RY(theta) then RZ(phi) on every qubit,
    then the CNOT ladder (1,2), ..., (n-1,n) (reference benchmark_circuit,
    benchmark/benchmarks.jl:14-31).  `noise=(p, gamma)` appends
    Depolarizing(p) then AmplitudeDamping(gamma) on every touched qubit after
    every gate."""
    if angles is None:
        angles = np.zeros(2 * n * layers)
    c = QCircuit()
    k = 0

    def add(g, qubits):
        c.append(g)
        if noise is not None:
            p, gamma = noise
            for q in qubits:
                c.append(Depolarizing(q, p))
                c.append(AmplitudeDamping(q, gamma))

    for _ in range(layers):
        for q in range(1, n + 1):
            add(RyGate(q, angles[k], isparas=True), (q,))
            add(RzGate(q, angles[k + 1], isparas=True), (q,))
            k += 2
        for q in range(1, n):
            add(CNOTGate(q, q + 1), (q, q + 1))
    return c


def haar_layers(n: int, depth: int, seed: int):
    """Haar 2x2 unitaries for every (layer, qubit), shape (depth, n, 2, 2)."""
    rng = np.random.default_rng(seed)
    return np.array([[haar_unitary(rng, 2) for _ in range(n)] for _ in range(depth)])


def random_circuit(n: int, depth: int, unitaries) -> QCircuit:
    """Bandwidth circuit: each layer a Haar 2x2 on every qubit, then CZ on the
    pairs (1,2),(3,4),... in even layers and (2,3),(4,5),... in odd layers."""
    c = QCircuit()
    for layer in range(depth):
        for q in range(1, n + 1):
            c.append(QuantumGate((q,), unitaries[layer, q - 1]))
        start = 1 if layer % 2 == 0 else 2
        for q in range(start, n, 2):
            c.append(CZGate(q, q + 1))
    return c


def tfim(n: int):
    return tfim_chain(n, 1.0)


def xxz(n: int):
    return xxz_chain(n, 0.5)


def batch_sizes(B: int, lo: int, hi: int, seed: int) -> np.ndarray:
    return np.random.default_rng(seed).integers(lo, hi + 1, size=B)


def c3_batch(B: int = 512, layers: int = 6, seed: int = 20260300):
    """The batched-gradient workload: B states with n ~ U{10..20} (seeded), each
    the ansatz with `layers` layers and its own angles, loss the XXZ chain.
    Returns (sizes, circuits, observables)."""
    sizes = [int(x) for x in batch_sizes(B, 10, 20, seed)]
    rng = np.random.default_rng(seed + 1)
    circs = [hea_circuit(n, layers, rng.uniform(0, 2 * np.pi, 2 * n * layers)) for n in sizes]
    return sizes, circs, [xxz(n) for n in sizes]
