"""Write the committed golden fixtures (run from the repository root:
`python tests/golden/make_golden.py`).

The reference is Julia-only and cannot run here, and its tests hold no literal
amplitude vectors, so the fixtures come from the CPU oracle (oracle/cpu_ref.c),
each cross-checked here against the independent dense restatement
(oracle/dense.py) before it is written:

  c0_hea12.npz   12-qubit, 10-layer RY/RZ + CNOT-ladder ansatz from |0...0>,
                 seeded angles; amplitudes (4096), TFIM energy, adjoint
                 gradient (240 slots).
  c1_haar26.npz  Haar 2x2 unitaries of the 26-qubit depth-20 bandwidth circuit.
  c2_noisy4.npz  4-qubit, 2-layer noisy ansatz (depolarizing 0.01, damping 0.02):
                 final rho, TFIM energy, adjoint gradient.
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from oracle import dense  # noqa: E402
from mossquantum_amd import lower  # noqa: E402
from mossquantum_amd.workloads import haar_layers, hea_angles, hea_circuit, tfim  # noqa: E402

OUT = Path(__file__).resolve().parent
SEED_C0, SEED_C1, SEED_C2 = 20260101, 20260126, 20260113


def c0():
    n, L = 12, 10
    ang = hea_angles(n, L, SEED_C0)
    lw = lower(hea_circuit(n, L, ang), n, False)
    value, grads, psi = oracle.adjoint(lw, tfim(n).groups(), want_state=True)
    E_dense, psi_dense = dense.hea_energy(n, L, ang, dense.operator_matrix(n, dense.tfim_terms(n)))
    assert np.max(np.abs(psi - psi_dense)) < 1e-12
    assert abs(value - E_dense) < 1e-10
    np.savez(OUT / "c0_hea12.npz", angles=ang, amplitudes=psi, energy=value, grads=grads, seed=SEED_C0)


def c1():
    U = haar_layers(26, 20, SEED_C1)
    for u in U.reshape(-1, 2, 2):
        assert np.allclose(u.conj().T @ u, np.eye(2), atol=1e-14)
    np.savez(OUT / "c1_haar26.npz", unitaries=U, seed=SEED_C1)


def c2():
    n, L = 4, 2
    ang = hea_angles(n, L, SEED_C2)
    lw = lower(hea_circuit(n, L, ang, noise=(0.01, 0.02)), n, True)
    value, grads, rho = oracle.adjoint(lw, tfim(n).groups(), want_state=True)
    # independent dense evolution
    from mossquantum_amd import AmplitudeDamping, Depolarizing
    r = np.zeros((1 << n, 1 << n), dtype=np.complex128)
    r[0, 0] = 1
    for pos, Um in dense.hea_dense_ops(n, L, ang):
        Uf = dense.embed(n, pos, Um)
        r = Uf @ r @ Uf.conj().T
        for q in pos:
            r = dense.kraus_apply(r, n, (q,), Depolarizing(q, 0.01).kraus_matrices())
            r = dense.kraus_apply(r, n, (q,), AmplitudeDamping(q, 0.02).kraus_matrices())
    assert np.max(np.abs(rho.reshape(r.shape, order="F") - r)) < 1e-12
    assert abs(value - np.real(np.trace(dense.operator_matrix(n, dense.tfim_terms(n)) @ r))) < 1e-12
    np.savez(OUT / "c2_noisy4.npz", angles=ang, rho=rho, energy=value, grads=grads, seed=SEED_C2)


if __name__ == "__main__":
    c0()
    c1()
    c2()
    print("fixtures written to", OUT)
