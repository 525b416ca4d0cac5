"""Pin the CPU oracle (oracle/cpu_ref.c) before trusting it.

Checks follow the reference's own test strategy (there are no literal
amplitude vectors in the reference tests):
  * brute-force dense embedding of every 1..5-qubit position pattern on 6 qubits
    (reference test/regression/kernels.jl:4-43);
  * density-matrix gates vs U rho U^dagger, channels vs sum_K K rho K^dagger
    (test/unit/noise_channels.jl:10-19), depolarizing (1-p) rho + p I/2
    (noise_channels.jl:24-26);
  * expectations vs <psi|H|psi> with a dense H, DM vs tr(H rho), pure vs
    DensityMatrix(psi) (test/unit/density_matrices.jl:4-60);
  * adjoint gradients vs closed forms (Ry: cos, -sin; amplitude damping:
    gamma + (1-gamma) cos, -(1-gamma) sin; test/unit/adjoint_objectives.jl:111-136),
    vs central finite differences (test/regression/gradients.jl), and the
    singular gamma=1 channel (gradient 0, gradients.jl:81-82).
Each check would fail for a dropped term, a wrong sign, a transposed operand
or a swapped index; `test_planted_mistakes_are_caught` plants such mistakes
in the oracle's inputs to show it.
"""
import itertools

import numpy as np
import pytest

import oracle
from oracle import dense
from mossquantum_amd import (AmplitudeDamping, CNOTGate, Depolarizing, QCircuit, QuantumGate, RyGate, RzGate,
                             lower, tfim_chain, xxz_chain)
from mossquantum_amd.gates import haar_unitary

RNG = np.random.default_rng(1234)


def rand_state(n, rng=RNG):
    v = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    return v / np.linalg.norm(v)


def rand_unitary(k, rng=RNG):
    return haar_unitary(rng, 1 << k)


def rand_rho(n, rng=RNG):
    A = rng.standard_normal((1 << n, 1 << n)) + 1j * rng.standard_normal((1 << n, 1 << n))
    rho = A @ A.conj().T
    return rho / np.trace(rho)


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_gate_all_positions_vs_embedding(k):
    n = 6
    for pos in itertools.combinations(range(1, n + 1), k):
        U = rand_unitary(k)
        psi = rand_state(n)
        want = dense.embed(n, pos, U) @ psi
        got = psi.copy()
        oracle.apply_gate(got, n, [p - 1 for p in pos], U)
        assert np.max(np.abs(got - want)) < 1e-12, pos


def test_tensor_apply_matches_embedding():
    n = 7
    for pos in [(1,), (7,), (2, 5), (6, 3), (1, 4, 7), (5, 2, 3)]:
        U = rand_unitary(len(pos))
        psi = rand_state(n)
        assert np.allclose(dense.apply_tensor(psi, n, list(pos), U), dense.embed(n, pos, U) @ psi, atol=1e-13)


def test_cz_specialisation():
    n = 6
    Zc = np.diag([1, 1, 1, -1]).astype(np.complex128)
    for a, b in itertools.combinations(range(1, n + 1), 2):
        psi = rand_state(n)
        got = psi.copy()
        oracle.lib().ref_apply_cz(got.ctypes.data_as(oracle.C.c_void_p), n, a - 1, b - 1)
        assert np.allclose(got, dense.embed(n, (a, b), Zc) @ psi, atol=1e-14)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_dm_gate_vs_conjugation(k):
    n = 4
    for pos in list(itertools.combinations(range(1, n + 1), k))[:6]:
        U = rand_unitary(k)
        rho = rand_rho(n)
        Uf = dense.embed(n, pos, U)
        want = Uf @ rho @ Uf.conj().T
        v = rho.reshape(-1, order="F").copy()
        oracle.dm_apply_gate(v, n, [p - 1 for p in pos], U)
        assert np.max(np.abs(v.reshape(rho.shape, order="F") - want)) < 1e-12


@pytest.mark.parametrize("k", [1, 2])
def test_channel_supermatrix_vs_kraus_sum(k):
    n = 4
    rng = np.random.default_rng(7)
    for pos in [(1,), (3,), (4,)] if k == 1 else [(1, 2), (4, 2), (2, 3)]:
        # random CPTP map from a Stinespring isometry, 3 Kraus operators
        d = 1 << k
        V = np.linalg.qr(rng.standard_normal((3 * d, d)) + 1j * rng.standard_normal((3 * d, d)))[0]
        kraus = [V[i * d:(i + 1) * d, :] for i in range(3)]
        from mossquantum_amd import QuantumMap
        qm = QuantumMap(pos, kraus, check_tp=True)
        S_ref = oracle.supermatrix(k, [dict(enumerate(qm._ordered_kraus))[i] for i in range(3)])
        assert np.allclose(S_ref, qm.ordered_supermat(), atol=1e-14)
        rho = rand_rho(n)
        want = dense.kraus_apply(rho, n, pos, kraus)
        v = rho.reshape(-1, order="F").copy()
        oracle.dm_apply_channel(v, n, [p - 1 for p in sorted(pos)], qm.ordered_supermat())
        assert np.max(np.abs(v.reshape(rho.shape, order="F") - want)) < 1e-12
        # dagger = adjoint map sum K^dag X K
        X = rand_rho(n)
        wantd = sum(dense.embed(n, pos, K).conj().T @ X @ dense.embed(n, pos, K) for K in kraus)
        v = X.reshape(-1, order="F").copy()
        oracle.dm_apply_channel(v, n, [p - 1 for p in sorted(pos)], qm.ordered_supermat(), dagger=True)
        assert np.max(np.abs(v.reshape(rho.shape, order="F") - wantd)) < 1e-12


def test_depolarizing_closed_form():
    n, p = 3, 0.3
    rho = rand_rho(n)
    ch = Depolarizing(2, p)
    v = rho.reshape(-1, order="F").copy()
    oracle.dm_apply_channel(v, n, [1], ch.ordered_supermat())
    # (1-p) rho + p tr_2(rho) (x) I/2 on qubit 2
    r = rho.reshape([2] * (2 * n), order="F")  # axes: row bits q1..q3, col bits q1..q3
    red = np.trace(r, axis1=1, axis2=4)  # partial trace over qubit 2
    mixed = np.zeros_like(r)
    for s in range(2):
        mixed[:, s, :, :, s, :] = red / 2
    want = (1 - p) * r + p * mixed
    assert np.allclose(v.reshape([2] * (2 * n), order="F"), want, atol=1e-13)


def test_expectations_vs_dense():
    for n in (3, 5, 6):
        psi = rand_state(n)
        for H, terms in ((tfim_chain(n), dense.tfim_terms(n)), (xxz_chain(n), dense.xxz_terms(n))):
            Hd = dense.operator_matrix(n, terms)
            groups = H.groups()
            assert abs(oracle.expect(psi, n, groups) - np.vdot(psi, Hd @ psi)) < 1e-12
            phi = rand_state(n)
            assert abs(oracle.cross_expect(phi, psi, n, groups) - np.vdot(phi, Hd @ psi)) < 1e-12
            assert np.allclose(oracle.apply_op(psi, n, groups), Hd @ psi, atol=1e-13)
            rho = rand_rho(n)
            assert abs(oracle.dm_expect(rho.reshape(-1, order="F"), n, groups) - np.trace(Hd @ rho)) < 1e-12
            # pure vs DensityMatrix(psi)
            pr = np.outer(psi, psi.conj()).reshape(-1, order="F")
            assert abs(oracle.dm_expect(pr, n, groups) - oracle.expect(psi, n, groups)) < 1e-12
            # left action on density rows: H rho
            assert np.allclose(oracle.apply_op(rho.reshape(-1, order="F"), 2 * n, groups),
                               (Hd @ rho).reshape(-1, order="F"), atol=1e-13)


def test_four_and_five_site_groups_vs_dense():
    """Groups of 4 and 5 sites (the reference's largest, backends/cpu/operators/
    generic.jl:1): expectation, cross expectation, H psi, tr(H rho) and H rho
    against the full Kronecker embedding of each group matrix; random
    (non-Hermitian) matrices and a Pauli-string operator."""
    n = 7
    rng = np.random.default_rng(45)

    def rm(d):
        return rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))

    groups = [((1, 2, 4, 6), rm(16)), ((2, 3, 5, 6, 7), rm(32)), ((3,), rm(2)), ((1, 7), rm(4))]
    Hd = sum(dense.embed(n, k, m) for k, m in groups)
    psi, phi = rand_state(n), rand_state(n)
    assert abs(oracle.expect(psi, n, groups) - np.vdot(psi, Hd @ psi)) < 1e-12
    assert abs(oracle.cross_expect(phi, psi, n, groups) - np.vdot(phi, Hd @ psi)) < 1e-12
    assert np.allclose(oracle.apply_op(psi, n, groups), Hd @ psi, atol=1e-12)
    rho = rand_rho(n)
    assert abs(oracle.dm_expect(rho.reshape(-1, order="F"), n, groups) - np.trace(Hd @ rho)) < 1e-12
    assert np.allclose(oracle.apply_op(rho.reshape(-1, order="F"), 2 * n, groups),
                       (Hd @ rho).reshape(-1, order="F"), atol=1e-12)
    # a Pauli-string operator: QubitsOperator group matrices against matrix(n, x)
    from mossquantum_amd import QubitsOperator, QubitsTerm
    H = QubitsOperator([QubitsTerm({1: "Z", 2: "X", 4: "Y", 6: "Z"}, 0.7),
                        QubitsTerm({2: "+", 3: "-", 5: "X", 6: "Y", 7: "Z"}, -1.3),
                        QubitsTerm({2: "Y", 3: "Y", 5: "Z", 6: "X", 7: "X"}, 0.4)])
    assert abs(oracle.expect(psi, n, H.groups()) - np.vdot(psi, H.matrix(n) @ psi)) < 1e-12


def test_non_hermitian_group():
    n = 4
    psi = rand_state(n)
    M = RNG.standard_normal((4, 4)) + 1j * RNG.standard_normal((4, 4))
    groups = [((1, 3), M)]
    want = np.vdot(psi, dense.embed(n, (1, 3), M) @ psi)
    assert abs(oracle.expect(psi, n, groups) - want) < 1e-12


def test_hea_program_matches_dense():
    n, L = 5, 3
    ang = RNG.uniform(0, 2 * np.pi, 2 * n * L)
    from mossquantum_amd.workloads import hea_circuit
    lw = lower(hea_circuit(n, L, ang), n, False)
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1
    oracle.run_program(psi, lw)
    E, want = dense.hea_energy(n, L, ang, dense.operator_matrix(n, dense.tfim_terms(n)))
    assert np.max(np.abs(psi - want)) < 1e-12


def _fd(f, x, h=1e-6):
    g = np.zeros_like(x)
    for i in range(x.size):
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        g[i] = (f(xp) - f(xm)) / (2 * h)
    return g


def test_adjoint_ry_closed_form():
    # <Z> after Ry(theta)|0> = cos(theta), derivative -sin(theta)
    for th in (0.3, 1.7, 4.0):
        lw = lower(QCircuit([RyGate(1, th)]), 1, False)
        v, g = oracle.adjoint(lw, [((1,), np.diag([1.0, -1.0]).astype(np.complex128))])
        assert abs(v - np.cos(th)) < 1e-14 and abs(g[0] + np.sin(th)) < 1e-14


def test_adjoint_damping_closed_form():
    # <Z> after Ry(theta) then AmplitudeDamping(gamma): gamma + (1-gamma) cos theta
    for th, gam in ((0.4, 0.2), (2.2, 0.7), (1.0, 1.0)):
        c = QCircuit([RyGate(1, th), AmplitudeDamping(1, gam)])
        lw = lower(c, 1, True)
        v, g = oracle.adjoint(lw, [((1,), np.diag([1.0, -1.0]).astype(np.complex128))])
        assert abs(v - (gam + (1 - gam) * np.cos(th))) < 1e-12
        assert abs(g[0] + (1 - gam) * np.sin(th)) < 1e-12  # gamma = 1 gives 0


def test_adjoint_pure_vs_fd():
    n, L = 4, 2
    ang = RNG.uniform(0, 2 * np.pi, 2 * n * L)
    from mossquantum_amd.workloads import hea_circuit
    H = xxz_chain(n)
    Hd = dense.operator_matrix(n, dense.xxz_terms(n))
    lw = lower(hea_circuit(n, L, ang), n, False)
    v, g = oracle.adjoint(lw, H.groups())
    f = lambda a: dense.hea_energy(n, L, a, Hd)[0]
    assert abs(v - f(ang)) < 1e-12
    assert np.max(np.abs(g - _fd(f, ang))) < 1e-8


def test_finite_difference_energies():
    """The FD helpers the device gradients are checked against: the
    extended-precision energy equals the dense energy; a one-qubit RY/RZ/RX
    chain's extended-precision central FD matches the closed-form derivative
    to 1e-12 (RY(t) on |0> under Z: E = cos t, dE = -sin t); on the ansatz it
    matches the oracle adjoint and, in double precision with compensated sums,
    to within the double FD's rounding floor; compensated and plain
    expectations agree."""
    from mossquantum_amd import RxGate
    from mossquantum_amd.workloads import hea_circuit
    t = 0.83
    Z = [((1,), np.diag([1.0, -1.0]).astype(np.complex128))]
    lw = lower(QCircuit([RyGate(1, t, isparas=True)]), 1, False)
    assert abs(float(oracle.energy_ld(lw, [t], Z)) - np.cos(t)) < 1e-15
    assert abs(oracle.central_fd(lw, [t], Z, [0])[0] + np.sin(t)) < 1e-12
    # X(rx) then Z(rz): <Z> = cos a, independent of the RZ angle
    lw = lower(QCircuit([RxGate(1, t, isparas=True), RzGate(1, 0.4, isparas=True)]), 1, False)
    fd = oracle.central_fd(lw, [t, 0.4], Z, [0, 1])
    assert abs(fd[0] + np.sin(t)) < 1e-12 and abs(fd[1]) < 1e-12
    n, L = 5, 2
    ang = RNG.uniform(0, 2 * np.pi, 2 * n * L)
    H = xxz_chain(n)
    Hd = dense.operator_matrix(n, dense.xxz_terms(n))
    lw = lower(hea_circuit(n, L, ang), n, False)
    assert abs(float(oracle.energy_ld(lw, ang, H.groups())) - dense.hea_energy(n, L, ang, Hd)[0]) < 1e-13
    v, g = oracle.adjoint(lw, H.groups())
    assert np.max(np.abs(oracle.central_fd(lw, ang, H.groups(), range(len(ang))) - g)) < 1e-11
    assert np.max(np.abs(oracle.central_fd(lw, ang, H.groups(), range(len(ang)), extended=False) - g)) < 1e-8
    psi = rand_state(n)
    assert abs(oracle.expect_compensated(psi, n, H.groups()) - oracle.expect(psi, n, H.groups())) < 1e-14


def test_adjoint_density_vs_fd():
    n, L = 3, 1
    ang = RNG.uniform(0, 2 * np.pi, 2 * n * L)
    from mossquantum_amd.workloads import hea_circuit
    H = tfim_chain(n)
    Hd = dense.operator_matrix(n, dense.tfim_terms(n))

    def energy(a):
        rho = np.zeros((1 << n, 1 << n), dtype=np.complex128)
        rho[0, 0] = 1
        for pos, U in dense.hea_dense_ops(n, L, a):
            Uf = dense.embed(n, pos, U)
            rho = Uf @ rho @ Uf.conj().T
            for q in pos:
                rho = dense.kraus_apply(rho, n, (q,), Depolarizing(q, 0.01).kraus_matrices())
                rho = dense.kraus_apply(rho, n, (q,), AmplitudeDamping(q, 0.02).kraus_matrices())
        return float(np.real(np.trace(Hd @ rho)))

    lw = lower(hea_circuit(n, L, ang, noise=(0.01, 0.02)), n, True)
    v, g = oracle.adjoint(lw, H.groups())
    assert abs(v - energy(ang)) < 1e-12
    assert np.max(np.abs(g - _fd(energy, ang))) < 1e-8


def test_probabilities_vs_direct():
    n = 5
    psi = rand_state(n)
    p = oracle.probabilities(psi, n, False, [3, 1])
    want = np.zeros(4)
    for i in range(1 << n):
        want[((i >> 2) & 1) | (((i >> 0) & 1) << 1)] += abs(psi[i]) ** 2
    assert np.allclose(p, want, atol=1e-15)


def test_planted_mistakes_are_caught():
    """A transposed gate, a conjugated gate or a dropped seed term changes the
    oracle's output by far more than the parity tolerances."""
    n = 5
    psi = rand_state(n)
    U = rand_unitary(2)
    good = psi.copy()
    oracle.apply_gate(good, n, [1, 3], U)
    for bad_U in (U.T, U.conj(), U[::-1]):
        bad = psi.copy()
        oracle.apply_gate(bad, n, [1, 3], bad_U)
        assert np.max(np.abs(bad - good)) > 1e-3
    H = xxz_chain(n)
    groups = H.groups()
    dropped = groups[:-1]
    assert abs(oracle.expect(good, n, groups) - oracle.expect(good, n, dropped)) > 1e-3
