"""HIP backend parity against the CPU oracle (GPU tests, call through the C ABI).

Tolerances (the plain complex128 method): amplitudes 1e-12, expectations
1e-10, gradients 1e-9.
"""
import ctypes as C
import itertools

import numpy as np
import pytest

import oracle
from oracle import dense
import mossquantum_amd as mq
from mossquantum_amd import lower
from mossquantum_amd.backend import plan_emulate
from mossquantum_amd.gates import haar_unitary
from mossquantum_amd.workloads import hea_circuit, random_circuit, haar_layers, tfim, xxz

pytestmark = pytest.mark.gpu
GOLD = __import__("pathlib").Path(__file__).resolve().parent / "golden"
AMP, EXP, GRAD = 1e-12, 1e-10, 1e-9


def rand_state(n, rng):
    v = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    return v / np.linalg.norm(v)


def rand_rho(n, rng):
    A = rng.standard_normal((1 << n, 1 << n)) + 1j * rng.standard_normal((1 << n, 1 << n))
    r = A @ A.conj().T
    return r / np.trace(r)


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_gate_positions(ctx, k):
    rng = np.random.default_rng(10 + k)
    n = 7
    for pos in itertools.combinations(range(1, n + 1), k):
        U = haar_unitary(rng, 1 << k)
        psi = rand_state(n, rng)
        sv = mq.StateVector.from_host(psi)
        mq.apply_(mq.QuantumGate(pos, U), sv)
        want = psi.copy()
        oracle.apply_gate(want, n, [p - 1 for p in pos], U)
        assert np.max(np.abs(sv.to_host() - want)) < AMP, pos


@pytest.mark.parametrize("k", [4, 5])
def test_wide_gates_matrix_core_path(ctx, k):
    """Four- and five-qubit gates on states with at least 16 blocks take the
    FP64 matrix-core kernel: random ordered target sets (low, high, adjacent,
    scattered, unsorted), on 12 qubits and, at the 16-block threshold, on k + 4."""
    rng = np.random.default_rng(40 + k)
    for n in (12, k + 4):
        psi = rand_state(n, rng)
        sv = mq.StateVector.from_host(psi)
        want = psi.copy()
        sets = [tuple(range(1, k + 1)), tuple(range(n - k + 1, n + 1))]
        sets += [tuple(int(p) + 1 for p in rng.permutation(n)[:k]) for _ in range(10)]
        for pos in sets:
            g = mq.QuantumGate(pos, haar_unitary(rng, 1 << k))
            mq.apply_(g, sv)
            oracle.apply_gate(want, n, [p - 1 for p in g.ordered_positions()], g.ordered_mat())
            assert np.max(np.abs(sv.to_host() - want)) < AMP, (n, pos)


def test_large_qubit_gates(ctx):
    rng = np.random.default_rng(3)
    n = 20
    psi = rand_state(n, rng)
    sv = mq.StateVector.from_host(psi)
    want = psi.copy()
    for pos in [(1,), (2,), (5,), (13,), (20,), (1, 20), (19, 3), (4, 11, 17)]:
        U = haar_unitary(rng, 1 << len(pos))
        g = mq.QuantumGate(pos, U)
        mq.apply_(g, sv)
        oracle.apply_gate(want, n, [p - 1 for p in g.ordered_positions()], g.ordered_mat())
    assert np.max(np.abs(sv.to_host() - want)) < AMP


def test_monomial_gates(ctx):
    """CZ, CNOT (both orientations), SWAP, iSWAP, X, S, T, TOFFOLI, FREDKIN, PHASE, Rz."""
    rng = np.random.default_rng(5)
    n = 9
    gates = [mq.CZGate(2, 7), mq.CNOTGate(3, 1), mq.CNOTGate(1, 8), mq.SWAPGate(4, 9), mq.iSWAPGate(5, 2),
             mq.XGate(6), mq.SGate(1), mq.TGate(9), mq.TOFFOLIGate(2, 5, 3), mq.FREDKINGate(8, 1, 4),
             mq.PHASEGate(3, 0.77), mq.RzGate(7, 1.3), mq.CPHASEGate((4, 6), 0.4)]
    psi = rand_state(n, rng)
    sv = mq.StateVector.from_host(psi)
    want = psi.copy()
    for g in gates:
        mq.apply_(g, sv)
        oracle.apply_gate(want, n, [p - 1 for p in g.ordered_positions()], g.ordered_mat())
    assert np.max(np.abs(sv.to_host() - want)) < AMP


@pytest.mark.parametrize("k", [1, 2, 3])
def test_dm_gates(ctx, k):
    rng = np.random.default_rng(20 + k)
    n = 5
    rho = rand_rho(n, rng)
    dm = mq.DensityMatrix.from_host(rho)
    want = rho.reshape(-1, order="F").copy()
    for pos in list(itertools.combinations(range(1, n + 1), k))[:5]:
        U = haar_unitary(rng, 1 << k)
        mq.apply_(mq.QuantumGate(pos, U), dm)
        oracle.dm_apply_gate(want, n, [p - 1 for p in pos], U)
    assert np.max(np.abs(dm.to_host().reshape(-1, order="F") - want)) < AMP


def test_channels_and_promotion(ctx):
    rng = np.random.default_rng(8)
    n = 5
    psi = rand_state(n, rng)
    sv = mq.StateVector.from_host(psi)
    st = mq.apply_(mq.AmplitudeDamping(3, 0.3), sv)  # SV -> DM promotion
    assert isinstance(st, mq.DensityMatrix)
    want = np.outer(psi, psi.conj()).reshape(-1, order="F")
    oracle.dm_apply_channel(want, n, [2], mq.AmplitudeDamping(3, 0.3).ordered_supermat())
    for ch in [mq.Depolarizing(1, 0.2), mq.Depolarizing((2, 5), 0.1), mq.PhaseDamping(4, 0.36)]:
        st = mq.apply_(ch, st)
        oracle.dm_apply_channel(want, n, [p - 1 for p in ch.ordered_positions()], ch.ordered_supermat())
    assert np.max(np.abs(st.to_host().reshape(-1, order="F") - want)) < AMP
    assert abs(mq.trace(st) - 1) < 1e-12


def test_channel_dagger(ctx):
    rng = np.random.default_rng(9)
    n = 4
    X = rand_rho(n, rng)
    ch = mq.Depolarizing((1, 3), 0.2)
    dm = mq.DensityMatrix.from_host(X)
    lib = ctx.lib
    pos = np.array([0, 2], dtype=np.int32)
    S = np.ascontiguousarray(ch.ordered_supermat().reshape(-1, order="F"))
    mq._lib.check(lib.mossq_dm_apply_channel(ctx.h, dm.storage.ptr, n, 2, mq._lib.iptr(pos),
                                             S.view(np.float64).ctypes.data_as(mq._lib._DP), 1))
    want = X.reshape(-1, order="F").copy()
    oracle.dm_apply_channel(want, n, [0, 2], ch.ordered_supermat(), dagger=True)
    assert np.max(np.abs(dm.to_host().reshape(-1, order="F") - want)) < AMP


def test_expectations(ctx):
    rng = np.random.default_rng(11)
    for n in (4, 9, 16):
        psi = rand_state(n, rng)
        sv = mq.StateVector.from_host(psi)
        for H in (tfim(n), xxz(n)):
            g = H.groups()
            assert abs(mq.expectation(H, sv) - oracle.expect(psi, n, g)) < EXP
            phi = rand_state(n, rng)
            assert abs(mq.cross_expectation(mq.StateVector.from_host(phi), H, sv)
                       - oracle.cross_expect(phi, psi, n, g)) < EXP
            out = mq.apply(H, sv).to_host()
            assert np.max(np.abs(out - oracle.apply_op(psi, n, g))) < AMP
    n = 6
    rho = rand_rho(n, rng)
    dm = mq.DensityMatrix.from_host(rho)
    H = tfim(n)
    assert abs(mq.expectation(H, dm) - oracle.dm_expect(rho.reshape(-1, order="F"), n, H.groups())) < EXP
    hr = mq.apply(H, dm).to_host()
    assert np.max(np.abs(hr.reshape(-1, order="F")
                         - oracle.apply_op(rho.reshape(-1, order="F"), 2 * n, H.groups()))) < AMP
    assert abs(mq.purity(dm) - np.real(np.vdot(rho, rho))) < EXP


def wide_operator():
    """Pauli strings on 4 and 5 sites (and a one-site term), one with ladder
    operators (non-Hermitian group), coefficients real."""
    return mq.QubitsOperator([mq.QubitsTerm({1: "Z", 2: "X", 4: "Y", 6: "Z"}, 0.7),
                              mq.QubitsTerm({2: "+", 3: "-", 5: "X", 6: "Y", 7: "Z"}, -1.3),
                              mq.QubitsTerm({2: "Y", 3: "Y", 5: "Z", 6: "X", 7: "X"}, 0.4),
                              mq.QubitsTerm({3: "X"}, 0.25)])


def test_four_and_five_site_groups(ctx):
    """Observables with 4- and 5-site groups (the reference's largest group
    width): expectation, cross expectation, H psi, tr(H rho), H rho, the
    adjoint value and gradient, and the batched engine's observe step, against
    the oracle."""
    rng = np.random.default_rng(46)
    H = wide_operator()
    g = H.groups()
    assert max(len(k) for k, _ in g) == 5
    for n in (7, 13, 18):
        psi, phi = rand_state(n, rng), rand_state(n, rng)
        sv = mq.StateVector.from_host(psi)
        assert abs(mq.expectation(H, sv) - oracle.expect(psi, n, g)) < EXP
        assert abs(mq.cross_expectation(mq.StateVector.from_host(phi), H, sv) - oracle.cross_expect(phi, psi, n, g)) < EXP
        assert np.max(np.abs(mq.apply(H, sv).to_host() - oracle.apply_op(psi, n, g))) < AMP
    n = 7
    rho = rand_rho(n, rng)
    dm = mq.DensityMatrix.from_host(rho)
    assert abs(mq.expectation(H, dm) - oracle.dm_expect(rho.reshape(-1, order="F"), n, g)) < EXP
    assert np.max(np.abs(mq.apply(H, dm).to_host().reshape(-1, order="F")
                         - oracle.apply_op(rho.reshape(-1, order="F"), 2 * n, g))) < AMP
    # adjoint: the seed (H + H^dagger) psi of a non-Hermitian observable, fused and unfused
    for n, fuse in ((8, False), (13, True)):
        tpl = hea_circuit(n, 2)
        th = rng.uniform(0, 2 * np.pi, tpl.nparameters())
        obj = mq.ExpectationObjective(tpl, H, mq.StateVector.zeros(n, ctx), backend_=mq.HIPBackend(ctx, fuse=fuse))
        v, gr = mq.value_and_gradient(mq.AdjointGradient(), obj, th)
        wv, wg = oracle.adjoint(lower(tpl.bind(th), n, False), g)
        assert abs(v - wv) < EXP and np.max(np.abs(gr - wg)) < GRAD
    # batched engine (gate-by-gate and tiled items) with the wide observable
    sizes = [9, 12, 14]
    circs = [hea_circuit(n, 2, rng.uniform(0, 2 * np.pi, 4 * n)) for n in sizes]
    vals, grads = mq.batched_adjoint(circs, [H] * 3, sizes, ctx, fuse=True)
    for i, n in enumerate(sizes):
        wv, wg = oracle.adjoint(lower(circs[i], n, False), g)
        assert abs(vals[i] - wv) < EXP and np.max(np.abs(grads[i] - wg)) < GRAD


def test_probabilities(ctx):
    rng = np.random.default_rng(12)
    n = 10
    psi = rand_state(n, rng)
    sv = mq.StateVector.from_host(psi)
    for qs in ([1], [3, 1], [10, 2, 5], list(range(1, 11))):
        raw = oracle.probabilities(psi, n, False, qs)
        assert np.max(np.abs(mq.probabilities(sv, qs) - raw / raw.sum())) < 1e-14
    rho = rand_rho(5, rng)
    raw = oracle.probabilities(rho.reshape(-1, order="F"), 5, True, [2, 4])
    assert np.max(np.abs(mq.probabilities(mq.DensityMatrix.from_host(rho), [2, 4]) - raw / raw.sum())) < 1e-14


def test_c0_golden(ctx):
    f = np.load(GOLD / "c0_hea12.npz")
    n, L = 12, 10
    c = hea_circuit(n, L, f["angles"])
    for fuse in (False, True):
        b = mq.HIPBackend(ctx, fuse=fuse)
        sv = mq.apply_(c, mq.StateVector.zeros(n), b)
        assert np.max(np.abs(sv.to_host() - f["amplitudes"])) < AMP, fuse
        assert abs(mq.expectation(tfim(n), sv).real - float(f["energy"])) < EXP


@pytest.mark.parametrize("n", [12, 14, 17])
def test_fused_random_circuit(ctx, n):
    U = haar_layers(n, 6, 77 + n)
    c = random_circuit(n, 6, U)
    psi0 = rand_state(n, np.random.default_rng(n))
    want = psi0.copy()
    oracle.run_program(want, lower(c, n, False))
    for fuse in (False, True):
        sv = mq.StateVector.from_host(psi0)
        mq.apply_(c, sv, mq.HIPBackend(ctx, fuse=fuse))
        assert np.max(np.abs(sv.to_host() - want)) < AMP, fuse
    p = mq.Program(c, n, False, ctx, fuse=True)
    assert p.sweeps() < len(list(c.flat()))


def test_fused_single_long_sweep(ctx):
    """A 12-qubit circuit is one sweep: at depth 40 that sweep runs ~370 tile
    ops over ~40 register layouts (long op lists, many transpositions)."""
    n = 12
    psi0 = rand_state(n, np.random.default_rng(5))
    for depth in (2, 40):
        c = random_circuit(n, depth, haar_layers(n, depth, 300 + depth))
        _, st = plan_emulate(c, n, False, psi0)
        assert st["sweeps"] == 1
        want = psi0.copy()
        oracle.run_program(want, lower(c, n, False))
        sv = mq.StateVector.from_host(psi0)
        mq.apply_(c, sv, mq.HIPBackend(ctx, fuse=True))
        assert np.max(np.abs(sv.to_host() - want)) < AMP, depth


def test_fused_density_program(ctx):
    n = 6  # 12 bits: one tile
    rng = np.random.default_rng(13)
    c = hea_circuit(n, 2, rng.uniform(0, 2 * np.pi, 2 * n * 2), noise=(0.01, 0.02))
    want = np.zeros(1 << (2 * n), dtype=np.complex128)
    want[0] = 1
    oracle.run_program(want, lower(c, n, True))
    for fuse in (False, True):
        dm = mq.apply_(c, mq.DensityMatrix.zeros(n), mq.HIPBackend(ctx, fuse=fuse))
        assert np.max(np.abs(dm.to_host().reshape(-1, order="F") - want)) < AMP, fuse
    n = 7  # 14 bits
    c = hea_circuit(n, 1, rng.uniform(0, 2 * np.pi, 2 * n), noise=(0.01, 0.02))
    want = np.zeros(1 << (2 * n), dtype=np.complex128)
    want[0] = 1
    oracle.run_program(want, lower(c, n, True))
    dm = mq.apply_(c, mq.DensityMatrix.zeros(n), mq.HIPBackend(ctx, fuse=True))
    assert np.max(np.abs(dm.to_host().reshape(-1, order="F") - want)) < AMP


def test_adjoint_c0(ctx):
    f = np.load(GOLD / "c0_hea12.npz")
    n, L = 12, 10
    tpl = hea_circuit(n, L)
    obj = mq.ExpectationObjective(tpl, tfim(n), mq.StateVector.zeros(n))
    v, g = mq.value_and_gradient(mq.AdjointGradient(), obj, f["angles"])
    assert abs(v - float(f["energy"])) < EXP
    assert np.max(np.abs(g - f["grads"])) < GRAD


def _mixed_param_circuit(n, layers, seed):
    """Rotations of every kind, CZ and CNOT ladders, a controlled rotation and a
    fixed two-qubit unitary per layer."""
    rng = np.random.default_rng(seed)
    c = mq.QCircuit()
    for layer in range(layers):
        for q in range(1, n + 1):
            c.append([mq.RxGate, mq.RyGate, mq.RzGate, mq.PHASEGate][(q + layer) % 4](q, rng.uniform(0, 6.3),
                                                                                    isparas=True))
        for q in range(1 + layer % 2, n, 2):
            c.append(mq.CZGate(q, q + 1) if layer % 3 else mq.CNOTGate(q, q + 1))
        c.append(mq.CRyGate((1, n), rng.uniform(0, 6.3), isparas=True))
        c.append(mq.QuantumGate((2, n - 1), haar_unitary(rng, 4)))
    return c


@pytest.mark.parametrize("n", [13, 16])
def test_fused_adjoint(ctx, n):
    """The fused reverse pass (adjoint tile kernel, gradient probes inside the
    sweep) against the step-by-step reverse pass and the oracle."""
    rng = np.random.default_rng(40 + n)
    H = tfim(n)
    hea = hea_circuit(n, 3)
    circs = [(hea, rng.uniform(0, 2 * np.pi, hea.nparameters()))]
    mix = _mixed_param_circuit(n, 3, n)
    circs.append((mix, np.array(mix.active_parameters(), dtype=float)))
    for tpl, th in circs:
        lw = lower(tpl.bind(th), n, False)
        want = lw.slot_gradients(oracle.adjoint(lw, H.groups())[1])
        for fuse in (True, False):
            obj = mq.ExpectationObjective(tpl, H, mq.StateVector.zeros(n), backend_=mq.HIPBackend(ctx, fuse=fuse))
            v, g = mq.value_and_gradient(mq.AdjointGradient(), obj, th)
            assert np.max(np.abs(g - want)) < GRAD, (fuse, n)


def _permutation_circuit(n, rng):
    """Rotations interleaved with every two-bit permutation the tile kernels
    move without arithmetic (CNOT either way, SWAP, X on one or both bits),
    some followed by controlled-Z flips, plus phase rotations (one-bit
    diagonals in adjoint plans)."""
    X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
    I2 = np.eye(2, dtype=np.complex128)
    ops = []
    for layer in range(2):
        for q in range(1, n + 1):
            ops += [mq.RyGate(q, rng.uniform(0, 6.3), isparas=True), mq.RzGate(q, rng.uniform(0, 6.3), isparas=True)]
        for q in range(1, n, 2):
            ops += [mq.CNOTGate(q, q + 1), mq.CNOTGate(q + 1, q), mq.SWAPGate(q, q + 1), mq.CZGate(q, q + 1)]
        for q in range(2, n, 3):
            ops += [mq.QuantumGate((q, q + 1), np.kron(X, I2)), mq.QuantumGate((q, q + 1), np.kron(I2, X)),
                    mq.QuantumGate((q, q + 1), np.kron(X, X))]
        ops.append(mq.PHASEGate(n, rng.uniform(0, 6.3), isparas=True))
    return mq.QCircuit(ops)


@pytest.mark.parametrize("n", [12, 14])
def test_fused_permutation_groups(ctx, n):
    """Forward and adjoint tile plans with permutation groups against the
    oracle (amplitudes, energy and gradient)."""
    rng = np.random.default_rng(60 + n)
    c = _permutation_circuit(n, rng)
    _, st = plan_emulate(c, n, False, None)
    assert st["permutations"] > 0
    psi0 = rand_state(n, rng)
    want = psi0.copy()
    oracle.run_program(want, lower(c, n, False))
    sv = mq.StateVector.from_host(psi0)
    mq.apply_(c, sv, mq.HIPBackend(ctx, fuse=True))
    assert np.max(np.abs(sv.to_host() - want)) < AMP
    th = np.array(c.active_parameters(), dtype=float)
    H = tfim(n)
    lw = lower(c, n, False)
    v_want, g_slots = oracle.adjoint(lw, H.groups())
    want_g = lw.slot_gradients(g_slots)
    obj = mq.ExpectationObjective(c, H, mq.StateVector.zeros(n), backend_=mq.HIPBackend(ctx, fuse=True))
    v, g = mq.value_and_gradient(mq.AdjointGradient(), obj, th)
    assert abs(v - v_want) < EXP and np.max(np.abs(g - want_g)) < GRAD


def _noisy_param_circuit(n, layers, seed):
    """Noisy HEA plus a parameterized two-qubit gate, a two-qubit channel, a
    dense two-qubit gate and a three-qubit gate (run unfused)."""
    rng = np.random.default_rng(seed)
    c = hea_circuit(n, layers, noise=(0.01, 0.02))
    c.append(mq.CRyGate((1, n), rng.uniform(0, 6.3), isparas=True))
    c.append(mq.Depolarizing((2, 3), 0.05))
    c.append(mq.QuantumGate((2, n - 1), haar_unitary(rng, 4)))
    c.append(mq.RxGate(3, rng.uniform(0, 6.3), isparas=True))
    c.append(mq.QuantumGate((1, 3, n), haar_unitary(rng, 8)))
    c.append(mq.AmplitudeDamping(n, 0.1))
    c.append(mq.RzGate(n, rng.uniform(0, 6.3), isparas=True))
    return c


@pytest.mark.parametrize("n", [6, 7])
def test_fused_density_adjoint(ctx, n):
    """The fused density adjoint (chunk plans cut at the parameterized gates,
    checkpoint recompute through LDS-tile sweeps) against the step-by-step
    density adjoint and the oracle, from |0><0| and from a pure input."""
    rng = np.random.default_rng(60 + n)
    H = tfim(n)
    tpl = _noisy_param_circuit(n, 3, n)
    th = np.array(tpl.active_parameters(), dtype=float)
    th[: 6 * n] = rng.uniform(0, 2 * np.pi, 6 * n)
    lw = lower(tpl.bind(th), n, True)
    want_v, want_g = oracle.adjoint(lw, H.groups())
    want = lw.slot_gradients(want_g)
    for fuse in (True, False):
        obj = mq.ExpectationObjective(tpl, H, mq.DensityMatrix.zeros(n), backend_=mq.HIPBackend(ctx, fuse=fuse))
        v, g = mq.value_and_gradient(mq.AdjointGradient(), obj, th)
        assert abs(v - want_v) < EXP, fuse
        assert np.max(np.abs(g - want)) < GRAD, (fuse, n)
    obj = mq.ExpectationObjective(tpl, H, mq.StateVector.zeros(n), backend_=mq.HIPBackend(ctx, fuse=True))
    v2, g2 = mq.value_and_gradient(mq.AdjointGradient(), obj, th)
    assert abs(v2 - want_v) < EXP and np.max(np.abs(g2 - want)) < GRAD


def test_fused_density_adjoint_segment_plans(ctx):
    """At 11 qubits (22 bits) each checkpoint segment's reverse pass runs as one
    plan with a fixed probe per parameter; against the step-by-step density
    adjoint, which the smaller cases pin to the oracle."""
    n = 11
    tpl = _noisy_param_circuit(n, 2, n)
    th = np.array(tpl.active_parameters(), dtype=float)
    th[: 4 * n] = np.random.default_rng(n).uniform(0, 2 * np.pi, 4 * n)
    res = {}
    for fuse in (True, False):
        obj = mq.ExpectationObjective(tpl, tfim(n), mq.DensityMatrix.zeros(n), backend_=mq.HIPBackend(ctx, fuse=fuse))
        res[fuse] = mq.value_and_gradient(mq.AdjointGradient(), obj, th)
    assert abs(res[True][0] - res[False][0]) < EXP
    assert np.max(np.abs(res[True][1] - res[False][1])) < GRAD


def test_adjoint_density_golden(ctx):
    f = np.load(GOLD / "c2_noisy4.npz")
    n, L = 4, 2
    tpl = hea_circuit(n, L, noise=(0.01, 0.02))
    obj = mq.ExpectationObjective(tpl, tfim(n), mq.DensityMatrix.zeros(n))
    v, g = mq.value_and_gradient(mq.AdjointGradient(), obj, f["angles"])
    assert abs(v - float(f["energy"])) < EXP
    assert np.max(np.abs(g - f["grads"])) < GRAD
    # pure input promoted at the first channel gives the same gradient
    obj2 = mq.ExpectationObjective(tpl, tfim(n), mq.StateVector.zeros(n))
    v2, g2 = mq.value_and_gradient(mq.AdjointGradient(), obj2, f["angles"])
    assert abs(v2 - v) < EXP and np.max(np.abs(g2 - g)) < GRAD


def test_adjoint_closed_forms(ctx):
    Z = mq.QubitsOperator([mq.QubitsTerm({1: "Z"})])
    for th in (0.3, 2.9):
        obj = mq.ExpectationObjective(mq.QCircuit([mq.RyGate(1, 0.0)]), Z, mq.StateVector.zeros(1))
        v, g = mq.value_and_gradient(mq.AdjointGradient(), obj, [th])
        assert abs(v - np.cos(th)) < 1e-14 and abs(g[0] + np.sin(th)) < 1e-14
        for gam in (0.25, 1.0):
            obj = mq.ExpectationObjective(mq.QCircuit([mq.RyGate(1, 0.0), mq.AmplitudeDamping(1, gam)]), Z,
                                          mq.DensityMatrix.zeros(1))
            v, g = mq.value_and_gradient(mq.AdjointGradient(), obj, [th])
            assert abs(v - (gam + (1 - gam) * np.cos(th))) < 1e-12
            assert abs(g[0] + (1 - gam) * np.sin(th)) < 1e-12


def test_batched_adjoint_vs_single(ctx):
    rng = np.random.default_rng(14)
    sizes = [10, 13, 11, 15, 10]
    circs, obs, want_v, want_g = [], [], [], []
    for n in sizes:
        ang = rng.uniform(0, 2 * np.pi, 2 * n * 2)
        c = hea_circuit(n, 2, ang)
        circs.append(c)
        obs.append(xxz(n))
        v, g = oracle.adjoint(lower(c, n, False), xxz(n).groups())
        want_v.append(v)
        want_g.append(g)
    vals, grads = mq.batched_adjoint(circs, obs, sizes, ctx)
    for i in range(len(sizes)):
        assert abs(vals[i] - want_v[i]) < EXP
        assert np.max(np.abs(grads[i] - want_g[i])) < GRAD


def test_batched_adjoint_uneven_sizes(ctx):
    """Many items with uneven qubit counts (chunked and single-chunk items) in
    one engine run; each item's value and gradient match the oracle, and match
    the item run alone."""
    rng = np.random.default_rng(15)
    sizes = [int(x) for x in rng.integers(10, 17, size=24)]
    circs, obs = [], []
    for n in sizes:
        circs.append(hea_circuit(n, 2, rng.uniform(0, 2 * np.pi, 4 * n)))
        obs.append(xxz(n))
    vals, grads = mq.batched_adjoint(circs, obs, sizes, ctx)
    for i in (0, 5, 11, 23):
        v, g = oracle.adjoint(lower(circs[i], sizes[i], False), obs[i].groups())
        assert abs(vals[i] - v) < EXP
        assert np.max(np.abs(grads[i] - g)) < GRAD
        v1, g1 = mq.batched_adjoint([circs[i]], [obs[i]], [sizes[i]], ctx)
        assert abs(v1[0] - vals[i]) < EXP and np.max(np.abs(g1[0] - grads[i])) < GRAD


def test_batched_adjoint_tiled_matches_stepped(ctx):
    """Items of 12+ qubits run as batched tile sweeps (forward and adjoint
    plans of every item advanced one sweep per launch); the gate-by-gate engine
    and the oracle give the same values and gradients."""
    rng = np.random.default_rng(16)
    sizes = [12, 10, 14, 13, 11, 16, 12, 15, 17, 12]
    circs = [hea_circuit(n, 3, rng.uniform(0, 2 * np.pi, 6 * n)) for n in sizes]
    obs = [xxz(n) for n in sizes]
    v_t, g_t = mq.batched_adjoint(circs, obs, sizes, ctx, fuse=True)
    v_s, g_s = mq.batched_adjoint(circs, obs, sizes, ctx, fuse=False)
    assert np.max(np.abs(v_t - v_s)) < EXP
    for a, b in zip(g_t, g_s):
        assert np.max(np.abs(a - b)) < GRAD
    for i in (0, 2, 5, 8):
        v, g = oracle.adjoint(lower(circs[i], sizes[i], False), obs[i].groups())
        assert abs(v_t[i] - v) < EXP
        assert np.max(np.abs(g_t[i] - g)) < GRAD
    # repeated calls give identical results (fixed-order reductions)
    v2, g2 = mq.batched_adjoint(circs, obs, sizes, ctx, fuse=True)
    assert np.array_equal(v2, v_t) and all(np.array_equal(a, b) for a, b in zip(g2, g_t))


def test_batched_adjoint_tiled_uneven_to_20_qubits(ctx):
    """A mixed batch up to 20 qubits (items of 1 to 256 tiles per sweep, and
    items below 12 qubits on the gate-by-gate engine in the same call): the
    tiled engine matches the gate-by-gate engine on every value and gradient,
    and the oracle on the smaller items."""
    rng = np.random.default_rng(18)
    sizes = [int(x) for x in rng.integers(10, 21, size=40)] + [20, 12, 11]
    circs = [hea_circuit(n, 2, rng.uniform(0, 2 * np.pi, 4 * n)) for n in sizes]
    obs = [xxz(n) for n in sizes]
    v_t, g_t = mq.batched_adjoint(circs, obs, sizes, ctx, fuse=True)
    v_s, g_s = mq.batched_adjoint(circs, obs, sizes, ctx, fuse=False)
    assert np.max(np.abs(v_t - v_s)) < EXP
    assert max(np.max(np.abs(a - b)) for a, b in zip(g_t, g_s)) < GRAD
    for i in [i for i, n in enumerate(sizes) if n <= 13][:3]:
        v, g = oracle.adjoint(lower(circs[i], sizes[i], False), obs[i].groups())
        assert abs(v_t[i] - v) < EXP and np.max(np.abs(g_t[i] - g)) < GRAD


def test_batched_expectation_tiled_with_initial_states(ctx):
    """Value-only batches of fused programs from given initial states match
    one expectation per item."""
    rng = np.random.default_rng(17)
    sizes = [12, 13, 10, 14]
    B = len(sizes)
    progs, dobs, states, want = [], [], [], []
    for n in sizes:
        c = hea_circuit(n, 2, rng.uniform(0, 2 * np.pi, 4 * n))
        st = mq.StateVector.from_host(rand_state(n, rng), ctx=ctx)
        progs.append(mq.Program(c, n, False, ctx, fuse=True))
        dobs.append(mq.DeviceObservable(tfim(n), n, ctx))
        states.append(st)
        ref = st.copy()
        mq.Program(c, n, False, ctx, fuse=False).run(ref.storage)
        want.append(mq.expectation(tfim(n), ref).real)
    ph = (C.c_void_p * B)(*[p.h.value for p in progs])
    oh = (C.c_void_p * B)(*[o.h.value for o in dobs])
    ih = (C.c_void_p * B)(*[s.storage.ptr.value for s in states])
    vals = np.zeros(B)
    mq._lib.check(ctx.lib.mossq_batch_expectation(ctx.h, B, ph, oh, ih, mq._lib.dptr(vals)))
    assert np.max(np.abs(vals - np.array(want))) < EXP


@pytest.mark.parametrize("sizes", [[10, 11, 1, 12, 3], [5]])
def test_batched_small_fused_programs_as_tiles(ctx, sizes):
    """Fused programs below 12 qubits run in the batched engine as their
    12-qubit form (the extra qubits idle in |0>): values from |0...0> (value
    only and with gradients) match the oracle, in batches with and without
    12-qubit items."""
    rng = np.random.default_rng(19 + len(sizes))
    B = len(sizes)
    circs = [hea_circuit(n, 3, rng.uniform(0, 2 * np.pi, 6 * n)) for n in sizes]
    obs = [tfim(n) for n in sizes]
    progs = [mq.Program(c, n, False, ctx, fuse=True) for c, n in zip(circs, sizes)]
    dobs = [mq.DeviceObservable(o, n, ctx) for o, n in zip(obs, sizes)]
    ph = (C.c_void_p * B)(*[p.h.value for p in progs])
    oh = (C.c_void_p * B)(*[o.h.value for o in dobs])
    vals = np.zeros(B)
    mq._lib.check(ctx.lib.mossq_batch_expectation(ctx.h, B, ph, oh, None, mq._lib.dptr(vals)))
    v_adj, g_adj = mq.batched_adjoint(circs, obs, sizes, ctx, fuse=True)
    for i, n in enumerate(sizes):
        v, g = oracle.adjoint(lower(circs[i], n, False), obs[i].groups())
        assert abs(vals[i] - v) < EXP and abs(v_adj[i] - v) < EXP, i
        assert np.max(np.abs(g_adj[i] - g)) < GRAD, i


def test_unsupported_before_payload(ctx):
    sv = mq.StateVector.zeros(3)
    big = mq.QubitsOperator([mq.QubitsTerm({1: "X", 2: "X", 3: "X"})])
    assert mq.supports(mq.backend(sv), big, sv, "expectation")
    wide = mq.QuantumGate((1, 2, 3), np.eye(8))
    assert mq.supports(mq.backend(sv), wide, sv, "apply")
    with pytest.raises(mq.UnsupportedOperation):
        mq.apply_(mq.QuantumGate((1, 2, 3, 4, 5, 6), np.eye(64)), mq.StateVector.zeros(6))  # wider than 5
    with pytest.raises(ValueError, match="between 1 and 3"):
        mq.apply_(mq.QuantumGate((1, 4), np.eye(4)), sv)  # position outside the register (ArgumentError)
    with pytest.raises(ValueError, match="between 1 and 3"):
        mq.expectation(mq.QubitsOperator([mq.QubitsTerm({4: "Z"})]), sv)
    # invalid arguments come back as status codes, not faults
    bad = np.array([2, 1], dtype=np.int32)
    eye = np.eye(4, dtype=np.complex128).reshape(-1)
    st = ctx.lib.mossq_sv_apply_gate(ctx.h, sv.storage.ptr, 3, 2, mq._lib.iptr(bad),
                                     eye.view(np.float64).ctypes.data_as(mq._lib._DP))
    assert st == mq._lib.MOSSQ_EINVAL


@pytest.mark.parametrize("n,fuse", [(13, True), (8, False), (8, True)])
def test_program_graph_replay(ctx, n, fuse):
    """A captured program run replays on its buffer: after each replay (with
    eager work and synchronisations in between) the buffer equals the eager
    path applied as many times, and the first replay equals the oracle."""
    rng = np.random.default_rng(70 + n)
    c = random_circuit(n, 6, haar_layers(n, 6, 70 + n))
    c.append(mq.QuantumGate((1, 3, n), haar_unitary(rng, 8)))  # an unfused three-qubit step
    prog = mq.Program(c, n, False, ctx, fuse=fuse)
    v = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    v /= np.linalg.norm(v)
    a = mq.StateVector.from_host(v, ctx)
    b = mq.StateVector.from_host(v, ctx)
    other = mq.StateVector.zeros(n, ctx)
    graph = prog.capture(a.storage)
    want = v.copy()
    oracle.run_program(want, lower(c, n, False))
    for r in range(3):
        graph.launch()
        prog.run(b.storage)
        mq.apply_(mq.HGate(1), other)  # unrelated eager work between replays
        ha = a.to_host()
        assert np.max(np.abs(ha - b.to_host())) < AMP, r
        if r == 0:
            assert np.max(np.abs(ha - want)) < AMP


def test_autotune_keeps_results(ctx):
    """The autotuner times the forward-plan variants and keeps one; the program
    still gives the oracle's amplitudes, and an unfused program has nothing to tune."""
    n, d = 15, 8
    c = random_circuit(n, d, haar_layers(n, d, 90))
    prog = mq.Program(c, n, False, ctx, fuse=True)
    variants, ms = prog.autotune(reps=2)
    assert variants >= 1 and ms > 0
    sv = mq.StateVector.zeros(n, ctx)
    prog.run(sv.storage)
    want = np.zeros(1 << n, dtype=np.complex128)
    want[0] = 1
    oracle.run_program(want, lower(c, n, False))
    assert np.max(np.abs(sv.to_host() - want)) < AMP
    assert mq.Program(c, n, False, ctx, fuse=False).autotune() == (0, 0.0)


def test_graph_replay_after_autotune_and_churn(ctx):
    """A graph captured before the autotuner replaces the program's plan keeps
    replaying the plan it recorded: the replaced plan's tables stay allocated
    while the graph lives, so the blocks that other programs, adjoints and
    tables take from the context's cache in between never overwrite them."""
    n, d = 15, 8
    c = random_circuit(n, d, haar_layers(n, d, 91))
    prog = mq.Program(c, n, False, ctx, fuse=True)
    sv = mq.StateVector.zeros(n, ctx)
    graph = prog.capture(sv.storage)
    variants, _ = prog.autotune(reps=1)
    assert variants >= 1
    # churn the block cache: programs with plans, adjoints with scratch, all freed
    for seed in range(6):
        cc = random_circuit(n, 3 + seed, haar_layers(n, 3 + seed, 200 + seed))
        p2 = mq.Program(cc, n, False, ctx, fuse=True)
        tmp = mq.StateVector.zeros(n, ctx)
        p2.run(tmp.storage)
        obj = mq.ExpectationObjective(hea_circuit(12, 2), tfim(12), mq.StateVector.zeros(12, ctx))
        mq.value_and_gradient(mq.AdjointGradient(), obj, np.linspace(0.1, 2.0, 48))
        del p2, tmp, obj
    want = np.zeros(1 << n, dtype=np.complex128)
    want[0] = 1
    oracle.run_program(want, lower(c, n, False))
    graph.launch()
    assert np.max(np.abs(sv.to_host() - want)) < AMP
    prog.run(sv.storage)  # the kept plan on the same buffer
    w2 = want.copy()
    oracle.run_program(w2, lower(c, n, False))
    assert np.max(np.abs(sv.to_host() - w2)) < AMP
    del graph
