"""Noise models (with_noise expansion, readout errors) on the host, following
the reference's noise tests (test/unit/noise_execution.jl:1-62,
test/unit/noise_channels.jl).  No device needed."""
import numpy as np
import pytest

import oracle
import mossquantum_amd as mq
from mossquantum_amd import lower
from mossquantum_amd.noise import apply_readout_errors, readout_register
from mossquantum_amd.workloads import hea_circuit


def per_qubit_noise(p=0.01, gamma=0.02):
    return mq.NoiseModel(rules=[mq.NoiseRule(
        mq.Gate, lambda g: [m for q in g.positions() for m in (mq.Depolarizing(q, p), mq.AmplitudeDamping(q, gamma))])])


def test_with_noise_matches_the_noisy_ansatz_builder():
    """The benchmark's noisy ansatz (Depolarizing then AmplitudeDamping on every
    touched qubit after every gate) is exactly with_noise of the clean ansatz."""
    n, L = 5, 3
    angles = np.random.default_rng(1).uniform(0, 2 * np.pi, 2 * n * L)
    got = lower(mq.with_noise(hea_circuit(n, L, angles), per_qubit_noise()).circuit, n, True)
    want = lower(hea_circuit(n, L, angles, noise=(0.01, 0.02)), n, True)
    (ga, gp), (wa, wp) = got.arrays(), want.arrays()
    assert len(got.records) == len(want.records) == (11 * n - 5) * L
    assert bytes(ga) == bytes(wa) and np.array_equal(gp, wp)


def test_expansion_structure():
    model = mq.NoiseModel(rules=[mq.NoiseRule(mq.XGate, lambda g: mq.AmplitudeDamping(g.positions()[0], 1.0))],
                          readout=mq.ReadoutError(1, p0_to_1=1, p1_to_0=1))
    c = mq.QCircuit([mq.QCircuit([mq.XGate(1)])])
    prog = mq.with_noise(c, model)
    assert len(prog.circuit) == 2 and len(c) == 1
    assert isinstance(prog.circuit.ops[0], mq.XGate) and isinstance(prog.circuit.ops[1], mq.QuantumMap)
    assert len(prog.readout) == 1
    with pytest.raises(ValueError):
        mq.with_noise(prog, model)
    before = mq.NoiseModel(rules=[mq.NoiseRule(mq.XGate, lambda g: mq.ResetChannel(1), when="before")])
    ops = mq.with_noise(c, before).circuit.ops
    assert isinstance(ops[0], mq.QuantumMap) and isinstance(ops[1], mq.XGate)
    staged = mq.NoiseModel(preparation=mq.ResetChannel(1, state=1), terminal=mq.ResetChannel(1))
    ops = mq.with_noise(mq.QCircuit([mq.HGate(1)]), staged).circuit.ops
    assert [type(o).__name__ for o in ops] == ["QuantumMap", "HGate", "QuantumMap"]
    exact = mq.NoiseModel(rules=[mq.NoiseRule(mq.XGate, lambda g: mq.ResetChannel(1), qubits=(1,))])
    assert len(mq.with_noise(mq.QCircuit([mq.XGate(2)]), exact).circuit) == 1
    assert len(mq.with_noise(mq.QCircuit([mq.XGate(1)]), exact).circuit) == 2
    timed = mq.NoiseModel(rules=[mq.NoiseRule(mq.Delay, lambda g: mq.ThermalRelaxation(g.position, 1, 2,
                                                                                        g.duration))])
    idle = mq.with_noise(mq.QCircuit([mq.XGate(1), mq.Delay(1, 2)]), timed).circuit
    assert [type(o).__name__ for o in idle.ops] == ["XGate", "QuantumMap"]
    fixed = mq.NoiseModel(rules=[mq.NoiseRule(mq.HGate, mq.BitFlip(1, 0.1))])
    assert len(mq.with_noise(mq.QCircuit([mq.HGate(1), mq.HGate(1)]), fixed).circuit) == 4


def test_validation():
    with pytest.raises(ValueError):
        mq.ReadoutError(1, [[0.8, 0.1], [0.2, 0.8]])
    with pytest.raises(ValueError):
        mq.ReadoutError(1, [[-0.1, 1.1], [0.2, 0.8]])
    with pytest.raises(ValueError):
        mq.ReadoutError((1, 2), np.eye(4), scope="all")
    err = mq.ReadoutError(1, p0_to_1=0.2, p1_to_0=0.4)
    with pytest.raises(ValueError):
        mq.NoiseModel(readout=[err, err])
    with pytest.raises(ValueError):
        mq.NoiseModel(preparation=mq.QuantumMap(1, [2 * np.eye(2)]))
    with pytest.raises(ValueError):
        mq.NoiseRule(mq.XGate, lambda g: None, when="during")
    bad = mq.NoiseModel(rules=[mq.NoiseRule(lambda g: 1, lambda g: mq.BitFlip(1, 0.1))])
    with pytest.raises(ValueError):
        mq.with_noise(mq.QCircuit([mq.XGate(1)]), bad)
    wrong = mq.NoiseModel(rules=[mq.NoiseRule(mq.XGate, lambda g: mq.XGate(1))])
    with pytest.raises(ValueError):
        mq.with_noise(mq.QCircuit([mq.XGate(1)]), wrong)
    with pytest.raises(ValueError):
        mq.Delay(1, -1.0)


def test_readout_error_distributions():
    err = mq.ReadoutError(1, p0_to_1=0.2, p1_to_0=0.4)
    zero = np.array([1.0, 0, 0, 0])
    exp = readout_register([err], [1], 2)
    assert exp == [1]
    p1 = np.array([1.0, 0.0])  # qubit-1 marginal of |00>
    assert np.allclose(apply_readout_errors(p1, exp, [err], 1), [0.8, 0.2])
    # correlated error copying bit 2 into bit 1; reading qubit 1 alone must
    # marginalize the whole block
    m = np.zeros((4, 4))
    for actual in range(4):
        m[actual, ((actual >> 1) & 1) * 3] = 1
    corr = mq.ReadoutError((1, 2), m)
    x2 = np.array([0.0, 0, 1, 0])  # |q2=1, q1=0>
    assert readout_register([corr], [1], 2) == [1, 2]
    assert apply_readout_errors(x2, [1, 2], [corr], 1).tolist() == [0, 1]
    x2_21 = np.array([0.0, 1, 0, 0])  # the same state on the register ordered (2, 1)
    assert apply_readout_errors(x2_21, [2, 1], [corr], 2).tolist() == [0, 0, 0, 1]
    assert apply_readout_errors(x2, [1, 2], [corr], 0).tolist() == [1]
    assert np.allclose(apply_readout_errors(zero, [1, 2], [err], 2), [0.8, 0.2, 0, 0])


def test_thermal_relaxation_closed_form():
    """An excited qubit keeps population exp(-t/T1); coherence decays as
    exp(-t/T2) (oracle channel application on the supermatrix)."""
    t, T1, T2 = 2.0, 1.0, 1.5
    ch = mq.ThermalRelaxation(1, T1, T2, t)
    rho = np.array([[0, 0], [0, 1]], dtype=np.complex128).reshape(-1, order="F")
    oracle.dm_apply_channel(rho, 1, [0], ch.ordered_supermat())
    assert abs(rho[3].real - np.exp(-t / T1)) < 1e-14
    plus = np.full(4, 0.5, dtype=np.complex128)
    oracle.dm_apply_channel(plus, 1, [0], ch.ordered_supermat())
    assert abs(abs(plus[2]) - 0.5 * np.exp(-t / T2)) < 1e-14
    reset = np.array([1, 0, 0, 0], dtype=np.complex128)
    oracle.dm_apply_channel(reset, 1, [0], mq.ResetChannel(1, state=1).ordered_supermat())
    assert np.allclose(reset, [0, 0, 0, 1])
    gad = mq.GeneralizedAmplitudeDamping(1, 1.0, 0.3)
    r = np.array([1, 0, 0, 0], dtype=np.complex128)
    oracle.dm_apply_channel(r, 1, [0], gad.ordered_supermat())
    assert np.allclose([r[0].real, r[3].real], [0.7, 0.3])
