"""Noise models (with_noise expansion) on the host, following the reference's
noise tests (test/unit/noise_execution.jl:1-62, test/unit/noise_channels.jl).
No device needed."""
import numpy as np
import pytest

import oracle
import mossquantum_amd as mq
from mossquantum_amd import lower
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
    model = mq.NoiseModel(rules=[mq.NoiseRule(mq.XGate, lambda g: mq.AmplitudeDamping(g.positions()[0], 1.0))])
    c = mq.QCircuit([mq.QCircuit([mq.XGate(1)])])
    prog = mq.with_noise(c, model)
    assert len(prog.circuit) == 2 and len(c) == 1
    assert isinstance(prog.circuit.ops[0], mq.XGate) and isinstance(prog.circuit.ops[1], mq.QuantumMap)
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
        mq.NoiseModel(readout=[object()])  # classical readout errors are outside this backend
    with pytest.raises(ValueError):
        mq.NoiseModel(preparation=mq.QuantumMap(1, [2 * np.eye(2)]))  # not trace preserving
    with pytest.raises(ValueError):
        mq.NoiseModel(terminal=[mq.BitFlip(1, 0.1), mq.QuantumMap(1, [np.diag([1.0, 0.5])])])
    with pytest.raises(ValueError):
        mq.NoiseRule(mq.XGate, lambda g: None, when="during")
    with pytest.raises(ValueError):
        mq.NoiseRule(mq.XGate, mq.BitFlip(1, 0.1), qubits=(1, 1))
    bad = mq.NoiseModel(rules=[mq.NoiseRule(lambda g: 1, lambda g: mq.BitFlip(1, 0.1))])
    with pytest.raises(ValueError):
        mq.with_noise(mq.QCircuit([mq.XGate(1)]), bad)
    wrong = mq.NoiseModel(rules=[mq.NoiseRule(mq.XGate, lambda g: mq.XGate(1))])
    with pytest.raises(ValueError):
        mq.with_noise(mq.QCircuit([mq.XGate(1)]), wrong)
    with pytest.raises(ValueError):
        mq.Delay(1, -1.0)


def test_rules_see_independent_copies_and_keep_order():
    """Factories and predicates receive copies (mutating them changes neither
    the circuit nor the expansion); rules apply in registration order; the
    input circuit is unchanged."""
    seen = []

    def predicate(g):
        seen.append(g)
        g.q = 99  # a mutation of the copy
        return True

    model = mq.NoiseModel(rules=[mq.NoiseRule(predicate, lambda g: [mq.PhaseFlip(1, 0.1)]),
                                 mq.NoiseRule(mq.Gate, lambda g: mq.BitFlip(1, 0.2))])
    c = mq.QCircuit([mq.XGate(1)])
    out = mq.with_noise(c, model).circuit.ops
    assert [type(o).__name__ for o in out] == ["XGate", "QuantumMap", "QuantumMap"]
    assert np.allclose(out[1].ordered_supermat(), mq.PhaseFlip(1, 0.1).ordered_supermat())
    assert np.allclose(out[2].ordered_supermat(), mq.BitFlip(1, 0.2).ordered_supermat())
    assert seen[0] is not c.ops[0] and c.ops[0].positions() == (1,)


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
