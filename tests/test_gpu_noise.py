"""Noise models executed on the device: with_noise programs through the
fused LDS-tile planner, idle relaxation (reference
test/unit/noise_execution.jl:1-62)."""
import numpy as np
import pytest

import oracle
import mossquantum_amd as mq
from mossquantum_amd import lower
from mossquantum_amd.workloads import hea_circuit, tfim

pytestmark = pytest.mark.gpu
AMP, EXP, GRAD = 1e-12, 1e-10, 1e-9


def test_noisy_program_matches_oracle(ctx):
    """A with_noise'd ansatz with mixed rules (before/after, per-qubit register
    filter, preparation and terminal maps) on a 7-qubit density matrix, fused
    and unfused, against the oracle."""
    n, L = 7, 2
    angles = np.random.default_rng(3).uniform(0, 2 * np.pi, 2 * n * L)
    model = mq.NoiseModel(
        rules=[mq.NoiseRule(mq.Gate, lambda g: [mq.Depolarizing(q, 0.01) for q in g.positions()]),
               mq.NoiseRule(mq.CNOTGate, lambda g: mq.PhaseDamping(g.positions()[1], 0.05), when="before"),
               mq.NoiseRule(mq.RyGate, lambda g: mq.ThermalRelaxation(3, 50.0, 70.0, 1.0), qubits=(3,))],
        preparation=[mq.ResetChannel(2, state=1, p=0.5)], terminal=[mq.BitFlip(q, 0.02) for q in (1, n)])
    prog = mq.with_noise(hea_circuit(n, L, angles), model).circuit
    want = np.zeros(1 << (2 * n), dtype=np.complex128)
    want[0] = 1
    oracle.run_program(want, lower(prog, n, True))
    for fuse in (True, False):
        dm = mq.apply_(prog, mq.DensityMatrix.zeros(n), mq.HIPBackend(ctx, fuse=fuse))
        assert np.max(np.abs(dm.to_host().reshape(-1, order="F") - want)) < AMP, fuse


def test_noisy_template_gradient(ctx):
    """Gradient through a with_noise'd template (the noise rules are instruction
    types with fixed maps, so rebinding keeps the expansion)."""
    n = 6
    model = mq.NoiseModel(rules=[mq.NoiseRule(mq.Gate, lambda g: [m for q in g.positions()
                                                                   for m in (mq.Depolarizing(q, 0.01),
                                                                             mq.AmplitudeDamping(q, 0.02))])])
    tpl = mq.with_noise(hea_circuit(n, 2), model).circuit
    th = np.random.default_rng(6).uniform(0, 2 * np.pi, tpl.nparameters())
    lw = lower(tpl.bind(th), n, True)
    want_v, want_g = oracle.adjoint(lw, tfim(n).groups())
    obj = mq.ExpectationObjective(tpl, tfim(n), mq.DensityMatrix.zeros(n))
    v, g = mq.value_and_gradient(mq.AdjointGradient(), obj, th)
    assert abs(v - want_v) < EXP
    assert np.max(np.abs(g - lw.slot_gradients(want_g))) < GRAD


def test_idle_relaxation(ctx):
    timed = mq.NoiseModel(rules=[mq.NoiseRule(mq.Delay, lambda g: mq.ThermalRelaxation(g.position, 1, 2,
                                                                                        g.duration))])
    idle = mq.with_noise(mq.QCircuit([mq.XGate(1), mq.Delay(1, 2)]), timed)
    st = mq.simulate(idle.circuit, mq.StateVector.zeros(1)).state
    assert abs(mq.probabilities(st)[1] - np.exp(-2)) < 1e-14
    # a bare Delay leaves the state unchanged
    st = mq.simulate(mq.QCircuit([mq.Delay(1, 2)]), mq.StateVector.zeros(1)).state
    assert mq.probabilities(st).tolist() == [1, 0]


def test_execution_capability_and_simulate(ctx):
    b = mq.HIPBackend(ctx)
    sv = mq.StateVector.zeros(3, ctx)
    c = mq.QCircuit([mq.HGate(1), mq.CNOTGate(1, 2)])
    step = mq.execution_capability_backend(b, sv, "simulate")
    assert step.status == "supported" and step.state_types == (mq.StateVector,)
    assert mq.execution_capability_backend(b, sv, "execute").status == "unsupported"
    model = mq.NoiseModel(rules=[mq.NoiseRule(mq.HGate, mq.BitFlip(1, 0.1))])
    res = mq.simulate(c, sv, b, noise_model=model)
    assert isinstance(res, mq.SimulationResult) and isinstance(res.state, mq.DensityMatrix)
    assert res.measurements == []
    want = mq.apply_(mq.with_noise(c, model).circuit, mq.StateVector.zeros(3, ctx), b)
    assert np.max(np.abs(res.state.to_host() - want.to_host())) < 1e-15
    assert np.array_equal(sv.to_host(), mq.StateVector.zeros(3, ctx).to_host())  # input untouched
