"""Static capability declarations of the HIP backend (reference
src/api/execution_capability.jl).  No device needed: the analysis only
inspects types and instruction shapes."""
import numpy as np
import pytest

import mossquantum_amd as mq
from mossquantum_amd.capability import _walk


def test_step_declarations():
    b = mq.HIPBackend.__new__(mq.HIPBackend)  # declarations need no context
    st = mq.execution_capability_step(b, mq.HGate(1), mq.StateVector)
    assert st.status == "supported" and st.state_types == (mq.StateVector,)
    st = mq.execution_capability_step(b, mq.AmplitudeDamping(1, 0.1), mq.StateVector)
    assert st.state_types == (mq.DensityMatrix,)
    st = mq.execution_capability_step(b, mq.AmplitudeDamping(1, 0.1), mq.BatchStateVector)
    assert st.state_types == (mq.BatchDensityMatrix,)
    assert mq.execution_capability_step(b, mq.Delay(1, 1.0), mq.DensityMatrix).state_types == (mq.DensityMatrix,)
    big = mq.QuantumGate(tuple(range(1, 7)), np.eye(64))
    assert mq.execution_capability_step(b, big, mq.StateVector).status == "unsupported"
    wide = mq.Depolarizing((1, 2), 0.1)
    assert mq.execution_capability_step(b, wide, mq.DensityMatrix).status == "supported"
    assert mq.execution_capability_step(object(), mq.HGate(1), mq.StateVector).status == "unknown"
    with pytest.raises(ValueError):
        mq.CapabilityStep("supported")
    with pytest.raises(ValueError):
        mq.CapabilityStep("maybe", (mq.StateVector,))


def test_walk_paths_and_promotion():
    b = mq.HIPBackend.__new__(mq.HIPBackend)
    big = mq.QuantumGate(tuple(range(1, 7)), np.eye(64))
    c = mq.QCircuit([mq.HGate(1), mq.QCircuit([mq.CNOTGate(1, 2), mq.BitFlip(2, 0.1)]), mq.RyGate(3, 0.2)])
    issues = []
    assert _walk(issues, b, c, [mq.StateVector], "simulate", "density_matrix_evolution", 3, ()) == [mq.DensityMatrix]
    assert issues == []
    c2 = mq.QCircuit([mq.HGate(1), mq.QCircuit([mq.XGate(1), big])])
    issues = []
    assert _walk(issues, b, c2, [mq.StateVector], "simulate", "density_matrix_evolution", 6, ()) == []
    assert issues[0].status == "unsupported" and issues[0].path == (2, 2)
    issues = []
    assert _walk(issues, b, mq.QCircuit([mq.HGate(4)]), [mq.StateVector], "apply", "density_matrix_evolution", 3,
                 ()) == []
    assert issues[0].path == (1,)
