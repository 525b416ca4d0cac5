"""The HIP backend's capability hooks (reference
src/api/execution_capability.jl:31-75).  No device needed: the declarations
only inspect types and instruction shapes."""
import numpy as np
import pytest

import mossquantum_amd as mq


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


def test_backend_modes():
    ctx_free = mq.HIPBackend.__new__(mq.HIPBackend)
    host_state = type("HostState", (), {"storage": np.zeros(4)})()
    assert mq.execution_capability_backend(ctx_free, host_state).status == "unsupported"
    assert mq.execution_capability_backend(object(), host_state).status == "unknown"
