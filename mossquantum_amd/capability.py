"""The HIP backend's two execution-capability hooks (reference
src/api/execution_capability.jl:31-75): which execution modes it runs, and
what each instruction does to the state representation.  The reference's own
capability analysis walks circuits with these hooks; the declarations never
touch a state's payload.  The backend runs the `apply` and `simulate` modes
for device-resident pure and mixed states (single or batch); shot execution
(`execute`, `estimate`) stays with the reference's CPU backend.
"""
from __future__ import annotations

from dataclasses import dataclass

from .backend import MAX_CHANNEL_QUBITS, MAX_GATE_QUBITS, HIPBackend
from .channels import QuantumMap
from .circuit import Delay
from .gates import Gate
from .states import BatchDensityMatrix, BatchStateVector, DensityMatrix, HIPArray, StateVector

_STATUSES = ("supported", "unsupported", "unknown")


@dataclass
class CapabilityStep:
    """One instruction's assessment: status and every resulting state type
    (reference CapabilityStep, execution_capability.jl:8-28)."""
    status: str
    state_types: tuple = ()
    reason: str = ""

    def __post_init__(self):
        if self.status not in _STATUSES:
            raise ValueError("invalid capability status")
        if (self.status == "supported") != bool(self.state_types):
            raise ValueError("only supported steps have resulting state types")


def _promoted(state_type):
    if state_type is StateVector:
        return DensityMatrix
    if state_type is BatchStateVector:
        return BatchDensityMatrix
    return state_type


def execution_capability_backend(b, state, mode="simulate", method="density_matrix_evolution") -> CapabilityStep:
    """reference execution_capability_backend (execution_capability.jl:45-62)
    for the HIP backend."""
    if not isinstance(b, HIPBackend):
        return CapabilityStep("unknown", reason="backend has not declared execution-mode support")
    if isinstance(state, (BatchStateVector, BatchDensityMatrix)):
        resident = all(isinstance(c.storage, HIPArray) for c in state.columns)
    else:
        resident = isinstance(getattr(state, "storage", None), HIPArray)
    if not resident:
        return CapabilityStep("unsupported", reason="the HIP backend requires device-resident HIPArray storage")
    if mode not in ("apply", "simulate"):
        return CapabilityStep("unsupported", reason=f"the HIP backend does not run {mode} (shot execution)")
    if method != "density_matrix_evolution":
        return CapabilityStep("unsupported", reason="the HIP backend runs density-matrix evolution only")
    return CapabilityStep("supported", (type(state),))


def execution_capability_step(b, op, state_type, mode="simulate",
                              method="density_matrix_evolution") -> CapabilityStep:
    """reference execution_capability_step (execution_capability.jl:64-75)
    for the HIP backend: gates up to 5 qubits keep the representation,
    channels up to 2 qubits promote a pure state to a density matrix, idle
    delays change nothing."""
    if not isinstance(b, HIPBackend):
        return CapabilityStep("unknown", reason=f"backend has not declared a static transition for {type(op).__name__}")
    if isinstance(op, Delay):
        return CapabilityStep("supported", (state_type,))
    if isinstance(op, Gate):
        if op.nqubits() > MAX_GATE_QUBITS:
            return CapabilityStep("unsupported", reason=f"gates act on at most {MAX_GATE_QUBITS} qubits")
        return CapabilityStep("supported", (state_type,))
    if isinstance(op, QuantumMap):
        if op.nqubits() > MAX_CHANNEL_QUBITS:
            return CapabilityStep("unsupported", reason=f"channels act on at most {MAX_CHANNEL_QUBITS} qubits")
        return CapabilityStep("supported", (_promoted(state_type),))
    return CapabilityStep("unsupported", reason=f"{type(op).__name__} is not supported in {mode}")
