"""Static execution-capability analysis: the HIP backend's declarations and
the walk that assembles them (reference src/api/execution_capability.jl).

An analysis never touches a state's payload, never expands noise factories
and never draws random numbers.  The HIP backend supports the `apply` and
`simulate` modes for device-resident pure and mixed states (single or batch);
shot execution (`execute`, `estimate`) stays with the reference's CPU
backend.
"""
from __future__ import annotations

from dataclasses import dataclass, field

from .backend import MAX_CHANNEL_QUBITS, MAX_GATE_QUBITS, HIPBackend
from .channels import QuantumMap
from .circuit import Delay, QCircuit
from .gates import Gate
from .states import BatchDensityMatrix, BatchStateVector, DensityMatrix, HIPArray, StateVector

_MODES = ("apply", "simulate", "execute", "estimate")
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


@dataclass
class CapabilityIssue:
    status: str
    path: tuple
    reason: str


@dataclass
class ExecutionCapability:
    status: str
    issues: list = field(default_factory=list)
    final_state_types: list = field(default_factory=list)


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


def _walk(issues, b, circuit, types, mode, method, n, path):
    for index, op in enumerate(circuit, start=1):
        itempath = path + (index,)
        if isinstance(op, QCircuit):
            types = _walk(issues, b, op, types, mode, method, n, itempath)
        else:
            if max(op.positions()) > n:
                issues.append(CapabilityIssue("unsupported", itempath, f"position outside a {n}-qubit register"))
                return []
            nxt = []
            for T in types:
                step = execution_capability_step(b, op, T, mode, method)
                if step.status == "supported":
                    nxt.extend(t for t in step.state_types if t not in nxt)
                else:
                    issues.append(CapabilityIssue(step.status, itempath, step.reason or "instruction capability not declared"))
            types = nxt
        if not types:
            return types
    return types


def execution_capability(b, circuit, state, mode="simulate", method="density_matrix_evolution",
                         noise_model=None) -> ExecutionCapability:
    """Static assessment of running `circuit` on `state` (reference
    execution_capability, execution_capability.jl:123-282, for the modes and
    options this backend has).  Noise models are not expanded here and make
    the report unknown, as in the reference."""
    from .noise import NoisyCircuit
    if mode not in _MODES:
        raise ValueError("invalid execution capability mode")
    if isinstance(circuit, NoisyCircuit):
        if mode == "apply":
            return ExecutionCapability("unsupported", [CapabilityIssue("unsupported", (),
                                                                       "NoisyCircuit has no apply entry point")])
        circuit = circuit.circuit
    issues: list = []
    modecap = execution_capability_backend(b, state, mode, method)
    if modecap.status != "supported":
        issues.append(CapabilityIssue(modecap.status, (), modecap.reason))
    if noise_model is not None:
        issues.append(CapabilityIssue("unknown", (), "noise models are not expanded during capability queries"))
        status = "unsupported" if any(i.status == "unsupported" for i in issues) else "unknown"
        return ExecutionCapability(status, issues, [])
    initial = list(modecap.state_types) if modecap.status == "supported" else [type(state)]
    types = _walk(issues, b, circuit, initial, mode, method, state.n, ())
    status = ("unsupported" if any(i.status == "unsupported" for i in issues)
              else "unknown" if any(i.status == "unknown" for i in issues) else "supported")
    return ExecutionCapability(status, issues, types)
