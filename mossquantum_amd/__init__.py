"""mossquantum_amd: MI355X-native dense simulation backend for MOSSQuantum.

The device library is libmossq_hip.so (HIP kernels for gfx950 behind the C ABI
of include/mossq_hip.h); this package is the host-side mirror of the
reference's backend, state, observable and gradient interfaces.
"""
from ._lib import Context, MossqError, MossqOOM, MossqUnsupported, default_context, device_count, load
from .backend import (DeviceObservable, HIPBackend, Program, SimulationResult, UnsupportedOperation, apply, apply_,
                      backend, cross_expectation, dot, expectation, probabilities, purity, simulate, simulate_circuit,
                      supports, trace, validate_positions)
from .capability import CapabilityStep, execution_capability_backend, execution_capability_step
from .channels import (AmplitudeDamping, BitFlip, BitPhaseFlip, Depolarizing, GeneralizedAmplitudeDamping,
                       MixedUnitary, PauliChannel, PhaseDamping, PhaseFlip, QuantumMap, ResetChannel,
                       ThermalRelaxation)
from .circuit import Delay, QCircuit, fuse_gates, lower
from .gates import (CNOTGate, CPHASEGate, CRxGate, CRyGate, CRzGate, CZGate, FREDKINGate, Gate, HGate, PHASEGate,
                    QuantumGate, RxGate, RyGate, RzGate, SGate, SWAPGate, TGate, TOFFOLIGate, XGate, YGate, ZGate,
                    iSWAPGate)
from .gradients import (AdjointGradient, BatchAdjointGraph, BatchExpectationObjective, ExpectationObjective,
                        ParameterMap, UnsupportedGradient, batched_adjoint, batched_adjoint_graph,
                        differentiable_value, evaluate, supports_gradient, value_and_gradient, value_and_pullback)
from .noise import NoiseModel, NoiseRule, NoisyCircuit, with_noise
from .operators import QubitsOperator, QubitsTerm, tfim_chain, xxz_chain
from .readout import counts, discrete_sample, draw_samples, measure_, sample
from .sampled_gradients import (FiniteDifference, GradientEstimate, ParameterShift, SPSAGradient, estimate_gradient,
                                parameter_shift_rule)
from .states import BatchDensityMatrix, BatchStateVector, DensityMatrix, HIPArray, StateVector

__all__ = [k for k in dir() if not k.startswith("_")]
