"""HIPBackend: the executor registered for HIPArray storage.

Mirrors the reference's backend protocol (src/api/backend_protocols.jl):
`backend(state)`, `supports(b, op, state, action)`, `apply_(op, state)` (the
reference `apply!`: mutates and returns the resulting state, which is a new
DensityMatrix when a channel meets a StateVector), `expectation(obs, state)`,
`apply(obs, state)` (independent O psi / O rho), `probabilities(state, qubits)`
and `simulate`.  Unsupported requests raise `UnsupportedOperation` before the
payload is touched; nothing falls back to the CPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check
from .channels import QuantumMap
from .circuit import Delay, QCircuit, lower
from .gates import Gate
from .operators import QubitsOperator, QubitsTerm
from .states import BatchDensityMatrix, BatchStateVector, DensityMatrix, HIPArray, StateVector

MAX_GATE_QUBITS = 5
MAX_CHANNEL_QUBITS = 2
MAX_OBS_SITES = 5


class UnsupportedOperation(Exception):
    """reference UnsupportedOperation(backend, op, state, action), backend_protocols.jl:16-36."""

    def __init__(self, backend, op, state, action):
        super().__init__(f"{type(backend).__name__} does not support {action} of "
                         f"{type(op).__name__} on {type(state).__name__}")
        self.backend, self.op, self.state, self.action = backend, op, state, action


class HIPBackend:
    """Executor for HIPArray-backed states.  fuse=True lets circuits run as
    fused LDS-tile sweeps; fuse=False runs one sweep per instruction."""

    def __init__(self, ctx: _lib.Context | None = None, fuse: bool = True):
        self.ctx = ctx or _lib.default_context()
        self.fuse = fuse

    def __repr__(self):
        return f"HIPBackend(device={self.ctx.device}, fuse={self.fuse})"


_default_backend: HIPBackend | None = None


def backend(state=None) -> HIPBackend:
    """The backend of a state's storage (only HIPArray storage is executable here)."""
    global _default_backend
    if state is not None:
        st = getattr(state, "storage", None)
        if state is not None and not isinstance(st, HIPArray) and not isinstance(
                state, (BatchStateVector, BatchDensityMatrix)):
            raise TypeError("no backend is registered for this storage")
    if _default_backend is None:
        _default_backend = HIPBackend()
    return _default_backend


# ------------------------------------------------------------------ support ----
def _obs_groups(obs):
    if isinstance(obs, QubitsTerm):
        obs = QubitsOperator([obs])
    if not isinstance(obs, QubitsOperator):
        return None
    return obs.groups()


def supports(b: HIPBackend, op, state, action: str = "apply") -> bool:
    """Whether the backend implements `action` for this operation and state
    (reference supports, src/api/backend_protocols.jl:79-95).  Like the
    reference, it describes capability only: register bounds are checked
    separately by `validate_positions`."""
    if not isinstance(b, HIPBackend):
        return False
    if action == "apply":
        if isinstance(op, Gate):
            return op.nqubits() <= MAX_GATE_QUBITS
        if isinstance(op, QuantumMap):
            return op.nqubits() <= MAX_CHANNEL_QUBITS
        if isinstance(op, Delay):
            return True
        if isinstance(op, QCircuit):
            return all(supports(b, o, state, "apply") for o in op.flat())
        if isinstance(op, (QubitsOperator, QubitsTerm)):
            g = _obs_groups(op)
            return all(len(k) <= MAX_OBS_SITES for k, _ in g)
        return False
    if action == "expectation":
        g = _obs_groups(op)
        return g is not None and all(len(k) <= MAX_OBS_SITES for k, _ in g)
    if action == "probabilities":
        return isinstance(state, (StateVector, DensityMatrix, BatchStateVector, BatchDensityMatrix))
    return False


def _positions(op):
    if isinstance(op, QCircuit):
        return [p for o in op.flat() for p in _positions(o)]
    if isinstance(op, (Gate, QuantumMap)):
        return list(op.positions())
    if isinstance(op, Delay):
        return [op.position]
    g = _obs_groups(op)
    if g is not None:
        return [p for k, _ in g for p in k]
    return []


def validate_positions(op, n: int) -> None:
    """Every qubit position of `op` is an integer in 1..n; raises ValueError
    (the reference's ArgumentError, src/quantum_circuits/instructions.jl:96-103,
    and for operators src/api/backend_protocols.jl:181-187)."""
    if n < 0:
        raise ValueError("the number of qubits must be nonnegative")
    for p in _positions(op):
        if not isinstance(p, (int, np.integer)):
            raise ValueError("qubit positions must be integers")
        if not 1 <= p <= n:
            raise ValueError(f"qubit position {p} must be between 1 and {n}")


def _require(b, op, state, action):
    """Capability first (UnsupportedOperation), then register bounds
    (ValueError), both before the payload is touched, in the reference's order
    (backend_protocols.jl:218-222)."""
    if not supports(b, op, state, action):
        raise UnsupportedOperation(b, op, state, action)
    if op is not None:
        validate_positions(op, state.n)


# ------------------------------------------------------------- observables ----
class DeviceObservable:
    """A QubitsOperator uploaded once (mossq_obs)."""

    def __init__(self, obs, n: int, ctx: _lib.Context):
        groups = _obs_groups(obs)
        if groups is None or not groups:
            raise ValueError("observable must be a non-empty QubitsOperator or QubitsTerm")
        ks = np.array([len(k) for k, _ in groups], dtype=np.int32)
        if ks.max() > MAX_OBS_SITES:
            raise ValueError(f"observable groups wider than {MAX_OBS_SITES} sites are not supported on the device")
        sites = np.array([p - 1 for k, _ in groups for p in k] or [0], dtype=np.int32)
        mats = np.concatenate([m.reshape(-1, order="F") for _, m in groups]).astype(np.complex128)
        self.ctx = ctx
        self.n = n
        h = C.c_void_p()
        check(ctx.lib.mossq_obs_create(ctx.h, n, len(groups), _lib.iptr(ks), _lib.iptr(sites),
                                       mats.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)), C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                self.ctx.lib.mossq_obs_destroy(self.h)
            self.h = None
        except Exception:
            pass


# ---------------------------------------------------------------- programs ----
class Program:
    """A lowered circuit on the device (mossq_program)."""

    def __init__(self, circuit: QCircuit | None, n: int, is_density: bool, ctx: _lib.Context, fuse: bool = False,
                 lowered=None, pool=None):
        """From a circuit, or from an existing lowering with an optional
        replacement pool (same records, e.g. rebound parameters)."""
        self.lowered = lowered if lowered is not None else lower(circuit, n, is_density)
        recs, base_pool = self.lowered.arrays()
        pool = base_pool if pool is None else np.ascontiguousarray(pool, dtype=np.complex128)
        self.ctx = ctx
        self.n = n
        self.is_density = is_density
        h = C.c_void_p()
        check(ctx.lib.mossq_program_create(ctx.h, n, int(is_density), len(self.lowered.records), recs,
                                           pool.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)),
                                           pool.size, int(fuse), C.byref(h)))
        self.h = h

    def sweeps(self) -> int:
        v = C.c_int(0)
        check(self.ctx.lib.mossq_program_sweeps(self.h, C.byref(v)))
        return v.value

    def run(self, storage: HIPArray):
        check(self.ctx.lib.mossq_program_run(self.ctx.h, self.h, storage.ptr))

    def autotune(self, reps: int = 3) -> tuple[int, float]:
        """Time every forward-plan variant on the device and keep the fastest;
        returns (variants timed, ms per run of the kept plan)."""
        nv = C.c_int(0)
        best = C.c_double(0)
        check(self.ctx.lib.mossq_program_autotune(self.ctx.h, self.h, int(reps), C.byref(nv), C.byref(best)))
        return nv.value, best.value

    def update(self, circuit: QCircuit | None = None, pool=None):
        """Rebind in place (mossq_program_update): new gate matrices for the
        same records, from a circuit of the same structure with other
        parameter values, or a pool for the current lowering.  Plans and
        device tables stay where they are, so graphs captured from this
        program (capture, BatchAdjointGraph) replay the new values.  Raises
        when the new values lower to another program or change a fused plan;
        the program is unchanged then."""
        lw = self.lowered if circuit is None else lower(circuit, self.n, self.is_density)
        recs, base_pool = lw.arrays()
        pool = base_pool if pool is None else np.ascontiguousarray(pool, dtype=np.complex128)
        check(self.ctx.lib.mossq_program_update(self.ctx.h, self.h, len(lw.records), recs,
                                                pool.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)),
                                                pool.size))
        self.lowered = lw

    def capture(self, storage: HIPArray) -> "ProgramGraph":
        """A HIP graph replaying this program on `storage` (one launch call per run)."""
        h = C.c_void_p()
        check(self.ctx.lib.mossq_program_capture(self.ctx.h, self.h, storage.ptr, C.byref(h)))
        return ProgramGraph(self, storage, h)

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                self.ctx.lib.mossq_program_destroy(self.h)
            self.h = None
        except Exception:
            pass


class ProgramGraph:
    """A captured run of a Program on one device buffer; `launch()` replays it.
    Holds the program and the buffer so both outlive the graph."""

    def __init__(self, program: Program, storage: HIPArray, h):
        self.program, self.storage, self.ctx, self.h = program, storage, program.ctx, h

    def launch(self):
        check(self.ctx.lib.mossq_graph_launch(self.ctx.h, self.h))

    def __del__(self):
        try:
            if self.h:
                self.ctx.lib.mossq_graph_destroy(self.h)
            self.h = None
        except Exception:
            pass


def update_many(programs, pools, ctx, threads: int = 0):
    """mossq_program_update_many: program i takes pools[i] for the records of
    its current lowering, all in one library call (host replanning on
    `threads` threads, 0 = one per hardware thread; one upload).  Returns each
    program's status (0: rebound; otherwise it kept its old values)."""
    m = len(programs)
    if m == 0:
        return []
    keep = [(p.lowered.arrays()[0], np.ascontiguousarray(pl, dtype=np.complex128)) for p, pl in zip(programs, pools)]
    ph = (C.c_void_p * m)(*[p.h.value for p in programs])
    nops = (C.c_int * m)(*[len(p.lowered.records) for p in programs])
    rp = (C.c_void_p * m)(*[C.addressof(r) for r, _ in keep])
    pp = (C.c_void_p * m)(*[pl.ctypes.data for _, pl in keep])
    lens = (C.c_uint64 * m)(*[pl.size for _, pl in keep])
    status = (C.c_int * m)()
    ctx.lib.mossq_program_update_many(ctx.h, m, ph, nops, rp, pp, lens, int(threads), status)
    return list(status)


def plan_emulate(circuit: QCircuit, n: int, is_density: bool, state: np.ndarray):
    """Run the fused plan of `circuit` on a host vector with the tile kernel's
    index arithmetic (no device).  state None: plan statistics only.
    Returns (new state, stats dict)."""
    lw = lower(circuit, n, is_density)
    recs, pool = lw.arrays()
    st = None if state is None else np.ascontiguousarray(np.array(state, dtype=np.complex128).reshape(-1, order="F"))
    stats = (C.c_int * 10)()
    check(_lib.load().mossq_plan_emulate(n, int(is_density), len(lw.records), recs,
                                         pool.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)), pool.size,
                                         None if st is None else st.view(np.float64).ctypes.data_as(
                                             C.POINTER(C.c_double)), stats))
    keys = ("sweeps", "unfused", "segments", "transitions", "tile_ops", "dense1", "dense2", "diag",
            "conflicted_transitions", "permutations")
    return st, dict(zip(keys, list(stats)))


def plan_emulate_adjoint(circuit: QCircuit, n: int, y: np.ndarray, lam: np.ndarray):
    """Run the fused reverse plan of the pure-state adjoint of `circuit` on host
    vectors (y = forward output, lam = seed) with the adjoint tile kernel's
    index arithmetic (no device).  Returns (y, lam, grads per lowered
    derivative slot, stats dict)."""
    lw = lower(circuit, n, False)
    recs, pool = lw.arrays()
    ys = np.ascontiguousarray(np.array(y, dtype=np.complex128))
    ls = np.ascontiguousarray(np.array(lam, dtype=np.complex128))
    grads = np.zeros(max(lw.nderiv, 1))
    stats = (C.c_int * 10)()
    dp = lambda a: a.view(np.float64).ctypes.data_as(C.POINTER(C.c_double))
    check(_lib.load().mossq_plan_emulate_adjoint(n, len(lw.records), recs, dp(pool), pool.size, dp(ys), dp(ls),
                                                 grads.ctypes.data_as(C.POINTER(C.c_double)), stats))
    keys = ("sweeps", "unfused", "segments", "transitions", "tile_ops", "dense1", "dense2", "diag",
            "conflicted_transitions", "permutations")
    return ys, ls, grads[:lw.nderiv], dict(zip(keys, list(stats)))


def plan_emulate_density_adjoint(circuit: QCircuit, n: int, rho: np.ndarray, delta: np.ndarray,
                                 whole_segments: int = -1):
    """Run the fused density-matrix adjoint of `circuit` (chunk plans cut at the
    parameterized gates, checkpoint schedule) on host vectors: rho = vec of the
    input density matrix, delta = vec of the seed at the circuit's end.
    whole_segments: 1 one reverse plan per checkpoint segment, 0 one per
    parameter, -1 as the device path chooses for this size.  rho None: build
    and count the plans only.  Returns (final rho, delta at the first
    parameterized gate, grads per lowered derivative slot, stats dict)."""
    lw = lower(circuit, n, True)
    recs, pool = lw.arrays()
    grads = np.zeros(max(lw.nderiv, 1))
    stats = (C.c_int * 6)()
    dp = lambda a: a.view(np.float64).ctypes.data_as(C.POINTER(C.c_double))
    if rho is None:  # plans and their counts only
        check(_lib.load().mossq_plan_emulate_density_adjoint(n, len(lw.records), recs, dp(pool), pool.size,
                                                             whole_segments, None, None, None, stats))
        keys = ("param_steps", "segment", "forward_sweeps", "reverse_sweeps", "unfused", "chunk_plans")
        return None, None, None, dict(zip(keys, list(stats)))
    r = np.ascontiguousarray(np.array(rho, dtype=np.complex128))
    d = np.ascontiguousarray(np.array(delta, dtype=np.complex128))
    check(_lib.load().mossq_plan_emulate_density_adjoint(n, len(lw.records), recs, dp(pool), pool.size,
                                                         whole_segments, dp(r), dp(d),
                                                         grads.ctypes.data_as(C.POINTER(C.c_double)), stats))
    keys = ("param_steps", "segment", "forward_sweeps", "reverse_sweeps", "unfused", "chunk_plans")
    return r, d, grads[:lw.nderiv], dict(zip(keys, list(stats)))


# ------------------------------------------------------------------- apply! ----
def _mat_ptr(m):
    a = np.ascontiguousarray(np.asarray(m, dtype=np.complex128).reshape(-1, order="F"))
    return a, a.view(np.float64).ctypes.data_as(C.POINTER(C.c_double))


def _pos_arr(positions):
    return np.array([p - 1 for p in positions], dtype=np.int32)


def apply_(op, state, b: HIPBackend | None = None):
    """reference apply!(op, state): mutate and return the resulting state."""
    b = b or backend(state)
    if isinstance(state, (BatchStateVector, BatchDensityMatrix)):
        return _apply_batch(op, state, b)
    _require(b, op, state, "apply")
    ctx = state.storage.ctx
    if isinstance(op, Gate):
        pos = _pos_arr(op.ordered_positions())
        keep, mp = _mat_ptr(op.ordered_mat())
        fn = ctx.lib.mossq_sv_apply_gate if isinstance(state, StateVector) else ctx.lib.mossq_dm_apply_gate
        check(fn(ctx.h, state.storage.ptr, state.n, len(pos), _lib.iptr(pos), mp))
        return state
    if isinstance(op, QuantumMap):
        if isinstance(state, StateVector):
            state = DensityMatrix.from_state(state)
        pos = _pos_arr(op.ordered_positions())
        keep, mp = _mat_ptr(op.ordered_supermat())
        check(ctx.lib.mossq_dm_apply_channel(ctx.h, state.storage.ptr, state.n, len(pos), _lib.iptr(pos), mp, 0))
        return state
    if isinstance(op, Delay):
        return state  # idle: the state is unchanged (noise_models.jl:44)
    if isinstance(op, QCircuit):
        flat = list(op.flat())
        if isinstance(state, StateVector) and any(isinstance(o, QuantumMap) for o in flat):
            first = next(i for i, o in enumerate(flat) if isinstance(o, QuantumMap))
            if first:
                Program(QCircuit(flat[:first]), state.n, False, ctx, b.fuse).run(state.storage)
            state = DensityMatrix.from_state(state)
            flat = flat[first:]
        if flat:
            Program(QCircuit(flat), state.n, isinstance(state, DensityMatrix), ctx, b.fuse).run(state.storage)
        return state
    raise UnsupportedOperation(b, op, state, "apply")


def _pow2_chunks(B):
    """(first column, log2 count) of the power-of-two column runs of B columns."""
    out, c0 = [], 0
    for j in reversed(range(B.bit_length())):
        if (B >> j) & 1:
            out.append((c0, j))
            c0 += 1 << j
    return out


def _apply_batch(op, batch, b):
    """apply! on a batch (reference backends/cpu/gates/batch.jl:22-82): on a
    (2^n, B) state-vector batch a gate or a channel-free circuit runs over each
    power-of-two run of 2^j columns as one vector of n + j qubits (columns are
    the high index bits, which no instruction touches), so every instruction is
    one launch per run; density-matrix batches and channels go column by column
    (a channel first promotes the whole batch)."""
    _require(b, op, batch.columns[0], "apply")
    flat = list(op.flat()) if isinstance(op, QCircuit) else [op]
    if isinstance(batch, BatchStateVector) and any(isinstance(o, QuantumMap) for o in flat):
        first = next(i for i, o in enumerate(flat) if isinstance(o, QuantumMap))
        if first:
            _apply_batch(QCircuit(flat[:first]), batch, b)
        dm = BatchDensityMatrix.from_states(batch)
        return _apply_batch(QCircuit(flat[first:]), dm, b)
    if isinstance(batch, BatchDensityMatrix) or not flat or all(isinstance(o, Delay) for o in flat):
        for c in batch.columns:
            apply_(op, c, b)
        return batch
    ctx, n, L = batch.storage.ctx, batch.n, 1 << batch.n
    for c0, j in _pow2_chunks(batch.B):
        run = StateVector(batch.storage.view(c0 * L, L << j), n + j)
        if isinstance(op, Gate):
            apply_(op, run, b)
        else:
            Program(QCircuit(flat), n + j, False, ctx, b.fuse).run(run.storage)
    return batch


class SimulationResult:
    """Final state and measurement records of one execution (reference
    SimulationResult, src/execution/simulation.jl:15-25).  Circuits here carry
    no mid-circuit measurements, so `measurements` stays empty."""

    def __init__(self, state, measurements=None):
        self.state = state
        self.measurements = list(measurements or [])


def simulate_circuit(b: HIPBackend, circuit: QCircuit, state, rng=None, records=None):
    """Backend bulk-execution hook (reference simulate_circuit!,
    src/execution/simulation.jl:166-175): the whole record-free circuit runs as
    one device program (fused LDS-tile sweeps when the backend fuses) on
    `state`, which it may replace (promotion at the first channel)."""
    return apply_(circuit, state, b)


def simulate(circuit: QCircuit, state, b: HIPBackend | None = None, rng=None, noise_model=None):
    """Run a circuit on a copy of `state` and return a SimulationResult
    (reference simulate, src/execution/simulation.jl:177-199); a noise model
    is expanded first (with_noise)."""
    from .noise import NoisyCircuit, with_noise
    if isinstance(circuit, NoisyCircuit):
        if noise_model is not None:
            raise ValueError("noise is already expanded")
        circuit = circuit.circuit
    elif noise_model is not None:
        circuit = with_noise(circuit, noise_model).circuit
    b = b or backend(state)
    if isinstance(state, (BatchStateVector, BatchDensityMatrix)):
        return SimulationResult(apply_(circuit, state.copy(), b))
    _require(b, circuit, state, "apply")
    records: list = []
    return SimulationResult(simulate_circuit(b, circuit, state.copy(), rng, records), records)


# -------------------------------------------------------------- expectation ----
def expectation(obs, state, b: HIPBackend | None = None):
    """<psi|O|psi> or tr(O rho) as a complex number; a vector for batches."""
    b = b or backend(state)
    if isinstance(state, (BatchStateVector, BatchDensityMatrix)):
        # one launch over every column (reference observables/batch_expectation.jl)
        _require(b, obs, state.columns[0], "expectation")
        ctx = state.storage.ctx
        dob = DeviceObservable(obs, state.n, ctx)
        out = np.zeros(2 * state.B)
        fn = (ctx.lib.mossq_sv_expectation_batch if isinstance(state, BatchStateVector)
              else ctx.lib.mossq_dm_expectation_batch)
        check(fn(ctx.h, dob.h, state.storage.ptr, state.B, _lib.dptr(out)))
        return out[0::2] + 1j * out[1::2]
    _require(b, obs, state, "expectation")
    ctx = state.storage.ctx
    dob = DeviceObservable(obs, state.n, ctx)
    out = np.zeros(2)
    fn = ctx.lib.mossq_sv_expectation if isinstance(state, StateVector) else ctx.lib.mossq_dm_expectation
    check(fn(ctx.h, dob.h, state.storage.ptr, _lib.dptr(out)))
    return complex(out[0], out[1])


def cross_expectation(phi: StateVector, obs, psi: StateVector):
    """<phi|O|psi> (reference expectation(state_c, m, state), observables/expectation.jl:48-66)."""
    ctx = psi.storage.ctx
    dob = DeviceObservable(obs, psi.n, ctx)
    out = np.zeros(2)
    check(ctx.lib.mossq_sv_cross_expectation(ctx.h, dob.h, phi.storage.ptr, psi.storage.ptr, _lib.dptr(out)))
    return complex(out[0], out[1])


def apply(obs, state, b: HIPBackend | None = None):
    """Independent O psi (or O rho, left action) (reference MOSSQuantum.apply)."""
    b = b or backend(state)
    _require(b, obs, state, "apply")
    ctx = state.storage.ctx
    dob = DeviceObservable(obs, state.n, ctx)
    out = state.storage.similar()
    if isinstance(state, StateVector):
        check(ctx.lib.mossq_sv_apply_operator(ctx.h, dob.h, state.storage.ptr, out.ptr))
        return StateVector(out, state.n)
    check(ctx.lib.mossq_dm_apply_operator(ctx.h, dob.h, state.storage.ptr, out.ptr))
    return DensityMatrix(out, state.n)


def trace(rho: DensityMatrix) -> complex:
    ctx = rho.storage.ctx
    out = np.zeros(2)
    check(ctx.lib.mossq_dm_trace(ctx.h, rho.storage.ptr, rho.n, _lib.dptr(out)))
    return complex(out[0], out[1])


def dot(a, b) -> complex:
    """sum conj(a) b over the storage (purity = real(dot(rho, rho)))."""
    ctx = a.storage.ctx
    out = np.zeros(2)
    check(ctx.lib.mossq_dot(ctx.h, a.storage.ptr, b.storage.ptr, a.storage.size, _lib.dptr(out)))
    return complex(out[0], out[1])


def purity(rho: DensityMatrix) -> float:
    return dot(rho, rho).real


# ------------------------------------------------------------ probabilities ----
def _probability_tolerance():
    return float(np.sqrt(np.finfo(np.float64).eps))


def _normalize(values, tol):
    if not np.all(np.isfinite(values)):
        raise ValueError("probabilities must be finite")
    if values.min() < -tol:
        raise ValueError("probabilities must be nonnegative")
    values = np.maximum(values, 0)
    total = values.sum()
    # Julia isapprox(total, 1; atol, rtol): |total - 1| <= max(atol, rtol * max(|total|, 1))
    if not abs(total - 1) <= max(tol, tol * max(abs(total), 1.0)):
        raise ValueError(f"state probabilities must sum to one; normalize the state first (sum = {total})")
    return values / total


def _device_p0(state, position):
    """Unnormalised probability that qubit `position` reads 0 (the device
    marginal; reference compute_probability_zero, execution/measurement.jl:156-192)."""
    ctx = state.storage.ctx
    q = np.array([int(position) - 1], dtype=np.int32)
    out = np.zeros(2)
    check(ctx.lib.mossq_probabilities(ctx.h, state.storage.ptr, state.n, int(isinstance(state, DensityMatrix)), 1,
                                      _lib.iptr(q), _lib.dptr(out)))
    return float(out[0])


def probabilities(state, qubits=None, b: HIPBackend | None = None):
    """Readout marginals, first requested qubit = least significant output bit
    (reference probabilities, src/execution/probabilities.jl:42-87).  The device
    sums the marginal; validation and normalisation follow the reference on the host."""
    b = b or backend(state)
    if isinstance(state, (BatchStateVector, BatchDensityMatrix)):
        return np.stack([probabilities(c, qubits, b) for c in state.columns], axis=1)
    _require(b, None, state, "probabilities")
    qubits = list(range(1, state.n + 1)) if qubits is None else [int(q) for q in qubits]
    if len(set(qubits)) != len(qubits) or not all(1 <= q <= state.n for q in qubits):
        raise ValueError("readout qubits must be distinct and in range")
    ctx = state.storage.ctx
    tol = _probability_tolerance()
    is_dm = isinstance(state, DensityMatrix)
    if is_dm:
        # every diagonal entry is checked before any marginal sums them
        # (reference probabilities.jl:55-64)
        bounds = np.zeros(3)
        check(ctx.lib.mossq_dm_diag_bounds(ctx.h, state.storage.ptr, state.n, _lib.dptr(bounds)))
        if bounds[2] == 0:
            raise ValueError("density-matrix diagonal must be finite")
        if bounds[0] > tol:
            raise ValueError("density-matrix diagonal must be real")
        if bounds[1] < -tol:
            raise ValueError("density-matrix probabilities must be nonnegative")
    # an empty readout register is the total probability as one entry
    q = np.array([x - 1 for x in (qubits or [1])], dtype=np.int32)
    out = np.zeros(1 << len(q))
    check(ctx.lib.mossq_probabilities(ctx.h, state.storage.ptr, state.n, int(is_dm), len(q), _lib.iptr(q),
                                      _lib.dptr(out)))
    if not qubits:
        out = np.array([out[0] + out[1]])
    return _normalize(out, tol)
