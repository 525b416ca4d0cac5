"""Gradient protocol and adjoint objectives on the HIP backend.

Mirrors reference src/api/gradient_methods.jl:14-138 (AdjointGradient,
UnsupportedGradient, evaluate, supports_gradient, value_and_gradient,
value_and_pullback, differentiable_value), src/api/objectives.jl:15-249
(ExpectationObjective, BatchExpectationObjective) and the AdjointGradient
drivers of src/differentiation/objectives.jl:183-295.  The reverse pass runs in
libmossq_hip (mossq_adjoint / mossq_batch_adjoint).

`differentiable_value` is the hook the reference's ChainRulesCore `rrule`
wraps (ext/MOSSQuantumChainRulesCoreExt.jl:7-28); here `TorchDifferentiableValue`
plays that role for torch.autograd.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import check
from .backend import DeviceObservable, HIPBackend, Program, backend
from .circuit import QCircuit, lower
from .states import BatchDensityMatrix, BatchStateVector, DensityMatrix, StateVector


class AbstractGradientMethod:
    pass


class AdjointGradient(AbstractGradientMethod):
    """Adjoint differentiation (reference gradient_methods.jl:14-18)."""


class UnsupportedGradient(Exception):
    def __init__(self, method, objective):
        super().__init__(f"{type(method).__name__} does not support parameter gradients for "
                         f"{type(objective).__name__}")


class ParameterMap:
    """Slot -> independent-parameter identity, 1-based (reference parameter_map.jl:24-116)."""

    def __init__(self, indices):
        idx = [int(i) for i in indices]
        count = max(idx, default=0)
        if any(i < 1 or i > len(idx) for i in idx) or len(set(idx)) != count:
            raise ValueError("parameter indices must use every integer in 1:count")
        self.indices = tuple(idx)
        self.count = count

    @classmethod
    def identity(cls, nslots):
        return cls(range(1, nslots + 1))

    def __len__(self):
        return len(self.indices)

    def expand(self, values):
        values = np.asarray(values, dtype=np.float64)
        if values.size != self.count:
            raise ValueError(f"expected {self.count} independent parameters, received {values.size}")
        if not np.all(np.isfinite(values)):
            raise ValueError("parameter values must be finite")
        return values[np.asarray(self.indices, dtype=np.int64) - 1]

    def accumulate(self, slot_grads):
        out = np.zeros(self.count)
        np.add.at(out, np.asarray(self.indices, dtype=np.int64) - 1, slot_grads)
        return out


class ExpectationObjective:
    """real(<observable>) after running the bound template on a copy of initial_state."""

    def __init__(self, template: QCircuit, observable, initial_state, parameter_map=None, backend_=None):
        self.template = template
        self.observable = observable
        self.initial_state = initial_state.copy()
        self.parameter_map = parameter_map or ParameterMap.identity(template.nparameters())
        if len(self.parameter_map) != template.nparameters():
            raise ValueError("parameter map does not match the template's active slots")
        self.backend = backend_ or backend(initial_state)


class BatchExpectationObjective:
    """Shared circuit/parameters/observable over a batch; reduction :mean/:sum/:weighted_sum."""

    def __init__(self, template, observable, initial_state, parameter_map=None, backend_=None, reduction="mean",
                 weights=None):
        if not isinstance(initial_state, (BatchStateVector, BatchDensityMatrix)):
            raise TypeError("a batch objective needs a batch state")
        if reduction not in ("mean", "sum", "weighted_sum"):
            raise ValueError("batch reduction must be mean, sum or weighted_sum")
        if reduction == "weighted_sum":
            if weights is None or len(weights) != initial_state.batch_size():
                raise ValueError("weighted_sum requires one weight per sample")
        elif weights is not None:
            raise ValueError("weights are only valid with weighted_sum")
        self.template = template
        self.observable = observable
        self.initial_state = initial_state.copy()
        self.parameter_map = parameter_map or ParameterMap.identity(template.nparameters())
        self.backend = backend_ or backend(initial_state)
        self.reduction = reduction
        self.weights = None if weights is None else np.asarray(weights, dtype=np.float64)

    def sample_weights(self):
        B = self.initial_state.batch_size()
        if self.reduction == "mean":
            return np.full(B, 1.0 / B)
        if self.reduction == "sum":
            return np.ones(B)
        return self.weights


def _objective_program(objective, flat, state, fuse):
    """The device program of the objective's template bound to slot values
    `flat`: the template is lowered once per (qubit count, representation)
    and later bindings only rewrite the parameterized gates' matrices.  The
    program itself is kept too and rebound in place (Program.update: its
    plans stay) when the values change; a rebinding the library refuses (new
    values that change a plan) builds a new program."""
    is_dm = isinstance(state, DensityMatrix) or objective.template.has_channel()
    ctx = state.storage.ctx
    cache = objective.__dict__.setdefault("_lowered", {})
    key = (state.n, is_dm)
    lw = cache.get(key)
    if lw is None:
        lw = lower(objective.template.bind(flat), state.n, is_dm)
        cache[key] = lw
        pool = None
    else:
        pool = lw.pool_with_slots(flat)
    progs = objective.__dict__.setdefault("_programs", {})
    pkey = (state.n, is_dm, bool(fuse))
    ent = progs.get(pkey)
    if ent is not None and ent[0].ctx is ctx and ent[0].lowered is lw:
        prog, bound = ent
        if np.array_equal(bound, flat):
            return prog
        try:
            prog.update(pool=pool if pool is not None else lw.arrays()[1])
            progs[pkey] = (prog, flat.copy())
            return prog
        except _lib.MossqError:
            pass
    prog = Program(None, state.n, is_dm, ctx, fuse=fuse, lowered=lw, pool=pool)
    progs[pkey] = (prog, flat.copy())
    return prog


def _objective_observable(objective, n, ctx):
    """The objective's observable on the device, created once per (qubit
    count, context)."""
    obs = objective.__dict__.setdefault("_device_obs", {})
    ent = obs.get(n)
    if ent is None or ent.ctx is not ctx:
        ent = obs[n] = DeviceObservable(objective.observable, n, ctx)
    return ent


def _adjoint_one(b: HIPBackend, prog: Program, observable, state, dob=None):
    """(value, slot gradients) of real(<obs>) for one program on one state
    (dob: the observable already on the device)."""
    ctx = state.storage.ctx
    dob = dob or DeviceObservable(observable, state.n, ctx)
    nd = prog.lowered.nderiv
    grads = np.zeros(max(nd, 1))
    value = C.c_double(0)
    check(ctx.lib.mossq_adjoint(ctx.h, prog.h, dob.h, state.storage.ptr, int(isinstance(state, DensityMatrix)),
                                None, C.byref(value), _lib.dptr(grads)))
    return value.value, prog.lowered.slot_gradients(grads[:nd])


def _adjoint_columns(prog: Program, observable, batch: BatchStateVector, dob=None):
    """(values[B], slot gradients per column) of real(<obs>) for one program
    on every column of a state-vector batch (mossq_batch_adjoint_init)."""
    ctx = batch.storage.ctx
    dob = dob or DeviceObservable(observable, batch.n, ctx)
    B, nd = batch.B, prog.lowered.nderiv
    offs = np.arange(B, dtype=np.int64) * nd
    vals = np.zeros(B)
    grads = np.zeros(max(B * nd, 1))
    ph = (C.c_void_p * B)(*([prog.h.value] * B))
    oh = (C.c_void_p * B)(*([dob.h.value] * B))
    ih = (C.c_void_p * B)(*[c.storage.ptr.value for c in batch.columns])
    check(ctx.lib.mossq_batch_adjoint_init(ctx.h, B, ph, oh, ih, offs.ctypes.data_as(C.POINTER(C.c_int64)),
                                           _lib.dptr(vals), _lib.dptr(grads)))
    return vals, [prog.lowered.slot_gradients(grads[o:o + nd]) for o in offs]


def _sampled(method):
    from .sampled_gradients import FiniteDifference, ParameterShift, SPSAGradient
    return isinstance(method, (ParameterShift, FiniteDifference, SPSAGradient))


def supports_gradient(method, objective, values=None) -> bool:
    if _sampled(method):
        from .sampled_gradients import supports_sampled_gradient
        return supports_sampled_gradient(method, objective, values)
    if not isinstance(method, AdjointGradient):
        return False
    if not isinstance(objective, (ExpectationObjective, BatchExpectationObjective)):
        return False
    return isinstance(objective.backend, HIPBackend)


def evaluate(objective, values):
    """Forward execution only (reference evaluate, api/objectives.jl:38-42)."""
    from .backend import apply_, expectation
    if callable(objective) and not isinstance(objective, (ExpectationObjective, BatchExpectationObjective)):
        return objective(values)
    bound = objective.template.bind(objective.parameter_map.expand(values))
    if isinstance(objective, BatchExpectationObjective):
        vals = np.array([expectation(objective.observable, apply_(bound, c.copy(), objective.backend)).real
                         for c in objective.initial_state.columns])
        return float(np.dot(objective.sample_weights(), vals))
    state = apply_(bound, objective.initial_state.copy(), objective.backend)
    return float(expectation(objective.observable, state, objective.backend).real)


def value_and_gradient(method, objective, values, rng=None):
    """(value, gradient in independent-parameter order).  AdjointGradient runs
    the device reverse pass; the sampled methods (ParameterShift,
    FiniteDifference, SPSAGradient) evaluate their shifted circuits together
    (rng drives SPSA's directions)."""
    if _sampled(method):
        from .sampled_gradients import estimate_gradient
        est = estimate_gradient(method, objective, values, rng)
        return est.value, est.gradient
    if not supports_gradient(method, objective):
        raise UnsupportedGradient(method, objective)
    pm = objective.parameter_map
    flat = pm.expand(values)
    b = objective.backend
    if isinstance(objective, BatchExpectationObjective):
        w = objective.sample_weights()
        init = objective.initial_state
        if isinstance(init, BatchStateVector) and not objective.template.has_channel():
            # every column in one batched engine call, each from its own state
            # (reference differentiation/batch.jl:35-66 reverses column by column);
            # programs the engine does not take (gates on three or more qubits
            # below the tiled size) go column by column below
            try:
                vals, slot_grads = _adjoint_columns(_objective_program(objective, flat, init.columns[0], b.fuse),
                                                    objective.observable, init,
                                                    _objective_observable(objective, init.n, init.storage.ctx))
                grad = sum(wi * pm.accumulate(g) for wi, g in zip(w, slot_grads))
                return float(np.dot(w, vals)), grad
            except _lib.MossqUnsupported:
                pass
        total, grad = 0.0, np.zeros(pm.count)
        for wi, col in zip(w, init.columns):
            v, g = _adjoint_one(b, _objective_program(objective, flat, col, b.fuse), objective.observable, col,
                                _objective_observable(objective, col.n, col.storage.ctx))
            total += wi * v
            grad += wi * pm.accumulate(g)
        return total, grad
    st = objective.initial_state
    v, g = _adjoint_one(b, _objective_program(objective, flat, st, b.fuse), objective.observable, st,
                        _objective_observable(objective, st.n, st.storage.ctx))
    return v, pm.accumulate(g)


def _check_gradient_input(values) -> np.ndarray:
    """Real, finite parameter values (reference _check_gradient_input,
    api/gradient_methods.jl:64-68)."""
    v = np.asarray(values)
    if v.ndim != 1 or not (np.issubdtype(v.dtype, np.floating) or np.issubdtype(v.dtype, np.integer)):
        raise TypeError("parameter values must be a real vector")
    v = v.astype(np.float64)
    if not np.all(np.isfinite(v)):
        raise ValueError("parameter values must be finite.")
    return v


def _check_objective_value(value) -> float:
    """A real scalar (reference _check_objective_value, gradient_methods.jl:70-73)."""
    if isinstance(value, (complex, np.complexfloating)) or not np.isscalar(value):
        raise ValueError("a differentiable objective must return a real scalar.")
    return float(value)


def _check_parameter_gradient(gradient, values) -> np.ndarray:
    """A real vector with one entry per parameter (reference
    _check_parameter_gradient, gradient_methods.jl:75-82)."""
    g = np.asarray(gradient)
    if g.ndim != 1 or np.iscomplexobj(g):
        raise ValueError("a parameter gradient must be a real vector.")
    if g.size != np.asarray(values).size:
        raise ValueError(f"expected {np.asarray(values).size} parameter gradients, received {g.size}.")
    return g.astype(np.float64)


def value_and_pullback(method, objective, values, rng=None):
    """(value, back) with back(seed) = seed * gradient (reference
    value_and_pullback, api/gradient_methods.jl:100-115): support is checked
    first, the gradient is computed once, and back returns a new array on
    every call, so repeated calls are independent and consume nothing."""
    if not supports_gradient(method, objective, values):
        raise UnsupportedGradient(method, objective)
    x = _check_gradient_input(values)
    v, g = value_and_gradient(method, objective, x, rng)
    v = _check_objective_value(v)
    saved = _check_parameter_gradient(g, x).copy()

    def back(seed):
        if isinstance(seed, (complex, np.complexfloating)) or not np.isscalar(seed):
            raise TypeError("the pullback seed must be a real scalar")
        return float(seed) * saved

    return v, back


def differentiable_value(method, objective, values):
    """The objective's value, forward execution only (reference
    differentiable_value, api/gradient_methods.jl:117-138); AD adapters
    attach the reverse rule (torch_differentiable_value here)."""
    x = _check_gradient_input(values)
    return _check_objective_value(evaluate(objective, x))


def batched_adjoint(circuits, observables, nqubits, ctx: _lib.Context | None = None, fuse: bool = True):
    """Independent pure-state programs (own qubit count, parameters, observable),
    all from |0...0>, differentiated in one library call (mossq_batch_adjoint).
    With fuse, items of 12 or more qubits run as batched LDS-tile sweeps; the
    others (and every item without fuse) run gate by gate.
    Returns (values[B], [slot gradient vector per item])."""
    ctx = ctx or _lib.default_context()
    B = len(circuits)
    progs = [Program(c, n, False, ctx, fuse=fuse) for c, n in zip(circuits, nqubits)]
    obs = [DeviceObservable(o, n, ctx) for o, n in zip(observables, nqubits)]
    offs = np.zeros(B, dtype=np.int64)
    tot = 0
    for i, p in enumerate(progs):
        offs[i] = tot
        tot += p.lowered.nderiv
    vals = np.zeros(B)
    grads = np.zeros(max(tot, 1))
    ph = (C.c_void_p * B)(*[p.h.value for p in progs])
    oh = (C.c_void_p * B)(*[o.h.value for o in obs])
    check(ctx.lib.mossq_batch_adjoint(ctx.h, B, ph, oh, offs.ctypes.data_as(C.POINTER(C.c_int64)),
                                      _lib.dptr(vals), _lib.dptr(grads)))
    out = [p.lowered.slot_gradients(grads[o:o + p.lowered.nderiv]) for p, o in zip(progs, offs)]
    return vals, out


class BatchAdjointGraph:
    """One batched adjoint run of device programs (the launches of
    mossq_batch_adjoint: tile sweeps, observe, any gate-by-gate items on the
    second stream) recorded as a HIP graph by mossq_batch_capture.  `run()`
    replays the whole batch with one launch call and returns (values[B],
    [slot gradient vector per item]).  Rebinding an item's program with
    Program.update between runs evaluates its new parameters with the same
    graph.  Holds the programs and observables so they outlive the graph."""

    def __init__(self, programs, observables, ctx: _lib.Context | None = None):
        self.ctx = ctx or programs[0].ctx
        self.programs, self.observables = list(programs), list(observables)
        B = len(self.programs)
        if B == 0:
            raise ValueError("a batch graph needs at least one program")
        self.offs = np.zeros(B, dtype=np.int64)
        tot = 0
        for i, p in enumerate(self.programs):
            self.offs[i] = tot
            tot += p.lowered.nderiv
        self.total = tot
        ph = (C.c_void_p * B)(*[p.h.value for p in self.programs])
        oh = (C.c_void_p * B)(*[o.h.value for o in self.observables])
        h = C.c_void_p()
        check(self.ctx.lib.mossq_batch_capture(self.ctx.h, B, ph, oh,
                                               self.offs.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(h)))
        self.h = h

    def rebind(self, items, workers: int | None = None):
        """New parameters for every program: items[i] is a circuit with
        program i's structure, or a pool for its current lowering (for
        instance `program.lowered.pool_with_slots(values)`).  Circuits are
        lowered here; one library call (mossq_program_update_many) replans
        the programs on `workers` host threads (default: one per hardware
        thread, at most 16) and uploads all new values at once.  Raises if any
        program refuses; the others keep their new values."""
        if len(items) != len(self.programs):
            raise ValueError("one circuit or pool per program")
        m = len(items)
        lws, keep = [], []
        progs, nops = (C.c_void_p * m)(), (C.c_int * m)()
        recp, poolp, lens = (C.c_void_p * m)(), (C.c_void_p * m)(), (C.c_uint64 * m)()
        for i, (p, it) in enumerate(zip(self.programs, items)):
            if isinstance(it, QCircuit):
                lw = lower(it, p.n, p.is_density)
                recs, pool = lw.arrays()
            else:
                lw = p.lowered
                recs, _ = lw.arrays()
                pool = np.ascontiguousarray(it, dtype=np.complex128)
            lws.append(lw)
            keep.append((recs, pool))  # alive through the call
            progs[i], nops[i] = p.h, len(lw.records)
            recp[i], poolp[i], lens[i] = C.addressof(recs), pool.ctypes.data, pool.size
        workers = workers or min(16, os.cpu_count() or 1)
        status = (C.c_int * m)()
        st = self.ctx.lib.mossq_program_update_many(self.ctx.h, m, progs, nops, recp, poolp, lens, workers, status)
        msg = self.ctx.lib.mossq_last_error() if st else b""
        for p, lw, s in zip(self.programs, lws, status):
            if s == _lib.MOSSQ_OK:
                p.lowered = lw
        if st != _lib.MOSSQ_OK:
            raise _lib.MossqError(st, msg.decode(errors="replace"))

    def run(self):
        vals = np.zeros(len(self.programs))
        grads = np.zeros(max(self.total, 1))
        check(self.ctx.lib.mossq_batch_graph_run(self.ctx.h, self.h, _lib.dptr(vals), _lib.dptr(grads)))
        return vals, [p.lowered.slot_gradients(grads[o:o + p.lowered.nderiv])
                      for p, o in zip(self.programs, self.offs)]

    def __del__(self):
        try:
            if self.h:
                self.ctx.lib.mossq_graph_destroy(self.h)
            self.h = None
        except Exception:
            pass


def batched_adjoint_graph(circuits, observables, nqubits, ctx: _lib.Context | None = None, fuse: bool = True):
    """batched_adjoint's programs captured once (BatchAdjointGraph); rebind
    item i with graph.programs[i].update(circuit) and replay with run()."""
    ctx = ctx or _lib.default_context()
    progs = [Program(c, n, False, ctx, fuse=fuse) for c, n in zip(circuits, nqubits)]
    obs = [DeviceObservable(o, n, ctx) for o, n in zip(observables, nqubits)]
    return BatchAdjointGraph(progs, obs, ctx)


_torch_rule = None


def torch_differentiable_value(method, objective, values):
    """AD adapter, the torch.autograd counterpart of the reference's
    ChainRulesCore rrule on differentiable_value
    (ext/MOSSQuantumChainRulesCoreExt.jl:7-28): the forward checks the input,
    prepares value_and_pullback once and returns the value (a 0-d tensor of
    `values`' dtype on its device); the backward feeds the output cotangent
    to the prepared pullback and returns seed * gradient shaped like
    `values` (the projection onto the input).  A zero cotangent gives a zero
    gradient without calling the pullback.  Only `values` is differentiable;
    the method and the objective are fixed configuration.  The prepared
    pullback is reusable, so backward passes with retain_graph are
    independent."""
    global _torch_rule
    if _torch_rule is None:
        import torch

        class _Rule(torch.autograd.Function):
            @staticmethod
            def forward(ctx, vals, m, obj):
                if vals.is_complex() or not vals.is_floating_point() or vals.dim() != 1:
                    raise TypeError("parameter values must be a real vector tensor")
                host = vals.detach().to("cpu", torch.float64).numpy()
                v, back = value_and_pullback(m, obj, host)
                ctx.back = back
                return torch.tensor(v, dtype=vals.dtype, device=vals.device)

            @staticmethod
            def backward(ctx, seed):
                s = float(seed.detach().to("cpu"))
                if s == 0.0:
                    g = np.zeros_like(ctx.back(1.0))
                else:
                    g = ctx.back(s)
                return torch.as_tensor(g, dtype=seed.dtype, device=seed.device), None, None

        _torch_rule = _Rule
    return _torch_rule.apply(values, method, objective)
