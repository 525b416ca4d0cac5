"""Sampled gradient methods over exact device expectations: parameter shifts,
central finite differences and SPSA (reference
src/differentiation/sampled_gradients.jl:1-435, for ExpectationObjective,
BatchExpectationObjective and plain functions).

All shifted or perturbed evaluations of one gradient run together: the
template is lowered once, each evaluation gets its own matrix pool with the
moved gates rebuilt, and pure-state objectives whose gates act on at most two
qubits go through one batched engine run (mossq_batch_expectation) instead of
one circuit at a time.  Other objectives (density matrices, wider gates) run
each evaluation as its own program on the device.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check
from .backend import DeviceObservable, Program, expectation, update_many
from .circuit import lower
from .gates import CPHASEGate, CRxGate, CRyGate, CRzGate, PHASEGate, RxGate, RyGate, RzGate
from .gradients import AbstractGradientMethod, BatchExpectationObjective, ExpectationObjective, UnsupportedGradient
from .states import DensityMatrix, StateVector


class ParameterShift(AbstractGradientMethod):
    """Exact gate-occurrence parameter shifts (sampled_gradients.jl:1-2)."""


class FiniteDifference(AbstractGradientMethod):
    """Central differences with `step` (sampled_gradients.jl:3-10)."""

    def __init__(self, step: float = 0.01):
        step = float(step)
        if not (np.isfinite(step) and step > 0):
            raise ValueError("step must be positive and finite")
        self.step = step


class SPSAGradient(AbstractGradientMethod):
    """Mean of `directions` Rademacher simultaneous perturbations with central
    differences (sampled_gradients.jl:11-19)."""

    def __init__(self, step: float = 0.1, directions: int = 8):
        self.step = FiniteDifference(step).step
        if isinstance(directions, bool) or int(directions) != directions or directions < 1:
            raise ValueError("directions must be a positive integer")
        self.directions = int(directions)


@dataclass
class GradientEstimate:
    """Gradient statistics (sampled_gradients.jl:117-134).  With exact
    expectations the sampling terms are zero and no shots are spent;
    `covariance` is SPSA's direction variability (None for one direction)."""
    value: float
    value_variance: float
    gradient: np.ndarray
    sampling_covariance: np.ndarray
    covariance: np.ndarray | None
    stderr: np.ndarray | None
    shots: int
    evaluations: int


_SHIFT2 = [(0.5, np.pi / 2), (-0.5, -np.pi / 2)]
_CA, _CB = (np.sqrt(2) + 1) / (4 * np.sqrt(2)), (1 - np.sqrt(2)) / (4 * np.sqrt(2))
_SHIFT4 = [(_CA, np.pi / 2), (-_CA, -np.pi / 2), (_CB, 3 * np.pi / 2), (-_CB, -3 * np.pi / 2)]


def parameter_shift_rule(gate, i: int):
    """(coefficient, shift) pairs for active slot i (1-based) of `gate`, or
    None when the gate family has no recipe (sampled_gradients.jl:21-40):
    rotations and phases have gaps 1/2 (two terms), controlled rotations gaps
    1/2 and 1 (four terms)."""
    if not 1 <= i <= gate.nparameters():
        raise IndexError("active parameter index out of range")
    if isinstance(gate, (CRxGate, CRyGate, CRzGate)):
        return list(_SHIFT4)
    if isinstance(gate, (RxGate, RyGate, RzGate, PHASEGate, CPHASEGate)):
        return list(_SHIFT2)
    return None


def _shift_recipes(circuit):
    recipes = []
    for op in circuit.flat():
        for i in range(1, op.nparameters() + 1):
            r = parameter_shift_rule(op, i)
            if r is None:
                return None
            if abs(sum(a for a, _ in r)) > 1e-12 * max(1.0, sum(abs(a) for a, _ in r)):
                raise ValueError("shift coefficients must sum to zero")
            recipes.append(r)
    return recipes


def _is_objective(objective):
    return isinstance(objective, (ExpectationObjective, BatchExpectationObjective))


def supports_sampled_gradient(method, objective, values=None) -> bool:
    """reference supports_gradient for the sampled methods
    (sampled_gradients.jl:178-196,264-279)."""
    if callable(objective) and not _is_objective(objective):
        return isinstance(method, (FiniteDifference, SPSAGradient))
    if not _is_objective(objective):
        return False
    if isinstance(method, SPSAGradient) and isinstance(objective, BatchExpectationObjective):
        return False
    if isinstance(method, ParameterShift):
        try:
            return _shift_recipes(objective.template) is not None
        except ValueError:
            return False
    return isinstance(method, (FiniteDifference, SPSAGradient))


# ------------------------------------------------------------ evaluation ----
_ENGINE_BYTES = 8 << 30  # device bytes of states per batched engine run


class _Evaluator:
    """Objective values at many slot vectors of one objective's template."""

    def __init__(self, objective, flat0):
        self.obj = objective
        self.bound = objective.template.bind(flat0)
        init = objective.initial_state
        self.columns = list(init.columns) if isinstance(objective, BatchExpectationObjective) else [init]
        self.weights = (objective.sample_weights() if isinstance(objective, BatchExpectationObjective)
                        else np.ones(1))
        self.n = init.n
        self.is_dm = isinstance(self.columns[0], DensityMatrix) or self.bound.has_channel()
        self.lw = lower(self.bound, self.n, self.is_dm)
        self.ctx = self.columns[0].storage.ctx
        self.engine = (not self.is_dm and all(isinstance(c, StateVector) for c in self.columns)
                       and all(r.k <= 2 for r in self.lw.records))

    def values(self, slot_vectors, moved=None):
        """Objective value per slot vector; moved[j] lists the slots that vector
        j changes (None: all)."""
        pools = [self.lw.pool_with_slots(v, None if moved is None else moved[j]) for j, v in enumerate(slot_vectors)]
        if self.engine:
            return self._engine_values(pools)
        out = np.zeros(len(pools))
        for j, pool in enumerate(pools):
            prog = Program(None, self.n, self.is_dm, self.ctx, self.obj.backend.fuse, lowered=self.lw, pool=pool)
            vals = []
            for col in self.columns:
                st = col.copy()
                if isinstance(st, StateVector) and self.is_dm:
                    st = DensityMatrix.from_state(st)
                prog.run(st.storage)
                vals.append(expectation(self.obj.observable, st, self.obj.backend).real)
            out[j] = float(np.dot(self.weights, vals))
        return out

    def _programs(self, pools):
        """Unfused programs holding `pools` (this lowering's records; fused
        ones measured slower here, their replanning costing more than their
        sweeps save).  The objective keeps them between calls: programs from an
        earlier call with the same records are rebound in one library call
        (update_many), and only the missing or refused ones are created."""
        fuse = False

        def make(pool):
            return Program(None, self.n, False, self.ctx, fuse, lowered=self.lw, pool=pool)

        recs, _ = self.lw.arrays()
        key = (self.n, bytes(recs), id(self.ctx), fuse)  # (the kept programs hold the context)
        kept = getattr(self.obj, "_shift_programs", None)
        progs = list(kept[1]) if kept is not None and kept[0] == key else []
        reuse = progs[:len(pools)]
        status = update_many(reuse, pools[:len(reuse)], self.ctx)
        for i, st in enumerate(status):
            if st == 0:
                reuse[i].lowered = self.lw
            else:  # refused (the new values change a step's kind): a program of its own
                reuse[i] = make(pools[i])
        out = reuse + [make(p) for p in pools[len(reuse):]]
        self.obj._shift_programs = (key, out if len(out) >= len(progs) else out + progs[len(out):])
        return out

    def _engine_values(self, pools):
        nc = len(self.columns)
        per_item = (1 << self.n) * 16
        chunk = max(1, _ENGINE_BYTES // (per_item * nc))
        dob = DeviceObservable(self.obj.observable, self.n, self.ctx)
        out = np.zeros(len(pools))
        every = self._programs(pools)
        for lo in range(0, len(pools), chunk):
            progs = every[lo:lo + chunk]
            items = [(p, col) for p in progs for col in self.columns]
            B = len(items)
            ph = (C.c_void_p * B)(*[p.h.value for p, _ in items])
            oh = (C.c_void_p * B)(*([dob.h.value] * B))
            ih = (C.c_void_p * B)(*[col.storage.ptr.value for _, col in items])
            vals = np.zeros(B)
            check(self.ctx.lib.mossq_batch_expectation(self.ctx.h, B, ph, oh, ih, _lib.dptr(vals)))
            out[lo:lo + len(progs)] = vals.reshape(len(progs), nc) @ self.weights
        return out


def _function_value(objective, values):
    v = float(objective(np.array(values, dtype=np.float64)))
    if not np.isfinite(v):
        raise ValueError("objective must be finite")
    return v


def _input(objective, values):
    x = np.array(values, dtype=np.float64).reshape(-1)
    if not np.all(np.isfinite(x)):
        raise ValueError("parameters must be finite")
    if _is_objective(objective) and x.size != objective.parameter_map.count:
        raise ValueError("independent parameter count disagrees")
    return x


_ZERO = object()


def _result(value, gradient, evaluations, covariance=_ZERO):
    """Exact expectations: zero sampling covariance; `covariance` defaults to
    a copy of it (parameter shifts, differences), SPSA passes its direction
    estimate or None."""
    n = gradient.size
    if covariance is _ZERO:
        covariance = np.zeros((n, n))
    stderr = None if covariance is None else np.sqrt(np.maximum(np.diag(covariance), 0))
    if not np.all(np.isfinite(gradient)):
        raise ValueError("gradient statistics overflowed")
    return GradientEstimate(value, 0.0, gradient, np.zeros((n, n)), covariance, stderr, 0, evaluations)


def estimate_gradient(method, objective, values, rng=None) -> GradientEstimate:
    """Gradient with evaluation accounting (reference estimate_gradient,
    sampled_gradients.jl:281-435)."""
    if not supports_sampled_gradient(method, objective):
        raise UnsupportedGradient(method, objective)
    x = _input(objective, values)
    rng = np.random.default_rng() if rng is None else rng
    if isinstance(method, ParameterShift):
        return _parameter_shift(objective, x)
    if isinstance(method, FiniteDifference):
        return _finite_difference(method, objective, x)
    return _spsa(method, objective, x, rng)


def _parameter_shift(objective, x):
    pm = objective.parameter_map
    flat = pm.expand(x)
    ev = _Evaluator(objective, flat)
    recipes = _shift_recipes(ev.bound)
    vectors, moved, terms = [flat], [[]], []
    for slot, recipe in enumerate(recipes):
        for weight, shift in recipe:
            v = flat.copy()
            v[slot] += shift
            vectors.append(v)
            moved.append([slot])
            terms.append((slot, weight))
    vals = ev.values(vectors, moved)
    slot_grad = np.zeros(flat.size)
    for (slot, weight), val in zip(terms, vals[1:]):
        slot_grad[slot] += weight * val
    return _result(float(vals[0]), pm.accumulate(slot_grad), len(vectors))


def _perturbed_values(objective, points):
    if callable(objective) and not _is_objective(objective):
        return np.array([_function_value(objective, p) for p in points])
    pm = objective.parameter_map
    flats = [pm.expand(p) for p in points]
    ev = _Evaluator(objective, flats[0])
    return ev.values(flats)


def _finite_difference(method, objective, x):
    h = method.step
    points = [x]
    for i in range(x.size):
        plus, minus = x.copy(), x.copy()
        plus[i] += h
        minus[i] -= h
        if plus[i] == minus[i]:
            raise ValueError("difference step is below parameter resolution")
        if not (np.all(np.isfinite(plus)) and np.all(np.isfinite(minus))):
            raise ValueError("difference step overflows parameter values")
        points += [plus, minus]
    vals = _perturbed_values(objective, points)
    grad = (vals[1::2] - vals[2::2]) / (2 * h)
    return _result(float(vals[0]), grad, len(points))


def _spsa(method, objective, x, rng):
    n, m, h = x.size, method.directions, method.step
    if n == 0:
        v = _perturbed_values(objective, [x])[0]
        return _result(float(v), np.zeros(0), 1)
    points, dirs = [x], []
    for _ in range(m):
        d = np.where(rng.random(n) < 0.5, 1.0, -1.0)
        plus, minus = x + h * d, x - h * d
        if not np.all(plus != minus):
            raise ValueError("SPSA step is below parameter resolution")
        points += [plus, minus]
        dirs.append(d)
    vals = _perturbed_values(objective, points)
    samples = np.stack([((vals[1 + 2 * j] - vals[2 + 2 * j]) / (2 * h)) * dirs[j] for j in range(m)], axis=1)
    grad = samples.sum(axis=1) / m
    centered = samples - grad[:, None]
    cov = centered @ centered.T / (m * (m - 1)) if m > 1 else None
    return _result(float(vals[0]), grad, len(points), cov)
