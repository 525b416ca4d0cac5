"""Request serving on top of libmossq_hip's continuous-batching server.

A `RequestTemplate` lowers one circuit shape (qubit count, layers, noise) once;
`bind(params)` rewrites only the rotation matrices in a copy of the matrix pool
(vectorised), so a request costs microseconds of host work.  `Server` wraps
mossq_server_*; `serve_poisson` replays a Poisson arrival stream in wall-clock
time and reports throughput and latency percentiles.

New capability: the reference evaluates one objective at a time
(src/differentiation/objectives.jl:183-218).
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib
from ._lib import check
from .backend import DeviceObservable
from .circuit import QCircuit, lower
from .gates import RxGate, RyGate, RzGate


class RequestTemplate:
    """One circuit shape; parameters are bound per request.

    Supports templates whose parametric gates are RX/RY/RZ (their derivative
    gates (dU)^dagger U are parameter independent, so only U changes)."""

    def __init__(self, template: QCircuit, n: int, is_density: bool, observable, ctx: _lib.Context):
        self.n, self.is_density, self.ctx = n, is_density, ctx
        self.lowered = lower(template, n, is_density)
        self.records, self.pool = self.lowered.arrays()
        self.nrec = len(self.lowered.records)
        self.obs = DeviceObservable(observable, n, ctx)
        offs = {"RxGate": [], "RyGate": [], "RzGate": []}
        slot = 0
        rec = 0
        order = []
        for op in template.flat():
            if hasattr(op, "nparameters") and op.nparameters():
                name = type(op).__name__
                if name not in offs:
                    raise ValueError(f"request templates bind RX/RY/RZ parameters, not {name}")
                order.append((name, slot))
                slot += 1
        # matrix offsets of the parametric gates, in slot order
        mats = [r.mat for r in self.lowered.records if r.dmat >= 0]
        if len(mats) != len(order):
            raise ValueError("multi-parameter gates are not supported in request templates")
        for (name, s), m in zip(order, mats):
            offs[name].append((s, m))
        self.offs = {k: (np.array([s for s, _ in v], dtype=np.int64), np.array([m for _, m in v], dtype=np.int64))
                     for k, v in offs.items()}
        self.nparams = slot
        del rec

    def bind(self, params) -> np.ndarray:
        """A copy of the matrix pool with U(theta) written for every rotation."""
        th = np.asarray(params, dtype=np.float64)
        if th.size != self.nparams:
            raise ValueError(f"expected {self.nparams} parameters")
        pool = self.pool.copy()
        c_ = np.cos(th / 2)
        s_ = np.sin(th / 2)
        sl, m = self.offs["RyGate"]
        if sl.size:  # [[c, -s], [s, c]] column-major
            pool[m] = c_[sl]
            pool[m + 1] = s_[sl]
            pool[m + 2] = -s_[sl]
            pool[m + 3] = c_[sl]
        sl, m = self.offs["RxGate"]
        if sl.size:  # [[c, -is], [-is, c]]
            pool[m] = c_[sl]
            pool[m + 1] = -1j * s_[sl]
            pool[m + 2] = -1j * s_[sl]
            pool[m + 3] = c_[sl]
        sl, m = self.offs["RzGate"]
        if sl.size:  # diag(e^{-i t/2}, e^{i t/2})
            pool[m] = np.exp(-0.5j * th[sl])
            pool[m + 1] = 0
            pool[m + 2] = 0
            pool[m + 3] = np.exp(0.5j * th[sl])
        return pool


class Server:
    """Continuous-batching energy + adjoint-gradient server (mossq_server)."""

    def __init__(self, ctx: _lib.Context, budget_bytes: int):
        self.ctx = ctx
        h = C.c_void_p()
        check(ctx.lib.mossq_server_create(ctx.h, int(budget_bytes), C.byref(h)))
        self.h = h
        self._buf = (C.c_int64 * 4096)()

    def close(self):
        if self.h:
            check(self.ctx.lib.mossq_server_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, tpl: RequestTemplate, pool: np.ndarray) -> int:
        t = C.c_int64(0)
        check(self.ctx.lib.mossq_server_submit(self.h, tpl.n, int(tpl.is_density), tpl.nrec, tpl.records,
                                               pool.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)),
                                               pool.size, tpl.obs.h, C.byref(t)))
        return t.value

    def memory(self) -> tuple[int, int]:
        """(bytes the requests hold now, cached blocks included; peak since creation)."""
        u, p = C.c_uint64(0), C.c_uint64(0)
        check(self.ctx.lib.mossq_server_memory(self.h, C.byref(u), C.byref(p)))
        return u.value, p.value

    def step(self):
        a, q = C.c_int(0), C.c_int(0)
        check(self.ctx.lib.mossq_server_step(self.h, C.byref(a), C.byref(q)))
        return a.value, q.value

    def completed(self):
        k = C.c_int(0)
        check(self.ctx.lib.mossq_server_completed(self.h, self._buf, len(self._buf), C.byref(k)))
        return list(self._buf[:k.value])

    def result(self, ticket: int, nderiv: int):
        done = C.c_int(0)
        v = C.c_double(0)
        g = np.zeros(max(nderiv, 1))
        check(self.ctx.lib.mossq_server_poll(self.h, ticket, C.byref(done), C.byref(v), _lib.dptr(g), g.size))
        if not done.value:
            return None
        return v.value, g[:nderiv]

    def run_until_done(self, tickets, nderivs, timeout_s=600.0):
        """Step until every ticket finished; {ticket: (value, gradient)}."""
        want = dict(zip(tickets, nderivs))
        out = {}
        t0 = time.perf_counter()
        while len(out) < len(want):
            self.step()
            for t in self.completed():
                if t in want:
                    out[t] = self.result(t, want[t])
            if time.perf_counter() - t0 > timeout_s:
                raise TimeoutError("server did not finish in time")
        return out


def serve_poisson(server: Server, requests, rate: float, seed: int = 0, idle_sleep: float = 2e-4, progress=None,
                  progress_every_s: float = 30.0):
    """Replay `requests` = [(template, pool)] arriving as a Poisson stream of
    `rate` per second (`pool` may be a callable returning it, bound at
    submission).  Latency = completion seen by the host - scheduled
    arrival.  `progress(message)`, when given, is called every
    `progress_every_s` seconds with the count finished.  Returns (results in
    request order, stats dict)."""
    rng = np.random.default_rng(seed)
    N = len(requests)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=N))
    done_at = np.full(N, np.nan)
    results = [None] * N
    by_ticket = {}
    i = 0
    start = time.perf_counter()
    finished = 0
    max_active = 0
    last_note = 0.0
    while finished < N:
        if progress is not None and time.perf_counter() - start - last_note >= progress_every_s:
            last_note = time.perf_counter() - start
            progress(f"{finished} of {N} requests finished after {last_note:.0f} s")
        now = time.perf_counter() - start
        while i < N and arrivals[i] <= now:
            tpl, pool = requests[i]
            by_ticket[server.submit(tpl, pool() if callable(pool) else pool)] = i
            i += 1
        active, queued = server.step()
        max_active = max(max_active, active + queued)
        for t in server.completed():
            j = by_ticket.pop(t)
            tpl = requests[j][0]
            results[j] = server.result(t, tpl.lowered.nderiv)
            done_at[j] = time.perf_counter() - start
            finished += 1
        if active == 0 and queued == 0 and i < N:
            wait = arrivals[i] - (time.perf_counter() - start)
            if wait > 0:
                time.sleep(min(wait, 0.01))
    total = time.perf_counter() - start
    lat = (done_at - arrivals) * 1e3
    stats = {"rate_req_s": rate, "requests": N, "wall_s": total, "throughput_req_s": N / total,
             "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
             "mean_ms": float(lat.mean()), "max_in_system": max_active}
    return results, stats
