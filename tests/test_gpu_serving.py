"""Continuous-batching server: each result is the same alone and in a mix
(staggered admissions, pure and density requests together), and matches the
oracle."""
import numpy as np
import pytest

import oracle
import mossquantum_amd as mq
from mossquantum_amd import lower
from mossquantum_amd.serving import RequestTemplate, Server, serve_poisson
from mossquantum_amd.workloads import hea_circuit, tfim

pytestmark = pytest.mark.gpu
EXP, GRAD = 1e-10, 1e-9


def make_requests(ctx, specs, seed):
    rng = np.random.default_rng(seed)
    tpls = {}
    reqs = []
    for kind, n, layers in specs:
        key = (kind, n, layers)
        if key not in tpls:
            noise = (0.01, 0.02) if kind == "dm" else None
            tpls[key] = RequestTemplate(hea_circuit(n, layers, noise=noise), n, kind == "dm", tfim(n), ctx)
        tpl = tpls[key]
        th = rng.uniform(0, 2 * np.pi, tpl.nparams)
        reqs.append((tpl, tpl.bind(th), th))
    return reqs


# 18 qubits takes the fused-adjoint worker path, a 6-qubit density request the fused
# density adjoint; the rest the engine or the density worker.
SPECS = [("sv", 8, 2), ("sv", 13, 2), ("dm", 4, 1), ("sv", 10, 3), ("sv", 15, 1), ("dm", 5, 1), ("sv", 9, 2),
         ("sv", 18, 2), ("dm", 6, 2)]


def test_server_alone_vs_mix(ctx):
    reqs = make_requests(ctx, SPECS, 7)
    srv = Server(ctx, 8 << 30)
    alone = []
    for tpl, pool, _ in reqs:
        t = srv.submit(tpl, pool)
        alone.append(srv.run_until_done([t], [tpl.lowered.nderiv])[t])
    # the mix: submissions staggered across engine iterations
    tickets = []
    for j, (tpl, pool, _) in enumerate(reqs):
        tickets.append(srv.submit(tpl, pool))
        for _ in range(j % 3):
            srv.step()
    mix = srv.run_until_done(tickets, [r[0].lowered.nderiv for r in reqs])
    for j, (tpl, pool, th) in enumerate(reqs):
        va, ga = alone[j]
        vm, gm = mix[tickets[j]]
        assert abs(va - vm) < EXP and np.max(np.abs(ga - gm), initial=0) < GRAD
        # oracle on the bound circuit
        circ = hea_circuit(tpl.n, len(th) // (2 * tpl.n), th, noise=(0.01, 0.02) if tpl.is_density else None)
        v, g = oracle.adjoint(lower(circ, tpl.n, tpl.is_density), tfim(tpl.n).groups())
        assert abs(vm - v) < EXP
        assert np.max(np.abs(gm - g)) < GRAD
    srv.close()


def test_server_budget_admission(ctx):
    """A budget that fits one 13-qubit state-vector request (engine path) at a
    time still completes a queue of them (first-in first-out admission)."""
    reqs = make_requests(ctx, [("sv", 13, 1)] * 4, 3)
    one = 2 * 16 * 2 ** 13 * 1.5
    srv = Server(ctx, int(one))
    tickets = [srv.submit(t, p) for t, p, _ in reqs]
    active, queued = srv.step()
    assert queued >= 2
    out = srv.run_until_done(tickets, [r[0].lowered.nderiv for r in reqs])
    assert len(out) == 4
    with pytest.raises(mq.MossqOOM):
        big = make_requests(ctx, [("sv", 18, 1)], 4)[0]
        srv.submit(big[0], big[1])
    srv.close()


def test_serve_poisson_small(ctx):
    reqs = make_requests(ctx, [("sv", 8, 1), ("sv", 10, 1), ("dm", 4, 1)] * 5, 11)
    srv = Server(ctx, 4 << 30)
    res, stats = serve_poisson(srv, [(t, p) for t, p, _ in reqs], rate=500.0, seed=1)
    assert all(r is not None for r in res)
    assert stats["p99_ms"] >= stats["p50_ms"] > 0
    srv.close()


def test_server_memory_stays_within_budget(ctx):
    """Fused state-vector requests of different sizes (worker contexts with
    block caches) and engine requests under a budget that holds a few of the
    largest at a time: the server's device memory, cached blocks included,
    never passes the budget, every request completes, and each result equals
    the same request served alone."""
    specs = [("sv", 16, 1), ("sv", 14, 2), ("sv", 17, 1), ("sv", 15, 1), ("sv", 10, 2), ("sv", 17, 1),
             ("sv", 14, 1), ("sv", 16, 2)]
    reqs = make_requests(ctx, specs, 21)
    budget = 40 << 20
    alone = []
    solo = Server(ctx, 8 << 30)
    for tpl, pool, _ in reqs:
        t = solo.submit(tpl, pool)
        alone.append(solo.run_until_done([t], [tpl.lowered.nderiv])[t])
    solo.close()
    srv = Server(ctx, budget)
    tickets = [srv.submit(t, p) for t, p, _ in reqs]
    out = srv.run_until_done(tickets, [r[0].lowered.nderiv for r in reqs])
    used, peak = srv.memory()
    assert 0 < peak <= budget, (peak, budget)
    assert used <= budget
    for j, t in enumerate(tickets):
        v, g = out[t]
        assert abs(v - alone[j][0]) < EXP and np.max(np.abs(g - alone[j][1])) < GRAD
    srv.close()
