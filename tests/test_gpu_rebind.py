"""Rebinding programs in place and replaying captured graphs (GPU tests,
through the C ABI):

  * mossq_program_update writes new gate matrices (the template's rotations at
    new angles, the device form of the reference's reset_parameters! /
    bind_parameters, src/quantum_circuits/instructions.jl:62-80,
    src/api/parameter_binding.jl:29) over a program's device tables; a plain
    run, the pure-state and density adjoints and captured graphs then give the
    oracle's results for the new angles;
  * mossq_batch_capture records one batched adjoint run (tile sweeps of every
    item, those below 12 qubits in their 12-qubit form); each replay, with eager work and
    synchronisations in between, equals an eager mossq_batch_adjoint call, and
    after rebinding items the replay equals an eager call on the new circuits;
  * refusals leave the program as it was: records of another program, and a
    live program graph that launches a changed monomial step on its own.
"""
import numpy as np
import pytest

import oracle
import mossquantum_amd as mq
from mossquantum_amd import lower
from mossquantum_amd.gradients import _adjoint_one
from mossquantum_amd.workloads import hea_circuit, tfim, xxz

pytestmark = pytest.mark.gpu
AMP, EXP, GRAD = 1e-12, 1e-10, 1e-9


def _angles(rng, n, layers):
    return rng.uniform(0.1, 2 * np.pi - 0.1, 2 * n * layers)


def _batch(rng, sizes, layers=2):
    return [hea_circuit(n, layers, _angles(rng, n, layers)) for n in sizes]


def _same(a, b):
    (va, ga), (vb, gb) = a, b
    assert np.max(np.abs(va - vb)) < EXP
    assert max(np.max(np.abs(x - y)) for x, y in zip(ga, gb)) < GRAD


def test_batch_graph_replays_equal_eager(ctx):
    """Items of 10-16 qubits in one graph (those below 12 in their 12-qubit form):
    three replays with unrelated eager work between them each equal the eager
    engine bit for bit (fixed-order reductions), and the oracle on small items."""
    rng = np.random.default_rng(301)
    sizes = [12, 10, 14, 13, 11, 16, 12]
    circs = _batch(rng, sizes)
    obs = [xxz(n) for n in sizes]
    graph = mq.batched_adjoint_graph(circs, obs, sizes, ctx, fuse=True)
    want = mq.batched_adjoint(circs, obs, sizes, ctx, fuse=True)
    for r in range(3):
        got = graph.run()
        assert np.array_equal(got[0], want[0]), r
        assert all(np.array_equal(a, b) for a, b in zip(got[1], want[1])), r
        # unrelated eager work on the same context between replays
        other = _batch(rng, [13, 10])
        mq.batched_adjoint(other, [tfim(13), tfim(10)], [13, 10], ctx, fuse=True)
        sv = mq.StateVector.zeros(12, ctx)
        mq.apply_(circs[0], sv, mq.HIPBackend(ctx, fuse=True))
    for i in (1, 0, 4):
        v, g = oracle.adjoint(lower(circs[i], sizes[i], False), obs[i].groups())
        assert abs(want[0][i] - v) < EXP and np.max(np.abs(want[1][i] - g)) < GRAD


def test_batch_graph_rebind(ctx):
    """Rebinding every item's program with new angles: the replay equals the
    eager engine on the new circuits (and the oracle on small items); rebinding
    some items back gives their first results again, the others keep theirs."""
    rng = np.random.default_rng(302)
    sizes = [13, 10, 12, 15, 11, 12]
    c1 = _batch(rng, sizes)
    c2 = _batch(rng, sizes)
    obs = [tfim(n) for n in sizes]
    graph = mq.batched_adjoint_graph(c1, obs, sizes, ctx, fuse=True)
    first = graph.run()
    _same(first, mq.batched_adjoint(c1, obs, sizes, ctx, fuse=True))
    for p, c in zip(graph.programs, c2):
        p.update(c)
    second = graph.run()
    _same(second, mq.batched_adjoint(c2, obs, sizes, ctx, fuse=True))
    for i in (1, 2, 4):
        v, g = oracle.adjoint(lower(c2[i], sizes[i], False), obs[i].groups())
        assert abs(second[0][i] - v) < EXP and np.max(np.abs(second[1][i] - g)) < GRAD
    back = (0, 1, 3)
    for i in back:
        graph.programs[i].update(c1[i])
    third = graph.run()
    for i in range(len(sizes)):
        ref = first if i in back else second
        assert abs(third[0][i] - ref[0][i]) < EXP, i
        assert np.max(np.abs(third[1][i] - ref[1][i])) < GRAD, i
    del graph


def test_batch_graph_parallel_rebind(ctx):
    """Rebinding every program on several host threads at once (circuits for
    some items, pools from the lowering's slot rebinding for the others) gives
    the same replay as the eager engine on the new circuits."""
    rng = np.random.default_rng(304)
    sizes = [12, 10, 13, 11, 14, 12, 15, 10]
    c1 = _batch(rng, sizes)
    angles = [_angles(rng, n, 2) for n in sizes]
    c2 = [hea_circuit(n, 2, a) for n, a in zip(sizes, angles)]
    obs = [xxz(n) for n in sizes]
    graph = mq.batched_adjoint_graph(c1, obs, sizes, ctx, fuse=True)
    graph.run()
    items = [c if i % 2 == 0 else graph.programs[i].lowered.pool_with_slots(angles[i])
             for i, c in enumerate(c2)]
    graph.rebind(items, workers=6)
    _same(graph.run(), mq.batched_adjoint(c2, obs, sizes, ctx, fuse=True))
    with pytest.raises(Exception):
        graph.rebind(c2[:-1])
    del graph


@pytest.mark.parametrize("n,fuse", [(14, True), (9, False)])
def test_program_update_run_and_adjoint(ctx, n, fuse):
    """After an update, a plain run and the pure-state adjoint of the same
    program give the oracle's amplitudes, value and gradient for the new angles."""
    rng = np.random.default_rng(303 + n)
    c1 = hea_circuit(n, 3, _angles(rng, n, 3))
    c2 = hea_circuit(n, 3, _angles(rng, n, 3))
    prog = mq.Program(c1, n, False, ctx, fuse=fuse)
    prog.update(c2)
    sv = mq.StateVector.zeros(n, ctx)
    prog.run(sv.storage)
    want = np.zeros(1 << n, dtype=np.complex128)
    want[0] = 1
    oracle.run_program(want, lower(c2, n, False))
    assert np.max(np.abs(sv.to_host() - want)) < AMP
    b = mq.HIPBackend(ctx, fuse=fuse)
    v, g = _adjoint_one(b, prog, xxz(n), mq.StateVector.zeros(n, ctx))
    v0, g0 = oracle.adjoint(lower(c2, n, False), xxz(n).groups())
    assert abs(v - v0) < EXP and np.max(np.abs(g - g0)) < GRAD


def test_program_update_density_adjoint(ctx):
    """A noisy density program (14 bits: fused forward and density adjoint
    plans) rebound to new angles gives the oracle's value and gradient."""
    n = 7
    rng = np.random.default_rng(305)
    c1 = hea_circuit(n, 2, _angles(rng, n, 2), noise=(0.01, 0.02))
    c2 = hea_circuit(n, 2, _angles(rng, n, 2), noise=(0.01, 0.02))
    prog = mq.Program(c1, n, True, ctx, fuse=True)
    b = mq.HIPBackend(ctx, fuse=True)
    v1, g1 = _adjoint_one(b, prog, tfim(n), mq.DensityMatrix.zeros(n, ctx))
    w1 = oracle.adjoint(lower(c1, n, True), tfim(n).groups())
    assert abs(v1 - w1[0]) < EXP and np.max(np.abs(g1 - w1[1])) < GRAD
    prog.update(c2)
    v2, g2 = _adjoint_one(b, prog, tfim(n), mq.DensityMatrix.zeros(n, ctx))
    w2 = oracle.adjoint(lower(c2, n, True), tfim(n).groups())
    assert abs(v2 - w2[0]) < EXP and np.max(np.abs(g2 - w2[1])) < GRAD


def test_program_graph_replays_updated_values(ctx):
    """A fused program's graph (rotations inside tile sweeps) replays the
    values of each update on its buffer."""
    n = 13
    rng = np.random.default_rng(306)
    prog = mq.Program(hea_circuit(n, 2, _angles(rng, n, 2)), n, False, ctx, fuse=True)
    sv = mq.StateVector.zeros(n, ctx)
    graph = prog.capture(sv.storage)
    for r in range(3):
        c = hea_circuit(n, 2, _angles(rng, n, 2))
        prog.update(c)
        mq.apply_(mq.HGate(1), sv)
        mq.apply_(mq.HGate(1), sv)  # eager work on the buffer between replays (H H = I)
        want = sv.to_host()
        oracle.run_program(want, lower(c, n, False))
        graph.launch()
        assert np.max(np.abs(sv.to_host() - want)) < 1e-11, r
    del graph


def test_program_update_refusals(ctx):
    """Records of another program are refused; so is a change of a monomial
    step that a live graph of an unfused program launches with its entries as
    kernel arguments (an R_z).  Each refusal leaves the program as it was, and
    the update goes through once the graph is gone."""
    n = 8
    rng = np.random.default_rng(307)
    c1 = hea_circuit(n, 1, _angles(rng, n, 1))
    c2 = hea_circuit(n, 1, _angles(rng, n, 1))
    prog = mq.Program(c1, n, False, ctx, fuse=False)
    want1 = np.zeros(1 << n, dtype=np.complex128)
    want1[0] = 1
    oracle.run_program(want1, lower(c1, n, False))

    def run():
        sv = mq.StateVector.zeros(n, ctx)
        prog.run(sv.storage)
        return sv.to_host()

    with pytest.raises(Exception):
        prog.update(hea_circuit(n, 2, _angles(rng, n, 2)))
    assert np.max(np.abs(run() - want1)) < AMP
    buf = mq.StateVector.zeros(n, ctx)
    graph = prog.capture(buf.storage)
    with pytest.raises(Exception):
        prog.update(c2)
    assert np.max(np.abs(run() - want1)) < AMP
    del graph
    prog.update(c2)
    want2 = np.zeros(1 << n, dtype=np.complex128)
    want2[0] = 1
    oracle.run_program(want2, lower(c2, n, False))
    assert np.max(np.abs(run() - want2)) < AMP


def test_update_many_statuses(ctx):
    """mossq_program_update_many on the fused programs of a captured batch
    (9-15 qubits; the 9-qubit one also as its 12-qubit form): the refused
    program (records of another structure) keeps its old values, every other
    program takes its new ones, each reports its own status, and the replay
    equals the eager engine on the resulting circuits."""
    import ctypes as C
    rng = np.random.default_rng(308)
    sizes = [13, 9, 14, 12, 15]
    c1 = _batch(rng, sizes)
    c2 = _batch(rng, sizes)
    c2[3] = hea_circuit(12, 3, _angles(rng, 12, 3))  # another structure: refused
    obs = [xxz(n) for n in sizes]
    graph = mq.batched_adjoint_graph(c1, obs, sizes, ctx, fuse=True)
    graph.run()
    progs, m = graph.programs, len(sizes)
    keep = [lower(c, n, False).arrays() for c, n in zip(c2, sizes)]
    ph = (C.c_void_p * m)(*[p.h for p in progs])
    nops = (C.c_int * m)(*[len(lower(c, n, False).records) for c, n in zip(c2, sizes)])
    rp = (C.c_void_p * m)(*[C.addressof(r) for r, _ in keep])
    pp = (C.c_void_p * m)(*[pl.ctypes.data for _, pl in keep])
    lens = (C.c_uint64 * m)(*[pl.size for _, pl in keep])
    status = (C.c_int * m)()
    st = ctx.lib.mossq_program_update_many(ctx.h, m, ph, nops, rp, pp, lens, 3, status)
    assert st != 0 and list(status) == [0, 0, 0, st, 0]
    c3 = [c1[i] if i == 3 else c2[i] for i in range(m)]
    for i, (p, n) in enumerate(zip(progs, sizes)):
        sv = mq.StateVector.zeros(n, ctx)
        p.run(sv.storage)
        want = np.zeros(1 << n, dtype=np.complex128)
        want[0] = 1
        oracle.run_program(want, lower(c3[i], n, False))
        assert np.max(np.abs(sv.to_host() - want)) < AMP, i
    _same(graph.run(), mq.batched_adjoint(c3, obs, sizes, ctx, fuse=True))
    for i in (1, 3):
        v0, g0 = oracle.adjoint(lower(c3[i], sizes[i], False), obs[i].groups())
        v, g = mq.batched_adjoint([c3[i]], [obs[i]], [sizes[i]], ctx, fuse=True)
        assert abs(v[0] - v0) < EXP and np.max(np.abs(g[0] - g0)) < GRAD, i
    del graph
