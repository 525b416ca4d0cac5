"""Uniform batches in one device buffer ((2^n, B) and (2^n, 2^n, B), the
reference's batch storage, src/states/batch_state_vector.jl:18-30 and
batch_density_matrix.jl:17-28), against the oracle column by column (GPU
tests, through the C ABI):

  * gates and channel-free circuits run over each power-of-two run of columns
    as one vector (one launch per instruction or fused sweep per run;
    reference backends/cpu/gates/batch.jl:22-82), at B = 5 (runs of 4 and 1)
    and B = 8; a 10-qubit batch's run of 4 columns is a 12-bit vector, so it
    takes the fused tile plan;
  * a channel promotes the whole batch to density matrices;
  * expectations over all columns in one launch (observables/
    batch_expectation.jl, batch_density_expectation.jl);
  * the batch objective's gradients over many initial states in one batched
    engine call (differentiation/batch.jl:35-66), with a program the engine
    does not take (a three-qubit gate below the tiled size) going column by
    column.
"""
import numpy as np
import pytest

import oracle
import mossquantum_amd as mq
from mossquantum_amd import lower
from mossquantum_amd.gates import haar_unitary
from mossquantum_amd.workloads import haar_layers, hea_circuit, random_circuit, tfim, xxz

pytestmark = pytest.mark.gpu
AMP, EXP, GRAD = 1e-12, 1e-10, 1e-9


def host_cols(n, B, rng):
    a = rng.standard_normal((1 << n, B)) + 1j * rng.standard_normal((1 << n, B))
    return a / np.linalg.norm(a, axis=0)


@pytest.mark.parametrize("B,fuse", [(5, True), (8, False), (1, True)])
def test_batch_gates_and_circuits(ctx, B, fuse):
    rng = np.random.default_rng(90 + B)
    n = 10
    a = host_cols(n, B, rng)
    batch = mq.BatchStateVector.from_host(a, ctx)
    assert batch.to_host().shape == (1 << n, B) and np.array_equal(batch.to_host(), a)
    c = random_circuit(n, 4, haar_layers(n, 4, 7))
    c.append(mq.QuantumGate((2, 5, 9), haar_unitary(rng, 8)))
    g = mq.QuantumGate((3, 10), haar_unitary(rng, 4))
    out = mq.apply_(c, batch, mq.HIPBackend(ctx, fuse=fuse))
    out = mq.apply_(g, out, mq.HIPBackend(ctx, fuse=fuse))
    assert out is batch
    got = out.to_host()
    for j in range(B):
        want = a[:, j].copy()
        oracle.run_program(want, lower(c, n, False))
        oracle.apply_gate(want, n, [p - 1 for p in g.ordered_positions()], g.ordered_mat())
        assert np.max(np.abs(got[:, j] - want)) < AMP, j
        assert np.array_equal(out.columns[j].to_host(), got[:, j])


def test_batch_channel_promotion_and_expectations(ctx):
    rng = np.random.default_rng(93)
    n, B = 5, 3
    a = host_cols(n, B, rng)
    batch = mq.BatchStateVector.from_host(a, ctx)
    c = mq.QCircuit([mq.HGate(1), mq.CNOTGate(1, 2), mq.AmplitudeDamping(2, 0.3), mq.RyGate(3, 0.7)])
    dm = mq.apply_(c, batch)
    assert isinstance(dm, mq.BatchDensityMatrix) and dm.to_host().shape == (1 << n, 1 << n, B)
    got = dm.to_host()
    for j in range(B):
        want = np.outer(a[:, j], a[:, j].conj()).reshape(-1, order="F")
        oracle.run_program(want, lower(c, n, True))
        assert np.max(np.abs(got[:, :, j].reshape(-1, order="F") - want)) < AMP
        e = oracle.dm_expect(want, n, tfim(n).groups())
        assert abs(mq.expectation(tfim(n), dm)[j] - e) < EXP
    # (apply_ ran the gates before the channel on `batch` in place, as apply! does)
    e = mq.expectation(xxz(n), mq.BatchStateVector.from_host(a, ctx))
    for j in range(B):
        assert abs(e[j] - oracle.expect(np.ascontiguousarray(a[:, j]), n, xxz(n).groups())) < EXP


def test_batch_objective_gradients(ctx):
    rng = np.random.default_rng(94)
    for n, fuse in ((8, False), (13, True)):
        B = 3
        a = host_cols(n, B, rng)
        batch = mq.BatchStateVector.from_host(a, ctx)
        tpl = hea_circuit(n, 2)
        th = rng.uniform(0, 2 * np.pi, tpl.nparameters())
        obj = mq.BatchExpectationObjective(tpl, xxz(n), batch, reduction="weighted_sum", weights=[0.5, -1.0, 2.0],
                                           backend_=mq.HIPBackend(ctx, fuse=fuse))
        v, g = mq.value_and_gradient(mq.AdjointGradient(), obj, th)
        per = [oracle.adjoint(lower(tpl.bind(th), n, False), xxz(n).groups(), state_in=a[:, j]) for j in range(B)]
        w = [0.5, -1.0, 2.0]
        assert abs(v - sum(wi * p[0] for wi, p in zip(w, per))) < EXP
        assert np.max(np.abs(g - sum(wi * p[1] for wi, p in zip(w, per)))) < GRAD
        assert np.array_equal(batch.to_host(), a)  # the initial states are not modified
    # a program the batched engine does not take: per-column adjoints
    n = 6
    tpl = mq.QCircuit([mq.RyGate(1, 0.0, isparas=True), mq.QuantumGate((1, 2, 3), haar_unitary(rng, 8)),
                       mq.RzGate(2, 0.0, isparas=True)])
    a = host_cols(n, 2, rng)
    obj = mq.BatchExpectationObjective(tpl, tfim(n), mq.BatchStateVector.from_host(a, ctx), reduction="sum")
    v, g = mq.value_and_gradient(mq.AdjointGradient(), obj, [0.3, 1.1])
    per = [oracle.adjoint(lower(tpl.bind([0.3, 1.1]), n, False), tfim(n).groups(), state_in=a[:, j]) for j in range(2)]
    assert abs(v - per[0][0] - per[1][0]) < EXP and np.max(np.abs(g - per[0][1] - per[1][1])) < GRAD
