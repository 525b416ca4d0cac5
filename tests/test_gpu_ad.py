"""The AD adapter and the pullback on the device objectives, and batch
expectations (GPU tests, through the C ABI), against the oracle.

  * torch_differentiable_value on the 12-qubit, 10-layer ansatz (240
    parameters): the value, and x.grad = seed * gradient through torch's chain
    rule (the rrule of ext/MOSSQuantumChainRulesCoreExt.jl:7-28 on
    differentiable_value);
  * value_and_pullback with AdjointGradient on a single and a batch objective
    (reduction mean and weighted_sum): back(seed) = seed * oracle gradient,
    reusable (reference api/gradient_methods.jl:100-115,
    differentiation/objectives.jl:183-295);
  * expectation on batch states (state vectors and density matrices): one
    value per column, each against the oracle (reference
    observables/batch_expectation.jl:1-179, batch_density_expectation.jl:4-63).
"""
import numpy as np
import pytest

import oracle
import mossquantum_amd as mq
from mossquantum_amd import lower
from mossquantum_amd.gradients import torch_differentiable_value
from mossquantum_amd.workloads import hea_circuit, tfim, xxz

pytestmark = pytest.mark.gpu
EXP, GRAD = 1e-10, 1e-9
GOLD = __import__("pathlib").Path(__file__).resolve().parent / "golden"


def test_torch_adapter_on_the_device_adjoint(ctx):
    torch = pytest.importorskip("torch")
    n, L = 12, 10
    th = np.load(GOLD / "c0_hea12.npz")["angles"]
    tpl = hea_circuit(n, L)
    obj = mq.ExpectationObjective(tpl, tfim(n), mq.StateVector.zeros(n, ctx))
    want_v, want_g = oracle.adjoint(lower(tpl.bind(th), n, False), tfim(n).groups())
    x = torch.tensor(th, dtype=torch.float64, requires_grad=True)
    y = torch_differentiable_value(mq.AdjointGradient(), obj, x)
    assert abs(y.detach().item() - want_v) < EXP
    (2.0 * y - 1.0).backward(retain_graph=True)
    assert np.max(np.abs(x.grad.numpy() - 2.0 * want_g)) < GRAD
    first = x.grad.clone()
    x.grad = None
    (2.0 * y).backward()
    assert torch.equal(x.grad, first)


def test_value_and_pullback_single_and_batch(ctx):
    rng = np.random.default_rng(31)
    n = 9
    tpl = hea_circuit(n, 3)
    th = rng.uniform(0, 2 * np.pi, tpl.nparameters())
    obj = mq.ExpectationObjective(tpl, xxz(n), mq.StateVector.zeros(n, ctx))
    v, back = mq.value_and_pullback(mq.AdjointGradient(), obj, th)
    wv, wg = oracle.adjoint(lower(tpl.bind(th), n, False), xxz(n).groups())
    assert abs(v - wv) < EXP
    assert np.max(np.abs(back(-0.5) + 0.5 * wg)) < GRAD and np.max(np.abs(back(1.0) - wg)) < GRAD
    assert mq.differentiable_value(mq.AdjointGradient(), obj, th) == pytest.approx(wv, abs=EXP)
    # a batch of three initial states
    cols = []
    for _ in range(3):
        psi = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
        cols.append(psi / np.linalg.norm(psi))
    batch = mq.BatchStateVector([mq.StateVector.from_host(c, ctx) for c in cols])
    per = [oracle.adjoint(lower(tpl.bind(th), n, False), xxz(n).groups(), state_in=c) for c in cols]
    for reduction, weights in (("mean", None), ("weighted_sum", [0.2, -1.0, 0.7])):
        bobj = mq.BatchExpectationObjective(tpl, xxz(n), batch, reduction=reduction, weights=weights)
        w = np.full(3, 1 / 3) if weights is None else np.array(weights)
        v, back = mq.value_and_pullback(mq.AdjointGradient(), bobj, th)
        assert abs(v - sum(wi * p[0] for wi, p in zip(w, per))) < EXP
        assert np.max(np.abs(back(2.0) - 2.0 * sum(wi * p[1] for wi, p in zip(w, per)))) < GRAD


def test_batch_expectations(ctx):
    rng = np.random.default_rng(32)
    n = 8
    cols = []
    for _ in range(4):
        psi = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
        cols.append(psi / np.linalg.norm(psi))
    batch = mq.BatchStateVector([mq.StateVector.from_host(c, ctx) for c in cols])
    for H in (tfim(n), xxz(n)):
        got = mq.expectation(H, batch)
        assert got.shape == (4,)
        for j, c in enumerate(cols):
            assert abs(got[j] - oracle.expect(c, n, H.groups())) < EXP
    m = 5
    rhos = []
    for _ in range(3):
        A = rng.standard_normal((1 << m, 1 << m)) + 1j * rng.standard_normal((1 << m, 1 << m))
        r = A @ A.conj().T
        rhos.append(r / np.trace(r))
    dbatch = mq.BatchDensityMatrix([mq.DensityMatrix.from_host(r, ctx) for r in rhos])
    got = mq.expectation(tfim(m), dbatch)
    for j, r in enumerate(rhos):
        assert abs(got[j] - oracle.dm_expect(r.reshape(-1, order="F"), m, tfim(m).groups())) < EXP
