"""The gradient protocol's host surface and the torch AD adapter, on the CPU
with a plain callable objective and central differences (the reference's
tests of the optional AD adapter, test/extensions/, and of
value_and_pullback, api/gradient_methods.jl:100-138).

  * value_and_pullback: back(seed) = seed * gradient; repeated calls are
    independent (a returned array's mutation does not reach later calls);
    non-finite inputs, unsupported methods and complex seeds are refused;
  * differentiable_value: the forward value only;
  * torch_differentiable_value (the counterpart of the rrule in
    ext/MOSSQuantumChainRulesCoreExt.jl:7-28): x.grad = seed * gradient
    through torch's chain rule, repeated backward passes with retain_graph
    give the same gradient each time, a zero cotangent gives zeros, the
    gradient takes the input's dtype and shape.
"""
import numpy as np
import pytest

import mossquantum_amd as mq
from mossquantum_amd.gradients import torch_differentiable_value

torch = pytest.importorskip("torch")


def f(x):
    return float(np.sum(np.sin(x)) + x[0] * x[1])


def df(x):
    g = np.cos(x)
    g[0] += x[1]
    g[1] += x[0]
    return g


X = np.array([0.3, -1.2, 2.5, 0.7])
FD = mq.FiniteDifference(step=1e-5)  # central differences: error ~ step^2 |f'''| / 6


def test_value_and_pullback_is_reusable_and_independent():
    v, back = mq.value_and_pullback(FD, f, X)
    assert v == f(X)
    g1 = back(1.0)
    assert np.max(np.abs(g1 - df(X))) < 1e-9
    g2 = back(2.5)
    assert np.allclose(g2, 2.5 * g1, rtol=0, atol=1e-15)
    kept = g1.copy()
    g1[:] = 99.0  # mutating one result leaves the pullback's saved gradient alone
    assert np.array_equal(back(1.0), kept)
    assert np.array_equal(back(0.0), np.zeros_like(g2))
    with pytest.raises(TypeError):
        back(1j)


def test_pullback_input_checks():
    with pytest.raises(ValueError, match="finite"):
        mq.value_and_pullback(FD, f, np.array([0.1, np.nan, 0.2, 0.3]))
    with pytest.raises(mq.UnsupportedGradient):
        mq.value_and_pullback(mq.AdjointGradient(), f, X)  # the adjoint needs an expectation objective
    with pytest.raises(TypeError):
        mq.value_and_pullback(FD, f, np.ones((2, 2)))


def test_differentiable_value_is_the_forward_value():
    assert mq.differentiable_value(FD, f, X) == f(X)
    with pytest.raises(ValueError):
        mq.differentiable_value(FD, f, [np.inf, 0, 0, 0])


def test_torch_adapter_chain_rule_and_repeated_backward():
    x = torch.tensor(X, dtype=torch.float64, requires_grad=True)
    y = torch_differentiable_value(FD, f, x)
    assert y.dtype == torch.float64 and y.dim() == 0 and y.detach().item() == f(X)
    loss = 3.0 * y + y ** 2
    loss.backward(retain_graph=True)
    want = (3.0 + 2 * f(X)) * df(X)
    first = x.grad.clone()
    assert np.max(np.abs(first.numpy() - want)) < 1e-8
    x.grad = None
    loss.backward()  # the prepared pullback runs again, independently
    assert torch.equal(x.grad, first)


def test_torch_adapter_zero_cotangent_dtype_and_shape():
    x = torch.tensor(X, dtype=torch.float32, requires_grad=True)
    y = torch_differentiable_value(FD, f, x)
    assert y.dtype == torch.float32
    (0.0 * y).backward()
    assert x.grad.dtype == torch.float32 and x.grad.shape == x.shape and not x.grad.any()
    with pytest.raises(TypeError):
        torch_differentiable_value(FD, f, torch.zeros((2, 2), dtype=torch.float64, requires_grad=True))
