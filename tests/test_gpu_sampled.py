"""Sampled gradients on the device: parameter shifts, finite differences and
SPSA over batched expectation runs, against the adjoint gradient and the
oracle (reference test/unit/sampled_gradients.jl)."""
import numpy as np
import pytest

import oracle
import mossquantum_amd as mq
from mossquantum_amd import lower
from mossquantum_amd.gates import haar_unitary
from mossquantum_amd.workloads import hea_circuit, tfim, xxz

pytestmark = pytest.mark.gpu
EXP, GRAD = 1e-10, 1e-9


def _mixed(n, seed):
    rng = np.random.default_rng(seed)
    c = hea_circuit(n, 2)
    c.append(mq.CRyGate((1, n), 0.4, isparas=True))
    c.append(mq.CPHASEGate((2, 3), 0.9, isparas=True))
    c.append(mq.RxGate(2, 1.3, isparas=True))
    return c


@pytest.mark.parametrize("n", [6, 11])
def test_parameter_shift_matches_adjoint(ctx, n):
    """Pure states through the batched engine (one run for all shifts)."""
    tpl = _mixed(n, n)
    th = np.random.default_rng(n).uniform(0, 2 * np.pi, tpl.nparameters())
    obj = mq.ExpectationObjective(tpl, tfim(n), mq.StateVector.zeros(n))
    v_adj, g_adj = mq.value_and_gradient(mq.AdjointGradient(), obj, th)
    est = mq.estimate_gradient(mq.ParameterShift(), obj, th)
    assert abs(est.value - v_adj) < EXP
    assert np.max(np.abs(est.gradient - g_adj)) < GRAD
    assert est.evaluations == 1 + 2 * (tpl.nparameters() - 1) + 4 and est.shots == 0
    v, g = mq.value_and_gradient(mq.ParameterShift(), obj, th)
    assert np.array_equal(g, est.gradient)


def test_parameter_shift_keeps_programs(ctx):
    """Repeated parameter-shift gradients of one objective rebind the shifted
    programs kept from the previous call (one mossq_program_update_many); angles
    at pi/2, whose shifts turn a rotation into a permutation, make some of them
    refuse and be rebuilt.  Every call equals the adjoint at its angles, and a
    call at earlier angles repeats that call's gradient bit for bit."""
    n = 9
    tpl = _mixed(n, 21)
    rng = np.random.default_rng(21)
    th1 = rng.uniform(0, 2 * np.pi, tpl.nparameters())
    th2 = th1 + 0.3
    th2[[0, 3, 5]] = np.pi / 2
    obj = mq.ExpectationObjective(tpl, xxz(n), mq.StateVector.zeros(n))
    first = mq.estimate_gradient(mq.ParameterShift(), obj, th1)
    kept = list(obj._shift_programs[1])
    for th in (th2, th1):
        est = mq.estimate_gradient(mq.ParameterShift(), obj, th)
        v_adj, g_adj = mq.value_and_gradient(mq.AdjointGradient(), obj, th)
        assert abs(est.value - v_adj) < EXP and np.max(np.abs(est.gradient - g_adj)) < GRAD
    assert np.array_equal(est.gradient, first.gradient)
    now = obj._shift_programs[1]
    assert len(now) == len(kept) and sum(a is b for a, b in zip(now, kept)) >= len(kept) - 12


def test_parameter_shift_per_program_paths(ctx):
    """A three-qubit gate (per-evaluation programs) and a noisy density matrix."""
    n = 6
    rng = np.random.default_rng(1)
    tpl = hea_circuit(n, 2)
    tpl.append(mq.QuantumGate((1, 3, 5), haar_unitary(rng, 8)))
    tpl.append(mq.RyGate(4, 0.2, isparas=True))
    th = rng.uniform(0, 2 * np.pi, tpl.nparameters())
    obj = mq.ExpectationObjective(tpl, xxz(n), mq.StateVector.zeros(n))
    _, g_adj = mq.value_and_gradient(mq.AdjointGradient(), obj, th)
    assert np.max(np.abs(mq.value_and_gradient(mq.ParameterShift(), obj, th)[1] - g_adj)) < GRAD
    noisy = hea_circuit(n, 2, noise=(0.01, 0.02))
    th = rng.uniform(0, 2 * np.pi, noisy.nparameters())
    dobj = mq.ExpectationObjective(noisy, tfim(n), mq.DensityMatrix.zeros(n))
    v_adj, g_adj = mq.value_and_gradient(mq.AdjointGradient(), dobj, th)
    est = mq.estimate_gradient(mq.ParameterShift(), dobj, th)
    assert abs(est.value - v_adj) < EXP and np.max(np.abs(est.gradient - g_adj)) < GRAD


def test_initial_state_and_parameter_map(ctx):
    """A non-trivial initial state reaches the engine, and shared parameters
    accumulate their slots' shifts."""
    n = 7
    rng = np.random.default_rng(2)
    v = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    v /= np.linalg.norm(v)
    tpl = hea_circuit(n, 2)
    pm = mq.ParameterMap([1 + (i % 5) for i in range(tpl.nparameters())])
    x = rng.uniform(0, 2 * np.pi, 5)
    obj = mq.ExpectationObjective(tpl, tfim(n), mq.StateVector.from_host(v), parameter_map=pm)
    v_adj, g_adj = mq.value_and_gradient(mq.AdjointGradient(), obj, x)
    v_ps, g_ps = mq.value_and_gradient(mq.ParameterShift(), obj, x)
    assert abs(v_ps - v_adj) < EXP and np.max(np.abs(g_ps - g_adj)) < GRAD
    lw = lower(tpl.bind(pm.expand(x)), n, False)
    want_v = oracle.adjoint(lw, tfim(n).groups(), state_in=v)[0]
    assert abs(v_ps - want_v) < EXP


def test_finite_difference_and_spsa(ctx):
    n = 8
    tpl = hea_circuit(n, 2)
    th = np.random.default_rng(3).uniform(0, 2 * np.pi, tpl.nparameters())
    obj = mq.ExpectationObjective(tpl, tfim(n), mq.StateVector.zeros(n))
    _, g_adj = mq.value_and_gradient(mq.AdjointGradient(), obj, th)
    fd = mq.estimate_gradient(mq.FiniteDifference(1e-4), obj, th)
    assert np.max(np.abs(fd.gradient - g_adj)) < 1e-6
    assert fd.evaluations == 1 + 2 * th.size
    one = mq.estimate_gradient(mq.SPSAGradient(step=1e-4, directions=1), obj, th, np.random.default_rng(9))
    # the direction is the first draw of the same generator; with one direction
    # the estimate is (d . g) d up to the O(h^2) difference error
    d = np.where(np.random.default_rng(9).random(th.size) < 0.5, 1.0, -1.0)
    assert np.max(np.abs(one.gradient - np.dot(d, g_adj) * d)) < 1e-5


def test_batch_objective(ctx):
    n = 6
    cols = [mq.StateVector.zeros(n), mq.apply_(mq.XGate(2), mq.StateVector.zeros(n))]
    tpl = hea_circuit(n, 2)
    th = np.random.default_rng(4).uniform(0, 2 * np.pi, tpl.nparameters())
    bobj = mq.BatchExpectationObjective(tpl, tfim(n), mq.BatchStateVector(cols), reduction="weighted_sum",
                                        weights=[0.25, 2.0])
    v_adj, g_adj = mq.value_and_gradient(mq.AdjointGradient(), bobj, th)
    v_ps, g_ps = mq.value_and_gradient(mq.ParameterShift(), bobj, th)
    assert abs(v_ps - v_adj) < EXP and np.max(np.abs(g_ps - g_adj)) < GRAD
    assert not mq.supports_gradient(mq.SPSAGradient(), bobj)
