"""Device adjoint gradients against central finite differences at h = 1e-6
(GPU tests; the reference's method, test/regression/gradients.jl:4-83 with
test/support/finite_differences.jl:8-46), at the gradient bar 1e-9:

  * the 12-qubit, 10-layer ansatz (all 240 parameters, TFIM), the adjoint of
    mossq_adjoint;
  * the 16- and 20-qubit items of the 512-state batch (XXZ), the gradients the
    batched engine returns for the whole batch (every parameter of the
    16-qubit item, 16 spread over the 20-qubit item's 240).

The FD energies are evaluated by the oracle in extended precision (angles,
gate matrices, state and sums), which keeps the quotient's rounding near
1e-13 (measured 2e-13 to 7e-13 against the device gradients).  The same
quotient in double precision (compensated sums, the steps actually
represented) is checked beside it: its floor is about eps * |H| / h (the
rounding of U(theta +- h) and of the state), measured 5e-11 to 2.1e-10.
"""
import numpy as np
import pytest

import oracle
import mossquantum_amd as mq
from mossquantum_amd import lower
from mossquantum_amd.workloads import c3_batch, hea_circuit, tfim

pytestmark = pytest.mark.gpu
GRAD = 1e-9
GOLD = __import__("pathlib").Path(__file__).resolve().parent / "golden"


@pytest.mark.timeout(300)
def test_c0_adjoint_against_central_fd(ctx):
    n, L = 12, 10
    th = np.load(GOLD / "c0_hea12.npz")["angles"]
    tpl = hea_circuit(n, L)
    obj = mq.ExpectationObjective(tpl, tfim(n), mq.StateVector.zeros(n, ctx))
    v, g = mq.value_and_gradient(mq.AdjointGradient(), obj, th)
    lw = lower(tpl.bind(th), n, False)
    slots = range(len(th))
    fd = oracle.central_fd(lw, th, tfim(n).groups(), slots, h=1e-6, extended=True)
    err = float(np.max(np.abs(g - fd)))
    assert err < GRAD, err
    fd_double = oracle.central_fd(lw, th, tfim(n).groups(), slots, h=1e-6, extended=False)
    assert float(np.max(np.abs(g - fd_double))) < GRAD


@pytest.mark.timeout(600)
def test_c3_batch_items_against_central_fd(ctx):
    sizes, circs, obs = c3_batch()
    vals, grads = mq.batched_adjoint(circs, obs, sizes, ctx, fuse=True)
    for n, nslots in ((16, None), (20, 16)):
        b = sizes.index(n)
        th = np.array(circs[b].active_parameters())
        lw = lower(circs[b], n, False)
        slots = range(len(th)) if nslots is None else np.linspace(0, len(th) - 1, nslots).astype(int)
        fd = oracle.central_fd(lw, th, obs[b].groups(), slots, h=1e-6, extended=True)
        err = float(np.max(np.abs(grads[b][list(slots)] - fd)))
        assert err < GRAD, (n, err)
        fd_double = oracle.central_fd(lw, th, obs[b].groups(), slots, h=1e-6, extended=False)
        assert float(np.max(np.abs(grads[b][list(slots)] - fd_double))) < GRAD
