"""fuse_gates (reference utilities/algorithms/gate_fusion.jl:10-62): adjacent
non-parameterized gates merge while their support stays within max_qubits;
parameterized gates, channels and nested circuits separate the blocks.  The
fused circuit is checked against the unfused one with the oracle (state
vectors and density matrices), and the merged matrices against the dense
Kronecker embedding.  No device needed."""
import numpy as np
import pytest

import oracle
from oracle import dense
import mossquantum_amd as mq
from mossquantum_amd import lower
from mossquantum_amd.circuit import _embed
from mossquantum_amd.gates import haar_unitary
from mossquantum_amd.workloads import haar_layers, random_circuit


def run(c, n, is_dm=False, seed=0):
    rng = np.random.default_rng(seed)
    if is_dm:
        A = rng.standard_normal((1 << n, 1 << n)) + 1j * rng.standard_normal((1 << n, 1 << n))
        r = A @ A.conj().T
        v = (r / np.trace(r)).reshape(-1, order="F")
    else:
        v = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
        v /= np.linalg.norm(v)
    oracle.run_program(v, lower(c, n, is_dm))
    return v


@pytest.mark.parametrize("kmax", [1, 2, 3, 4, 5])
def test_fused_circuit_matches_unfused(kmax):
    n, d = 7, 4
    c = random_circuit(n, d, haar_layers(n, d, 40 + kmax))
    c.append(mq.QuantumGate((2, 5, 7), haar_unitary(np.random.default_rng(kmax), 8)))
    fg = mq.fuse_gates(c, kmax)
    ops = list(fg.flat())
    assert all(o.nqubits() <= max(kmax, 3) for o in ops)  # wider gates pass through unmerged
    assert len(ops) <= len(list(c.flat()))
    if kmax >= 2:
        assert len(ops) < len(list(c.flat()))
    assert np.max(np.abs(run(fg, n) - run(c, n))) < 1e-12


def test_parameterized_gates_and_channels_separate_blocks():
    n = 4
    rng = np.random.default_rng(5)
    c = mq.QCircuit([mq.HGate(1), mq.CNOTGate(1, 2), mq.RyGate(2, 0.4, isparas=True), mq.XGate(3),
                     mq.AmplitudeDamping(3, 0.1), mq.CZGate(3, 4), mq.QuantumGate((1, 4), haar_unitary(rng, 4))])
    fg = mq.fuse_gates(c, 5)
    kinds = [type(o).__name__ for o in fg.flat()]
    # H and CNOT merge; the parameterized RY stays; X stands alone before the
    # channel; CZ and the two-qubit gate after it merge
    assert kinds == ["QuantumGate", "RyGate", "XGate", "QuantumMap", "QuantumGate"]
    ops = list(fg.flat())
    assert tuple(ops[0].positions()) == (1, 2) and tuple(ops[4].positions()) == (1, 3, 4)
    assert np.max(np.abs(run(fg, n, True) - run(c, n, True))) < 1e-12
    with pytest.raises(ValueError):
        mq.fuse_gates(c, 6)


def test_embedding_matches_dense_kronecker():
    rng = np.random.default_rng(9)
    for pos, support in (((2,), [1, 2, 4]), ((1, 4), [1, 3, 4]), ((3, 5), [2, 3, 5, 7]), ((1, 2, 3), [1, 2, 3])):
        U = haar_unitary(rng, 1 << len(pos))
        g = mq.QuantumGate(pos, U)
        got = _embed(g, support)
        # dense: g embedded in len(support) qubits relabelled 1..m in support order
        local = tuple(support.index(p) + 1 for p in g.ordered_positions())
        want = dense.embed(len(support), local, g.ordered_mat())
        assert np.max(np.abs(got - want)) < 1e-15
