"""The fused-sweep planner, executed on the host with the tile kernel's index
arithmetic (mossq_plan_emulate), against the CPU oracle's one-sweep-per-gate
execution.  No device needed."""
import numpy as np
import pytest

import oracle
import mossquantum_amd as mq
from mossquantum_amd import lower
from mossquantum_amd.backend import plan_emulate, plan_emulate_adjoint
from mossquantum_amd.gates import haar_unitary
from mossquantum_amd.workloads import haar_layers, hea_circuit, random_circuit

AMP = 1e-12


def rand_state(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    return v / np.linalg.norm(v)


def check(circ, n, is_density, psi0):
    want = np.array(psi0, dtype=np.complex128).copy()
    oracle.run_program(want, lower(circ, n, is_density))
    got, stats = plan_emulate(circ, n, is_density, psi0)
    assert np.max(np.abs(got - want)) < AMP
    return stats


@pytest.mark.parametrize("n,depth", [(12, 3), (13, 4), (15, 6)])
def test_random_circuit_plans(n, depth):
    c = random_circuit(n, depth, haar_layers(n, depth, 100 + n))
    st = check(c, n, False, rand_state(n, n))
    assert st["sweeps"] < len(list(c.flat()))
    assert st["unfused"] == 0
    assert st["conflicted_transitions"] == 0  # every transposition got a bank-conflict-free slot map


def test_depth20_plan_uses_few_sweeps():
    n, d = 14, 20
    c = random_circuit(n, d, haar_layers(n, d, 7))
    st = check(c, n, False, rand_state(n, 1))
    assert st["sweeps"] <= 2 * d


def test_ansatz_with_cnot_and_rotations():
    n = 13
    rng = np.random.default_rng(3)
    c = hea_circuit(n, 3, rng.uniform(0, 2 * np.pi, 2 * n * 3))
    check(c, n, False, rand_state(n, 2))


def test_mixed_instruction_kinds():
    """Diagonals on tile/thread/global bits, non-±1 diagonals, 3-qubit gates
    (run unfused), permutations and a wide diagonal."""
    n = 14
    rng = np.random.default_rng(4)
    ops = [mq.QuantumGate((1,), haar_unitary(rng)), mq.CZGate(1, 14), mq.RzGate(13, 0.3),
           mq.CPHASEGate((2, 12), 0.7), mq.QuantumGate((3, 9, 11), haar_unitary(rng, 8)), mq.CNOTGate(14, 2),
           mq.SWAPGate(5, 13), mq.QuantumGate((6, 14), haar_unitary(rng, 4)), mq.TOFFOLIGate(1, 7, 12),
           mq.QuantumGate((4,), np.diag([1j, -1])), mq.CZGate(8, 9), mq.QuantumGate((8,), haar_unitary(rng)),
           mq.QuantumGate((9,), haar_unitary(rng)), mq.CZGate(8, 9), mq.ZGate(3)]
    c = mq.QCircuit(ops)
    st = check(c, n, False, rand_state(n, 3))
    assert st["unfused"] >= 1


def test_permutation_groups():
    """Two-bit permutations (CNOT either way, SWAP, X on one or both bits) run
    as register moves, alone and followed by controlled-Z sign flips; the
    planner keeps one-qubit gates out of them."""
    n = 13
    rng = np.random.default_rng(5)
    X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
    I2 = np.eye(2, dtype=np.complex128)
    ops = []
    for layer in range(3):
        for q in range(1, n + 1):
            ops.append(mq.QuantumGate((q,), haar_unitary(rng)))
        for q in range(1, n, 2):
            ops += [mq.CNOTGate(q, q + 1), mq.CNOTGate(q + 1, q), mq.SWAPGate(q, q + 1)]
        for q in range(2, n, 3):
            ops += [mq.QuantumGate((q, q + 1), np.kron(X, I2)), mq.QuantumGate((q, q + 1), np.kron(I2, X)),
                    mq.QuantumGate((q, q + 1), np.kron(X, X)), mq.CZGate(q, q + 1)]
        ops += [mq.CNOTGate(1, 2), mq.CZGate(1, 2), mq.CZGate(2, 5)]
    c = mq.QCircuit(ops)
    st = check(c, n, False, rand_state(n, 5))
    assert st["permutations"] > 0
    # forward and reverse (adjoint) plans of an ansatz keep CNOTs as permutations
    circ = hea_circuit(n, 2, rng.uniform(0, 2 * np.pi, 4 * n))
    _, st = plan_emulate(circ, n, False, None)
    assert st["permutations"] > 0
    _, _, _, st = plan_emulate_adjoint(circ, n, rand_state(n, 6), rand_state(n, 7))
    assert st["permutations"] >= n - 1


@pytest.mark.parametrize("n", [6, 7])
def test_density_programs(n):
    rng = np.random.default_rng(n)
    c = hea_circuit(n, 2, rng.uniform(0, 2 * np.pi, 4 * n), noise=(0.01, 0.02))
    c.append(mq.Depolarizing((1, 3), 0.05))  # two-qubit channel: unfused
    c.append(mq.CZGate(2, n))
    A = rng.standard_normal((1 << n, 1 << n)) + 1j * rng.standard_normal((1 << n, 1 << n))
    rho = A @ A.conj().T
    rho /= np.trace(rho)
    check(c, n, True, rho.reshape(-1, order="F"))


def test_small_registers_rejected():
    with pytest.raises(mq.MossqError):
        plan_emulate(mq.QCircuit([mq.XGate(1)]), 5, False, rand_state(5, 0))


def adjoint_inputs(circ, n, H):
    """Forward output y = C|0> and seed (H + H^dag) y = 2 H y, from the oracle."""
    y = np.zeros(1 << n, dtype=np.complex128)
    y[0] = 1
    oracle.run_program(y, lower(circ, n, False))
    return y, 2 * oracle.apply_op(y, n, H.groups())


def mixed_param_circuit(n, layers, seed):
    """Rotations of every kind, CZ and CNOT ladders, controlled rotations."""
    rng = np.random.default_rng(seed)
    c = mq.QCircuit()
    for layer in range(layers):
        for q in range(1, n + 1):
            c.append([mq.RxGate, mq.RyGate, mq.RzGate, mq.PHASEGate][(q + layer) % 4](q, rng.uniform(0, 6.3),
                                                                                    isparas=True))
        for q in range(1 + layer % 2, n, 2):
            c.append(mq.CZGate(q, q + 1) if layer % 3 else mq.CNOTGate(q, q + 1))
        c.append(mq.CRyGate((1, n), rng.uniform(0, 6.3), isparas=True))
        c.append(mq.QuantumGate((2, n - 1), haar_unitary(rng, 4)))
    return c


@pytest.mark.parametrize("n,layers", [(12, 3), (13, 2), (14, 2)])
def test_adjoint_plan_matches_oracle(n, layers):
    """The fused reverse plan (probes of Re<y|D|l> inside LDS tiles) gives the
    oracle's adjoint gradients, for the HEA ansatz and a mixed circuit."""
    from mossquantum_amd.backend import plan_emulate_adjoint
    from mossquantum_amd.workloads import tfim
    rng = np.random.default_rng(n)
    H = tfim(n)
    for circ in (hea_circuit(n, layers, rng.uniform(0, 2 * np.pi, 2 * n * layers)),
                 mixed_param_circuit(n, layers, n)):
        y, lam = adjoint_inputs(circ, n, H)
        _, want = oracle.adjoint(lower(circ, n, False), H.groups())
        y1, _, g, st = plan_emulate_adjoint(circ, n, y, lam)
        assert st["sweeps"] >= 1 and st["conflicted_transitions"] == 0
        assert np.max(np.abs(g - want)) < 1e-9
        # the reverse pass takes y back to the initial state |0...0>
        e0 = np.zeros(1 << n)
        e0[0] = 1
        assert np.max(np.abs(y1 - e0)) < 1e-12


def dense_operator(n, H):
    """H as a dense matrix, column by column from the oracle's apply_op."""
    dim = 1 << n
    M = np.zeros((dim, dim), dtype=np.complex128)
    for j in range(dim):
        e = np.zeros(dim, dtype=np.complex128)
        e[j] = 1
        M[:, j] = oracle.apply_op(e, n, H.groups())
    return M


def noisy_param_circuit(n, layers, seed):
    """The noisy HEA plus a parameterized two-qubit gate, a two-qubit channel
    (unfused), a dense two-qubit gate (row/column split) and a three-qubit gate
    (unfused row and column passes)."""
    rng = np.random.default_rng(seed)
    c = hea_circuit(n, layers, rng.uniform(0, 2 * np.pi, 2 * n * layers), noise=(0.01, 0.02))
    c.append(mq.CRyGate((1, n), rng.uniform(0, 6.3), isparas=True))
    c.append(mq.Depolarizing((2, 3), 0.05))
    c.append(mq.QuantumGate((2, n - 1), haar_unitary(rng, 4)))
    c.append(mq.RxGate(3, rng.uniform(0, 6.3), isparas=True))
    c.append(mq.QuantumGate((1, 3, n), haar_unitary(rng, 8)))
    c.append(mq.CZGate(2, n))
    c.append(mq.AmplitudeDamping(n, 0.1))
    c.append(mq.RzGate(n, rng.uniform(0, 6.3), isparas=True))
    c.append(mq.Depolarizing(1, 0.03))
    return c


@pytest.mark.parametrize("n,layers", [(6, 2), (7, 3)])
def test_density_adjoint_plan_matches_oracle(n, layers):
    """The fused density adjoint (chunk plans cut at the parameterized gates,
    checkpoints every `segment` parameters, recompute per segment) gives the
    oracle's gradients and final state, from |0><0| and from a mixed input."""
    from mossquantum_amd.backend import plan_emulate_density_adjoint
    from mossquantum_amd.workloads import tfim
    H = tfim(n)
    Hd = dense_operator(n, H)
    seed = Hd.conj().T.reshape(-1, order="F")  # vec(H^dagger)
    rng = np.random.default_rng(10 + n)
    A = rng.standard_normal((1 << n, 1 << n)) + 1j * rng.standard_normal((1 << n, 1 << n))
    mixed = A @ A.conj().T
    mixed /= np.trace(mixed)
    zero = np.zeros((1 << n, 1 << n), dtype=np.complex128)
    zero[0, 0] = 1
    circ = noisy_param_circuit(n, layers, n)
    lw = lower(circ, n, True)
    for rho0 in (zero, mixed):
        r0 = rho0.reshape(-1, order="F")
        _, want, want_state = oracle.adjoint(lw, H.groups(), state_in=r0, want_state=True)
        # reverse plans per parameter, and one per checkpoint segment (several
        # fixed probes per plan; the segment holding the two-qubit
        # parameterized gate falls back to per-parameter plans)
        for whole in (0, 1):
            rho, _, g, st = plan_emulate_density_adjoint(circ, n, r0, seed, whole)
            assert st["param_steps"] == lw.nderiv and st["unfused"] >= 2
            assert np.max(np.abs(rho - want_state)) < AMP
            assert np.max(np.abs(g - want)) < 1e-9, whole


def test_density_adjoint_plan_checkpoint_segments():
    """Segment boundaries of the checkpoint schedule: one parameter, a
    parameter first and last, adjacent parameters (empty reverse chunks)."""
    from mossquantum_amd.backend import plan_emulate_density_adjoint
    from mossquantum_amd.workloads import tfim
    n = 6
    H = tfim(n)
    seed = dense_operator(n, H).conj().T.reshape(-1, order="F")
    zero = np.zeros(1 << (2 * n), dtype=np.complex128)
    zero[0] = 1
    rng = np.random.default_rng(5)
    cases = [
        [mq.HGate(1), mq.RyGate(2, 0.4, isparas=True), mq.Depolarizing(2, 0.1), mq.CNOTGate(2, 3)],
        [mq.RyGate(1, 0.7, isparas=True), mq.RzGate(1, 0.2, isparas=True), mq.AmplitudeDamping(1, 0.2),
         mq.CNOTGate(1, 2), mq.RxGate(2, 1.1, isparas=True)],
        [mq.RyGate(q, rng.uniform(0, 6.3), isparas=True) for q in range(1, n + 1)] * 3 + [mq.Depolarizing(3, 0.2)],
    ]
    for ops in cases:
        circ = mq.QCircuit(ops)
        lw = lower(circ, n, True)
        _, want, want_state = oracle.adjoint(lw, H.groups(), state_in=zero, want_state=True)
        for whole in (0, 1):
            rho, _, g, st = plan_emulate_density_adjoint(circ, n, zero, seed, whole)
            assert st["param_steps"] == lw.nderiv
            assert np.max(np.abs(rho - want_state)) < AMP
            assert np.max(np.abs(g - want)) < 1e-9


def _launch_tables(ntiles):
    import ctypes as C
    from mossquantum_amd import _lib
    lib = _lib.load()
    nt = np.ascontiguousarray(ntiles, dtype=np.int32)
    total = int(nt.sum())
    items = np.zeros(8, dtype=np.int32)
    ent = np.zeros(max(total, 1), dtype=np.uint32)
    L, ne = C.c_int(0), C.c_uint64(0)
    _lib.check(lib.mossq_batch_launch_tables(nt.size, _lib.iptr(nt), items.size, _lib.iptr(items), ent.size,
                                             ent.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(L), C.byref(ne)))
    return items[:L.value], ent[:ne.value]


@pytest.mark.parametrize("m", [1, 32767, 32768, 65536, 65537, 131073])
def test_batched_launch_tables_pack_items_and_tiles(m):
    """The batched tile sweeps' workgroup tables (entry = item << 16 | tile,
    unsigned): launches hold at most 65536 items, every (item, tile) of the
    launch appears exactly once, round robin, and the item field decodes to
    the item inside its launch at 32767, 32768 and 65536 items."""
    rng = np.random.default_rng(m)
    nt = rng.choice([1, 2, 16, 64], size=m, p=[0.85, 0.1, 0.04, 0.01]).astype(np.int32)
    nt[m // 2] = 65536  # a 28-qubit item: its tile field uses all 16 bits
    items, ent = _launch_tables(nt)
    assert items.sum() == m and items.max() <= 65536 and len(items) == -(-m // 65536)
    assert ent.size == int(nt.sum())
    off, b0 = 0, 0
    for L in items:
        tiles = nt[b0:b0 + L]
        e = ent[off:off + int(tiles.sum())]
        item, tile = e >> 16, e & 0xFFFF
        assert item.max() == L - 1 and item.min() == 0
        # every (item, tile) once
        key = item.astype(np.int64) * 65536 + tile
        assert np.unique(key).size == e.size
        assert np.all(tile < tiles[item])
        # round robin: the first L entries are tile 0 of items 0..L-1
        assert np.array_equal(item[:L], np.arange(L)) and not tile[:L].any()
        off += e.size
        b0 += L


def test_batched_launch_tables_reject_oversized_items():
    import ctypes as C
    from mossquantum_amd import _lib
    lib = _lib.load()
    nt = np.array([65537], dtype=np.int32)  # a 29-qubit item does not fit the 16-bit tile field
    items = np.zeros(1, dtype=np.int32)
    ent = np.zeros(65537, dtype=np.uint32)
    L, ne = C.c_int(0), C.c_uint64(0)
    st = lib.mossq_batch_launch_tables(1, _lib.iptr(nt), 1, _lib.iptr(items), ent.size,
                                       ent.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(L), C.byref(ne))
    assert st == _lib.MOSSQ_EINVAL
