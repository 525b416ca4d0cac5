"""Parity oracle for the HIP backend -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may import
this package, and only as the checker / the timed CPU baseline.  The product
(mossquantum_amd, libmossq_hip.so) never imports or links it.

Two independent restatements live here:
  * cpu_ref.c (via ctypes, `Ref`): the reference CPU backend's algorithms --
    per-gate block sweeps, two-pass density gates, supermatrix channels,
    per-group expectations, the adjoint reverse pass with per-channel
    checkpoints -- each function citing the reference file it follows;
  * dense.py: textbook dense linear algebra (full Kronecker embeddings and
    tensor contractions) used to pin cpu_ref on small registers.

Parity status of each function is listed in dense.py and the tests.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "libmossq_oracle.so"


class RefOp(C.Structure):
    _fields_ = [("kind", C.c_int), ("k", C.c_int), ("pos", C.c_int * 5), ("mat", C.c_int), ("dmat", C.c_int),
                ("reserved", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            from . import build
            build.build()
        L = C.CDLL(str(LIB))
        P = C.c_void_p
        I = C.c_int
        DP = C.POINTER(C.c_double)
        IP = C.POINTER(C.c_int)
        sig = {
            "ref_num_threads": (I, []),
            "ref_set_threads": (None, [I]),
            "ref_apply_gate": (None, [P, I, I, IP, P]),
            "ref_apply_cz": (None, [P, I, I, I]),
            "ref_dm_apply_gate": (None, [P, I, I, IP, P]),
            "ref_supermatrix": (None, [I, I, P, P]),
            "ref_dm_apply_channel": (None, [P, I, I, IP, P, I]),
            "ref_cross_expect_groups": (None, [P, P, I, I, IP, IP, P, DP]),
            "ref_expect_groups": (None, [P, I, I, IP, IP, P, DP]),
            "ref_apply_op": (None, [P, P, I, I, IP, IP, P]),
            "ref_dm_expect_groups": (None, [P, I, I, IP, IP, P, DP]),
            "ref_run_program": (None, [P, I, I, I, C.POINTER(RefOp), P]),
            "ref_adjoint": (I, [I, I, I, C.POINTER(RefOp), P, I, IP, IP, P, P, P, DP, DP]),
            "ref_probabilities": (None, [P, I, I, I, IP, DP]),
            "ref_expect_groups_compensated": (None, [P, I, I, IP, IP, P, DP]),
            "ref_energy_ld": (I, [I, I, C.POINTER(RefOp), P, I, IP, IP, P, P]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _p(a):
    if not a.flags.c_contiguous:
        raise ValueError("the oracle reads contiguous arrays (pass np.ascontiguousarray of a view)")
    return a.ctypes.data_as(C.c_void_p)


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def set_threads(n: int):
    lib().ref_set_threads(int(n))


def num_threads() -> int:
    return lib().ref_num_threads()


def groups_arrays(groups):
    """[(sites 1-based, matrix)] -> (ks, sites 0-based, mats column-major)."""
    ks = np.array([len(k) for k, _ in groups], dtype=np.int32)
    sites = np.array([p - 1 for k, _ in groups for p in k] or [0], dtype=np.int32)
    mats = np.ascontiguousarray(np.concatenate([m.reshape(-1, order="F") for _, m in groups])).astype(np.complex128)
    return ks, sites, mats


def apply_gate(v, nbits, positions0, U):
    pos = np.array(positions0, dtype=np.int32)
    U = np.ascontiguousarray(np.asarray(U, dtype=np.complex128).reshape(-1, order="F"))
    lib().ref_apply_gate(_p(v), nbits, len(pos), _ip(pos), _p(U))


def dm_apply_gate(rho, n, positions0, U):
    pos = np.array(positions0, dtype=np.int32)
    U = np.ascontiguousarray(np.asarray(U, dtype=np.complex128).reshape(-1, order="F"))
    lib().ref_dm_apply_gate(_p(rho), n, len(pos), _ip(pos), _p(U))


def dm_apply_channel(rho, n, positions0, S, dagger=False):
    pos = np.array(positions0, dtype=np.int32)
    S = np.ascontiguousarray(np.asarray(S, dtype=np.complex128).reshape(-1, order="F"))
    lib().ref_dm_apply_channel(_p(rho), n, len(pos), _ip(pos), _p(S), int(dagger))


def supermatrix(k, kraus):
    ks = np.ascontiguousarray(np.concatenate([np.asarray(K, dtype=np.complex128).reshape(-1, order="F")
                                              for K in kraus]))
    L = 1 << (2 * k)
    S = np.zeros(L * L, dtype=np.complex128)
    lib().ref_supermatrix(k, len(kraus), _p(ks), _p(S))
    return S.reshape((L, L), order="F")


def expect(psi, n, groups):
    ks, sites, mats = groups_arrays(groups)
    out = np.zeros(2)
    lib().ref_expect_groups(_p(psi), n, len(ks), _ip(ks), _ip(sites), _p(mats), _dp(out))
    return complex(out[0], out[1])


def cross_expect(phi, psi, n, groups):
    ks, sites, mats = groups_arrays(groups)
    out = np.zeros(2)
    lib().ref_cross_expect_groups(_p(phi), _p(psi), n, len(ks), _ip(ks), _ip(sites), _p(mats), _dp(out))
    return complex(out[0], out[1])


def apply_op(psi, nbits, groups):
    ks, sites, mats = groups_arrays(groups)
    out = np.zeros_like(psi)
    lib().ref_apply_op(_p(out), _p(psi), nbits, len(ks), _ip(ks), _ip(sites), _p(mats))
    return out


def dm_expect(rho_vec, n, groups):
    ks, sites, mats = groups_arrays(groups)
    out = np.zeros(2)
    lib().ref_dm_expect_groups(_p(rho_vec), n, len(ks), _ip(ks), _ip(sites), _p(mats), _dp(out))
    return complex(out[0], out[1])


def _records(lowered):
    recs = (RefOp * max(len(lowered.records), 1))()
    for i, r in enumerate(lowered.records):
        recs[i].kind, recs[i].k, recs[i].mat, recs[i].dmat = r.kind, r.k, r.mat, r.dmat
        for j in range(5):
            recs[i].pos[j] = r.pos[j]
    pool = np.ascontiguousarray(np.concatenate(lowered.pool) if lowered.pool else np.zeros(1, np.complex128))
    return recs, pool


def run_program(state, lowered):
    """Apply a lowered circuit (mossquantum_amd.circuit.Lowered) to a host vector in place."""
    recs, pool = _records(lowered)
    lib().ref_run_program(_p(state), lowered.n, int(lowered.is_density), len(lowered.records), recs, _p(pool))


def adjoint(lowered, groups, state_in=None, want_state=False):
    """(value, device-order gradients[, final state]) of real(<H>) (cpu_ref ref_adjoint)."""
    recs, pool = _records(lowered)
    ks, sites, mats = groups_arrays(groups)
    n = lowered.n
    length = 1 << (2 * n if lowered.is_density else n)
    out_state = np.zeros(length, dtype=np.complex128) if want_state else None
    inp = None if state_in is None else np.ascontiguousarray(state_in, dtype=np.complex128)
    grads = np.zeros(max(lowered.nderiv, 1))
    value = np.zeros(1)
    st = lib().ref_adjoint(n, int(lowered.is_density), len(lowered.records), recs, _p(pool), len(ks), _ip(ks),
                           _ip(sites), _p(mats), None if inp is None else _p(inp),
                           None if out_state is None else _p(out_state), _dp(value), _dp(grads))
    if st != 0:
        raise MemoryError("oracle adjoint allocation failed")
    g = grads[:lowered.nderiv]
    return (value[0], g, out_state) if want_state else (value[0], g)


def probabilities(state, n, is_dm, qubits1):
    q = np.array([x - 1 for x in qubits1], dtype=np.int32)
    out = np.zeros(1 << len(q))
    lib().ref_probabilities(_p(state), n, int(is_dm), len(q), _ip(q), _dp(out))
    return out


def expect_compensated(psi, n, groups):
    """<psi|H|psi> with compensated (Neumaier) sums (cpu_ref ref_expect_groups_compensated)."""
    ks, sites, mats = groups_arrays(groups)
    out = np.zeros(2)
    lib().ref_expect_groups_compensated(_p(psi), n, len(ks), _ip(ks), _ip(sites), _p(mats), _dp(out))
    return complex(out[0], out[1])


_ROTATIONS = ("RxGate", "RyGate", "RzGate", "PHASEGate")


def _pool_ld(lowered, slot_values):
    """The program's matrix pool in extended precision, every rotation matrix
    (RX, RY, RZ, PHASE) rebuilt from its angle in extended precision."""
    pool = np.concatenate(lowered.pool).astype(np.clongdouble)
    th = np.asarray(slot_values, dtype=np.longdouble)
    for _, op, s0, off in lowered.param_gates:
        kind = type(op).__name__
        if kind not in _ROTATIONS or op.nqubits() != 1:
            raise ValueError(f"extended-precision energies take one-qubit rotations, not {kind}")
        t = th[s0]
        c, s = np.cos(t / 2), np.sin(t / 2)
        j = np.clongdouble(1j)
        if kind == "RxGate":  # column-major [[c, -is], [-is, c]]
            vals = (c, -j * s, -j * s, c)
        elif kind == "RyGate":
            vals = (c, s, -s, c)
        elif kind == "RzGate":
            vals = (c - j * s, 0, 0, c + j * s)
        else:
            vals = (1, 0, 0, np.cos(t) + j * np.sin(t))
        for q, x in enumerate(vals):
            pool[off + q] = x
    return np.ascontiguousarray(pool)


def energy_ld(lowered, slot_values, groups):
    """Re<H> of a state-vector program from |0...0> in extended precision
    (cpu_ref ref_energy_ld), as np.longdouble."""
    recs, _ = _records(lowered)
    pool = _pool_ld(lowered, slot_values)
    ks, sites, mats = groups_arrays(groups)
    out = np.zeros(2, dtype=np.longdouble)
    st = lib().ref_energy_ld(lowered.n, len(lowered.records), recs, _p(pool), len(ks), _ip(ks), _ip(sites),
                             _p(mats), _p(out))
    if st != 0:
        raise ValueError("extended-precision energies take gate-only state-vector programs")
    return out[0]


def central_fd(lowered, slot_values, groups, slots, h=1e-6, extended=True):
    """Central finite differences (E(theta + h e_j) - E(theta - h e_j)) / 2h of
    Re<H> for the active slots `slots` of a state-vector program from
    |0...0> (reference test/support/finite_differences.jl:8-46 with h = 1e-6).
    extended=True evaluates energies and angles in extended precision;
    extended=False in double, with compensated expectation sums and the
    quotient taken over the steps actually represented, (theta_j + h) - theta_j
    and theta_j - (theta_j - h)."""
    th = np.asarray(slot_values, dtype=np.float64)
    out = np.zeros(len(slots))
    for i, j in enumerate(slots):
        if extended:
            t = th.astype(np.longdouble)
            hh = np.longdouble(h)
            tp, tm = t.copy(), t.copy()
            tp[j] += hh
            tm[j] -= hh
            out[i] = float((energy_ld(lowered, tp, groups) - energy_ld(lowered, tm, groups)) / (2 * hh))
        else:
            tp, tm = th.copy(), th.copy()
            tp[j] += h
            tm[j] -= h
            e = []
            for t in (tp, tm):
                psi = np.zeros(1 << lowered.n, dtype=np.complex128)
                psi[0] = 1
                pool = lowered.pool_with_slots(t)
                _run_with_pool(psi, lowered, pool)
                e.append(expect_compensated(psi, lowered.n, groups).real)
            out[i] = (e[0] - e[1]) / ((tp[j] - th[j]) + (th[j] - tm[j]))
    return out


def _run_with_pool(state, lowered, pool):
    recs, _ = _records(lowered)
    pool = np.ascontiguousarray(pool, dtype=np.complex128)
    lib().ref_run_program(_p(state), lowered.n, int(lowered.is_density), len(lowered.records), recs, _p(pool))
