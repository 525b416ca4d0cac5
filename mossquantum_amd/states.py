"""Device-resident state containers (reference src/states/).

`HIPArray` is the storage family: a device buffer owned through a context.  Like
the reference's device-backend model (test/fixtures/MOSSQuantumDemoArrayBackend:
no scalar indexing, no implicit Array conversion), it refuses implicit host
access; `to_host()` is the one explicit download.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check


class HIPArray:
    """complex128 device buffer of `size` elements (independent allocations never alias)."""

    def __init__(self, size: int, ctx: _lib.Context | None = None):
        self.ctx = ctx or _lib.default_context()
        self.size = int(size)
        self.nbytes = self.size * 16
        p = C.c_void_p()
        check(self.ctx.lib.mossq_alloc(self.ctx.h, self.nbytes, C.byref(p)))
        self.ptr = p

    def free(self):
        if getattr(self, "ptr", None) and self.ctx.h:
            check(self.ctx.lib.mossq_free(self.ctx.h, self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    # explicit transfers only
    def to_host(self) -> np.ndarray:
        out = np.empty(self.size, dtype=np.complex128)
        check(self.ctx.lib.mossq_copy_d2h(self.ctx.h, out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes))
        return out

    def copy_from_host(self, a) -> "HIPArray":
        a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128).reshape(-1, order="F"))
        if a.size != self.size:
            raise ValueError("size mismatch")
        check(self.ctx.lib.mossq_copy_h2d(self.ctx.h, self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes))
        return self

    def copy(self) -> "HIPArray":
        out = HIPArray(self.size, self.ctx)
        check(self.ctx.lib.mossq_copy_d2d(self.ctx.h, out.ptr, self.ptr, self.nbytes))
        return out

    def similar(self) -> "HIPArray":
        return HIPArray(self.size, self.ctx)

    def __array__(self, *a, **k):
        raise TypeError("HIPArray has no implicit host conversion; call to_host()")

    def __getitem__(self, i):
        raise TypeError("HIPArray does not support scalar indexing")


class StateVector:
    """Pure state of n qubits: 2^n amplitudes, little-endian (reference state_vector.jl:12-47)."""

    def __init__(self, storage: HIPArray, n: int):
        if storage.size != 1 << n:
            raise ValueError("storage length must be 2^n")
        self.storage = storage
        self.n = n

    @classmethod
    def zeros(cls, n: int, ctx=None):
        s = HIPArray(1 << n, ctx)
        check(s.ctx.lib.mossq_sv_init_zero(s.ctx.h, s.ptr, n))
        return cls(s, n)

    @classmethod
    def from_host(cls, amps, ctx=None):
        amps = np.asarray(amps, dtype=np.complex128)
        n = int(round(np.log2(amps.size)))
        if amps.size != 1 << n:
            raise ValueError("length must be a power of two")
        return cls(HIPArray(amps.size, ctx).copy_from_host(amps), n)

    def nqubits(self):
        return self.n

    def copy(self):
        return StateVector(self.storage.copy(), self.n)

    def to_host(self):
        return self.storage.to_host()


class DensityMatrix:
    """Mixed state: vec(rho), rho[i,j] at i + 2^n j (reference density_matrix.jl:12-106)."""

    def __init__(self, storage: HIPArray, n: int):
        if storage.size != 1 << (2 * n):
            raise ValueError("storage length must be 4^n")
        self.storage = storage
        self.n = n

    @classmethod
    def zeros(cls, n: int, ctx=None):
        s = HIPArray(1 << (2 * n), ctx)
        check(s.ctx.lib.mossq_dm_init_zero(s.ctx.h, s.ptr, n))
        return cls(s, n)

    @classmethod
    def from_state(cls, sv: StateVector):
        s = HIPArray(1 << (2 * sv.n), sv.storage.ctx)
        check(s.ctx.lib.mossq_dm_from_sv(s.ctx.h, s.ptr, sv.storage.ptr, sv.n))
        return cls(s, sv.n)

    @classmethod
    def from_host(cls, rho, ctx=None):
        rho = np.asarray(rho, dtype=np.complex128)
        n = int(round(np.log2(rho.shape[0])))
        return cls(HIPArray(rho.size, ctx).copy_from_host(rho.reshape(-1, order="F")), n)

    def nqubits(self):
        return self.n

    def copy(self):
        return DensityMatrix(self.storage.copy(), self.n)

    def to_host(self):
        return self.storage.to_host().reshape((1 << self.n, 1 << self.n), order="F")


class BatchStateVector:
    """B pure states of one qubit count, columns of a (2^n, B) array (reference batch_state_vector.jl:18-80)."""

    def __init__(self, columns: list[StateVector]):
        if not columns:
            raise ValueError("a batch needs at least one state")
        n = columns[0].n
        if any(c.n != n for c in columns):
            raise ValueError("batch states must share a qubit count")
        self.columns = columns
        self.n = n

    @classmethod
    def zeros(cls, n, B, ctx=None):
        return cls([StateVector.zeros(n, ctx) for _ in range(B)])

    def batch_size(self):
        return len(self.columns)

    def copy(self):
        return BatchStateVector([c.copy() for c in self.columns])

    def to_host(self):
        return np.stack([c.to_host() for c in self.columns], axis=1)


class BatchDensityMatrix:
    """B density matrices of one qubit count (reference batch_density_matrix.jl:17-134)."""

    def __init__(self, columns: list[DensityMatrix]):
        if not columns:
            raise ValueError("a batch needs at least one state")
        n = columns[0].n
        if any(c.n != n for c in columns):
            raise ValueError("batch states must share a qubit count")
        self.columns = columns
        self.n = n

    @classmethod
    def zeros(cls, n, B, ctx=None):
        return cls([DensityMatrix.zeros(n, ctx) for _ in range(B)])

    def batch_size(self):
        return len(self.columns)

    def copy(self):
        return BatchDensityMatrix([c.copy() for c in self.columns])

    def to_host(self):
        return np.stack([c.to_host() for c in self.columns], axis=2)
