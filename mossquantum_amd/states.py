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

    def view(self, offset: int, size: int) -> "HIPArray":
        """Entries [offset, offset + size) of this buffer, not owning them
        (it keeps this buffer alive): the columns of a batch."""
        if offset < 0 or size < 0 or offset + size > self.size:
            raise ValueError("view outside the buffer")
        v = HIPArray.__new__(HIPArray)
        v.ctx, v.size, v.nbytes = self.ctx, int(size), int(size) * 16
        v.ptr = C.c_void_p(self.ptr.value + 16 * int(offset))
        v._owner = self
        return v

    def free(self):
        if getattr(self, "_owner", None) is not None:  # a view: the owner frees
            self.ptr = None
            return
        if getattr(self, "ptr", None) and self.ctx.h:
            check(self.ctx.lib.mossq_free(self.ctx.h, self.ptr))
        self.ptr = None

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


class _Batch:
    """B states of one qubit count in one device buffer, column b at entries
    [b L, (b + 1) L) (L = 2^n amplitudes, or 4^n for density matrices): the
    reference's (2^n, B) / (2^n, 2^n, B) storage (batch_state_vector.jl:18-30,
    batch_density_matrix.jl:17-28).  `columns` are views of that buffer."""

    _single = None  # StateVector or DensityMatrix
    _len_bits = 1   # L = 2^(_len_bits n)

    def __init__(self, columns_or_storage, n: int | None = None, B: int | None = None):
        if isinstance(columns_or_storage, HIPArray):
            storage = columns_or_storage
            if n is None or B is None or B < 1 or storage.size != B << (self._len_bits * n):
                raise ValueError("batch storage must hold B columns of the state size")
        else:
            cols = list(columns_or_storage)
            if not cols:
                raise ValueError("a batch needs at least one state")
            n = cols[0].n
            if any(not isinstance(c, self._single) or c.n != n for c in cols):
                raise ValueError("batch states must share a qubit count and kind")
            B = len(cols)
            L = 1 << (self._len_bits * n)
            storage = HIPArray(B * L, cols[0].storage.ctx)
            for b, c in enumerate(cols):
                check(storage.ctx.lib.mossq_copy_d2d(storage.ctx.h, C.c_void_p(storage.ptr.value + 16 * b * L),
                                                     c.storage.ptr, L * 16))
        self.storage, self.n, self.B = storage, int(n), int(B)
        L = 1 << (self._len_bits * self.n)
        self.columns = [self._single(storage.view(b * L, L), self.n) for b in range(self.B)]

    def batch_size(self):
        return self.B

    def nqubits(self):
        return self.n

    def copy(self):
        return type(self)(self.storage.copy(), self.n, self.B)


class BatchStateVector(_Batch):
    """B pure states of one qubit count, columns of one (2^n, B) device array
    (reference batch_state_vector.jl:18-80)."""

    _single = StateVector
    _len_bits = 1

    @classmethod
    def zeros(cls, n, B, ctx=None):
        s = HIPArray(B << n, ctx)
        out = cls(s, n, B)
        for c in out.columns:
            check(s.ctx.lib.mossq_sv_init_zero(s.ctx.h, c.storage.ptr, n))
        return out

    @classmethod
    def from_host(cls, amps, ctx=None):
        """From a (2^n, B) host array."""
        a = np.asarray(amps, dtype=np.complex128)
        n = int(round(np.log2(a.shape[0])))
        return cls(HIPArray(a.size, ctx).copy_from_host(a), n, a.shape[1])

    def to_host(self):
        return self.storage.to_host().reshape((1 << self.n, self.B), order="F")


class BatchDensityMatrix(_Batch):
    """B density matrices of one qubit count in one (2^n, 2^n, B) device
    array (reference batch_density_matrix.jl:17-134)."""

    _single = DensityMatrix
    _len_bits = 2

    @classmethod
    def zeros(cls, n, B, ctx=None):
        s = HIPArray(B << (2 * n), ctx)
        out = cls(s, n, B)
        for c in out.columns:
            check(s.ctx.lib.mossq_dm_init_zero(s.ctx.h, c.storage.ptr, n))
        return out

    @classmethod
    def from_states(cls, batch: BatchStateVector):
        """psi_b psi_b^dagger for every column (the promotion at a batch's first channel)."""
        s = HIPArray(batch.B << (2 * batch.n), batch.storage.ctx)
        out = cls(s, batch.n, batch.B)
        for c, sv in zip(out.columns, batch.columns):
            check(s.ctx.lib.mossq_dm_from_sv(s.ctx.h, c.storage.ptr, sv.storage.ptr, batch.n))
        return out

    def to_host(self):
        N = 1 << self.n
        return self.storage.to_host().reshape((N, N, self.B), order="F")
