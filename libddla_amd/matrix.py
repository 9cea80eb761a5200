"""Device matrix with a ScaLAPACK descriptor (mirrors python/ddla/matrix.py).

On the 1 x 1 grid the local matrix is the whole matrix: column-major, leading
dimension LLD = desc[8] >= max(1, m).  Memory comes from ddla_c_malloc on the
handle (device memory), or wraps a caller-owned device buffer (for example a
torch tensor) without copying.
"""
import ctypes

import numpy as np

from ._capi import LLD_, check_status, lib

_TYPE_CHARS = {np.dtype(np.float32): "s", np.dtype(np.float64): "d", np.dtype(np.complex64): "c",
               np.dtype(np.complex128): "z"}


def descinit(handle, m, n, mb, nb, rsrc=0, csrc=0, lld=-1) -> np.ndarray:
    desc = np.zeros(9, dtype=np.int32)
    check_status(lib.ddla_c_descinit(desc.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), handle.ptr, m, n, mb, nb,
                                     rsrc, csrc, lld), "ddla_c_descinit")
    return desc


class DistributedMatrix:
    def __init__(self, handle, m: int, n: int, mb: int = 64, nb: int = 64, dtype=np.float64, lld: int = -1,
                 rsrc: int = 0, csrc: int = 0, data_ptr: int = None, owner=None):
        self.handle = handle
        self.dtype = np.dtype(dtype)
        if self.dtype not in _TYPE_CHARS:
            raise TypeError(f"unsupported dtype {self.dtype}; float32, float64, complex64 and complex128 are built")
        self.type_char = _TYPE_CHARS[self.dtype]
        self.m, self.n, self.mb, self.nb = m, n, mb, nb
        self.desc = descinit(handle, m, n, mb, nb, rsrc, csrc, lld)
        self.lld = int(self.desc[LLD_])
        self.nbytes = self.lld * n * self.dtype.itemsize
        self._owner = owner
        if data_ptr is not None:
            self._ptr = ctypes.c_void_p(data_ptr)
            self._owned = False
        else:
            self._ptr = ctypes.c_void_p()
            check_status(lib.ddla_c_malloc(handle.ptr, ctypes.byref(self._ptr), self.nbytes), "ddla_c_malloc")
            self._owned = True

    def free(self):
        if getattr(self, "_owned", False) and self._ptr.value:
            lib.ddla_c_free(self.handle.ptr, self._ptr)
            self._ptr = ctypes.c_void_p()
            self._owned = False

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def data_ptr(self) -> int:
        return self._ptr.value or 0

    @property
    def desc_ptr(self):
        return self.desc.ctypes.data_as(ctypes.POINTER(ctypes.c_int))

    @property
    def shape(self):
        return (self.m, self.n)

    def to_host(self) -> np.ndarray:
        """The whole local buffer, LLD x n, Fortran order (padding included)."""
        out = np.empty((self.lld, self.n), dtype=self.dtype, order="F")
        if self.nbytes:
            check_status(lib.ddla_c_memcpy(self.handle.ptr, out.ctypes.data, self._ptr, self.nbytes), "ddla_c_memcpy")
            self.handle.synchronize()
        return out

    def from_host(self, host: np.ndarray):
        """Upload an LLD x n (or m x n, padded with zeros) host array."""
        buf = np.zeros((self.lld, self.n), dtype=self.dtype, order="F")
        buf[: host.shape[0], : host.shape[1]] = host
        if self.nbytes:
            check_status(lib.ddla_c_memcpy(self.handle.ptr, self._ptr, buf.ctypes.data, self.nbytes), "ddla_c_memcpy")
            self.handle.synchronize()
        return self

    def upload_raw(self, full: np.ndarray):
        """Upload an LLD x n Fortran array verbatim (padding included)."""
        full = np.asfortranarray(full, dtype=self.dtype)
        assert full.shape == (self.lld, self.n)
        if self.nbytes:
            check_status(lib.ddla_c_memcpy(self.handle.ptr, self._ptr, full.ctypes.data, self.nbytes), "ddla_c_memcpy")
            self.handle.synchronize()
        return self

    def to_global_matrix(self) -> np.ndarray:
        return np.array(self.to_host()[: self.m, : self.n])

    @classmethod
    def from_global_matrix(cls, handle, arr: np.ndarray, mb: int = 64, nb: int = 64, lld: int = -1):
        arr = np.asarray(arr)
        mat = cls(handle, arr.shape[0], arr.shape[1], mb, nb, arr.dtype, lld)
        return mat.from_host(arr)

    @classmethod
    def from_torch(cls, handle, tensor, m: int, n: int, mb: int = 64, nb: int = 64):
        """Wrap a contiguous torch tensor of shape (n, lld) as an m x n column-major matrix (no copy)."""
        import torch

        if not tensor.is_contiguous() or tensor.dim() != 2 or tensor.shape[0] != n:
            raise ValueError("expected a contiguous (n, lld) tensor holding a column-major matrix")
        dtype = {torch.float32: np.float32, torch.float64: np.float64, torch.complex64: np.complex64,
                 torch.complex128: np.complex128}[tensor.dtype]
        return cls(handle, m, n, mb, nb, dtype, lld=tensor.shape[1], data_ptr=tensor.data_ptr(), owner=tensor)

    def __repr__(self):
        return f"DistributedMatrix({self.m}x{self.n}, {self.dtype}, lld={self.lld})"
