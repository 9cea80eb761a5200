"""Batched SPD solves of uneven order over the C ABI (ddla_{s,d}_pposv_vbatched).

VBatch packs the matrices A_i (order n_i) and right-hand sides B_i (n_i x
nrhs) of a batch in one device arena, column-major with leading dimension
n_i + pad, and keeps the device arrays of orders, leading dimensions and
pointers that the library takes, plus the per-matrix info and log-determinant
outputs.  Every compute call runs the hand-written gfx950 kernels.
"""
import ctypes

import numpy as np

from ._capi import char, check_status, lib

_TYPE = {np.dtype(np.float64): "d", np.dtype(np.float32): "s"}


class _Device:
    """One device allocation owned by a handle."""

    def __init__(self, handle, nbytes):
        self.handle = handle
        self.ptr = ctypes.c_void_p()
        check_status(lib.ddla_c_malloc(handle.ptr, ctypes.byref(self.ptr), max(int(nbytes), 8)), "ddla_c_malloc")

    def upload(self, host: np.ndarray, offset: int = 0):
        host = np.ascontiguousarray(host)
        if host.nbytes:
            check_status(lib.ddla_c_memcpy(self.handle.ptr, ctypes.c_void_p(self.ptr.value + offset),
                                           ctypes.c_void_p(host.ctypes.data), host.nbytes), "ddla_c_memcpy")

    def download(self, out: np.ndarray, offset: int = 0):
        if out.nbytes:
            check_status(lib.ddla_c_memcpy(self.handle.ptr, ctypes.c_void_p(out.ctypes.data),
                                           ctypes.c_void_p(self.ptr.value + offset), out.nbytes), "ddla_c_memcpy")
        return out

    def free(self):
        if self.ptr.value:
            lib.ddla_c_free(self.handle.ptr, self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class VBatch:
    """A device-resident batch of n_i x n_i matrices and n_i x nrhs right-hand sides."""

    def __init__(self, handle, sizes, nrhs: int = 1, dtype=np.float64, pad: int = 0):
        self.handle = handle
        self.dtype = np.dtype(dtype)
        if self.dtype not in _TYPE:
            raise TypeError("float32 and float64 are built")
        self.type_char = _TYPE[self.dtype]
        self.sizes = np.asarray(sizes, dtype=np.int32)
        self.batch = int(self.sizes.size)
        self.nrhs = int(nrhs)
        self.ld = np.maximum(1, self.sizes + pad).astype(np.int32)
        es = self.dtype.itemsize
        a_elems = self.ld.astype(np.int64) * np.maximum(self.sizes, 0)
        b_elems = self.ld.astype(np.int64) * self.nrhs
        self.a_off = np.concatenate([[0], np.cumsum(a_elems)[:-1]]).astype(np.int64) if self.batch else np.zeros(0)
        self.b_off = np.concatenate([[0], np.cumsum(b_elems)[:-1]]).astype(np.int64) if self.batch else np.zeros(0)
        self.A = _Device(handle, int(a_elems.sum()) * es)
        self.B = _Device(handle, int(b_elems.sum()) * es)
        self.meta = _Device(handle, 8 * 4 * max(self.batch, 1))
        self.ptrs = _Device(handle, 8 * 2 * max(self.batch, 1))
        self.info = _Device(handle, 4 * max(self.batch, 1))
        self.logdet = _Device(handle, 8 * max(self.batch, 1))
        meta = np.concatenate([self.sizes, self.ld, self.ld]).astype(np.int32)
        self.meta.upload(meta)
        ptrs = np.concatenate([self.A.ptr.value + self.a_off * es, self.B.ptr.value + self.b_off * es]).astype(
            np.uint64)
        self.ptrs.upload(ptrs)
        b4 = 4 * self.batch
        base = self.meta.ptr.value
        self.d_n, self.d_lda, self.d_ldb = (ctypes.c_void_p(base), ctypes.c_void_p(base + b4),
                                            ctypes.c_void_p(base + 2 * b4))
        self.d_A = ctypes.c_void_p(self.ptrs.ptr.value)
        self.d_B = ctypes.c_void_p(self.ptrs.ptr.value + 8 * self.batch)
        handle.synchronize()

    # host <-> device
    def set_matrix(self, i: int, a: np.ndarray):
        n, ld = int(self.sizes[i]), int(self.ld[i])
        full = np.zeros((n, ld), dtype=self.dtype)  # column-major: row j of `full` is column j
        full[:, :n] = np.asarray(a, dtype=self.dtype).T
        self.A.upload(full, int(self.a_off[i]) * self.dtype.itemsize)

    def set_rhs(self, i: int, b: np.ndarray):
        n, ld = int(self.sizes[i]), int(self.ld[i])
        full = np.zeros((self.nrhs, ld), dtype=self.dtype)
        full[:, :n] = np.asarray(b, dtype=self.dtype).reshape(n, self.nrhs).T
        self.B.upload(full, int(self.b_off[i]) * self.dtype.itemsize)

    def set_raw(self, i: int, a_full: np.ndarray):
        """Upload matrix i's whole ld x n column-major storage (padding included)."""
        self.A.upload(np.asfortranarray(a_full, dtype=self.dtype).T, int(self.a_off[i]) * self.dtype.itemsize)

    def get_raw(self, i: int) -> np.ndarray:
        n, ld = int(self.sizes[i]), int(self.ld[i])
        out = np.empty((n, ld), dtype=self.dtype)
        self.handle.synchronize()
        return self.A.download(out, int(self.a_off[i]) * self.dtype.itemsize).T

    def get_matrix(self, i: int) -> np.ndarray:
        n = int(self.sizes[i])
        return np.array(self.get_raw(i)[:n, :n])

    def get_rhs(self, i: int) -> np.ndarray:
        n, ld = int(self.sizes[i]), int(self.ld[i])
        out = np.empty((self.nrhs, ld), dtype=self.dtype)
        self.handle.synchronize()
        return np.array(self.B.download(out, int(self.b_off[i]) * self.dtype.itemsize).T[:n])

    def get_info(self) -> np.ndarray:
        self.handle.synchronize()
        return self.info.download(np.empty(self.batch, dtype=np.int32))

    def get_logdet(self) -> np.ndarray:
        self.handle.synchronize()
        return self.logdet.download(np.empty(self.batch, dtype=np.float64))

    def generate(self, seed: int, delta: float):
        """A_i = G_i G_i^T / n_i + delta I and B_i ~ N(0, 1) on the GPU (Philox streams 2i and 2i + 1)."""
        es = self.dtype.itemsize
        a_ptrs = (ctypes.c_void_p * max(self.batch, 1))(*[int(self.A.ptr.value + o * es) for o in self.a_off])
        b_ptrs = (ctypes.c_void_p * max(self.batch, 1))(*[int(self.B.ptr.value + o * es) for o in self.b_off])
        n = (ctypes.c_int * max(self.batch, 1))(*[int(x) for x in self.sizes])
        ld = (ctypes.c_int * max(self.batch, 1))(*[int(x) for x in self.ld])
        work = _Device(self.handle, 16 * max(1, self.max_n) ** 2)
        st = lib.ddla_c_spd_generate_batch(self.handle.ptr, self.batch, n, seed, delta,
                                           1 if self.dtype == np.float32 else 0, a_ptrs, ld, self.nrhs, b_ptrs, ld,
                                           ctypes.c_void_p(work.ptr.value))
        check_status(st, "ddla_c_spd_generate_batch")
        self.handle.synchronize()
        work.free()

    def _fn(self, name):
        return getattr(lib, f"ddla_{self.type_char}_{name}")

    def _common(self, uplo):
        return (self.handle.ptr, char(uplo), self.d_n, self.d_A, self.d_lda, self.nrhs, self.d_B, self.d_ldb)

    @property
    def max_n(self) -> int:
        return int(self.sizes.max()) if self.batch else 0


def pposv_vbatched(vb: VBatch, uplo: str = "L", max_n: int = None):
    """Fused potrf + potrs + logdet of every member (asynchronous)."""
    max_n = vb.max_n if max_n is None else max_n
    st = vb._fn("pposv_vbatched")(*vb._common(uplo), ctypes.c_void_p(vb.info.ptr.value),
                                  ctypes.c_void_p(vb.logdet.ptr.value), vb.batch, max_n)
    check_status(st, f"ddla_{vb.type_char}_pposv_vbatched")


def pposv_vbatched_workspace_size(vb: VBatch, max_n: int = None) -> int:
    max_n = vb.max_n if max_n is None else max_n
    b = ctypes.c_longlong()
    check_status(vb._fn("pposv_vbatched_workspace_size")(vb.handle.ptr, vb.batch, max_n, ctypes.byref(b)),
                 "pposv_vbatched_workspace_size")
    return b.value


def pposv_vbatched_run(vb: VBatch, work_ptr: int, work_bytes: int, uplo: str = "L", max_n: int = None):
    max_n = vb.max_n if max_n is None else max_n
    st = vb._fn("pposv_vbatched_run")(*vb._common(uplo), ctypes.c_void_p(vb.info.ptr.value),
                                      ctypes.c_void_p(vb.logdet.ptr.value), vb.batch, max_n,
                                      ctypes.c_void_p(work_ptr), work_bytes)
    check_status(st, f"ddla_{vb.type_char}_pposv_vbatched_run")


def pposv_vbatched_tune(vb: VBatch, uplo: str = "L", max_n: int = None):
    """Times the schedules on copies of the batch; returns (workgroups per CU, largest first)."""
    max_n = vb.max_n if max_n is None else max_n
    occ, ordered = ctypes.c_int(), ctypes.c_int()
    st = vb._fn("pposv_vbatched_tune")(*vb._common(uplo), vb.batch, max_n, ctypes.byref(occ), ctypes.byref(ordered))
    check_status(st, f"ddla_{vb.type_char}_pposv_vbatched_tune")
    return occ.value, bool(ordered.value)


def vbatched_schedule(handle, dtype=np.float64):
    """The batched schedule the handle holds for dtype: {"tuned", "workgroups_per_cu", "largest_first",
    "split_blocks", "left_looking", "step_columns"} (members of more than split_blocks 64-row blocks take the
    blocked path with gemmVbatched trailing updates, 16 meaning none, in steps of step_columns = 64 or 128
    columns; left_looking: every member runs on the left-looking persistent kernel instead, and split_blocks
    is unused)."""
    t, occ, ordered, split, kernel = (ctypes.c_int() for _ in range(5))
    item = np.dtype(dtype).itemsize
    check_status(lib.ddla_c_vbatched_schedule(handle.ptr, item, ctypes.byref(t), ctypes.byref(occ),
                                              ctypes.byref(ordered), ctypes.byref(split)),
                 "ddla_c_vbatched_schedule")
    check_status(lib.ddla_c_vbatched_kernel(handle.ptr, item, ctypes.byref(kernel)), "ddla_c_vbatched_kernel")
    return {"tuned": bool(t.value), "workgroups_per_cu": occ.value, "largest_first": bool(ordered.value),
            "split_blocks": split.value, "left_looking": kernel.value == 1,
            "step_columns": 128 if kernel.value == 2 else 64}


def set_vbatched_schedule(handle, dtype=np.float64, workgroups_per_cu: int = 1, largest_first: bool = True,
                          split_blocks: int = 16, left_looking: bool = False, step_columns: int = 64):
    """Pins the batched schedule for dtype without timing (see vbatched_schedule)."""
    if step_columns not in (64, 128) or (left_looking and step_columns != 64):
        raise ValueError("step_columns is 64 or 128, and 64 with the left-looking kernel")
    item = np.dtype(dtype).itemsize
    check_status(lib.ddla_c_vbatched_set_schedule(handle.ptr, item, workgroups_per_cu, 1 if largest_first else 0,
                                                  split_blocks),
                 "ddla_c_vbatched_set_schedule")
    kernel = 1 if left_looking else (2 if step_columns == 128 else 0)
    if kernel:
        check_status(lib.ddla_c_vbatched_set_kernel(handle.ptr, item, kernel), "ddla_c_vbatched_set_kernel")


class BatchedGraph:
    """One batched call captured into a hipGraph (replay with launch())."""

    def __init__(self, vb: VBatch, uplo: str = "L", max_n: int = None):
        max_n = vb.max_n if max_n is None else max_n
        self.handle = vb.handle
        self.ptr = ctypes.c_void_p()
        st = vb._fn("pposv_vbatched_graph")(*vb._common(uplo), ctypes.c_void_p(vb.info.ptr.value),
                                            ctypes.c_void_p(vb.logdet.ptr.value), vb.batch, max_n,
                                            ctypes.byref(self.ptr))
        check_status(st, f"ddla_{vb.type_char}_pposv_vbatched_graph")

    def launch(self):
        check_status(lib.ddla_c_graph_launch(self.ptr, self.handle.ptr), "ddla_c_graph_launch")

    def destroy(self):
        if self.ptr.value:
            check_status(lib.ddla_c_graph_destroy(self.ptr), "ddla_c_graph_destroy")
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
