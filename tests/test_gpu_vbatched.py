"""GPU parity of gemmVbatched (C ABI ddla_{s,d}_gemm_vbatched) against the
oracle GEMM on the closed-form inputs of tests/test_gemm_vbatched.cpp:26-43:
all op combinations, beta in {0, 1, x} (beta 0 never reads C), zero k and
zero alpha (scale only), padded leading dimensions, invalid sizes rejected.
Tolerances 1e-12 (double) and 1e-5 (float), as the reference's test."""
import ctypes

import numpy as np
import pytest

import libddla_amd as dd
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def value(start, count):
    return ((np.arange(start, start + count) % 13) - 6) / 7.0


class DeviceBuf:
    def __init__(self, handle, host):
        self.handle = handle
        self.host = np.ascontiguousarray(host)
        self.ptr = ctypes.c_void_p()
        nbytes = max(self.host.nbytes, 8)
        dd.check_status(dd.lib.ddla_c_malloc(handle.ptr, ctypes.byref(self.ptr), nbytes))
        if self.host.nbytes:
            dd.check_status(dd.lib.ddla_c_memcpy(handle.ptr, self.ptr, self.host.ctypes.data, self.host.nbytes))
        handle.synchronize()

    def download(self):
        out = np.empty_like(self.host)
        if out.nbytes:
            dd.check_status(dd.lib.ddla_c_memcpy(self.handle.ptr, out.ctypes.data, self.ptr, out.nbytes))
        self.handle.synchronize()
        return out

    def __del__(self):
        dd.lib.ddla_c_free(self.handle.ptr, self.ptr)


def run_case(handle, dtype, ta, tb, sizes, alpha, beta, pad=0, zero_c_nan=False):
    pfx = "d" if dtype == np.float64 else "s"
    mats, bufs = [], []
    for b, (m, n, k) in enumerate(sizes):
        ar, ac = (m, k) if ta == "N" else (k, m)
        br, bc = (k, n) if tb == "N" else (n, k)
        lda, ldb, ldc = max(ar, 1) + pad, max(br, 1) + pad, max(m, 1) + pad
        a = np.zeros((lda, max(ac, 1)), dtype=dtype, order="F")
        a[:ar, :ac] = value(17 * b, ar * ac).reshape(ac, ar).T
        bb = np.zeros((ldb, max(bc, 1)), dtype=dtype, order="F")
        bb[:br, :bc] = value(31 * b + 5, br * bc).reshape(bc, br).T
        c = np.full((ldc, max(n, 1)), np.nan if zero_c_nan else 0.0, dtype=dtype, order="F")
        if not zero_c_nan:
            c[:m, :n] = value(7 * b + 3, m * n).reshape(n, m).T
        mats.append((m, n, k, a, bb, c, lda, ldb, ldc))
        bufs.append((DeviceBuf(handle, a.T.copy()), DeviceBuf(handle, bb.T.copy()), DeviceBuf(handle, c.T.copy())))
    ints = lambda idx: DeviceBuf(handle, np.array([x[idx] for x in mats], dtype=np.int32))
    dm, dn, dk, dlda, dldb, dldc = ints(0), ints(1), ints(2), ints(6), ints(7), ints(8)
    ptrs = lambda j: DeviceBuf(handle, np.array([b[j].ptr.value for b in bufs], dtype=np.uint64))
    pa, pb, pc = ptrs(0), ptrs(1), ptrs(2)
    ctype = ctypes.c_double if dtype == np.float64 else ctypes.c_float
    al, be = ctype(alpha), ctype(beta)
    st = getattr(dd.lib, f"ddla_{pfx}_gemm_vbatched")(handle.ptr, ta.encode(), tb.encode(), dm.ptr, dn.ptr, dk.ptr,
                                                      ctypes.byref(al), pa.ptr, dlda.ptr, pb.ptr, dldb.ptr,
                                                      ctypes.byref(be), pc.ptr, dldc.ptr, len(sizes))
    dd.check_status(st, "gemm_vbatched")
    handle.synchronize()
    tol = 1e-12 if dtype == np.float64 else 1e-5
    for (m, n, k, a, bb, c, lda, ldb, ldc), (_, _, dcb) in zip(mats, bufs):
        got = dcb.download().T
        cin = np.where(np.isnan(c), 0, c) if zero_c_nan else c
        want = orc.gemm(ta, tb, m, n, k, alpha, a.astype(np.float64), bb.astype(np.float64), beta,
                        cin.astype(np.float64))
        assert not np.isnan(got[:m, :n]).any()
        assert np.abs(got[:m, :n] - want[:m, :n]).max() <= tol * max(1.0, np.abs(want).max())
        # padding rows untouched
        if ldc > m:
            assert np.array_equal(np.isnan(got[m:, :n]), np.isnan(c[m:, :n]))


SIZES = [(5, 7, 3), (64, 64, 64), (100, 37, 129), (1, 1, 1), (130, 70, 20), (33, 200, 65)]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ta", ["N", "T", "C"])
@pytest.mark.parametrize("tb", ["N", "T", "c"])
def test_vbatched_ops(handle, dtype, ta, tb):
    run_case(handle, dtype, ta.upper(), tb.upper(), SIZES, 1.25, 0.5)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_vbatched_beta_zero_never_reads_c(handle, dtype):
    run_case(handle, dtype, "N", "T", SIZES, -0.75, 0.0, zero_c_nan=True)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_vbatched_beta_one_and_padding(handle, dtype):
    run_case(handle, dtype, "T", "N", SIZES, 1.0, 1.0, pad=5)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_vbatched_zero_k_and_zero_alpha(handle, dtype):
    run_case(handle, dtype, "N", "N", [(9, 4, 0), (3, 3, 0)], 1.0, 2.0)
    run_case(handle, dtype, "N", "N", [(9, 4, 6), (3, 3, 2)], 0.0, 0.5)


def test_vbatched_rejects_bad_dimensions(handle):
    """A negative size at batch 2 is reported by index before any GEMM runs
    (src/BLAS/gemm_vbatched.cpp:74-160)."""
    sizes = [(8, 8, 8), (8, 8, 8), (4, -1, 2), (8, 8, 8)]
    m = DeviceBuf(handle, np.array([s[0] for s in sizes], dtype=np.int32))
    n = DeviceBuf(handle, np.array([s[1] for s in sizes], dtype=np.int32))
    k = DeviceBuf(handle, np.array([s[2] for s in sizes], dtype=np.int32))
    ld = DeviceBuf(handle, np.full(4, 8, dtype=np.int32))
    ptrs = DeviceBuf(handle, np.zeros(4, dtype=np.uint64))
    one = ctypes.c_double(1.0)
    st = dd.lib.ddla_d_gemm_vbatched(handle.ptr, b"N", b"N", m.ptr, n.ptr, k.ptr, ctypes.byref(one), ptrs.ptr,
                                     ld.ptr, ptrs.ptr, ld.ptr, ctypes.byref(one), ptrs.ptr, ld.ptr, 4)
    assert st == 6
    assert b"batch 2" in dd.lib.ddla_c_error_string(6)


def test_vbatched_zero_batch_is_noop(handle):
    one = ctypes.c_double(1.0)
    st = dd.lib.ddla_d_gemm_vbatched(handle.ptr, b"N", b"N", None, None, None, ctypes.byref(one), None, None, None,
                                     None, ctypes.byref(one), None, None, 0)
    assert st == 0
    st = dd.lib.ddla_d_gemm_vbatched(handle.ptr, b"Q", b"N", None, None, None, ctypes.byref(one), None, None, None,
                                     None, ctypes.byref(one), None, None, 1)
    assert st == 6
