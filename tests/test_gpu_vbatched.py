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


def _mat(dtype, rows, cols, seed, ld=None):
    ld = max(rows, 1) if ld is None else ld
    a = np.zeros((ld, max(cols, 1)), dtype=dtype, order="F")
    a[:rows, :cols] = value(seed, rows * cols).reshape(cols, rows).T
    return a


def _ptrs(handle, bufs):
    return DeviceBuf(handle, np.array([b.ptr.value for b in bufs], dtype=np.uint64))


def _ints(handle, xs):
    return DeviceBuf(handle, np.array(xs, dtype=np.int32))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("c0_left", [True, False])
def test_vbatched2s_segments_reuse_temporaries(handle, dtype, c0_left):
    """src/BLAS/gemm_vbatched.cpp:420-523: stage 0 writes C0 = a0 A0 B0^T,
    stage 1 C1 = a1 C0 AB1 + b1 C1 (C0_left) or a1 AB1^T C0 + b1 C1; the three
    segments share one pool of C0 buffers, so each segment's stage 1 must run
    before the next segment's stage 0 overwrites them."""
    pfx = "d" if dtype == np.float64 else "s"
    segs = [2, 3, 1]
    batch = sum(segs)
    rng = np.random.RandomState(3)
    dims = [(int(rng.randint(1, 70)), int(rng.randint(1, 70)), int(rng.randint(1, 40)), int(rng.randint(1, 50)))
            for _ in range(batch)]  # (m0, n0, k0, other)
    a0s, b0s, ab1s, c1s, keep = [], [], [], [], []
    m1, n1, k1, lda1, ldb1, ldc1, ldc0 = [], [], [], [], [], [], []
    pool_ld, pool_cols = 80, 80
    pool = [DeviceBuf(handle, np.full((pool_cols, pool_ld), np.nan, dtype=dtype)) for _ in range(max(segs))]
    c0_ptrs = []
    slot = [i for s in segs for i in range(s)]
    for b, (m0, n0, k0, o) in enumerate(dims):
        a0 = _mat(dtype, m0, k0, 11 * b + 1)          # op N
        b0 = _mat(dtype, n0, k0, 13 * b + 2)          # op T: B0 is n0 x k0
        c0 = a0[:m0, :k0].astype(np.float64) @ b0[:n0, :k0].astype(np.float64).T * 0.5
        if c0_left:   # C1 (m0 x o) = C0 (m0 x n0) * AB1 (n0 x o), ops N N
            ab1 = _mat(dtype, n0, o, 17 * b + 3)
            c1 = _mat(dtype, m0, o, 19 * b + 4, ld=m0 + 2)
            want = 1.5 * c0 @ ab1[:n0, :o].astype(np.float64) - 0.25 * c1[:m0, :o].astype(np.float64)
            m1.append(m0), n1.append(o), k1.append(n0), lda1.append(pool_ld), ldb1.append(ab1.shape[0])
        else:         # C1 (o x n0) = AB1^T (o x m0) * C0 (m0 x n0), ops T N
            ab1 = _mat(dtype, m0, o, 17 * b + 3)
            c1 = _mat(dtype, o, n0, 19 * b + 4, ld=o + 2)
            want = 1.5 * ab1[:m0, :o].astype(np.float64).T @ c0 - 0.25 * c1[:o, :n0].astype(np.float64)
            m1.append(o), n1.append(n0), k1.append(m0), lda1.append(ab1.shape[0]), ldb1.append(pool_ld)
        ldc1.append(c1.shape[0])
        ldc0.append(pool_ld)
        bufs = (DeviceBuf(handle, a0.T.copy()), DeviceBuf(handle, b0.T.copy()), DeviceBuf(handle, ab1.T.copy()),
                DeviceBuf(handle, c1.T.copy()))
        a0s.append(bufs[0]), b0s.append(bufs[1]), ab1s.append(bufs[2]), c1s.append(bufs[3])
        keep.append((want, c1.shape))
        c0_ptrs.append(pool[slot[b]].ptr.value)
    d = lambda xs: _ints(handle, xs)
    dm0, dn0, dk0 = d([x[0] for x in dims]), d([x[1] for x in dims]), d([x[2] for x in dims])
    dlda0, dldb0 = d([_mat(dtype, x[0], x[2], 0).shape[0] for x in dims]), d([max(x[1], 1) for x in dims])
    dldc0 = d(ldc0)
    dm1, dn1, dk1, dlda1, dldb1, dldc1 = d(m1), d(n1), d(k1), d(lda1), d(ldb1), d(ldc1)
    pa0, pb0, pab1, pc1 = _ptrs(handle, a0s), _ptrs(handle, b0s), _ptrs(handle, ab1s), _ptrs(handle, c1s)
    pc0 = DeviceBuf(handle, np.array(c0_ptrs, dtype=np.uint64))
    ctype = ctypes.c_double if dtype == np.float64 else ctypes.c_float
    al0, be0, al1, be1 = ctype(0.5), ctype(0.0), ctype(1.5), ctype(-0.25)
    seg = (ctypes.c_int * len(segs))(*segs)
    fn = getattr(dd.lib, f"ddla_{pfx}_gemm_vbatched2s")
    args = [handle.ptr, b"N", b"T", dm0.ptr, dn0.ptr, dk0.ptr, ctypes.byref(al0), pa0.ptr, dlda0.ptr, pb0.ptr,
            dldb0.ptr, ctypes.byref(be0), pc0.ptr, dldc0.ptr,
            b"N" if c0_left else b"T", b"N", dm1.ptr, dn1.ptr, dk1.ptr, ctypes.byref(al1), pab1.ptr, dlda1.ptr,
            dldb1.ptr, ctypes.byref(be1), pc1.ptr, dldc1.ptr, 1 if c0_left else 0, batch, seg, len(segs)]
    dd.check_status(fn(*args), "gemm_vbatched2s")
    handle.synchronize()
    tol = 1e-12 if dtype == np.float64 else 2e-5
    for (want, shape), buf in zip(keep, c1s):
        got = buf.download().T
        r, c = want.shape
        assert np.abs(got[:r, :c] - want).max() <= tol * max(1.0, np.abs(want).max())
        assert np.array_equal(got[r:, :c], np.zeros_like(got[r:, :c]))  # padding rows of C1 untouched
    # contract violations (host-side checks, no kernel runs)
    bad_beta = list(args)
    bad_beta[11] = ctypes.byref(ctype(1.0))
    assert fn(*bad_beta) == 6 and b"beta" in dd.lib.ddla_c_error_string(6)
    bad_seg = list(args)
    bad_seg[-2] = (ctypes.c_int * 2)(2, 3)
    bad_seg[-1] = 2
    assert fn(*bad_seg) == 6 and b"segment" in dd.lib.ddla_c_error_string(6)
    zero_seg = list(args)
    zero_seg[-2] = (ctypes.c_int * 4)(2, 0, 3, 1)
    zero_seg[-1] = 4
    assert fn(*zero_seg) == 6
