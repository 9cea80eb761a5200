"""GPU parity of the bottom-right Cholesky factorization, potrf_bottom_right
and ppotrf_bottom_right (the reference's include/ddla/ddla.h:1090, :1115; C
entries fortran/ddla_c_api.cpp:1191-1258), against the oracle's
column-by-column recurrence from the last column up
(oracle_?potrf_bottom_right).

Cases follow tests/test_potrf_bottom_right.cpp of the reference: the known
factor of its make_factor (A formed as U U^T or L^T L) is recovered within
128 eps n; the opposite triangle and the rows between the order and the
leading dimension keep their sentinels bit for bit; orders 1, 70, 127-129, 257
and, beyond the reference's, multi-panel orders 1100 and 2100 (ragged last
panel).  Against the oracle the factor agrees to 1e-12 (double) / 1e-4 (float)
relative to max |factor|.  At order 8192 (no oracle) the factor reproduces A
(relative 1e-14) and equals the ordinary factorization of the order-reversed
matrix, reversed back.  The solve of test_backend_solvers.cpp:237-258 (U y =
b, U^T x = y through ptrtrs) recovers x.
"""
import ctypes
import functools

import numpy as np
import pytest
import torch

import libddla_amd as dd
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def known_factor(uplo, n):
    """tests/test_potrf_bottom_right.cpp:96-118 (real part)."""
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    f = 0.004 * (((3 * i + 5 * j) % 11) - 5)
    f = np.where(i == j, 1.5 + 0.002 * (i % 23), f)
    return np.triu(f) if uplo == "U" else np.tril(f)


def product(uplo, f):
    return f @ f.T if uplo == "U" else f.T @ f


def other_triangle(uplo, n):
    return np.tril(np.ones((n, n), bool), -1) if uplo == "U" else np.triu(np.ones((n, n), bool), 1)


def local_call(handle, dtype, uplo, n, buf, lda):
    """potrf_bottom_right through the C ABI on an lda x n column-major buffer."""
    t = torch.tensor(np.ascontiguousarray(buf.T), device="cuda")
    info = ctypes.c_int(-1)
    fn = dd.lib.ddla_d_potrf_bottom_right if dtype == np.float64 else dd.lib.ddla_s_potrf_bottom_right
    st = fn(handle.ptr, uplo.encode(), n, ctypes.c_void_p(t.data_ptr()), lda, ctypes.byref(info))
    torch.cuda.synchronize()
    return st, info.value, t.cpu().numpy().T


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [1, 70, 127, 128, 129, 257, 1100, 2100])
@pytest.mark.parametrize("uplo", ["U", "L"])
def test_known_factor_and_untouched_storage(handle, uplo, n, dtype):
    print(f"case bottom_right known factor uplo={uplo} n={n} {np.dtype(dtype).name}", flush=True)
    lda = n + 3
    f = known_factor(uplo, n)
    a = product(uplo, f)
    buf = np.full((lda, n), -911.0)
    buf[:n, :n] = a
    other = other_triangle(uplo, n)
    sentinel = -73.0 - 0.013 * np.arange(n)[:, None] + 0.017 * np.arange(n)[None, :]
    buf[:n, :n][other] = sentinel[other]
    buf = buf.astype(dtype)
    st, info, out = local_call(handle, dtype, uplo, n, buf, lda)
    assert st == 0 and info == 0
    keep = ~other
    got = out[:n, :n]
    assert np.abs(got[keep] - f[keep]).max() <= 128 * np.finfo(dtype).eps * n
    assert np.array_equal(got[other], buf[:n, :n][other])
    assert np.array_equal(out[n:, :], buf[n:, :])
    if n <= 1100:
        want, winfo = orc.potrf_bottom_right(uplo, buf[:n, :n])
        assert winfo == 0
        err = np.abs(got[keep] - want[keep]).max() / np.abs(want[keep]).max()
        assert err < (1e-12 if dtype == np.float64 else 1e-4), err


@pytest.mark.parametrize("uplo", ["U", "L"])
def test_descriptor_call_and_solve(handle, uplo):
    """ppotrf_bottom_right on a descriptor (seeded SPD, order 1100, 16 columns
    of B), then the two ptrtrs solves of the reference's solver test."""
    print(f"case bottom_right descriptor + solve uplo={uplo}", flush=True)
    n, nrhs = 1100, 16
    a = orc.spd_matrix(n, seed=6, delta=0.1)
    b = orc.normal_matrix(6, 1, n, nrhs)
    A = dd.DistributedMatrix.from_global_matrix(handle, a, mb=64, nb=64)
    _, info = dd.ppotrf_bottom_right(A, uplo=uplo)
    assert info == 0
    want, _ = orc.potrf_bottom_right(uplo, a)
    tri = np.triu if uplo == "U" else np.tril
    got = tri(A.to_global_matrix())
    assert np.abs(got - tri(want)).max() < 1e-12 * np.abs(want).max()
    B = dd.DistributedMatrix.from_global_matrix(handle, b, mb=64, nb=64)
    if uplo == "U":  # A = U U^T: U y = b, U^T x = y
        dd.ptrtrs(A, B, uplo="U", trans="N")
        dd.ptrtrs(A, B, uplo="U", trans="T")
    else:  # A = L^T L: L^T y = b, L x = y
        dd.ptrtrs(A, B, uplo="L", trans="T")
        dd.ptrtrs(A, B, uplo="L", trans="N")
    x = B.to_global_matrix()
    x_ref = np.linalg.solve(a, b)
    assert np.abs(x - x_ref).max() < 1e-10 * np.abs(x_ref).max()


@functools.lru_cache(maxsize=1)
def spd_8192():
    return orc.spd_matrix(8192, seed=1, delta=0.1)


@pytest.mark.parametrize("uplo", ["U", "L"])
def test_full_size_reversal_identity(handle, uplo):
    """Order 8192: the factor reproduces A, and equals ppotrf of J A J (J the
    order reversal) reversed back by the anti-transpose."""
    print(f"case bottom_right n=8192 uplo={uplo}", flush=True)
    n = 8192
    a = spd_8192()
    A = dd.DistributedMatrix.from_global_matrix(handle, a, mb=512, nb=512)
    _, info = dd.ppotrf_bottom_right(A, uplo=uplo)
    assert info == 0
    tri = np.triu if uplo == "U" else np.tril
    f = tri(A.to_global_matrix())
    R = dd.DistributedMatrix.from_global_matrix(handle, np.ascontiguousarray(a[::-1, ::-1]), mb=512, nb=512)
    _, rinfo = dd.ppotrf(R, uplo=uplo)
    assert rinfo == 0
    r = tri(R.to_global_matrix())
    anti = r.T[::-1, ::-1]  # M(i, j) -> M(n-1-j, n-1-i)
    assert np.abs(f - anti).max() <= 1e-13 * np.abs(f).max()
    ft = torch.tensor(f, device="cuda")
    prod = ft @ ft.T if uplo == "U" else ft.T @ ft
    res = (prod - torch.tensor(a, device="cuda")).abs().max().item()
    assert res < 1e-14 * n * np.abs(a).max()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("uplo", ["U", "L"])
def test_failing_row(handle, uplo, dtype):
    """info is the first row (1-based) of the trailing submatrix that is not
    positive definite: with negative pivots planted in rows 3 and 700 of 1000
    the bottom-right order meets row 700 first (info 701); a last pivot gives
    n, a first one 1."""
    print(f"case bottom_right failing row uplo={uplo} {np.dtype(dtype).name}", flush=True)
    n = 1000
    for rows, want in (((3, 700), 701), ((n - 1,), n), ((0,), 1)):
        a = np.eye(n) * 4.0 + 0.001
        for r in rows:
            a[r, r] = -1.0
        A = dd.DistributedMatrix.from_global_matrix(handle, a.astype(dtype), mb=64, nb=64)
        _, info = dd.ppotrf_bottom_right(A, uplo=uplo)
        _, oinfo = orc.potrf_bottom_right(uplo, a.astype(dtype))
        assert info == want == oinfo


def test_argument_errors(handle):
    """Exact-case uplo and lda >= n at the local entry (std::runtime_error ->
    status 6); descriptor rules at the distributed one; n = 0 does nothing."""
    t = torch.eye(8, dtype=torch.float64, device="cuda")
    info = ctypes.c_int(-1)
    p = ctypes.c_void_p(t.data_ptr())
    assert dd.lib.ddla_d_potrf_bottom_right(handle.ptr, b"u", 8, p, 8, ctypes.byref(info)) == 6
    assert dd.lib.ddla_d_potrf_bottom_right(handle.ptr, b"U", 8, p, 7, ctypes.byref(info)) == 6
    assert dd.lib.ddla_d_potrf_bottom_right(handle.ptr, b"U", -1, p, 8, ctypes.byref(info)) == 6
    assert dd.lib.ddla_d_potrf_bottom_right(handle.ptr, b"U", 0, None, 1, ctypes.byref(info)) == 0
    assert info.value == 0
    assert dd.lib.ddla_d_potrf_bottom_right(handle.ptr, b"U", 8, p, 8, None) != 0
    assert dd.lib.ddla_d_potrf_bottom_right(handle.ptr, b"U", 8, p, 8, ctypes.byref(info)) == 0
    assert info.value == 0 and torch.equal(t, torch.eye(8, dtype=torch.float64, device="cuda"))
    A = dd.DistributedMatrix.from_global_matrix(handle, np.eye(8), mb=4, nb=4)
    desc = A.desc_ptr
    assert dd.lib.ddla_d_ppotrf_bottom_right(handle.ptr, b"U", 9, ctypes.c_void_p(A.data_ptr), desc,
                                             ctypes.byref(info)) == 6
    assert dd.lib.ddla_d_ppotrf_bottom_right(handle.ptr, b"X", 8, ctypes.c_void_p(A.data_ptr), desc,
                                             ctypes.byref(info)) == 6
    B = dd.DistributedMatrix(handle, 8, 8, mb=4, nb=8)
    assert dd.lib.ddla_d_ppotrf_bottom_right(handle.ptr, b"U", 8, ctypes.c_void_p(B.data_ptr), B.desc_ptr,
                                             ctypes.byref(info)) == 6
