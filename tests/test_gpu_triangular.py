"""GPU parity of the triangular solves ptrtrs / trsm (include/ddla/ddla.h:96,
:1620 of the reference; C entries fortran/ddla_c_api.cpp:351-367,
:1451-1468) against scipy's LAPACK-backed solve_triangular in FP64.  The
triangle that is not referenced holds NaN, and so does the diagonal for
diag 'U', so a read of either shows up in the result.  Tolerances: 1e-10
(double) / 1e-4 (float) relative to max |X|, for diagonals in [2, 3] and
off-diagonal entries of size 0.3 / sqrt(order).
"""
import ctypes

import numpy as np
import pytest
from scipy.linalg import solve_triangular

import libddla_amd as dd

pytestmark = pytest.mark.gpu


def tri_matrix(order, uplo, diag, seed):
    rng = np.random.RandomState(seed)
    a = rng.standard_normal((order, order)) * 0.3 / np.sqrt(order)
    a[np.diag_indices(order)] = 2.0 + rng.uniform(size=order)
    keep = np.tril(np.ones((order, order), bool)) if uplo == "L" else np.triu(np.ones((order, order), bool))
    stored = np.where(keep, a, np.nan)
    math = np.where(keep, a, 0.0)
    if diag == "U":
        stored[np.diag_indices(order)] = np.nan
        math[np.diag_indices(order)] = 1.0
    return stored, math


def reference(math, b, side, uplo, trans, alpha):
    t = 0 if trans == "N" else 1
    if side == "L":
        x = solve_triangular(math, b, trans=t, lower=(uplo == "L"))
    else:  # X op(A) = B  <=>  op(A)^T X^T = B^T
        x = solve_triangular(math, b.T, trans=1 - t, lower=(uplo == "L")).T
    return alpha * x


def upload(h, a, dtype):
    return dd.DistributedMatrix.from_global_matrix(h, np.asarray(a).astype(dtype), mb=64, nb=64)


def run_case(handle, dtype, side, uplo, trans, diag, m, n, alpha, api):
    order = m if side == "L" else n
    stored, math = tri_matrix(order, uplo, diag, seed=m * 7 + n + ord(uplo) + ord(trans) + ord(diag))
    b = np.random.RandomState(m + 3 * n).standard_normal((m, n))
    A = upload(handle, stored, dtype)
    B = upload(handle, b, dtype)
    if api == "ptrtrs":
        dd.ptrtrs(A, B, side=side, uplo=uplo, trans=trans, diag=diag)
        alpha = 1.0
    else:
        dd.trsm(A, B, side=side.lower(), uplo=uplo.lower(), trans=trans.lower(), diag=diag.lower(), alpha=alpha)
    x = B.to_global_matrix()
    want = reference(math, b, side, uplo, trans, alpha)
    err = np.abs(x - want).max() / max(1e-300, np.abs(want).max())
    assert err < (1e-10 if dtype == np.float64 else 1e-4), err


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("diag", ["N", "U"])
@pytest.mark.parametrize("trans", ["N", "T", "C"])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("side", ["L", "R"])
def test_ptrtrs_all_flags(handle, dtype, side, uplo, trans, diag):
    """Every side x uplo x trans x diag combination, ragged order (130) and
    a ragged second dimension (37)."""
    m, n = (130, 37) if side == "L" else (37, 130)
    run_case(handle, dtype, side, uplo, trans, diag, m, n, 1.0, "ptrtrs")


@pytest.mark.parametrize("side,uplo,trans,diag", [("L", "L", "N", "N"), ("L", "U", "T", "U"), ("R", "L", "T", "N"),
                                                   ("R", "U", "N", "U")])
def test_trsm_alpha_and_lower_case_flags(handle, side, uplo, trans, diag):
    """trsm scales by alpha and takes lower-case flags (src/blas_enums.h)."""
    m, n = (200, 19) if side == "L" else (19, 200)
    run_case(handle, np.float64, side, uplo, trans, diag, m, n, -2.5, "trsm")


@pytest.mark.parametrize("side,uplo,trans", [("L", "L", "N"), ("L", "U", "N"), ("L", "L", "T"), ("R", "U", "T")])
def test_many_columns_recursive(handle, side, uplo, trans):
    """Order 2100 with 300 right-hand sides: the solve splits the order
    recursively (GEMM off-diagonal blocks, sweep leaves)."""
    print(f"case triangular many-rhs side={side} uplo={uplo} trans={trans}", flush=True)
    m, n = (2100, 300) if side == "L" else (300, 2100)
    run_case(handle, np.float64, side, uplo, trans, "N", m, n, 1.0, "ptrtrs")


def test_trsm_alpha_zero_reads_nothing(handle):
    """alpha == 0: B := 0 without reading A (A all NaN here)."""
    A = upload(handle, np.full((50, 50), np.nan), np.float64)
    B = upload(handle, np.ones((50, 7)), np.float64)
    dd.trsm(A, B, alpha=0.0)
    assert (B.to_global_matrix() == 0).all()


def test_flag_and_shape_errors(handle):
    """At the C ABI ptrtrs wants upper-case flags (src/PBLAS/ptrtrs.cpp:34-50)
    while trsm takes either case (src/blas_enums.h:12-30); both reject bad
    flags and sizes beyond the descriptors with an error status.  (The Python
    wrappers upper-case every flag.)"""
    A = upload(handle, np.eye(8), np.float64)
    B = upload(handle, np.full((8, 2), 3.0), np.float64)
    st = dd.lib.ddla_d_ptrtrs(handle.ptr, b"l", b"L", b"N", b"N", 8, 2, ctypes.c_void_p(A.data_ptr), A.desc_ptr,
                              ctypes.c_void_p(B.data_ptr), B.desc_ptr)
    assert st != 0
    alpha = ctypes.c_double(2.0)
    st = dd.lib.ddla_d_trsm(handle.ptr, b"l", b"u", b"t", b"n", 8, 2, ctypes.byref(alpha),
                            ctypes.c_void_p(A.data_ptr), A.lld, ctypes.c_void_p(B.data_ptr), B.lld)
    assert st == 0
    assert (B.to_global_matrix() == 6.0).all()
    with pytest.raises(RuntimeError):
        dd.ptrtrs(A, B, diag="X")
    with pytest.raises(RuntimeError):
        dd.trsm(A, B, trans="Q")
    C = upload(handle, np.ones((9, 2)), np.float64)
    with pytest.raises(RuntimeError):
        dd.ptrtrs(A, C)  # m = 9 exceeds A's order 8
    # the identity leaves B as it is
    Z = upload(handle, np.ones((8, 1)), np.float64)
    dd.ptrtrs(A, Z)
    assert (Z.to_global_matrix() == 1).all()
