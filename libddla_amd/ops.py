"""Cholesky-path routines over DistributedMatrix (mirrors python/ddla/ops.py:113-143).

Every call goes through the C ABI of libddla_mi355x.so and runs the
hand-written gfx950 kernels; there is no host fallback.
"""
import ctypes

import numpy as np

from ._capi import char, check_status, lib


def _fn(mat, name):
    return getattr(lib, f"ddla_{mat.type_char}_{name}")


def ppotrf(A, uplo: str = "L"):
    """Cholesky factorization in place; returns (A, info)."""
    info = ctypes.c_int(0)
    st = _fn(A, "ppotrf")(A.handle.ptr, char(uplo), A.n, ctypes.c_void_p(A.data_ptr), 1, 1, A.desc_ptr,
                          ctypes.byref(info))
    check_status(st, f"ddla_{A.type_char}_ppotrf")
    return A, info.value


def ppotrf_bottom_right(A, uplo: str = "U"):
    """Cholesky from the bottom-right corner in place (the reference's
    ppotrf_bottom_right): uplo U gives A = U U^T, uplo L gives A = L^T L;
    returns (A, info), info = first row (1-based) of the trailing submatrix
    that is not positive definite."""
    info = ctypes.c_int(0)
    st = _fn(A, "ppotrf_bottom_right")(A.handle.ptr, char(uplo), A.n, ctypes.c_void_p(A.data_ptr), A.desc_ptr,
                                       ctypes.byref(info))
    check_status(st, f"ddla_{A.type_char}_ppotrf_bottom_right")
    return A, info.value


def potrf_bottom_right(A, uplo: str = "U"):
    """The local routine (the reference's potrf_bottom_right) on A's local
    matrix and leading dimension; returns (A, info) as ppotrf_bottom_right."""
    info = ctypes.c_int(0)
    st = _fn(A, "potrf_bottom_right")(A.handle.ptr, char(uplo), A.n, ctypes.c_void_p(A.data_ptr), A.lld,
                                      ctypes.byref(info))
    check_status(st, f"ddla_{A.type_char}_potrf_bottom_right")
    return A, info.value


def ppotrs(A, B, uplo: str = "L", side: str = "L", trans: str = "N"):
    """Solve with the factor from ppotrf: A X = B (side L) or X A = B (side R), B overwritten."""
    n = A.n
    nrhs = B.n if char(side) == b"L" else B.m
    st = _fn(A, "ppotrs")(A.handle.ptr, char(side), char(uplo), char(trans), n, nrhs, ctypes.c_void_p(A.data_ptr),
                          A.desc_ptr, ctypes.c_void_p(B.data_ptr), B.desc_ptr)
    check_status(st, f"ddla_{A.type_char}_ppotrs")
    return B


def pposv(A, B, uplo: str = "L", side: str = "L", trans: str = "N"):
    """Factor and solve (the reference's pposv driver); B is untouched when info != 0."""
    info = ctypes.c_int(0)
    n = A.n
    nrhs = B.n if char(side) == b"L" else B.m
    st = _fn(A, "pposv")(A.handle.ptr, char(side), char(uplo), char(trans), n, nrhs, ctypes.c_void_p(A.data_ptr),
                         1, 1, A.desc_ptr, ctypes.c_void_p(B.data_ptr), 1, 1, B.desc_ptr, ctypes.byref(info))
    check_status(st, f"ddla_{A.type_char}_pposv")
    return B, info.value


def ptrtrs(A, B, side: str = "L", uplo: str = "L", trans: str = "N", diag: str = "N"):
    """B := op(A)^-1 B (side L) or B op(A)^-1 (side R) for a triangular A (the reference's ptrtrs); B overwritten."""
    m, n = B.m, B.n
    st = _fn(A, "ptrtrs")(A.handle.ptr, char(side), char(uplo), char(trans), char(diag), m, n,
                          ctypes.c_void_p(A.data_ptr), A.desc_ptr, ctypes.c_void_p(B.data_ptr), B.desc_ptr)
    check_status(st, f"ddla_{A.type_char}_ptrtrs")
    return B


def trsm(A, B, side: str = "L", uplo: str = "L", trans: str = "N", diag: str = "N", alpha: float = 1.0):
    """B := alpha op(A)^-1 B (side L) or alpha B op(A)^-1 (side R) on the local matrices (the reference's trsm)."""
    m, n = B.m, B.n
    a = (ctypes.c_double if A.type_char == "d" else ctypes.c_float)(alpha)
    st = _fn(A, "trsm")(A.handle.ptr, char(side), char(uplo), char(trans), char(diag), m, n, ctypes.byref(a),
                        ctypes.c_void_p(A.data_ptr), A.lld, ctypes.c_void_p(B.data_ptr), B.lld)
    check_status(st, f"ddla_{A.type_char}_trsm")
    return B


def ppotrf_workspace_size(A, uplo: str = "L"):
    bb, hb = ctypes.c_longlong(), ctypes.c_longlong()
    check_status(_fn(A, "ppotrf_workspace_size")(A.handle.ptr, char(uplo), A.n, A.desc_ptr, ctypes.byref(bb),
                                                 ctypes.byref(hb)), "ppotrf_workspace_size")
    return bb.value, hb.value


def ppotrf_run(A, uplo, backend_work: int, backend_bytes: int, host_work: int = 0, host_bytes: int = 0):
    info = ctypes.c_int(0)
    st = _fn(A, "ppotrf_run")(A.handle.ptr, char(uplo), A.n, ctypes.c_void_p(A.data_ptr), 1, 1, A.desc_ptr,
                              ctypes.byref(info), ctypes.c_void_p(backend_work), backend_bytes,
                              ctypes.c_void_p(host_work), host_bytes)
    check_status(st, f"ddla_{A.type_char}_ppotrf_run")
    return info.value


def ppotrs_workspace_size(A, B, side: str = "L"):
    bb, hb = ctypes.c_longlong(), ctypes.c_longlong()
    nrhs = B.n if char(side) == b"L" else B.m
    check_status(_fn(A, "ppotrs_workspace_size")(A.handle.ptr, char(side), A.n, nrhs, A.desc_ptr, B.desc_ptr,
                                                 ctypes.byref(bb), ctypes.byref(hb)), "ppotrs_workspace_size")
    return bb.value, hb.value


def pposv_workspace_size(A, B, uplo: str = "L", side: str = "L"):
    bb, hb = ctypes.c_longlong(), ctypes.c_longlong()
    nrhs = B.n if char(side) == b"L" else B.m
    check_status(_fn(A, "pposv_workspace_size")(A.handle.ptr, char(side), char(uplo), A.n, nrhs, A.desc_ptr,
                                                B.desc_ptr, ctypes.byref(bb), ctypes.byref(hb)),
                 "pposv_workspace_size")
    return bb.value, hb.value


def ppotrs_run(A, B, uplo, backend_work: int, backend_bytes: int, side: str = "L", trans: str = "N"):
    """The solve on caller scratch (from ppotrs_workspace_size).  Returns synchronized; raises
    DdlaExceptionError if a triangular-solve sweep exceeded its wait limit."""
    nrhs = B.n if char(side) == b"L" else B.m
    st = _fn(A, "ppotrs_run")(A.handle.ptr, char(side), char(uplo), char(trans), A.n, nrhs,
                              ctypes.c_void_p(A.data_ptr), A.desc_ptr, ctypes.c_void_p(B.data_ptr), B.desc_ptr,
                              ctypes.c_void_p(backend_work), backend_bytes, ctypes.c_void_p(0), 0)
    check_status(st, f"ddla_{A.type_char}_ppotrs_run")
    return B


def pposv_run(A, B, uplo, backend_work: int, backend_bytes: int, side: str = "L", trans: str = "N"):
    info = ctypes.c_int(0)
    nrhs = B.n if char(side) == b"L" else B.m
    st = _fn(A, "pposv_run")(A.handle.ptr, char(side), char(uplo), char(trans), A.n, nrhs,
                             ctypes.c_void_p(A.data_ptr), 1, 1, A.desc_ptr, ctypes.c_void_p(B.data_ptr), 1, 1,
                             B.desc_ptr, ctypes.byref(info), ctypes.c_void_p(backend_work), backend_bytes,
                             ctypes.c_void_p(0), 0)
    check_status(st, f"ddla_{A.type_char}_pposv_run")
    return info.value


FACTOR_FP32 = 0
FACTOR_BF16X3 = 1  # trailing updates as three bfloat16 split products (FP32 accumulation)
FACTOR_BF16X6 = 2  # six split products, about FP32 accuracy
FACTOR_FP16X3 = 3  # three half-precision split products of row-scaled panels (~2^-22), at bf16x3's cost


def _mixed_args(A, B, uplo):
    if A.type_char != "d" or B.type_char != "d":
        raise TypeError("pposv_mixed solves FP64 systems (float64 A and B)")
    return char(uplo), A.n, B.n


def pposv_mixed(A, B, uplo: str = "L", factor: int = FACTOR_FP32, tol: float = None, max_iter: int = 30):
    """FP64 A X = B by an FP32 factorization plus FP64 iterative refinement (LAPACK dsposv semantics).

    factor selects the trailing updates of the FP32 factorization: FACTOR_FP32
    (FP32 MFMA), FACTOR_BF16X3 or FACTOR_BF16X6 (split bfloat16 products,
    FP32 accumulation), FACTOR_FP16X3 (split half-precision products of
    row-scaled panels, FP32 accumulation).

    Returns (B, iter, info): iter >= 0 is the number of corrections applied;
    -3 (FP32 factor failed) or -(max_iter + 1) (no convergence) mean the FP64
    pposv ran instead, and info is its result.  tol defaults to sqrt(n) * eps(FP64).
    """
    c_uplo, n, nrhs = _mixed_args(A, B, uplo)
    if tol is None:
        tol = max(1.0, float(n)) ** 0.5 * float(np.finfo(np.float64).eps)
    it, info = ctypes.c_int(0), ctypes.c_int(0)
    st = lib.ddla_d_pposv_mixed(A.handle.ptr, c_uplo, n, nrhs, ctypes.c_void_p(A.data_ptr), 1, 1, A.desc_ptr,
                                ctypes.c_void_p(B.data_ptr), 1, 1, B.desc_ptr, factor, tol, max_iter,
                                ctypes.byref(it), ctypes.byref(info))
    check_status(st, "ddla_d_pposv_mixed")
    return B, it.value, info.value


def pposv_mixed_workspace_size(A, B, uplo: str = "L"):
    c_uplo, n, nrhs = _mixed_args(A, B, uplo)
    bb, hb = ctypes.c_longlong(), ctypes.c_longlong()
    check_status(lib.ddla_d_pposv_mixed_workspace_size(A.handle.ptr, c_uplo, n, nrhs, A.desc_ptr, B.desc_ptr,
                                                       ctypes.byref(bb), ctypes.byref(hb)),
                 "ddla_d_pposv_mixed_workspace_size")
    return bb.value, hb.value


def pposv_mixed_run(A, B, uplo, backend_work: int, backend_bytes: int, factor: int = FACTOR_FP32,
                    tol: float = None, max_iter: int = 30):
    c_uplo, n, nrhs = _mixed_args(A, B, uplo)
    if tol is None:
        tol = max(1.0, float(n)) ** 0.5 * float(np.finfo(np.float64).eps)
    it, info = ctypes.c_int(0), ctypes.c_int(0)
    st = lib.ddla_d_pposv_mixed_run(A.handle.ptr, c_uplo, n, nrhs, ctypes.c_void_p(A.data_ptr), A.desc_ptr,
                                    ctypes.c_void_p(B.data_ptr), B.desc_ptr, factor, tol, max_iter,
                                    ctypes.byref(it), ctypes.byref(info), ctypes.c_void_p(backend_work),
                                    backend_bytes)
    check_status(st, "ddla_d_pposv_mixed_run")
    return it.value, info.value


def diag_logdet(A, n: int = None):
    """Principal log of the product of the diagonal: complex (sum log|d|, phase)."""
    n = A.m if n is None else n
    res = np.zeros(2, dtype=np.finfo(A.dtype).dtype)  # the real type of A's elements
    st = _fn(A, "diag_logdet")(A.handle.ptr, n, ctypes.c_void_p(A.data_ptr), A.lld,
                               ctypes.c_void_p(res.ctypes.data))
    check_status(st, f"ddla_{A.type_char}_diag_logdet")
    return complex(float(res[0]), float(res[1]))


def pdiag_logdet(A, n: int = None):
    """diag_logdet of the leading n x n block, through A's descriptor (the reference's pdiag_logdet)."""
    n = A.m if n is None else n
    res = np.zeros(2, dtype=np.finfo(A.dtype).dtype)
    st = _fn(A, "pdiag_logdet")(A.handle.ptr, n, ctypes.c_void_p(A.data_ptr), A.desc_ptr,
                                ctypes.c_void_p(res.ctypes.data))
    check_status(st, f"ddla_{A.type_char}_pdiag_logdet")
    return complex(float(res[0]), float(res[1]))
