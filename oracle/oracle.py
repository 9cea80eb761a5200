"""ctypes front end of the CPU oracle (oracle/ddla_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The shipped library never
imports it.
"""
import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    u64p = ctypes.POINTER(ctypes.c_uint64)
    dp, fp, ip = ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)
    lib.oracle_philox4x64_10.argtypes = [u64p, u64p, u64p]
    lib.oracle_fill_normal.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int64, dp]
    for p in ("d", "s"):
        getattr(lib, f"oracle_{p}potrf").argtypes = [ctypes.c_char, ctypes.c_int, dp, ctypes.c_int, ip]
        getattr(lib, f"oracle_{p}potrf_bottom_right").argtypes = [ctypes.c_char, ctypes.c_int, dp, ctypes.c_int, ip]
        getattr(lib, f"oracle_{p}potrs").argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int, dp,
                                                     ctypes.c_int, dp, ctypes.c_int]
        getattr(lib, f"oracle_{p}diag_logdet").argtypes = [ctypes.c_int, dp, ctypes.c_int,
                                                           ctypes.POINTER(ctypes.c_double), ip]
    lib.oracle_dgemm.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_double, dp, ctypes.c_int, dp, ctypes.c_int, ctypes.c_double, dp,
                                 ctypes.c_int]
    lib.oracle_sgemm.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_float, fp, ctypes.c_int, fp, ctypes.c_int, ctypes.c_float, fp,
                                 ctypes.c_int]
    for p in ("z", "c"):
        getattr(lib, f"oracle_{p}potrf").argtypes = [ctypes.c_char, ctypes.c_int, dp, ctypes.c_int, ip]
        getattr(lib, f"oracle_{p}potrs").argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int, dp,
                                                     ctypes.c_int, dp, ctypes.c_int]
        getattr(lib, f"oracle_{p}diag_logdet").argtypes = [ctypes.c_int, dp, ctypes.c_int,
                                                           ctypes.POINTER(ctypes.c_double),
                                                           ctypes.POINTER(ctypes.c_double), ip]
        getattr(lib, f"oracle_{p}gemm").argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int,
                                                    ctypes.c_int, dp, dp, ctypes.c_int, ctypes.c_int, dp,
                                                    ctypes.c_int, ctypes.c_int, dp, dp, ctypes.c_int]
    return lib


lib = _load()


def _pfx(a):
    return {np.dtype(np.float64): "d", np.dtype(np.float32): "s", np.dtype(np.complex128): "z",
            np.dtype(np.complex64): "c"}[np.dtype(a.dtype)]


def philox(ctr, key):
    c = (ctypes.c_uint64 * 4)(*ctr)
    k = (ctypes.c_uint64 * 2)(*key)
    out = (ctypes.c_uint64 * 4)()
    lib.oracle_philox4x64_10(c, k, out)
    return [int(x) for x in out]


def fill_normal(seed, stream, first, count):
    out = np.empty(count, dtype=np.float64)
    lib.oracle_fill_normal(seed, stream, first, count, out.ctypes.data)
    return out


def normal_matrix(seed, stream, m, n):
    """m x n column-major matrix of the project's seeded normals (element i + j*m)."""
    return fill_normal(seed, stream, 0, m * n).reshape((n, m)).T.copy()


def spd_matrix(n, seed, delta):
    """A = G G^T / n + delta I, G the n x n normals of stream 0 (formed in float64 on the CPU)."""
    g = normal_matrix(seed, 0, n, n)
    a = g @ g.T / n
    a[np.diag_indices(n)] += delta
    return a


def potrf(uplo, a, lda=None):
    """Factor a copy of the Fortran-ordered array a; returns (factor, info)."""
    f = np.array(a, dtype=a.dtype, order="F", copy=True)
    n = f.shape[1]
    info = ctypes.c_int(0)
    getattr(lib, f"oracle_{_pfx(f)}potrf")(uplo.encode(), n, f.ctypes.data, f.shape[0], ctypes.byref(info))
    return f, info.value


def potrf_bottom_right(uplo, a):
    """Bottom-right factorization of a copy of a (A = U U^T for uplo U, L^T L
    for L); returns (factor, info)."""
    f = np.array(a, dtype=a.dtype, order="F", copy=True)
    n = f.shape[1]
    info = ctypes.c_int(0)
    getattr(lib, f"oracle_{_pfx(f)}potrf_bottom_right")(uplo.encode(), n, f.ctypes.data, f.shape[0],
                                                         ctypes.byref(info))
    return f, info.value


def potrs(side, uplo, factor, b):
    """Solve with the factor; b is copied (Fortran order) and returned."""
    f = np.asfortranarray(factor)
    x = np.array(b, dtype=b.dtype, order="F", copy=True)
    n = f.shape[1]
    nrhs = x.shape[1] if side == "L" else x.shape[0]
    getattr(lib, f"oracle_{_pfx(f)}potrs")(side.encode(), uplo.encode(), n, nrhs, f.ctypes.data, f.shape[0],
                                           x.ctypes.data, x.shape[0])
    return x


def diag_logdet(a, n=None):
    """(magnitude, flags) with flags 1 nonfinite | 2 zero | 4 odd negative count; a complex a gives
    (magnitude, phase, flags), the phase summed over the entries that are not real."""
    if np.iscomplexobj(a):
        a = np.asfortranarray(a)
        n = a.shape[1] if n is None else n
        mag, phase, flags = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_int(0)
        getattr(lib, f"oracle_{_pfx(a)}diag_logdet")(n, a.ctypes.data, max(1, a.shape[0]), ctypes.byref(mag),
                                                     ctypes.byref(phase), ctypes.byref(flags))
        return mag.value, phase.value, flags.value
    a = np.asfortranarray(a)
    n = a.shape[1] if n is None else n
    mag = ctypes.c_double(0)
    flags = ctypes.c_int(0)
    getattr(lib, f"oracle_{_pfx(a)}diag_logdet")(n, a.ctypes.data, max(1, a.shape[0]), ctypes.byref(mag),
                                                 ctypes.byref(flags))
    return mag.value, flags.value


def finish_logdet(mag, flags, phase=0.0):
    """The reference's result rules (src/diag_logdet_impl.h:104-113, src/diag_logdet.cpp:52-54)."""
    if flags & 1:
        raise ValueError("nonfinite diagonal")
    if flags & 2:
        return complex(-np.inf, 0.0)
    phase = math.remainder(phase + ((flags >> 2) & 1) * math.pi, 2 * math.pi)
    if phase <= -math.pi:
        phase = math.pi
    return complex(mag, phase)


def gemm(ta, tb, m, n, k, alpha, a, b, beta, c):
    """C := alpha op(A) op(B) + beta C on a copy of c.  A complex c may take one real operand."""
    c = np.array(c, order="F", copy=True)
    a = np.asfortranarray(a)
    b = np.asfortranarray(b)
    if np.iscomplexobj(c):
        rt = np.float64 if c.dtype == np.complex128 else np.float32
        al = np.array([complex(alpha).real, complex(alpha).imag], dtype=rt)
        be = np.array([complex(beta).real, complex(beta).imag], dtype=rt)
        a = a.astype(c.dtype if np.iscomplexobj(a) else rt, order="F", copy=False)
        b = b.astype(c.dtype if np.iscomplexobj(b) else rt, order="F", copy=False)
        getattr(lib, f"oracle_{_pfx(c)}gemm")(ta.encode(), tb.encode(), m, n, k, al.ctypes.data, a.ctypes.data,
                                              int(np.iscomplexobj(a)), max(1, a.shape[0]), b.ctypes.data,
                                              int(np.iscomplexobj(b)), max(1, b.shape[0]), be.ctypes.data,
                                              c.ctypes.data, max(1, c.shape[0]))
        return c
    fn = lib.oracle_dgemm if c.dtype == np.float64 else lib.oracle_sgemm
    fn(ta.encode(), tb.encode(), m, n, k, alpha, a.ctypes.data, max(1, a.shape[0]), b.ctypes.data,
       max(1, b.shape[0]), beta, c.ctypes.data, max(1, c.shape[0]))
    return c


def backward_error(a, x, b):
    """Normwise backward error ||A X - B||_inf / (||A||_inf ||X||_inf), in float64."""
    r = a @ x - b
    return np.abs(r).sum(axis=1).max() / (np.abs(a).sum(axis=1).max() * np.abs(x).sum(axis=1).max())
