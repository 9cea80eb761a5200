"""ddla_d_pposv_noalloc: FP64 pposv on the library's static device scratch.

It runs the kernels of ddla_d_pposv on a static array instead of the handle's
scratch, so its factor and solution must equal ddla_d_pposv's bit for bit on
the same inputs.  Two calls in a row, the second smaller (the array keeps the
first call's contents), both triangles and both sides; a failed pivot; the
size limit; n = 0.  Each result is also checked against the oracle.
"""
import ctypes

import numpy as np
import pytest

import libddla_amd as dd
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def upload(h, a, nb=64):
    return dd.DistributedMatrix.from_global_matrix(h, np.asarray(a, dtype=np.float64), mb=nb, nb=nb)


def call(entry, h, side, uplo, n, nrhs, A, B):
    info = ctypes.c_int(-7)
    st = entry(h.ptr, side.encode(), uplo.encode(), b"N", n, nrhs, ctypes.c_void_p(A.data_ptr), 1, 1, A.desc_ptr,
               ctypes.c_void_p(B.data_ptr), 1, 1, B.desc_ptr, ctypes.byref(info))
    return st, info.value


def solve_both(h, a, b, side, uplo):
    """(factor, solution, info) of ddla_d_pposv_noalloc and of ddla_d_pposv on copies of a, b."""
    n = a.shape[0]
    nrhs = b.shape[1] if side == "L" else b.shape[0]
    out = []
    for entry in (dd.lib.ddla_d_pposv_noalloc, dd.lib.ddla_d_pposv):
        A, B = upload(h, a), upload(h, b)
        st, info = call(entry, h, side, uplo, n, nrhs, A, B)
        assert st == 0, dd.lib.ddla_c_error_string(st)
        out.append((A.to_global_matrix(), B.to_global_matrix(), info))
    return out


@pytest.mark.parametrize("side,uplo", [("L", "L"), ("R", "U")])
def test_two_calls_match_pposv(handle, side, uplo):
    tri = np.tril if uplo == "L" else np.triu
    for n, nrhs, seed in ((3000, 16, 41), (700, 3, 42)):  # the second call is smaller
        print(f"case pposv_noalloc side={side} uplo={uplo} n={n} nrhs={nrhs}", flush=True)
        a = orc.spd_matrix(n, seed=seed, delta=0.1)
        b = orc.normal_matrix(seed, 1, n, nrhs)
        if side == "R":
            b = np.ascontiguousarray(b.T)
        (fa, xa, ia), (fb, xb, ib) = solve_both(handle, a, b, side, uplo)
        assert ia == ib == 0
        assert np.array_equal(tri(fa), tri(fb))
        assert np.array_equal(fa - tri(fa), a - tri(a))  # other triangle untouched
        assert np.array_equal(xa, xb)
        x = xa if side == "L" else xa.T
        rhs = b if side == "L" else b.T
        assert orc.backward_error(a, x, rhs) < 0.05 * n * EPS
        ref, rinfo = orc.potrf(uplo, a)
        assert rinfo == 0
        assert np.linalg.norm(tri(fa) - tri(ref)) <= 1e-12 * np.linalg.norm(tri(ref))


def test_failed_pivot_leaves_b(handle):
    print("case pposv_noalloc planted pivot", flush=True)
    n, nrhs = 900, 2
    a = orc.spd_matrix(n, seed=43, delta=0.1)
    a[611, 611] = -1.0
    b = orc.normal_matrix(43, 1, n, nrhs)
    (fa, xa, ia), (fb, xb, ib) = solve_both(handle, a, b, "L", "L")
    assert ia == ib == 612
    assert np.array_equal(xa, b) and np.array_equal(xb, b)


def test_size_limit_and_empty(handle):
    print("case pposv_noalloc size limit", flush=True)
    n, nrhs = 2000, 4200  # side 'R': B^T staging alone is 8 n nrhs = 67.2 MB > 64 MiB
    a = orc.spd_matrix(n, seed=44, delta=0.1)
    A = upload(handle, a)
    B = upload(handle, np.ones((nrhs, n)))
    st, _ = call(dd.lib.ddla_d_pposv_noalloc, handle, "R", "L", n, nrhs, A, B)
    assert st == 6
    msg = dd.lib.ddla_c_error_string(6).decode()
    assert "static scratch" in msg and str(64 << 20) in msg
    assert np.array_equal(A.to_global_matrix(), a)  # rejected before any work
    st, info = call(dd.lib.ddla_d_pposv_noalloc, handle, "L", "L", 0, 1, A, B)
    assert st == 0 and info == 0
    # a valid call after the rejection still works
    b = orc.normal_matrix(44, 1, n, 3)
    (fa, xa, ia), (fb, xb, ib) = solve_both(handle, a, b, "L", "L")
    assert ia == ib == 0 and np.array_equal(xa, xb)
