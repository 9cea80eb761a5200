"""GPU parity of the complex (Hermitian) path: ppotrf / ppotrs / pposv,
diag_logdet and gemmVbatched(2s) for std::complex<float> ('c') and
std::complex<double> ('z'), through the C ABI of libddla_mi355x.so, against
the oracle, the scipy LAPACK fixtures (zpotrf / cpotrf / zpotrs / cpotrs) and
the reference's closed forms (tests/api_grid_test_common.h:490-512).

Tolerances: factors 3e-12 (z) / 2e-5 (c) relative to the factor's largest
entry; closed-form solutions 5e-9 (z) / 2e-3 (c) as the reference's solver
tests; logdet 3e-12 (z) / 2e-5 (c) as tests/test_diag_logdet.cpp:33; GEMM
1e-12 (z) / 1e-5 (c) relative to the largest entry.
"""
import ctypes
import math
import os

import numpy as np
import pytest

import libddla_amd as dd
from oracle import oracle as orc
from test_gpu_vbatched import DeviceBuf

pytestmark = pytest.mark.gpu

GOLD = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lapack_cases.npz"))
CTYPES = [np.complex128, np.complex64]
FTOL = {np.complex128: 3e-12, np.complex64: 2e-5}
XTOL = {np.complex128: 5e-9, np.complex64: 2e-3}


def tri(uplo, a):
    return np.tril(a) if uplo == "L" else np.triu(a)


def triangular_l_value(n):
    """tests/api_grid_test_common.h:502-507."""
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    off = 0.04 * ((i + j) % 4 + 1) + 1j * 0.01 * ((i - j) % 3)
    return np.where(i < j, 0j, np.where(i == j, 2.0 + 0.05 * i + 0j, off))


def hpd_value(n):
    """tests/api_grid_test_common.h:490-500."""
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    lo, hi = np.minimum(i, j), np.maximum(i, j)
    val = 0.01 * ((lo + 2 * hi) % 5 - 2) + 1j * 0.006 * ((3 * lo + hi) % 7 - 3)
    return np.where(i == j, 5.0 + 0.2 * n + 0.05 * i + 0j, np.where(i < j, val, np.conj(val)))


def x_value(rows, cols):
    """tests/api_grid_test_common.h:509-512."""
    i, j = np.meshgrid(np.arange(rows), np.arange(cols), indexing="ij")
    return (0.2 + 0.03 * i - 0.02 * j) + 1j * (0.04 * (i + 1) + 0.01 * j)


def random_hpd(n, seed, dtype=np.complex128):
    rng = np.random.RandomState(seed)
    m = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    a = m @ m.conj().T / n + 0.5 * np.eye(n)
    return ((a + a.conj().T) / 2).astype(dtype)


def gold_keys():
    return sorted(k[len("potrf_in_"):] for k in GOLD.files if k.startswith("potrf_in_") and k[9] in "zc")


@pytest.mark.parametrize("key", gold_keys())
def test_fixture_factor_and_solve(handle, key):
    """scipy LAPACK results; the unused triangle is NaN and stays so, padding rows stay NaN."""
    tag, uplo, n, pad = key.split("_")
    n = int(n)
    print(f"case fixture {key}", flush=True)
    full = GOLD[f"potrf_in_{key}"]
    dtype = full.dtype.type
    A = dd.DistributedMatrix(handle, n, n, 64, 64, dtype, lld=full.shape[0]).upload_raw(full)
    _, info = dd.ppotrf(A, uplo)
    assert info == 0
    got = A.to_host()
    want = GOLD[f"potrf_out_{key}"]
    assert np.abs(tri(uplo, got[:n, :n]) - want).max() <= FTOL[dtype] * max(1.0, np.abs(want).max())
    other = np.triu(np.ones((n, n), bool), 1) if uplo == "L" else np.tril(np.ones((n, n), bool), -1)
    assert np.isnan(got[:n, :n][other]).all()
    assert np.isnan(got[n:, :]).all()
    b = GOLD[f"potrs_b_{key}"]
    B = dd.DistributedMatrix.from_global_matrix(handle, b, mb=64, nb=64)
    dd.ppotrs(A, B, uplo)
    x = B.to_global_matrix()
    xw = GOLD[f"potrs_x_{key}"]
    tol = 1e-11 if dtype == np.complex128 else 2e-4
    assert np.abs(x - xw).max() <= tol * max(1.0, np.abs(xw).max())


@pytest.mark.parametrize("dtype", CTYPES)
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 129, 300])
def test_closed_form_factor(handle, n, uplo, dtype):
    """A = L L^H from triangular_l_value: ppotrf returns L (uplo L) or L^H (uplo U)."""
    print(f"case closed-form factor n={n} uplo={uplo} {np.dtype(dtype).name}", flush=True)
    l = triangular_l_value(n)
    a = (l @ l.conj().T).astype(dtype)
    A = dd.DistributedMatrix.from_global_matrix(handle, a, mb=64, nb=64)
    _, info = dd.ppotrf(A, uplo)
    assert info == 0
    got = tri(uplo, A.to_global_matrix())
    want = l if uplo == "L" else l.conj().T
    assert np.abs(got - want).max() <= FTOL[dtype] * 10 * np.abs(l).max()


@pytest.mark.parametrize("dtype", CTYPES)
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("n", [40, 200])
def test_closed_form_pposv(handle, n, side, uplo, dtype):
    """B = A X (side L) or X A (side R) with hpd_value / x_value; pposv recovers X
    (tests/api_grid_test_common.h:539-563)."""
    print(f"case closed-form pposv n={n} side={side} uplo={uplo} {np.dtype(dtype).name}", flush=True)
    nrhs = 7
    a = hpd_value(n)
    x_true = x_value(n, nrhs) if side == "L" else x_value(nrhs, n)
    b = a @ x_true if side == "L" else x_true @ a
    A = dd.DistributedMatrix.from_global_matrix(handle, a.astype(dtype), mb=64, nb=64)
    B = dd.DistributedMatrix.from_global_matrix(handle, b.astype(dtype), mb=64, nb=64)
    _, info = dd.pposv(A, B, uplo, side)
    assert info == 0
    x = B.to_global_matrix()
    assert np.abs(x - x_true).max() < XTOL[dtype]
    # against the oracle's solve with the oracle's factor
    f, _ = orc.potrf(uplo, a.astype(dtype))
    xo = orc.potrs(side, uplo, f, b.astype(dtype))
    tol = 1e-12 if dtype == np.complex128 else 1e-4
    assert np.abs(x - xo).max() <= tol * np.abs(xo).max()


@pytest.mark.parametrize("dtype", CTYPES)
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_random_hpd_against_oracle(handle, uplo, dtype):
    """Order 700 (eleven 64-column panels with trailing updates, ragged last block), padded lld."""
    n, nrhs = 700, 33
    print(f"case random HPD n={n} uplo={uplo} {np.dtype(dtype).name}", flush=True)
    a = random_hpd(n, 11, dtype)
    f_or, info_or = orc.potrf(uplo, a)
    assert info_or == 0
    A = dd.DistributedMatrix.from_global_matrix(handle, a, mb=64, nb=64, lld=n + 5)
    _, info = dd.ppotrf(A, uplo)
    assert info == 0
    got = tri(uplo, A.to_global_matrix())
    want = tri(uplo, f_or)
    rel = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert rel <= (1e-12 if dtype == np.complex128 else 1e-5), rel
    rng = np.random.RandomState(12)
    for side in ("L", "R"):
        shape = (n, nrhs) if side == "L" else (nrhs, n)
        b = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)
        B = dd.DistributedMatrix.from_global_matrix(handle, b, mb=64, nb=64)
        dd.ppotrs(A, B, uplo, side)
        x = B.to_global_matrix()
        xo = orc.potrs(side, uplo, f_or, b)
        tol = 1e-11 if dtype == np.complex128 else 2e-4
        assert np.abs(x - xo).max() <= tol * np.abs(xo).max(), side


@pytest.mark.parametrize("dtype", CTYPES)
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_pivot_info_and_untouched_b(handle, uplo, dtype):
    """A non-positive real pivot at row 71 (second panel) gives info 71; pposv leaves B as it was."""
    n = 150
    a = random_hpd(n, 13, dtype)
    a[70, 70] = -50.0 + 3j
    _, info_or = orc.potrf(uplo, a)
    assert info_or == 71
    A = dd.DistributedMatrix.from_global_matrix(handle, a, mb=64, nb=64)
    b = (np.arange(n * 2).reshape(n, 2) * (1 - 0.5j)).astype(dtype)
    B = dd.DistributedMatrix.from_global_matrix(handle, b, mb=64, nb=64)
    _, info = dd.pposv(A, B, uplo)
    assert info == 71
    assert np.array_equal(B.to_global_matrix(), b)
    # NaN pivot in the first panel
    a = random_hpd(n, 14, dtype)
    a[5, 5] = np.nan
    A = dd.DistributedMatrix.from_global_matrix(handle, a, mb=64, nb=64)
    assert dd.ppotrf(A, uplo)[1] == 6


@pytest.mark.parametrize("dtype", CTYPES)
def test_three_phase_pposv(handle, dtype):
    """ppotrf / pposv workspace_size + run with caller-owned device workspace."""
    import torch

    n, nrhs, uplo = 260, 5, "U"
    a = hpd_value(n)
    x_true = x_value(n, nrhs)
    A = dd.DistributedMatrix.from_global_matrix(handle, a.astype(dtype), mb=64, nb=64)
    B = dd.DistributedMatrix.from_global_matrix(handle, (a @ x_true).astype(dtype), mb=64, nb=64)
    need, _ = dd.pposv_workspace_size(A, B, uplo)
    work = torch.empty(max(need, 1), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    info = dd.pposv_run(A, B, uplo, work.data_ptr(), need)
    assert info == 0
    assert np.abs(B.to_global_matrix() - x_true).max() < XTOL[dtype]


def logdet_cases(dtype):
    rt = np.float64 if dtype == np.complex128 else np.float32
    fmax, tiny = float(np.finfo(rt).max), float(np.finfo(rt).smallest_subnormal)
    z = complex(math.cos(0.8 * math.pi), math.sin(0.8 * math.pi))
    rng = np.random.RandomState(21)
    return {"2,3,5": [2, 3, 5], "-2,3,5": [-2, 3, 5], "-2,-3,5": [-2, -3, 5], "0,-3,5": [0, -3, 5],
            "max": [complex(fmax, fmax), complex(tiny, -tiny)], "polar": [z, z, z],
            "ones11": [1 + i % 3 for i in range(11)], "neg1025": [-1.0] * 1025, "empty": [],
            "random": list(rng.standard_normal(300) + 1j * rng.standard_normal(300))}


@pytest.mark.parametrize("dtype", CTYPES)
@pytest.mark.parametrize("case", ["2,3,5", "-2,3,5", "-2,-3,5", "0,-3,5", "max", "polar", "ones11", "neg1025",
                                  "empty", "random"])
def test_diag_logdet_cases(handle, case, dtype):
    """tests/test_diag_logdet.cpp:78-110 on a NaN-poisoned array (lda = n + 3), local and through
    a descriptor; the input is not modified."""
    d = [complex(dtype(v)) for v in logdet_cases(dtype)[case]]
    n = len(d)
    lda = max(n, 1) + 3
    host = np.full((lda, max(n, 1)), np.nan + 1j * np.nan, dtype=dtype)
    for i, v in enumerate(d):
        host[i, i] = v
    mag, phase, flags = orc.diag_logdet(host, n)
    want = orc.finish_logdet(mag, flags, phase)
    tol = 3e-12 if dtype == np.complex128 else 2e-5
    pi_r = float((np.float64 if dtype == np.complex128 else np.float32)(math.pi))  # pi rounded to the result type
    M = dd.DistributedMatrix(handle, max(n, 1), max(n, 1), 64, 64, dtype, lld=lda).upload_raw(host)
    for got in (dd.diag_logdet(M, n), dd.pdiag_logdet(M, n)):
        if math.isinf(want.real):
            assert math.isinf(got.real) and got.real < 0 and got.imag == 0
        else:
            assert abs(got.real - want.real) < tol * (1 + abs(want.real))
            assert abs(math.remainder(got.imag - want.imag, 2 * math.pi)) < tol
            assert -pi_r < got.imag <= pi_r
    if case == "neg1025":
        assert abs(got.imag - math.pi) < 1e-6 and abs(got.real) < 1e-9
    after = M.to_host()
    assert np.array_equal(np.isnan(after), np.isnan(host)) and np.array_equal(after[~np.isnan(host)],
                                                                              host[~np.isnan(host)])


@pytest.mark.parametrize("dtype", CTYPES)
@pytest.mark.parametrize("bad", [complex(1, np.inf), complex(np.nan, 0)])
def test_diag_logdet_nonfinite_raises(handle, dtype, bad):
    host = np.diag([1.0 + 0j, bad, 2.0]).astype(dtype)
    M = dd.DistributedMatrix.from_global_matrix(handle, host)
    with pytest.raises(RuntimeError):
        dd.diag_logdet(M)


@pytest.mark.parametrize("dtype", CTYPES)
def test_logdet_of_factor_is_half_logdet(handle, dtype):
    """2 Re logdet(diag L) = log det A (numpy slogdet) for an HPD A; the phase is 0."""
    a = random_hpd(300, 17, dtype)
    A = dd.DistributedMatrix.from_global_matrix(handle, a, mb=64, nb=64)
    assert dd.ppotrf(A, "L")[1] == 0
    got = dd.diag_logdet(A)
    sign, want = np.linalg.slogdet(a.astype(np.complex128))
    assert abs(2 * got.real - want) < (1e-10 if dtype == np.complex128 else 2e-3) * abs(want)
    assert abs(got.imag) < 1e-12


# ------------------------------------------------------------------------------------------------
# gemmVbatched, complex and mixed real/complex
# ------------------------------------------------------------------------------------------------
SIZES = [(5, 7, 3), (64, 64, 64), (100, 37, 129), (1, 1, 1), (130, 70, 20), (0, 4, 3)]


def cmat(rng, dtype, rows, cols, ld=None, real=False):
    ld = max(rows, 1) if ld is None else ld
    rt = np.float64 if dtype == np.complex128 else np.float32
    a = np.zeros((ld, max(cols, 1)), dtype=rt if real else dtype, order="F")
    v = rng.standard_normal((rows, cols))
    if not real:
        v = v + 1j * rng.standard_normal((rows, cols))
    a[:rows, :cols] = v
    return a


def scalar(dtype, v):
    rt = np.float64 if dtype == np.complex128 else np.float32
    return np.array([complex(v).real, complex(v).imag], dtype=rt)


def run_vbatched(handle, dtype, ta, tb, sizes, alpha, beta, real_operand=None, pad=0, beta_zero_nan=False):
    pfx = "z" if dtype == np.complex128 else "c"
    rng = np.random.RandomState(len(sizes) * 7 + ord(ta) + 3 * ord(tb))
    mats, bufs = [], []
    for m, n, k in sizes:
        ar, ac = (m, k) if ta == "N" else (k, m)
        br, bc = (k, n) if tb == "N" else (n, k)
        a = cmat(rng, dtype, ar, ac, max(ar, 1) + pad, real=real_operand == 0)
        b = cmat(rng, dtype, br, bc, max(br, 1) + pad, real=real_operand == 1)
        c = cmat(rng, dtype, m, n, max(m, 1) + pad)
        if beta_zero_nan:
            c[:] = np.nan
        mats.append((m, n, k, a, b, c))
        bufs.append((DeviceBuf(handle, a.T.copy()), DeviceBuf(handle, b.T.copy()), DeviceBuf(handle, c.T.copy())))
    ints = lambda f: DeviceBuf(handle, np.array([f(x) for x in mats], dtype=np.int32))
    dm, dn, dk = ints(lambda x: x[0]), ints(lambda x: x[1]), ints(lambda x: x[2])
    dlda, dldb, dldc = ints(lambda x: x[3].shape[0]), ints(lambda x: x[4].shape[0]), ints(lambda x: x[5].shape[0])
    ptrs = lambda j: DeviceBuf(handle, np.array([b[j].ptr.value for b in bufs], dtype=np.uint64))
    pa, pb, pc = ptrs(0), ptrs(1), ptrs(2)
    al, be = scalar(dtype, alpha), scalar(dtype, beta)
    if real_operand is None:
        st = getattr(dd.lib, f"ddla_{pfx}_gemm_vbatched")(handle.ptr, ta.encode(), tb.encode(), dm.ptr, dn.ptr,
                                                          dk.ptr, al.ctypes.data, pa.ptr, dlda.ptr, pb.ptr,
                                                          dldb.ptr, be.ctypes.data, pc.ptr, dldc.ptr, len(sizes))
    else:
        st = getattr(dd.lib, f"ddla_{pfx}_gemm_vbatched_mixed")(handle.ptr, real_operand, ta.encode(), tb.encode(),
                                                                dm.ptr, dn.ptr, dk.ptr, al.ctypes.data, pa.ptr,
                                                                dlda.ptr, pb.ptr, dldb.ptr, be.ctypes.data, pc.ptr,
                                                                dldc.ptr, len(sizes))
    dd.check_status(st, "complex gemm_vbatched")
    handle.synchronize()
    tol = 1e-12 if dtype == np.complex128 else 1e-5
    for (m, n, k, a, b, c), (_, _, dc) in zip(mats, bufs):
        got = dc.download().T
        cin = np.where(np.isnan(c), 0, c)
        want = orc.gemm(ta, tb, m, n, k, alpha, a, b, beta, cin)
        if m and n:
            assert not np.isnan(got[:m, :n]).any()
            assert np.abs(got[:m, :n] - want[:m, :n]).max() <= tol * max(1.0, np.abs(want).max())
        if c.shape[0] > m and not beta_zero_nan:
            assert np.array_equal(got[m:, :n], c[m:, :n])


@pytest.mark.parametrize("dtype", CTYPES)
@pytest.mark.parametrize("ta", ["N", "T", "C"])
@pytest.mark.parametrize("tb", ["N", "T", "C"])
def test_complex_vbatched_ops(handle, dtype, ta, tb):
    run_vbatched(handle, dtype, ta, tb, SIZES, 0.75 - 1.25j, 0.5 + 0.25j)


@pytest.mark.parametrize("dtype", CTYPES)
def test_complex_vbatched_special_cases(handle, dtype):
    run_vbatched(handle, dtype, "C", "N", SIZES, 1.0j, 0.0, beta_zero_nan=True, pad=3)
    run_vbatched(handle, dtype, "N", "N", [(9, 4, 0), (3, 3, 2)], 1.0, 2.0 - 1j)
    run_vbatched(handle, dtype, "N", "N", [(9, 4, 6), (3, 3, 2)], 0.0, 0.5j)


@pytest.mark.parametrize("dtype", CTYPES)
@pytest.mark.parametrize("real_operand", [0, 1])
@pytest.mark.parametrize("t", ["N", "T", "C"])
def test_mixed_real_complex_vbatched(handle, dtype, real_operand, t):
    """include/ddla/ddla.h:1878-1916: one real operand; 'C' on it is 'T'."""
    ta, tb = (t, "C") if real_operand == 0 else ("T", t)
    run_vbatched(handle, dtype, ta, tb, SIZES, 1.5 - 0.5j, -0.25j, real_operand=real_operand, pad=2)


@pytest.mark.parametrize("dtype", CTYPES)
@pytest.mark.parametrize("mixed", [False, True])
def test_complex_vbatched2s(handle, dtype, mixed):
    """Two stages per segment (include/ddla/ddla.h:1857-1877, 1917-1960): C0 = a0 A0 B0^H, then
    C1 = a1 C0 AB1 + b1 C1; mixed: A0 real and AB1 real."""
    pfx = "z" if dtype == np.complex128 else "c"
    rng = np.random.RandomState(5 + mixed)
    segs = [2, 1, 2]
    batch = sum(segs)
    pool_ld = 70
    pool = [DeviceBuf(handle, np.full((70, pool_ld), np.nan + 1j * np.nan, dtype=dtype)) for _ in range(2)]
    slot = [i for s in segs for i in range(s)]
    keep, dims = [], []
    b_a0, b_b0, b_ab1, b_c1 = [], [], [], []
    a0_, b0_, al0, al1, be1 = None, None, 0.5 + 0.5j, 1.5, -0.25 + 1j
    for bi in range(batch):
        m0, n0, k0, o = (int(x) for x in rng.randint(1, 60, size=4))
        a0 = cmat(rng, dtype, m0, k0, real=mixed)
        b0 = cmat(rng, dtype, n0, k0)  # op 'C'
        ab1 = cmat(rng, dtype, n0, o, real=mixed)
        c1 = cmat(rng, dtype, m0, o, ld=m0 + 2)
        c0 = al0 * (a0[:m0, :k0].astype(np.complex128) @ b0[:n0, :k0].astype(np.complex128).conj().T)
        want = al1 * (c0 @ ab1[:n0, :o]) + be1 * c1[:m0, :o]
        dims.append((m0, n0, k0, o, a0.shape[0], b0.shape[0], ab1.shape[0], c1.shape[0]))
        keep.append(want)
        b_a0.append(DeviceBuf(handle, a0.T.copy()))
        b_b0.append(DeviceBuf(handle, b0.T.copy()))
        b_ab1.append(DeviceBuf(handle, ab1.T.copy()))
        b_c1.append(DeviceBuf(handle, c1.T.copy()))
    d = lambda f: DeviceBuf(handle, np.array([f(x) for x in dims], dtype=np.int32))
    ptrs = lambda bs: DeviceBuf(handle, np.array([b.ptr.value for b in bs], dtype=np.uint64))
    pc0 = DeviceBuf(handle, np.array([pool[slot[i]].ptr.value for i in range(batch)], dtype=np.uint64))
    dm0, dn0, dk0 = d(lambda x: x[0]), d(lambda x: x[1]), d(lambda x: x[2])
    dlda0, dldb0, dldc0 = d(lambda x: x[4]), d(lambda x: x[5]), d(lambda x: pool_ld)
    dm1, dn1, dk1 = d(lambda x: x[0]), d(lambda x: x[3]), d(lambda x: x[1])
    dlda1, dldb1, dldc1 = d(lambda x: pool_ld), d(lambda x: x[6]), d(lambda x: x[7])
    s_al0, s_be0, s_al1, s_be1 = scalar(dtype, al0), scalar(dtype, 0), scalar(dtype, al1), scalar(dtype, be1)
    seg = (ctypes.c_int * len(segs))(*segs)
    # the pointer arrays stay referenced until the call has finished (a DeviceBuf frees on collection)
    pa0, pb0, pab1, pc1 = ptrs(b_a0), ptrs(b_b0), ptrs(b_ab1), ptrs(b_c1)
    tail = [b"N", b"N", dm1.ptr, dn1.ptr, dk1.ptr, s_al1.ctypes.data, pab1.ptr, dlda1.ptr, dldb1.ptr,
            s_be1.ctypes.data, pc1.ptr, dldc1.ptr, 1, batch, seg, len(segs)]
    head = [b"N", b"C", dm0.ptr, dn0.ptr, dk0.ptr, s_al0.ctypes.data, pa0.ptr, dlda0.ptr, pb0.ptr,
            dldb0.ptr, s_be0.ctypes.data, pc0.ptr, dldc0.ptr]
    if mixed:
        st = getattr(dd.lib, f"ddla_{pfx}_gemm_vbatched2s_mixed")(handle.ptr, 0, 1, *head, *tail)
    else:
        st = getattr(dd.lib, f"ddla_{pfx}_gemm_vbatched2s")(handle.ptr, *head, *tail)
    dd.check_status(st, "complex gemm_vbatched2s")
    handle.synchronize()
    tol = 1e-12 if dtype == np.complex128 else 2e-5
    for want, buf, x in zip(keep, b_c1, dims):
        got = buf.download().T
        r, c = want.shape
        assert np.abs(got[:r, :c] - want).max() <= tol * max(1.0, np.abs(want).max())
