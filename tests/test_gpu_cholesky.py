"""GPU parity of ppotrf / ppotrs / pposv / diag_logdet against the oracle, the
LAPACK fixtures and the reference's own closed-form tests.  Every call goes
through the C ABI of libddla_mi355x.so.

Tolerances: float64 factor 3e-12 (analytic factor, tests/test_ppotrf_grids.cpp:82)
and 1e-12 relative Frobenius against the CPU factor (the config-2 bar); float32
factor 2e-5; solutions 5e-9 (double) / 2e-3 (float) as tests/test_backend_solvers.cpp.
"""
import ctypes
import math
import os

import numpy as np
import pytest

import libddla_amd as dd
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

GOLD = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lapack_cases.npz"))
EPS = {np.float64: np.finfo(np.float64).eps, np.float32: np.finfo(np.float32).eps}


def upload(h, a, nb=64, lld=None, dtype=None):
    a = np.asarray(a)
    dtype = dtype or a.dtype
    return dd.DistributedMatrix.from_global_matrix(h, a.astype(dtype), mb=nb, nb=nb, lld=lld if lld else -1)


def analytic_factor(n):
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    l = np.where(i == j, 2 + 0.01 * i, 0.003 * (1 + (i + j) % 5))
    return np.tril(l)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_n1_known_answer(handle, dtype):
    A = upload(handle, np.array([[4.0]]), dtype=dtype)
    _, info = dd.ppotrf(A, "L")
    assert info == 0
    assert A.to_global_matrix()[0, 0] == 2.0


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("n", [1, 2, 17, 38, 129, 257, 600])
def test_analytic_factor_padding_and_poison(handle, dtype, uplo, n):
    """tests/test_ppotrf_grids.cpp:28-87: A = L L^T from a known L, NaN in the
    unused triangle, a larger parent matrix and padded LLD preserved bitwise."""
    print(f"case analytic n={n} uplo={uplo} {np.dtype(dtype).name}", flush=True)
    l = analytic_factor(n)
    a = l @ l.T
    parent = n + 3
    lld = parent + 3
    full = np.full((lld, parent), -731.0, dtype=dtype, order="F")
    mask = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    full[:n, :n] = np.where(mask, a, np.nan)
    M = dd.DistributedMatrix(handle, parent, parent, 8, 8, dtype, lld=lld).upload_raw(full)
    # factor only the leading n x n block of the parent descriptor
    info = ctypes.c_int(-1)
    fn = dd.lib.ddla_d_ppotrf if dtype == np.float64 else dd.lib.ddla_s_ppotrf
    dd.check_status(fn(handle.ptr, uplo.encode(), n, ctypes.c_void_p(M.data_ptr), 1, 1, M.desc_ptr,
                       ctypes.byref(info)))
    assert info.value == 0
    out = M.to_host()
    want = l if uplo == "L" else l.T
    got = out[:n, :n]
    err = (np.abs(got - want) / (1 + np.abs(want)))[mask].max()
    assert err < (3e-12 if dtype == np.float64 else 2e-5)
    active = np.zeros(full.shape, bool)
    active[:n, :n] = mask
    inactive_outside = np.ones(full.shape, bool)
    inactive_outside[:n, :n] = False
    assert out[inactive_outside].tobytes() == full[inactive_outside].tobytes()


def fixture_keys():
    return sorted(k[len("potrf_in_"):] for k in GOLD.files if k.startswith("potrf_in_"))


@pytest.mark.parametrize("key", fixture_keys())
def test_lapack_fixture(handle, key):
    tag, uplo, n, pad = key.split("_")
    n, pad = int(n), int(pad)
    full = GOLD[f"potrf_in_{key}"]
    dtype = full.dtype
    A = dd.DistributedMatrix(handle, n, n, 64, 64, dtype, lld=n + pad).upload_raw(full)
    _, info = dd.ppotrf(A, uplo)
    assert info == 0
    out = A.to_host()
    want = GOLD[f"potrf_out_{key}"]
    mask = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    tol = 1e-13 if tag == "d" else 2e-6
    assert np.abs(out[:n, :n][mask] - want[mask]).max() <= tol * max(1, np.abs(want).max())
    assert np.isnan(out[:n, :n][~mask]).all() and np.isnan(out[n:, :]).all()
    B = upload(handle, GOLD[f"potrs_b_{key}"])
    dd.ppotrs(A, B, uplo)
    x = B.to_global_matrix()
    xw = GOLD[f"potrs_x_{key}"]
    tol = 1e-12 if tag == "d" else 1e-4
    assert np.abs(x - xw).max() <= tol * max(1, np.abs(xw).max())


@pytest.mark.parametrize("n", [1000, 1500])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_factor_matches_oracle_seeded(handle, n, uplo):
    """Config-2 generator (A = G G^T / n + 0.1 I, seed 1) at oracle-sized n;
    1e-12 relative Frobenius, the config-2 bar."""
    print(f"case oracle n={n} uplo={uplo}", flush=True)
    a = orc.spd_matrix(n, seed=1, delta=0.1)
    A = upload(handle, a, nb=512)
    _, info = dd.ppotrf(A, uplo)
    assert info == 0
    tri = np.tril if uplo == "L" else np.triu
    got = tri(A.to_global_matrix())
    ref, rinfo = orc.potrf(uplo, a)
    assert rinfo == 0
    ref = tri(ref)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-12


def test_pposv_n2048_nrhs16(handle):
    """n=2048, nrhs=16, A = G G^T / n + I (seed 0), B seed-0 normals
    (stream 1); backward error ||AX-B|| / (||A|| ||X|| n eps) and the oracle's
    solution."""
    n, nrhs = 2048, 16
    a = orc.spd_matrix(n, seed=0, delta=1.0)
    b = orc.normal_matrix(0, 1, n, nrhs)
    A = upload(handle, a, nb=128)
    B = upload(handle, b, nb=128)
    X, info = dd.pposv(A, B, "L")
    assert info == 0
    x = X.to_global_matrix()
    berr = orc.backward_error(a, x, b) / (n * EPS[np.float64])
    assert berr < 0.05, berr
    ref, _ = orc.potrf("L", a)
    x_ref = orc.potrs("L", "L", ref, b)
    assert np.abs(x - x_ref).max() / np.abs(x_ref).max() < 1e-12


def hpd_real(n):
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    lo, hi = np.minimum(i, j), np.maximum(i, j)
    return np.where(i == j, 5.0 + 0.2 * n + 0.05 * i, 0.01 * ((lo + 2 * hi) % 5 - 2))


def x_real(n, nrhs):
    i, j = np.meshgrid(np.arange(n), np.arange(nrhs), indexing="ij")
    return 0.2 + 0.03 * i - 0.02 * j


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "C"])
def test_ppotrs_flags_closed_form_and_readonly_factor(handle, dtype, side, uplo, trans):
    """tests/test_api_grid_ppotrs.cpp:16-54: all side x uplo x trans; the factor
    is read-only (bitwise) across two solves."""
    n, nrhs = 200, 7
    a = hpd_real(n)
    x_true = x_real(n, nrhs) if side == "L" else x_real(n, nrhs).T
    b = a @ x_true if side == "L" else x_true @ a
    A = upload(handle, a, nb=16, dtype=dtype)
    _, info = dd.ppotrf(A, uplo)
    assert info == 0
    before = A.to_host()
    for _ in range(2):
        B = upload(handle, b, nb=16, dtype=dtype)
        dd.ppotrs(A, B, uplo, side=side, trans=trans)
        x = B.to_global_matrix()
        assert np.abs(x - x_true).max() < (5e-9 if dtype == np.float64 else 2e-3)
    assert before.tobytes(order="F") == A.to_host().tobytes(order="F")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_ppotrs_many_rhs_recursive_split(handle, dtype, side, uplo):
    """Many right-hand sides (n = 2100, 300 columns): the solve splits the
    order recursively (GEMM off-diagonal updates, ragged 64-row leaves, two
    levels) in both storage orientations; closed-form solution as above."""
    n, nrhs = 2100, 300
    print(f"case many-rhs n={n} nrhs={nrhs} side={side} uplo={uplo} {np.dtype(dtype).name}", flush=True)
    a = hpd_real(n)
    x_true = x_real(n, nrhs) if side == "L" else x_real(n, nrhs).T
    b = a @ x_true if side == "L" else x_true @ a
    A = upload(handle, a, dtype=dtype)
    _, info = dd.ppotrf(A, uplo)
    assert info == 0
    B = upload(handle, b, dtype=dtype)
    dd.ppotrs(A, B, uplo, side=side)
    x = B.to_global_matrix()
    assert np.abs(x - x_true).max() < (5e-9 if dtype == np.float64 else 2e-3)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_pposv_final_negative_pivot(handle, dtype, uplo):
    """python/tests/test_cholesky.py:79-92: info = n and B unchanged."""
    n = 9
    a = np.eye(n)
    a[-1, -1] = -1
    b = np.ones((n, 2))
    A = upload(handle, a, nb=3, dtype=dtype)
    B = upload(handle, b, nb=3, dtype=dtype)
    _, info = dd.pposv(A, B, uplo)
    assert info == n
    assert np.array_equal(B.to_global_matrix(), b.astype(dtype))


@pytest.mark.parametrize("bad", [5, 63, 64, 600, 1100])
def test_info_reports_first_failed_pivot_across_panels(handle, bad):
    """A pivot that fails inside a later 64-block or 512-panel stops the
    factorization there: info = bad + 1, like LAPACK."""
    n = 1200
    a = orc.spd_matrix(n, seed=5, delta=0.5)
    a[bad, bad] = -1e3
    A = upload(handle, a, nb=512)
    _, info = dd.ppotrf(A, "L")
    _, rinfo = orc.potrf("L", a)
    assert info == rinfo == bad + 1


def test_workspace_query_and_exact_guards(handle):
    """tests/test_ppotrf_grids.cpp:47-85: ppotrf_run on exactly the queried
    scratch leaves 0x57 guard bytes intact and matches the public driver."""
    n = 700
    a = orc.spd_matrix(n, seed=2, delta=0.1)
    A1 = upload(handle, a, nb=64)
    A2 = upload(handle, a, nb=64)
    bb, hb = dd.ppotrf_workspace_size(A1, "L")
    assert bb > 0 and hb == 0
    guard = 256
    work = ctypes.c_void_p()
    dd.check_status(dd.lib.ddla_c_malloc(handle.ptr, ctypes.byref(work), bb + 2 * guard))
    dd.check_status(dd.lib.ddla_c_memset(handle.ptr, work, 0x57, bb + 2 * guard))
    info = dd.ppotrf_run(A1, "L", work.value + guard, bb)
    assert info == 0
    edge = np.zeros(bb + 2 * guard, dtype=np.uint8)
    dd.check_status(dd.lib.ddla_c_memcpy(handle.ptr, edge.ctypes.data, work, bb + 2 * guard))
    handle.synchronize()
    assert (edge[:guard] == 0x57).all() and (edge[guard + bb:] == 0x57).all()
    _, info2 = dd.ppotrf(A2, "L")
    assert info2 == 0
    assert np.array_equal(np.tril(A1.to_global_matrix()), np.tril(A2.to_global_matrix()))
    dd.lib.ddla_c_free(handle.ptr, work)


def test_pposv_run_matches_driver(handle):
    n, nrhs = 300, 4
    a = orc.spd_matrix(n, seed=3, delta=0.2)
    b = orc.normal_matrix(3, 1, n, nrhs)
    A1, B1 = upload(handle, a), upload(handle, b)
    A2, B2 = upload(handle, a), upload(handle, b)
    bb, _ = dd.pposv_workspace_size(A1, B1, "L")
    work = ctypes.c_void_p()
    dd.check_status(dd.lib.ddla_c_malloc(handle.ptr, ctypes.byref(work), bb))
    info = dd.pposv_run(A1, B1, "L", work.value, bb)
    handle.synchronize()
    _, info2 = dd.pposv(A2, B2, "L")
    assert info == info2 == 0
    assert np.array_equal(B1.to_global_matrix(), B2.to_global_matrix())
    A3, B3 = upload(handle, a), upload(handle, b)
    with pytest.raises(dd.DdlaExceptionError, match="workspace"):
        dd.pposv_run(A3, B3, "L", work.value, bb - 256)
    dd.lib.ddla_c_free(handle.ptr, work)


def test_invalid_arguments_raise(handle):
    A = upload(handle, np.eye(8), nb=4)
    with pytest.raises(dd.DdlaExceptionError):
        dd.ppotrf(A, "X")
    info = ctypes.c_int(0)
    st = dd.lib.ddla_d_ppotrf(handle.ptr, b"L", 8, ctypes.c_void_p(A.data_ptr), 2, 1, A.desc_ptr, ctypes.byref(info))
    assert st == 6 and b"offsets" in dd.lib.ddla_c_error_string(6)
    st = dd.lib.ddla_d_ppotrf(handle.ptr, b"L", 9, ctypes.c_void_p(A.data_ptr), 1, 1, A.desc_ptr, ctypes.byref(info))
    assert st == 6
    bad = A.desc.copy()
    bad[5] = 2  # MB != NB
    st = dd.lib.ddla_d_ppotrf(handle.ptr, b"L", 8, ctypes.c_void_p(A.data_ptr), 1, 1,
                              bad.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ctypes.byref(info))
    assert st == 6
    info.value = -5
    st = dd.lib.ddla_d_ppotrf(handle.ptr, b"L", 0, None, 1, 1, A.desc_ptr, ctypes.byref(info))
    assert st == 0 and info.value == 0


def logdet_cases():
    return [[2, 3, 5], [-2, 3, 5], [-2, -3, 5], [0, -3, 5], [1 + i % 3 for i in range(11)], [-1.0] * 1025, []]


@pytest.mark.parametrize("d", logdet_cases())
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_diag_logdet_cases(handle, d, dtype):
    """tests/test_diag_logdet.cpp:76-111 against the oracle's sum."""
    n = len(d)
    lda = max(1, n) + 3
    full = np.full((lda, max(n, 1)), np.nan, dtype=dtype, order="F")
    for i, x in enumerate(d):
        full[i, i] = x
    M = dd.DistributedMatrix(handle, max(n, 1), max(n, 1), 4, 4, dtype, lld=lda).upload_raw(full)
    want = orc.finish_logdet(*orc.diag_logdet(full.astype(np.float64), n))
    # diag_logdet on the leading dimension, pdiag_logdet through the descriptor
    for got in (dd.diag_logdet(M, n), dd.pdiag_logdet(M, n)):
        if math.isinf(want.real):
            assert math.isinf(got.real) and got.real < 0 and got.imag == 0
        else:
            tol = 3e-12 if dtype == np.float64 else 2e-5
            assert abs(got.real - want.real) < tol * (1 + abs(want.real))
            assert abs(math.remainder(got.imag - want.imag, 2 * math.pi)) < tol
    assert M.to_host().tobytes(order="F") == full.tobytes(order="F")


def test_pdiag_logdet_descriptor_checks(handle):
    """src/PBLAS/pdiag_logdet.cpp:17-29: an order beyond the descriptor is
    rejected."""
    M = upload(handle, np.diag([2.0, 3.0, 4.0]))
    with pytest.raises(RuntimeError):
        dd.pdiag_logdet(M, 4)


@pytest.mark.parametrize("bad", [np.inf, np.nan])
def test_diag_logdet_nonfinite_raises(handle, bad):
    M = upload(handle, np.diag([1.0, bad]))
    with pytest.raises(dd.DdlaExceptionError, match="nonfinite"):
        dd.diag_logdet(M)


def test_cholesky_logdet_matches_slogdet(handle):
    n = 900
    a = orc.spd_matrix(n, seed=4, delta=0.3)
    A = upload(handle, a, nb=512)
    _, info = dd.ppotrf(A, "L")
    assert info == 0
    z = dd.diag_logdet(A)
    sign, ld = np.linalg.slogdet(a)
    assert sign > 0 and abs(2 * z.real - ld) < 1e-10 * abs(ld) and z.imag == 0


@pytest.mark.parametrize("n", [8192])
def test_full_size_pposv_properties(handle, n):
    """FP64 pposv at n = 8192, nrhs = 64: backward error, and sampled entries of L L^T
    against A (a size-independent reconstruction check)."""
    print(f"case full-size pposv n={n}", flush=True)
    nrhs = 64
    a = orc.spd_matrix(n, seed=1, delta=0.1)
    b = orc.normal_matrix(1, 1, n, nrhs)
    A = upload(handle, a, nb=512)
    B = upload(handle, b, nb=512)
    X, info = dd.pposv(A, B, "L")
    assert info == 0
    x = X.to_global_matrix()
    berr = orc.backward_error(a, x, b) / (n * EPS[np.float64])
    assert berr < 0.05, berr
    l = np.tril(A.to_global_matrix())
    rng = np.random.RandomState(0)
    i = rng.randint(0, n, 3000)
    j = rng.randint(0, n, 3000)
    rec = np.einsum("ij,ij->i", l[i, :], l[j, :])
    assert np.abs(rec - a[i, j]).max() < 1e-13 * n


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("n", [1, 130, 1100])
def test_local_potrf_device_info(handle, n, uplo, dtype):
    """The local potrf (include/ddla/ddla.h:2160 of the reference; C entry
    ddla_{s,d}_potrf): asynchronous on the handle's stream with a device info
    word.  The factor against the oracle (1e-12 / 1e-4 relative to max |L|),
    the NaN-poisoned other triangle and the padding rows (lda = n + 3) left bit
    for bit, info 0; a planted negative pivot reports its 1-based row."""
    import torch

    print(f"case local potrf n={n} uplo={uplo} {np.dtype(dtype).name}", flush=True)
    lda = n + 3
    a = orc.spd_matrix(n, seed=n + 11, delta=0.2)
    keep = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    buf = np.full((lda, n), -5.0)
    buf[:n, :n] = np.where(keep, a, np.nan)
    buf = buf.astype(dtype)
    t = torch.tensor(np.ascontiguousarray(buf.T), device="cuda")
    info = torch.full((1,), -7, dtype=torch.int32, device="cuda")
    fn = dd.lib.ddla_d_potrf if dtype == np.float64 else dd.lib.ddla_s_potrf
    torch.cuda.synchronize()
    assert fn(handle.ptr, uplo.encode(), n, ctypes.c_void_p(t.data_ptr()), lda, ctypes.c_void_p(info.data_ptr())) == 0
    handle.synchronize()
    assert int(info.item()) == 0
    out = t.cpu().numpy().T
    want, winfo = orc.potrf(uplo, a.astype(dtype))
    assert winfo == 0
    err = np.abs(out[:n, :n][keep] - want[keep]).max() / np.abs(want[keep]).max()
    assert err < (1e-12 if dtype == np.float64 else 1e-4), err
    assert np.isnan(out[:n, :n][~keep]).all()
    assert (out[n:, :] == buf[n:, :]).all()
    if n > 1:
        bad = np.where(keep, a, 0.0).astype(dtype)
        r = n // 2
        bad[r, r] = -1.0
        t2 = torch.tensor(np.ascontiguousarray(np.pad(bad, ((0, 3), (0, 0))).T), device="cuda")
        torch.cuda.synchronize()
        assert fn(handle.ptr, uplo.encode(), n, ctypes.c_void_p(t2.data_ptr()), lda,
                  ctypes.c_void_p(info.data_ptr())) == 0
        handle.synchronize()
        _, oinfo = orc.potrf(uplo, bad)
        assert int(info.item()) == oinfo and 1 <= oinfo <= r + 1


def test_local_potrf_argument_rules(handle):
    """n = 0 sets info to 0 without touching A; lda < n, a null info word and
    a bad uplo are errors (status 6)."""
    import torch

    info = torch.full((1,), -7, dtype=torch.int32, device="cuda")
    t = torch.eye(4, dtype=torch.float64, device="cuda")
    pi, pt = ctypes.c_void_p(info.data_ptr()), ctypes.c_void_p(t.data_ptr())
    torch.cuda.synchronize()
    assert dd.lib.ddla_d_potrf(handle.ptr, b"L", 0, None, 1, pi) == 0
    handle.synchronize()
    assert int(info.item()) == 0
    assert dd.lib.ddla_d_potrf(handle.ptr, b"L", 4, pt, 3, pi) == 6
    assert dd.lib.ddla_d_potrf(handle.ptr, b"L", 4, pt, 4, None) == 6
    assert dd.lib.ddla_d_potrf(handle.ptr, b"X", 4, pt, 4, pi) == 6


def test_leading_dimension_beyond_buffer_offsets(handle):
    """Leading dimensions above 2^20 leave the GEMM tile's buffer-load path
    (its 32-bit per-thread offsets) for the general one, which computes in the
    same order: the local FP64 potrf at order 1536 (three 512-column panels,
    128 x 128 trailing tiles) with lda = 2^20 + 64 must match lda = n bit for
    bit, and leave the padding rows alone (about 13 GB of device memory)."""
    import torch

    n, big = 1536, (1 << 20) + 64
    print(f"case potrf n={n} lda={big}", flush=True)
    a = orc.spd_matrix(n, seed=97, delta=0.3)
    info = torch.full((1,), -7, dtype=torch.int32, device="cuda")
    fn = dd.lib.ddla_d_potrf
    small = torch.tensor(np.ascontiguousarray(a.T), device="cuda")  # column-major n x n, lda = n
    torch.cuda.synchronize()
    assert fn(handle.ptr, b"L", n, ctypes.c_void_p(small.data_ptr()), n, ctypes.c_void_p(info.data_ptr())) == 0
    handle.synchronize()
    assert int(info.item()) == 0
    want = small.cpu().numpy().T
    wide = torch.full((n, big), -5.0, dtype=torch.float64, device="cuda")  # column-major big x n
    wide[:, :n] = torch.tensor(np.ascontiguousarray(a.T), device="cuda")
    torch.cuda.synchronize()
    assert fn(handle.ptr, b"L", n, ctypes.c_void_p(wide.data_ptr()), big, ctypes.c_void_p(info.data_ptr())) == 0
    handle.synchronize()
    assert int(info.item()) == 0
    got = wide[:, :n].cpu().numpy().T
    low = np.tril(np.ones((n, n), bool))
    assert (got[low] == want[low]).all()
    assert bool((wide[:, n:n + 4096] == -5.0).all()) and bool((wide[:, big - 64:] == -5.0).all())
    del wide
    torch.cuda.empty_cache()
