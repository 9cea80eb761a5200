"""GPU tests of the mixed-precision pposv (FP32 factor, FP64 refinement).

The reference has no mixed solver; what pins it is the FP64 problem itself:
the normwise backward error ||B - A X||_inf / (||A||_inf ||X||_inf) of every
column, evaluated here in extended precision, must meet the stopping tolerance
the call was given, and the forward error against the oracle's FP64 LAPACK
solution must stay within the conditioning bound kappa * eps (a factor of 100
of slack for both solutions' own errors).  The fallback paths are forced with
matrices whose FP32 rounding is singular and with max_iter = 0.
"""
import ctypes

import numpy as np
import pytest

import libddla_amd as dd
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

EPS = float(np.finfo(np.float64).eps)


def upload(h, a, lld=None):
    return dd.DistributedMatrix.from_global_matrix(h, np.asarray(a, dtype=np.float64), mb=64, nb=64,
                                                   lld=lld if lld else -1)


def backward_errors(a, x, b):
    """Per-column ||b - A x||_inf / (||A||_inf ||x||_inf) in long double."""
    al, xl, bl = (np.asarray(v, dtype=np.longdouble) for v in (a, x, b))
    r = bl - al @ xl
    anorm = np.abs(al).sum(axis=1).max()
    return (np.abs(r).max(axis=0) / (anorm * np.abs(xl).max(axis=0))).astype(np.float64)


def oracle_solution(a, b):
    l, info = orc.potrf("L", a)
    assert info == 0
    return orc.potrs("L", "L", l, b)


def stored(a, uplo):
    """Only the referenced triangle is meaningful: poison the other with NaN."""
    mask = np.tril(np.ones(a.shape, bool)) if uplo == "L" else np.triu(np.ones(a.shape, bool))
    return np.where(mask, a, np.nan)


FACTORS = [dd.FACTOR_FP32, dd.FACTOR_BF16X3, dd.FACTOR_BF16X6, dd.FACTOR_FP16X3]


@pytest.mark.parametrize("factor", FACTORS)
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("delta", [4e-2, 4e-4, 4e-6])
def test_mixed_converges_to_fp64_accuracy(handle, uplo, delta, factor):
    n, nrhs, tol = 777, 5, 1e-14
    print(f"case mixed n={n} delta={delta} uplo={uplo} factor={factor}", flush=True)
    a = orc.spd_matrix(n, 2, delta)
    b = orc.normal_matrix(2, 1, n, nrhs)
    A = upload(handle, stored(a, uplo), lld=n + 5)
    before = A.to_host().tobytes(order="F")
    B = upload(handle, b)
    _, it, info = dd.pposv_mixed(A, B, uplo, factor=factor, tol=tol)
    x = B.to_global_matrix()
    assert info == 0
    assert 0 <= it <= 30
    assert A.to_host().tobytes(order="F") == before  # A untouched when refinement converged
    assert backward_errors(a, x, b).max() <= tol
    kappa = np.linalg.cond(a)
    xo = oracle_solution(a, b)
    fwd = np.abs(x - xo).max() / np.abs(xo).max()
    assert fwd <= 100 * kappa * EPS, (fwd, kappa)


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_split_factor_modes_three_panels(handle, uplo):
    """n = 1500 (three 512-column panels; the second trailing update starts 512
    rows into the split panel, and the last tiles are ragged): every factor
    mode reaches 1e-14; bf16x6 (about FP32 accuracy) needs at most one
    correction more than FP32, and bf16x3 (about 2^-16) at least as many as
    bf16x6."""
    n, nrhs, tol = 1500, 3, 1e-14
    a = orc.spd_matrix(n, 8, 4e-3)
    b = orc.normal_matrix(8, 1, n, nrhs)
    its = {}
    for factor in FACTORS:
        print(f"case mixed n={n} uplo={uplo} factor={factor}", flush=True)
        A, B = upload(handle, stored(a, uplo)), upload(handle, b)
        _, it, info = dd.pposv_mixed(A, B, uplo, factor=factor, tol=tol)
        assert info == 0 and 1 <= it <= 30, (factor, it)
        assert backward_errors(a, B.to_global_matrix(), b).max() <= tol
        its[factor] = it
    assert its[dd.FACTOR_BF16X6] <= its[dd.FACTOR_FP32] + 1, its
    assert its[dd.FACTOR_BF16X3] >= its[dd.FACTOR_BF16X6], its
    assert its[dd.FACTOR_FP16X3] <= its[dd.FACTOR_BF16X3], its


@pytest.mark.parametrize("factor", FACTORS)
@pytest.mark.parametrize("kind", ["row_scaled", "block_diagonal"])
def test_scaled_and_block_diagonal_systems(handle, kind, factor):
    """n = 1100 (panels of 512, 512, 76 columns).  row_scaled: A = D S D with
    D = diag(2^k), k uniform in [-24, 24] (exact in FP64 and FP32), so panel
    rows span 2^48 in magnitude -- far outside half precision's range, which
    fp16x3's per-row power-of-two scaling must absorb.  block_diagonal: two
    SPD blocks of 700 and 400, so whole rows of the first panel below its
    diagonal block are zero (their scale exponent is 0).  Every mode
    converges to the 1e-14 backward error."""
    n, nrhs, tol = 1100, 3, 1e-14
    print(f"case mixed {kind} n={n} factor={factor}", flush=True)
    s = orc.spd_matrix(n, 5, 0.1)
    if kind == "row_scaled":
        k = np.random.RandomState(6).randint(-24, 25, size=n)
        d = np.ldexp(1.0, k)
        a = s * d[:, None] * d[None, :]
    else:
        a = np.zeros((n, n))
        a[:700, :700] = s[:700, :700]
        a[700:, 700:] = s[700:, 700:]
    b = orc.normal_matrix(5, 1, n, nrhs)
    A, B = upload(handle, stored(a, "L")), upload(handle, b)
    _, it, info = dd.pposv_mixed(A, B, "L", factor=factor, tol=tol)
    # the normwise test can pass with no correction when the scaling makes
    # the largest rows dominate
    assert info == 0 and 0 <= it <= 30, (it, info)
    assert backward_errors(a, B.to_global_matrix(), b).max() <= tol


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("factor", FACTORS)
def test_full_order(handle, factor, uplo):
    """n = 16384 (32 panels, trailing updates over 15872 rows): A = G G^T / n
    + 4e-4 I generated on the GPU (seed 2), four columns of B; the backward
    error, in FP64 on the GPU (the residual's own rounding is ~1e-17 against
    the 1e-14 target), meets the tolerance, and A is left untouched."""
    import torch

    print(f"case mixed full order n=16384 factor={factor} uplo={uplo}", flush=True)
    n, nrhs, tol = 16384, 4, 1e-14
    a0 = torch.empty((n, n), dtype=torch.float64, device="cuda")
    g = torch.empty_like(a0)
    dd.check_status(dd.lib.ddla_d_spd_generate(handle.ptr, n, 2, 4e-4, ctypes.c_void_p(a0.data_ptr()), n,
                                               ctypes.c_void_p(g.data_ptr())))
    handle.synchronize()
    del g
    gen = torch.Generator(device="cuda").manual_seed(2)
    b0 = torch.randn((nrhs, n), generator=gen, dtype=torch.float64, device="cuda")  # column-major n x nrhs
    a, b = a0.clone(), b0.clone()
    torch.cuda.synchronize()
    A = dd.DistributedMatrix.from_torch(handle, a, n, n, 512, 512)
    B = dd.DistributedMatrix.from_torch(handle, b, n, nrhs, 512, 512)
    _, it, info = dd.pposv_mixed(A, B, uplo, factor=factor, tol=tol)
    assert info == 0 and 1 <= it <= 30
    assert torch.equal(a, a0)  # A untouched
    r = b0 - b @ a0
    berr = (r.abs().amax(1) / (a0.abs().sum(1).max() * b.abs().amax(1))).max().item()
    assert berr <= tol, berr


def test_iterations_grow_with_condition(handle):
    n, nrhs = 512, 3
    its = []
    for delta in (4e-2, 4e-6):
        a = orc.spd_matrix(n, 2, delta)
        A, B = upload(handle, a), upload(handle, orc.normal_matrix(2, 1, n, nrhs))
        _, it, info = dd.pposv_mixed(A, B, "L", tol=1e-14)
        assert info == 0 and it >= 1
        its.append(it)
    assert its[1] >= its[0]


def test_default_tolerance_is_dsposv(handle):
    """Default tol = sqrt(n) eps (LAPACK dsposv's CTE with BWDMAX = 1)."""
    n = 300
    a = orc.spd_matrix(n, 5, 0.1)
    b = orc.normal_matrix(5, 1, n, 2)
    A, B = upload(handle, a), upload(handle, b)
    _, it, info = dd.pposv_mixed(A, B)
    assert info == 0 and it >= 1
    assert backward_errors(a, B.to_global_matrix(), b).max() <= np.sqrt(n) * EPS


def test_fp32_singular_falls_back_to_fp64(handle):
    """ones * ones^T + 1e-10 I is PD in FP64 but singular once rounded to FP32:
    the FP32 factor fails (iter = -3) and the FP64 pposv solves it."""
    n, nrhs = 200, 2
    a = np.ones((n, n)) + 1e-10 * np.eye(n)
    b = orc.normal_matrix(7, 1, n, nrhs)
    A, B = upload(handle, a), upload(handle, b)
    _, it, info = dd.pposv_mixed(A, B, "L")
    assert it == -3 and info == 0
    x = B.to_global_matrix()
    xo = oracle_solution(a, b)
    assert np.abs(x - xo).max() / np.abs(xo).max() <= 100 * np.linalg.cond(a) * EPS
    # A now holds the FP64 factor, as after pposv
    l, _ = orc.potrf("L", a)
    got = np.tril(A.to_global_matrix())
    assert np.abs(got - np.tril(l)).max() <= 1e-6


def test_no_convergence_falls_back(handle):
    """max_iter = 0 cannot reach 1e-14 from the FP32 solution: iter = -1 and the FP64 pposv result is returned."""
    n, nrhs = 400, 4
    a = orc.spd_matrix(n, 3, 4e-4)
    b = orc.normal_matrix(3, 1, n, nrhs)
    A, B = upload(handle, a), upload(handle, b)
    _, it, info = dd.pposv_mixed(A, B, "L", tol=1e-14, max_iter=0)
    assert it == -1 and info == 0
    A2, B2 = upload(handle, a), upload(handle, b)
    dd.pposv(A2, B2, "L")
    assert B.to_global_matrix().tobytes() == B2.to_global_matrix().tobytes()
    assert A.to_global_matrix().tobytes() == A2.to_global_matrix().tobytes()


def test_not_pd_reports_fp64_info(handle):
    """A final negative pivot: FP32 and FP64 both fail; info comes from the FP64 pposv and B is untouched."""
    n = 130
    a = orc.spd_matrix(n, 4, 0.5)
    a[-1, -1] = -1.0
    b = orc.normal_matrix(4, 1, n, 1)
    A, B = upload(handle, a), upload(handle, b)
    _, it, info = dd.pposv_mixed(A, B, "L")
    assert it == -3 and info == n
    assert B.to_global_matrix().tobytes() == b.tobytes()


def test_trivial_sizes(handle):
    A, B = upload(handle, [[4.0]]), upload(handle, [[2.0, -6.0]])
    _, it, info = dd.pposv_mixed(A, B, "L")
    assert info == 0 and it >= 0
    assert np.array_equal(B.to_global_matrix(), [[0.5, -1.5]])
    # n = 0 returns at once
    A0 = dd.DistributedMatrix(handle, 0, 0, 64, 64, np.float64)
    B0 = dd.DistributedMatrix(handle, 0, 3, 64, 64, np.float64)
    _, it, info = dd.pposv_mixed(A0, B0)
    assert (it, info) == (0, 0)


def test_workspace_api_matches_driver(handle):
    n, nrhs = 333, 6
    a = orc.spd_matrix(n, 6, 1e-3)
    b = orc.normal_matrix(6, 1, n, nrhs)
    A1, B1 = upload(handle, a), upload(handle, b)
    _, it1, info1 = dd.pposv_mixed(A1, B1, "U", tol=1e-14)
    A2, B2 = upload(handle, a), upload(handle, b)
    bb, hb = dd.pposv_mixed_workspace_size(A2, B2, "U")
    assert bb > n * n * 4 and hb == 0
    W = dd.DistributedMatrix(handle, bb // 8 + 1, 1, 64, 64, np.float64)
    it2, info2 = dd.pposv_mixed_run(A2, B2, "U", W.data_ptr, bb, tol=1e-14)
    assert (it1, info1) == (it2, info2)
    assert B1.to_global_matrix().tobytes() == B2.to_global_matrix().tobytes()
    with pytest.raises(dd.DdlaExceptionError):
        dd.pposv_mixed_run(upload(handle, a), upload(handle, b), "U", W.data_ptr, bb - 1)


def test_invalid_arguments(handle):
    a = orc.spd_matrix(64, 1, 1.0)
    A, B = upload(handle, a), upload(handle, np.ones((64, 1)))
    for factor in (4, -1):
        with pytest.raises(dd.DdlaExceptionError):
            dd.pposv_mixed(A, B, factor=factor)
    with pytest.raises(dd.DdlaExceptionError):
        dd.pposv_mixed(A, B, tol=0.0)
    with pytest.raises(dd.DdlaExceptionError):
        dd.pposv_mixed(A, B, max_iter=-1)
    with pytest.raises(dd.DdlaExceptionError):
        dd.pposv_mixed(A, B, uplo="X")
    A32 = dd.DistributedMatrix.from_global_matrix(handle, a.astype(np.float32))
    with pytest.raises(TypeError):
        dd.pposv_mixed(A32, B)
    it, info = ctypes.c_int(), ctypes.c_int()
    st = dd.lib.ddla_d_pposv_mixed(handle.ptr, b"L", 64, 1, ctypes.c_void_p(A.data_ptr), 1, 1, A.desc_ptr,
                                   ctypes.c_void_p(B.data_ptr), 1, 1, B.desc_ptr, 0, 1e-14, 30, None, None)
    assert st == 2  # null iter/info
