"""Pins the CPU oracle (oracle/ddla_oracle.c) against facts that do not come from it.

Each oracle function is checked against at least one of: library output
(numpy's Philox, scipy's LAPACK fixtures), the reference's closed-form test
matrices, brute force on tiny inputs, or invariants; and the planted-mistake
tests show that a plausible error in a factor or a solution fails these checks.
"""
import math
import os

import numpy as np
import pytest

from oracle import oracle as orc

GOLD = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lapack_cases.npz"))
M64 = (1 << 64) - 1


def numpy_philox_block(ctr, key):
    """numpy's Philox increments its 256-bit counter before generating a block."""
    c = sum(int(w) << (64 * i) for i, w in enumerate(ctr))
    c = (c - 1) % (1 << 256)
    counter = [(c >> (64 * i)) & M64 for i in range(4)]
    bg = np.random.Philox(key=np.array(key, dtype=np.uint64), counter=np.array(counter, dtype=np.uint64))
    return [int(x) for x in bg.random_raw(4)]


@pytest.mark.parametrize("key", [(0, 0), (7, 11), (M64, 5)])
@pytest.mark.parametrize("ctr", [(0, 0, 0, 0), (1, 0, 0, 0), (5, 9, 0, 0), (M64, 3, 7, 2)])
def test_philox_matches_numpy(ctr, key):
    assert orc.philox(ctr, key) == numpy_philox_block(ctr, key)


def test_fill_normal_is_box_muller_of_pinned_words():
    seed, stream = 3, 1
    vals = orc.fill_normal(seed, stream, 5, 40)
    for t, v in enumerate(vals):
        e = 5 + t
        w = numpy_philox_block((e >> 2, stream, 0, 0), (seed, 0))
        word, pair = e & 3, (e & 3) >> 1
        u1 = ((w[2 * pair] >> 11) + 1) * 2.0 ** -53
        u2 = (w[2 * pair + 1] >> 11) * 2.0 ** -53
        r = math.sqrt(-2.0 * math.log(u1))
        z = r * (math.sin(2 * math.pi * u2) if word & 1 else math.cos(2 * math.pi * u2))
        assert v == pytest.approx(z, rel=1e-15, abs=1e-300)


def test_fill_normal_moments():
    x = orc.fill_normal(0, 0, 0, 400_000)
    assert abs(x.mean()) < 0.006
    assert abs(x.var() - 1) < 0.01
    assert abs((x ** 4).mean() - 3) < 0.05


def _tri(uplo, a):
    return np.tril(a) if uplo == "L" else np.triu(a)


def fixture_keys():
    return sorted(k[len("potrf_in_"):] for k in GOLD.files if k.startswith("potrf_in_"))


@pytest.mark.parametrize("key", fixture_keys())
def test_potrf_matches_lapack_fixture(key):
    tag, uplo, n, pad = key.split("_")
    n = int(n)
    a = GOLD[f"potrf_in_{key}"]
    f, info = orc.potrf(uplo, a)
    assert info == 0
    want = GOLD[f"potrf_out_{key}"]
    got = _tri(uplo, f[:n, :n])
    tol = 1e-13 if tag in "dz" else 2e-6
    assert np.abs(got - want).max() <= tol * max(1.0, np.abs(want).max())
    # the other triangle and the padding are neither read nor written: they stay NaN
    other = np.triu(np.ones((n, n), bool), 1) if uplo == "L" else np.tril(np.ones((n, n), bool), -1)
    assert np.isnan(f[:n, :n][other]).all()
    assert np.isnan(f[n:, :]).all()


@pytest.mark.parametrize("key", fixture_keys())
def test_potrs_matches_lapack_fixture(key):
    tag, uplo, n, pad = key.split("_")
    factor = GOLD[f"potrf_out_{key}"]
    b = GOLD[f"potrs_b_{key}"]
    x = orc.potrs("L", uplo, factor, b)
    want = GOLD[f"potrs_x_{key}"]
    tol = 1e-12 if tag in "dz" else 1e-4
    assert np.abs(x - want).max() <= tol * max(1.0, np.abs(want).max())


def analytic_factor(n):
    """Real part of factor(i, j) in tests/test_ppotrf_grids.cpp:17-20."""
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    l = 0.003 * (1 + (i + j) % 5)
    l = np.where(i == j, 2 + 0.01 * i, l)
    return np.tril(l)


@pytest.mark.parametrize("n", [1, 2, 17, 38, 129])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_potrf_recovers_analytic_factor(n, uplo):
    l = analytic_factor(n)
    a = l @ l.T
    f, info = orc.potrf(uplo, a)
    assert info == 0
    got = np.tril(f) if uplo == "L" else np.triu(f).T
    assert (np.abs(got - l) / (1 + np.abs(l))).max() < 3e-12


def test_potrf_n1_known_answer():
    f, info = orc.potrf("L", np.array([[4.0]]))
    assert info == 0 and f[0, 0] == 2.0


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_potrf_final_negative_pivot_reports_n(dtype, uplo):
    a = np.eye(9, dtype=dtype)
    a[-1, -1] = -1
    _, info = orc.potrf(uplo, a)
    assert info == 9


def test_potrf_nan_pivot_reports_its_index():
    a = np.eye(5)
    a[2, 2] = np.nan
    _, info = orc.potrf("L", a)
    assert info == 3


def test_potrf_reconstruction_and_planted_mistakes():
    rng = np.random.RandomState(1)
    n = 60
    m = rng.standard_normal((n, n))
    a = m @ m.T / n + np.eye(n)
    f, info = orc.potrf("L", a)
    l = np.tril(f)
    tol = 200 * np.finfo(float).eps * n * n  # tests/test_backend_lapack.cpp:313-334
    assert np.abs(a - l @ l.T).max() <= tol
    # a dropped term of the update, a transposed factor and a wrong sign are all caught
    bad = l.copy()
    bad[30, 10] += 1e-9
    assert np.abs(a - bad @ bad.T).max() > tol
    assert np.abs(a - l.T @ l).max() > tol
    bad = l.copy()
    bad[40, 3] = -bad[40, 3]
    assert np.abs(a - bad @ bad.T).max() > tol


def hpd_real(n):
    """Real part of hpd_value (tests/api_grid_test_common.h:490-500)."""
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    lo, hi = np.minimum(i, j), np.maximum(i, j)
    a = 0.01 * ((lo + 2 * hi) % 5 - 2)
    return np.where(i == j, 5.0 + 0.2 * n + 0.05 * i, a)


def x_real(n, nrhs):
    """Real part of x_value (tests/api_grid_test_common.h:509-512)."""
    i, j = np.meshgrid(np.arange(n), np.arange(nrhs), indexing="ij")
    return 0.2 + 0.03 * i - 0.02 * j


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("side", ["L", "R"])
def test_potrs_closed_form_solution(uplo, side):
    n, nrhs = 40, 5
    a = hpd_real(n)
    f, info = orc.potrf(uplo, a)
    assert info == 0
    x_true = x_real(n, nrhs) if side == "L" else x_real(n, nrhs).T
    b = a @ x_true if side == "L" else x_true @ a
    x = orc.potrs(side, uplo, f, b)
    assert np.abs(x - x_true).max() < 5e-9
    # planted mistake: solving with the transposed factor misses the closed form
    wrong = orc.potrs(side, uplo, f.T.copy(), b)
    assert np.abs(wrong - x_true).max() > 1e-6


def logdet_reference(d):
    """Long-double sum of logs and phase (tests/test_diag_logdet.cpp:21-30)."""
    mag = np.longdouble(0)
    negatives = 0
    for x in d:
        if x == 0:
            return complex(-np.inf, 0)
        mag += np.log(np.longdouble(abs(x)))
        negatives += x < 0
    # for real entries the summed phase atan2(0, x) is exactly pi per negative entry
    return complex(float(mag), math.pi if negatives % 2 else 0.0)


def phase_close(a, b, tol):
    return abs(math.remainder(a - b, 2 * math.pi)) < tol


@pytest.mark.parametrize("d", [[2, 3, 5], [-2, 3, 5], [-2, -3, 5], [0, -3, 5],
                               [np.finfo(float).max, np.finfo(float).tiny * 2 ** -52],
                               [1 + i % 3 for i in range(11)], [-1.0] * 1025, []])
def test_logdet_cases(d):
    n = len(d)
    a = np.full((n + 3, max(n, 1)), np.nan)
    for i, x in enumerate(d):
        a[i, i] = x
    mag, flags = orc.diag_logdet(a, n)
    got = orc.finish_logdet(mag, flags)
    want = logdet_reference(d)
    if math.isinf(want.real):
        assert math.isinf(got.real) and got.real < 0 and got.imag == 0
    else:
        assert abs(got.real - want.real) < 3e-12 * (1 + abs(want.real))
        assert phase_close(got.imag, want.imag, 3e-12)


@pytest.mark.parametrize("d", [[2, 3, 5], [-2, 3, 5], [0, -3, 5], [np.finfo(np.float32).max, np.finfo(np.float32).tiny],
                               [1.1 + i % 3 for i in range(11)], [-0.7] * 1025, []])
def test_logdet_cases_fp32(d):
    """oracle_sdiag_logdet (the float instantiation): the same rules on float32
    entries, against the long-double sum of the float32 values."""
    n = len(d)
    a = np.full((n + 3, max(n, 1)), np.nan, dtype=np.float32)
    for i, x in enumerate(d):
        a[i, i] = x
    mag, flags = orc.diag_logdet(a, n)
    got = orc.finish_logdet(mag, flags)
    want = logdet_reference([float(np.float32(x)) for x in d])
    if math.isinf(want.real):
        assert math.isinf(got.real) and got.real < 0 and got.imag == 0
    else:
        assert abs(got.real - want.real) < 3e-12 * (1 + abs(want.real))
        assert phase_close(got.imag, want.imag, 3e-12)
    bad = a.copy()
    if n:
        bad[0, 0] = np.inf
        assert orc.diag_logdet(bad, n)[1] & 1


@pytest.mark.parametrize("bad", [np.inf, np.nan, -np.inf])
def test_logdet_nonfinite_flag(bad):
    a = np.diag([1.0, bad])
    mag, flags = orc.diag_logdet(a)
    assert flags & 1
    with pytest.raises(ValueError):
        orc.finish_logdet(mag, flags)


def closed_form(count, shape):
    """value(i) = ((i % 13) - 6) / 7 of tests/test_gemm_vbatched.cpp:26-43, column-major fill."""
    v = ((np.arange(count) % 13) - 6) / 7.0
    return v.reshape(shape[::-1]).T.copy()


@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
def test_gemm_against_numpy(ta, tb):
    m, n, k = 7, 5, 9
    a = closed_form(m * k, (m, k) if ta == "N" else (k, m))
    b = closed_form(k * n, (k, n) if tb == "N" else (n, k))
    c = closed_form(m * n, (m, n)) + 0.25
    opa = a if ta == "N" else a.T
    opb = b if tb == "N" else b.T
    got = orc.gemm(ta, tb, m, n, k, 1.5, a, b, -0.5, c)
    assert np.abs(got - (1.5 * opa @ opb - 0.5 * c)).max() < 1e-13


@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
def test_sgemm_against_numpy(ta, tb):
    """oracle_sgemm (the float instantiation) on the same closed-form inputs,
    against numpy's product of the float32 operands in float64."""
    m, n, k = 7, 5, 9
    a = closed_form(m * k, (m, k) if ta == "N" else (k, m)).astype(np.float32)
    b = closed_form(k * n, (k, n) if tb == "N" else (n, k)).astype(np.float32)
    c = (closed_form(m * n, (m, n)) + 0.25).astype(np.float32)
    opa = (a if ta == "N" else a.T).astype(np.float64)
    opb = (b if tb == "N" else b.T).astype(np.float64)
    got = orc.gemm(ta, tb, m, n, k, 1.5, a, b, -0.5, c)
    assert got.dtype == np.float32
    assert np.abs(got - (1.5 * opa @ opb - 0.5 * c)).max() < 1e-5
    # alpha 0 and k 0 scale C only; beta 0 never reads C
    nan = np.full((m, n), np.nan, dtype=np.float32)
    assert (orc.gemm(ta, tb, m, n, k, 0.0, np.full_like(a, np.nan), b, 2.0, c) == 2.0 * c).all()
    assert np.abs(orc.gemm(ta, tb, m, n, k, 1.0, a, b, 0.0, nan) - opa @ opb).max() < 1e-5


def test_gemm_special_cases():
    a = np.full((3, 2), np.nan)
    b = np.full((2, 4), np.nan)
    c = np.ones((3, 4))
    assert (orc.gemm("N", "N", 3, 4, 2, 0.0, a, b, 2.0, c) == 2.0).all()  # alpha 0: A, B unread
    assert (orc.gemm("N", "N", 3, 4, 0, 1.0, a, b, 3.0, c) == 3.0).all()  # k 0: scale only
    a = np.ones((3, 2))
    b = np.ones((2, 4))
    c = np.full((3, 4), np.nan)
    assert (orc.gemm("N", "N", 3, 4, 2, 1.0, a, b, 0.0, c) == 2.0).all()  # beta 0: C unread


def bottom_right_factor(uplo, n):
    """The factor of tests/test_potrf_bottom_right.cpp:96-118 (real part):
    diagonal 1.5 + 0.002 (i % 23), off-diagonal 0.004 (((3i + 5j) % 11) - 5),
    on the uplo triangle."""
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    f = 0.004 * (((3 * i + 5 * j) % 11) - 5)
    f = np.where(i == j, 1.5 + 0.002 * (i % 23), f)
    return np.triu(f) if uplo == "U" else np.tril(f)


def bottom_right_product(uplo, f):
    """A = U U^T (uplo U) or L^T L (uplo L)."""
    return f @ f.T if uplo == "U" else f.T @ f


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [1, 2, 70, 129])
@pytest.mark.parametrize("uplo", ["U", "L"])
def test_potrf_bottom_right_recovers_known_factor(uplo, n, dtype):
    """The reference test's acceptance: factor error <= 128 eps n, other
    triangle untouched (here: a sentinel)."""
    f = bottom_right_factor(uplo, n)
    a = bottom_right_product(uplo, f).astype(dtype)
    other = np.tril(np.ones((n, n), bool), -1) if uplo == "U" else np.triu(np.ones((n, n), bool), 1)
    a[other] = -73.0
    got, info = orc.potrf_bottom_right(uplo, a)
    assert info == 0
    keep = ~other
    assert np.abs(got[keep] - f[keep]).max() <= 128 * np.finfo(dtype).eps * n
    assert (got[other] == -73.0).all()


@pytest.mark.parametrize("uplo", ["U", "L"])
def test_potrf_bottom_right_is_cholesky_of_the_reversed_matrix(uplo):
    """With J the order reversal, J A J = V^T V (scipy, upper) gives U = J V^T J,
    and J A J = W W^T (lower) gives L = J W^T J."""
    import scipy.linalg

    a = orc.spd_matrix(97, seed=5, delta=0.1)
    rev = a[::-1, ::-1]
    if uplo == "U":
        want = scipy.linalg.cholesky(rev, lower=False).T[::-1, ::-1]
    else:
        want = scipy.linalg.cholesky(rev, lower=True).T[::-1, ::-1]
    got, info = orc.potrf_bottom_right(uplo, a)
    assert info == 0
    tri = np.triu if uplo == "U" else np.tril
    assert np.abs(tri(got) - want).max() < 1e-13 * np.abs(want).max()
    assert np.allclose(bottom_right_product(uplo, tri(got)), a, rtol=0, atol=1e-13)


@pytest.mark.parametrize("uplo", ["U", "L"])
def test_potrf_bottom_right_reports_the_highest_failing_row(uplo):
    """Two planted negative pivots (rows 3 and 11 of 16): the bottom-right
    order meets row 11 first, so info = 12 (the forward potrf reports 4)."""
    a = np.eye(16) * 4.0
    a[3, 3] = a[11, 11] = -1.0
    _, info = orc.potrf_bottom_right(uplo, a)
    assert info == 12
    _, finfo = orc.potrf(uplo, a)
    assert finfo == 4
    g, info = orc.potrf_bottom_right(uplo, np.array([[9.0]]))
    assert info == 0 and g[0, 0] == 3.0


# --------------------------------------------------------------------------
# Complex (Hermitian) oracle functions.
# --------------------------------------------------------------------------
def hpd_value(n):
    """hpd_value of tests/api_grid_test_common.h:490-500 (Hermitian, diagonally dominant)."""
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    lo, hi = np.minimum(i, j), np.maximum(i, j)
    val = 0.01 * ((lo + 2 * hi) % 5 - 2) + 1j * 0.006 * ((3 * lo + hi) % 7 - 3)
    a = np.where(i < j, val, np.conj(val))
    return np.where(i == j, 5.0 + 0.2 * n + 0.05 * i + 0j, a)


def triangular_l_value(n):
    """triangular_l_value of tests/api_grid_test_common.h:502-507 (lower, real positive diagonal)."""
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    off = 0.04 * ((i + j) % 4 + 1) + 1j * 0.01 * ((i - j) % 3)
    return np.where(i < j, 0j, np.where(i == j, 2.0 + 0.05 * i + 0j, off))


def x_value(rows, cols):
    """x_value of tests/api_grid_test_common.h:509-512."""
    i, j = np.meshgrid(np.arange(rows), np.arange(cols), indexing="ij")
    return (0.2 + 0.03 * i - 0.02 * j) + 1j * (0.04 * (i + 1) + 0.01 * j)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", [1, 2, 40, 129])
def test_hpotrf_recovers_known_factor(n, dtype):
    """A = L L^H from the reference's closed-form L: the unique Cholesky factor with a
    positive real diagonal is L itself (uplo 'L') and L^H (uplo 'U')."""
    l = triangular_l_value(n)
    a = (l @ l.conj().T).astype(dtype)
    tol = 1e-13 if dtype == np.complex128 else 2e-6
    scale = np.abs(l).max()
    for uplo, want in (("L", l), ("U", l.conj().T)):
        f, info = orc.potrf(uplo, a)
        assert info == 0
        assert np.abs(_tri(uplo, f) - want).max() <= tol * 10 * scale
    # planted mistake: the factor of the transposed (conjugated) matrix is conj(L), not L
    f, _ = orc.potrf("L", a.conj().copy())
    if n > 1:
        assert np.abs(np.tril(f) - l).max() > 1e-3


def test_hpotrf_upper_is_conjugate_transpose_of_lower():
    rng = np.random.RandomState(5)
    m = rng.standard_normal((37, 37)) + 1j * rng.standard_normal((37, 37))
    a = m @ m.conj().T / 37 + np.eye(37)
    a = (a + a.conj().T) / 2  # exactly Hermitian
    fl, _ = orc.potrf("L", a)
    fu, _ = orc.potrf("U", a)
    # conj commutes exactly with complex +, - and *: the two triangles agree bit for bit
    assert np.array_equal(np.triu(fu), np.tril(fl).conj().T)


def test_hpotrf_ignores_the_imaginary_part_of_the_diagonal():
    l = triangular_l_value(9)
    a = l @ l.conj().T
    noisy = a + 1j * np.diag(np.linspace(-3, 3, 9))
    f0, _ = orc.potrf("L", a)
    f1, _ = orc.potrf("L", noisy)
    assert np.array_equal(np.tril(f0), np.tril(f1))
    assert (np.diag(f1).imag == 0).all()


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_hpotrf_pivot_info(uplo, dtype):
    l = triangular_l_value(12)
    a = (l @ l.conj().T).astype(dtype)
    bad = a.copy()
    bad[11, 11] = -1.0 + 5j  # the real part decides: -1 - sum |l|^2 < 0
    assert orc.potrf(uplo, bad)[1] == 12
    bad = a.copy()
    bad[4, 4] = np.nan
    assert orc.potrf(uplo, bad)[1] == 5
    bad = a.copy()
    bad[0, 0] = 0.0
    f, info = orc.potrf(uplo, bad)
    assert info == 1 and f[0, 0] == 0


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("side", ["L", "R"])
def test_hpotrs_closed_form_solution(side, uplo, dtype):
    """B = A X (side L) or X A (side R) with the reference's hpd_value / x_value; the solve
    recovers X (tests/api_grid_test_common.h:524-563 build_rhs_side / check_solution)."""
    n, nrhs = 40, 5
    a = hpd_value(n)
    x_true = x_value(n, nrhs) if side == "L" else x_value(nrhs, n)
    b = a @ x_true if side == "L" else x_true @ a
    f, info = orc.potrf(uplo, a.astype(dtype))
    assert info == 0
    x = orc.potrs(side, uplo, f, b.astype(dtype))
    tol = 1e-12 if dtype == np.complex128 else 2e-5
    assert np.abs(x - x_true).max() < tol
    # planted mistake: the transposed factor (conjugation dropped) misses the closed form
    wrong = orc.potrs(side, uplo, f.T.copy(), b.astype(dtype))
    assert np.abs(wrong - x_true).max() > 1e-4


def complex_logdet_reference(d):
    """tests/test_diag_logdet.cpp:21-30: sum of log hypot and of atan2, phase reduced by remainder.
    The magnitude is summed in 50-digit decimal from the exact |z|^2, which neither overflows
    at the largest float nor loses digits at subnormals (where a binary hypot would)."""
    import decimal

    ctx = decimal.Context(prec=50)
    mag, phases = decimal.Decimal(0), []
    for z in d:
        x, y = decimal.Decimal(z.real), decimal.Decimal(z.imag)
        r2 = ctx.add(ctx.multiply(x, x), ctx.multiply(y, y))
        if r2 == 0:
            return complex(-np.inf, 0)
        mag = ctx.add(mag, ctx.divide(ctx.ln(r2), 2))
        phases.append(math.atan2(z.imag, z.real))
    # fsum: the correctly rounded sum of the terms (the C++ test sums them in long double)
    return complex(float(mag), math.remainder(math.fsum(phases), 2 * math.pi))


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("case", ["2,3,5", "-2,3,5", "-2,-3,5", "0,-3,5", "max", "polar", "ones11", "neg1025",
                                  "empty", "random"])
def test_complex_logdet_cases(case, dtype):
    """The cases of tests/test_diag_logdet.cpp:78-110 for complex types, on a NaN-poisoned
    array with lda = n + 3."""
    rt = np.float64 if dtype == np.complex128 else np.float32
    fmax, tiny = np.finfo(rt).max, np.finfo(rt).smallest_subnormal
    d = {"2,3,5": [2, 3, 5], "-2,3,5": [-2, 3, 5], "-2,-3,5": [-2, -3, 5], "0,-3,5": [0, -3, 5],
         "max": [complex(fmax, fmax), complex(tiny, -tiny)],
         "polar": [complex(math.cos(0.8 * math.pi), math.sin(0.8 * math.pi))] * 3,
         "ones11": [1 + i % 3 for i in range(11)], "neg1025": [-1.0] * 1025, "empty": [],
         "random": list(np.random.RandomState(3).standard_normal(50) + 1j * np.random.RandomState(4).standard_normal(50))}[case]
    d = [complex(dtype(z)) for z in d]
    n = len(d)
    a = np.full((n + 3, max(n, 1)), np.nan + 1j * np.nan, dtype=dtype)
    for i, z in enumerate(d):
        a[i, i] = z
    mag, phase, flags = orc.diag_logdet(a, n)
    got = orc.finish_logdet(mag, flags, phase)
    want = complex_logdet_reference(d)
    if math.isinf(want.real):
        assert math.isinf(got.real) and got.real < 0 and got.imag == 0
    else:
        assert abs(got.real - want.real) < 3e-12 * (1 + abs(want.real))
        assert phase_close(got.imag, want.imag, 3e-12)
    if case == "neg1025":
        assert got == complex(0, math.pi)
    if case == "random":
        assert abs(got.real - np.log(np.abs(np.array(d))).sum()) < 1e-12
        assert phase_close(got.imag, float(np.angle(np.array(d)).sum()), 1e-12)


@pytest.mark.parametrize("bad", [complex(1, np.inf), complex(np.nan, 0), complex(-np.inf, 2)])
def test_complex_logdet_nonfinite_flag(bad):
    a = np.diag([1.0 + 0j, bad])
    mag, phase, flags = orc.diag_logdet(a)
    assert flags & 1
    with pytest.raises(ValueError):
        orc.finish_logdet(mag, flags, phase)


OPS = ["N", "T", "C"]


def _op(x, t):
    return x if t == "N" else (x.T if t == "T" else x.conj().T)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("ta", OPS)
@pytest.mark.parametrize("tb", OPS)
def test_complex_gemm_against_numpy(ta, tb, dtype):
    rng = np.random.RandomState(8)
    m, n, k = 7, 5, 9
    cplx = lambda *s: (rng.standard_normal(s) + 1j * rng.standard_normal(s)).astype(dtype)
    a = cplx(*((m, k) if ta == "N" else (k, m)))
    b = cplx(*((k, n) if tb == "N" else (n, k)))
    c = cplx(m, n)
    alpha, beta = 0.7 - 1.3j, -0.4 + 0.2j
    got = orc.gemm(ta, tb, m, n, k, alpha, a, b, beta, c)
    want = alpha * (_op(a.astype(np.complex128), ta) @ _op(b.astype(np.complex128), tb)) + beta * c
    tol = 1e-13 if dtype == np.complex128 else 2e-6
    assert np.abs(got - want).max() < tol * np.abs(want).max() * 10


@pytest.mark.parametrize("real_side", ["A", "B"])
@pytest.mark.parametrize("t", OPS)
def test_mixed_real_complex_gemm(real_side, t):
    """One real operand (include/ddla/ddla.h:1878-1960): 'C' on the real operand equals 'T'."""
    rng = np.random.RandomState(9)
    m, n, k = 6, 4, 11
    c = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
    if real_side == "A":
        a = rng.standard_normal((m, k) if t == "N" else (k, m))
        b = rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n))
        got = orc.gemm(t, "N", m, n, k, 1.5j, a, b, 2.0, c)
        want = 1.5j * (_op(a, t) @ b) + 2.0 * c
    else:
        a = rng.standard_normal((m, k)) + 1j * rng.standard_normal((m, k))
        b = rng.standard_normal((k, n) if t == "N" else (n, k))
        got = orc.gemm("N", t, m, n, k, 1.5j, a, b, 2.0, c)
        want = 1.5j * (a @ _op(b, t)) + 2.0 * c
    assert np.abs(got - want).max() < 1e-13 * np.abs(want).max() * 10


def test_complex_gemm_special_cases():
    """src/vbatched/gemm_template_kernel_vbatched.hpp:46-105: alpha == 0 or k == 0 never reads A, B;
    beta == 0 never reads C."""
    nanc = np.full((4, 4), np.nan + 1j * np.nan)
    c = np.arange(16).reshape(4, 4) * (1 - 1j)
    got = orc.gemm("N", "N", 4, 4, 4, 0, nanc, nanc, 2 - 1j, c)
    assert np.array_equal(got, (2 - 1j) * c)
    got = orc.gemm("N", "N", 4, 4, 0, 1, nanc, nanc, 0.5, c)
    assert np.array_equal(got, 0.5 * c)
    eye = np.eye(4) + 0j
    got = orc.gemm("N", "C", 4, 4, 4, 1j, eye, eye, 0, nanc)
    assert np.array_equal(got, 1j * eye)
