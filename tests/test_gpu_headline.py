"""Parity of the timed path at its real sizes.

* FP64 ppotrf + pposv at n = 16384 (uplo L and U) and n = 32768 (uplo L) on
  A = G G^T / n + 0.1 I (seed 1) with B n x 64 (seed 1, stream 1): the factor
  against the reference's CPU path -- LAPACK dpotrf on the host
  (src/Cholesky/potrf.cpp:46-49 of the reference; scipy's OpenBLAS here) --
  within 1e-12 relative Frobenius over the factored triangle, and the
  solution's normwise backward error ||A X - B||_inf / (||A||_inf ||X||_inf).
  These orders run the FP64 factorization's 768-column panels.
* The mixed-precision pposv at n = 32768, nrhs = 64, seed 2, delta in {4e-2,
  4e-4, 4e-6} (condition numbers about 1e2, 1e4, 1e6) in every factor mode:
  converged without the FP64 fallback, backward error <= 1e-14, and the forward
  error against the host LAPACK dpotrf + dpotrs solution within 1e-13 kappa
  (first-order perturbation: both solutions sit within kappa times their
  backward errors, <= 1e-14 each, of the exact one; the rest is the gap between
  the infinity-norm and 2-norm condition numbers).
* The 768-column panels on small matrices through ddla_c_set_panel_width:
  n = 2000 (panels 768, 768 and a short 464), both triangles, FP64 and FP32,
  against the oracle; planted negative pivots in block column 11 of the first
  and the second panel report the oracle's info, through ppotrf and through
  the local potrf's device info word.

A, B and G are generated on the GPU (ddla_d_spd_generate / ddla_c_fill_normal,
pinned against the oracle in test_gpu_runtime.py); the host copies feed LAPACK.
"""
import ctypes

import numpy as np
import pytest

import libddla_amd as dd
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def generate(h, n, seed, delta, nrhs):
    """A (n x n, symmetric) and B (column-major n x nrhs, as an (nrhs, n) tensor) on the GPU."""
    import torch

    a0 = torch.empty((n, n), dtype=torch.float64, device="cuda")
    g = torch.empty_like(a0)
    torch.cuda.synchronize()
    dd.check_status(dd.lib.ddla_d_spd_generate(h.ptr, n, seed, delta, ctypes.c_void_p(a0.data_ptr()), n,
                                               ctypes.c_void_p(g.data_ptr())))
    b0 = torch.empty((nrhs, n), dtype=torch.float64, device="cuda")
    dd.check_status(dd.lib.ddla_c_fill_normal(h.ptr, seed, 1, 0, n * nrhs, ctypes.c_void_p(b0.data_ptr())))
    h.synchronize()
    del g
    return a0, b0


def host_fortran(t):
    """Column-major host copy of a device tensor holding a column-major matrix."""
    return t.cpu().numpy().T  # F-contiguous view: element [i, j] = data[i + j * ld]


def triangle_rel_frobenius(got, ref, lower, step=2048):
    """||tri(got) - tri(ref)||_F / ||tri(ref)||_F over the factored triangle, by column blocks."""
    n = ref.shape[0]
    num = den = 0.0
    for j0 in range(0, n, step):
        j1 = min(n, j0 + step)
        if lower:
            g, r = got[j0:, j0:j1], ref[j0:, j0:j1]
            keep = np.ones(g.shape, bool)
            keep[: j1 - j0] = np.tril(np.ones((j1 - j0, j1 - j0), bool))
        else:
            g, r = got[:j1, j0:j1], ref[:j1, j0:j1]
            keep = np.ones(g.shape, bool)
            keep[j0:j1] = np.triu(np.ones((j1 - j0, j1 - j0), bool))
        d = np.where(keep, g - r, 0.0)
        num += float(np.einsum("ij,ij->", d, d))
        rr = np.where(keep, r, 0.0)
        den += float(np.einsum("ij,ij->", rr, rr))
    return (num / den) ** 0.5


def backward_error_dev(a0, b0, x):
    """max over columns of ||b - A x||_inf / (||A||_inf ||x||_inf), FP64 on the GPU (x, b0: (nrhs, n))."""
    r = b0 - x @ a0
    return (r.abs().amax(1) / (a0.abs().sum(1).max() * x.abs().amax(1))).max().item()


@pytest.mark.parametrize("n,uplo", [(16384, "L"), (16384, "U"), (32768, "L")])
def test_fp64_factor_and_solve_vs_host_lapack(handle, n, uplo):
    from scipy.linalg import lapack

    print(f"case headline FP64 ppotrf/pposv n={n} uplo={uplo}", flush=True)
    nrhs = 64
    a0, b0 = generate(handle, n, 1, 0.1, nrhs)
    a, b = a0.clone(), b0.clone()
    A = dd.DistributedMatrix.from_torch(handle, a, n, n, 512, 512)
    B = dd.DistributedMatrix.from_torch(handle, b, n, nrhs, 512, 512)
    _, info = dd.pposv(A, B, uplo)
    assert info == 0
    berr = backward_error_dev(a0, b0, b)
    assert berr < 0.05 * n * EPS, berr  # config 1's bar, ||AX-B|| / (||A|| ||X|| n eps) < 0.05
    got = host_fortran(a)
    del a, A
    ref = host_fortran(a0)  # F-contiguous and writable: LAPACK factors it in place
    ref, rinfo = lapack.dpotrf(ref, lower=1 if uplo == "L" else 0, clean=0, overwrite_a=1)
    assert rinfo == 0
    rel = triangle_rel_frobenius(got, ref, uplo == "L")
    print(f"  factor vs host LAPACK: {rel:.3e}; backward error {berr:.3e}", flush=True)
    assert rel <= 1e-12, rel


@pytest.mark.parametrize("delta,kappa", [(4e-2, 1e2), (4e-4, 1e4), (4e-6, 1e6)])
def test_mixed_config_at_full_size(handle, delta, kappa):
    import torch
    from scipy.linalg import lapack

    n, nrhs, tol = 32768, 64, 1e-14
    print(f"case headline mixed pposv n={n} nrhs={nrhs} delta={delta:g}", flush=True)
    a0, b0 = generate(handle, n, 2, delta, nrhs)
    ah = host_fortran(a0)
    bh = host_fortran(b0)
    c, rinfo = lapack.dpotrf(ah, lower=1, clean=0, overwrite_a=1)
    assert rinfo == 0
    x_ref, rinfo = lapack.dpotrs(c, bh, lower=1)
    assert rinfo == 0
    del ah, c
    a, b = a0.clone(), b0.clone()
    A = dd.DistributedMatrix.from_torch(handle, a, n, n, 512, 512)
    B = dd.DistributedMatrix.from_torch(handle, b, n, nrhs, 512, 512)
    for factor in (dd.FACTOR_FP32, dd.FACTOR_BF16X3, dd.FACTOR_BF16X6, dd.FACTOR_FP16X3):
        a.copy_(a0)
        b.copy_(b0)
        torch.cuda.synchronize()
        _, it, info = dd.pposv_mixed(A, B, "L", factor=factor, tol=tol)
        assert info == 0 and 1 <= it <= 30, (factor, it, info)
        berr = backward_error_dev(a0, b0, b)
        x = host_fortran(b)
        fwd = np.abs(x - x_ref).max() / np.abs(x_ref).max()
        print(f"  factor {factor}: {it} corrections, backward error {berr:.3e}, forward error {fwd:.3e}",
              flush=True)
        assert berr <= tol, (factor, berr)
        assert fwd <= 1e-13 * kappa, (factor, fwd)


@pytest.fixture
def wide_panels():
    previous = dd.lib.ddla_c_set_panel_width(768)
    assert previous == 0
    yield
    assert dd.lib.ddla_c_set_panel_width(0) == 768


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_wide_panels_at_small_order(handle, wide_panels, uplo, dtype):
    """Panels of 768 columns (12 block columns, the FP64 factorization's width
    from order 16384) at n = 2000: two full panels and a short one of 464
    columns with a ragged last block."""
    print(f"case 768-column panels n=2000 uplo={uplo} {np.dtype(dtype).name}", flush=True)
    n, nrhs = 2000, 5
    a = orc.spd_matrix(n, seed=21, delta=0.1)
    b = orc.normal_matrix(21, 1, n, nrhs)
    A = dd.DistributedMatrix.from_global_matrix(handle, a.astype(dtype), mb=512, nb=512)
    B = dd.DistributedMatrix.from_global_matrix(handle, b.astype(dtype), mb=512, nb=512)
    X, info = dd.pposv(A, B, uplo)
    assert info == 0
    tri = np.tril if uplo == "L" else np.triu
    got = tri(A.to_global_matrix())
    ref, rinfo = orc.potrf(uplo, a.astype(dtype))
    assert rinfo == 0
    ref = tri(ref)
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert rel <= (1e-12 if dtype == np.float64 else 2e-5), rel
    x = X.to_global_matrix().astype(np.float64)
    berr = orc.backward_error(a, x, b)
    assert berr <= (0.05 * n * EPS if dtype == np.float64 else 1e-5), berr


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("row", [709, 1479])
def test_wide_panel_planted_pivot(handle, wide_panels, uplo, row):
    """A negative diagonal entry in block column 11 of the first (row 709) or
    the second (row 1479) 768-column panel: ppotrf's host info and the local
    potrf's device info both equal the oracle's (the 1-based row)."""
    import torch

    print(f"case 768-column panels planted pivot row={row} uplo={uplo}", flush=True)
    n = 2000
    a = orc.spd_matrix(n, seed=22, delta=0.1)
    a[row, row] = -1.0
    _, oinfo = orc.potrf(uplo, a)
    assert oinfo == row + 1
    A = dd.DistributedMatrix.from_global_matrix(handle, a, mb=512, nb=512)
    _, info = dd.ppotrf(A, uplo)
    assert info == oinfo
    t = torch.tensor(np.ascontiguousarray(a.T), device="cuda")
    d_info = torch.full((1,), -7, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    assert dd.lib.ddla_d_potrf(handle.ptr, uplo.encode(), n, ctypes.c_void_p(t.data_ptr()), n,
                               ctypes.c_void_p(d_info.data_ptr())) == 0
    handle.synchronize()
    assert int(d_info.item()) == oinfo
