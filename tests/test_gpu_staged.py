"""The staged triangular solve (512-row stages) that ppotrs / pposv take from
order 8192 with fewer than 256 right-hand sides.

Its arithmetic differs from the 64-row sweeps (explicit inverses of the
512 x 512 diagonal blocks), so it is pinned against them and against the
problem itself:
* FP64: the factor is the same computation (bit for bit equal to
  ddla_d_pposv_noalloc's, which always takes the 64-row sweeps), and the
  solutions agree within 1e-11 relative (kappa about 40); the normwise
  backward error ||A X - B||_inf / (||A||_inf ||X||_inf) meets config 1's bar
  0.05 n eps.  Orders on and off the stage grid (8192, 8200: a last stage of 8
  rows, 9000), one and two right-hand-side slabs (70 columns), both triangles
  and both sides.
* FP32: backward error against the FP32 system within 0.05 n eps32.
* The mixed-precision solver's corrections through the staged solve.
* The wait limit forced from the host: the staged sweeps give up, every
  workgroup exits, and the call reports the wait error.

A and B are generated on the GPU (ddla_d_spd_generate / ddla_c_fill_normal).
"""
import ctypes

import numpy as np
import pytest

import libddla_amd as dd

pytestmark = pytest.mark.gpu

EPS = float(np.finfo(np.float64).eps)
EPS32 = float(np.finfo(np.float32).eps)


def generate(h, n, seed, delta, rows, cols):
    """A (n x n, column-major as an (n, n) tensor) and B (rows x cols column-major, as a (cols, rows) tensor)."""
    import torch

    a0 = torch.empty((n, n), dtype=torch.float64, device="cuda")
    g = torch.empty_like(a0)
    torch.cuda.synchronize()
    dd.check_status(dd.lib.ddla_d_spd_generate(h.ptr, n, seed, delta, ctypes.c_void_p(a0.data_ptr()), n,
                                               ctypes.c_void_p(g.data_ptr())))
    b0 = torch.empty((cols, rows), dtype=torch.float64, device="cuda")
    dd.check_status(dd.lib.ddla_c_fill_normal(h.ptr, seed, 1, 0, rows * cols, ctypes.c_void_p(b0.data_ptr())))
    h.synchronize()
    del g
    return a0, b0


def backward_error(a, x, b, side):
    """max over columns (side L) or rows (side R) of ||b - A x|| / (||A|| ||x||), infinity norms, on the GPU.

    Tensors hold column-major matrices transposed: a (n, n) symmetric, x and b (cols, rows)."""
    if side == "L":
        r = b - x @ a  # row j of r = column j of B - A X
        return (r.abs().amax(1) / (a.abs().sum(1).max() * x.abs().amax(1))).max().item()
    r = b - a @ x  # side R: X A = B, row i of B is column i of the tensor's transpose
    return (r.abs().amax(0) / (a.abs().sum(1).max() * x.abs().amax(0))).max().item()


@pytest.mark.parametrize("n,nrhs,uplo,side", [(8192, 64, "L", "L"), (8200, 70, "U", "L"), (9000, 1, "L", "L"),
                                              (8200, 3, "U", "R")])
def test_fp64_staged_matches_64_row_sweeps(handle, n, nrhs, uplo, side):
    import torch

    print(f"case staged FP64 n={n} nrhs={nrhs} uplo={uplo} side={side}", flush=True)
    rows, cols = (n, nrhs) if side == "L" else (nrhs, n)
    a0, b0 = generate(handle, n, 31, 0.1, rows, cols)
    out = []
    for entry in (dd.lib.ddla_d_pposv, dd.lib.ddla_d_pposv_noalloc):
        a, b = a0.clone(), b0.clone()
        A = dd.DistributedMatrix.from_torch(handle, a, n, n, 512, 512)
        B = dd.DistributedMatrix.from_torch(handle, b, rows, cols, 512, 512)
        info = ctypes.c_int(-7)
        torch.cuda.synchronize()
        st = entry(handle.ptr, side.encode(), uplo.encode(), b"N", n, nrhs, ctypes.c_void_p(a.data_ptr()), 1, 1,
                   A.desc_ptr, ctypes.c_void_p(b.data_ptr()), 1, 1, B.desc_ptr, ctypes.byref(info))
        assert st == 0, dd.lib.ddla_c_error_string(st)
        assert info.value == 0
        out.append((a, b))
    (fa, xa), (fb, xb) = out
    assert torch.equal(fa, fb)
    rel = ((xa - xb).abs().max() / xb.abs().max()).item()
    berr = backward_error(a0, xa, b0, side)
    print(f"  staged vs 64-row sweeps {rel:.2e}, backward error {berr:.2e}", flush=True)
    assert rel <= 1e-11, rel
    assert berr <= 0.05 * n * EPS, berr


def test_fp32_staged_backward_error(handle):
    import torch

    n, nrhs = 8200, 70
    print(f"case staged FP32 n={n} nrhs={nrhs}", flush=True)
    a0, b0 = generate(handle, n, 32, 0.1, n, nrhs)
    a32, b32 = a0.float(), b0.float()
    a, b = a32.clone(), b32.clone()
    A = dd.DistributedMatrix.from_torch(handle, a, n, n, 512, 512)
    B = dd.DistributedMatrix.from_torch(handle, b, n, nrhs, 512, 512)
    torch.cuda.synchronize()
    _, info = dd.pposv(A, B, "U")
    assert info == 0
    berr = backward_error(a32.double(), b.double(), b32.double(), "L")
    print(f"  backward error {berr:.2e}", flush=True)
    assert berr <= 0.05 * n * EPS32, berr


@pytest.mark.parametrize("factor", [dd.FACTOR_FP32, dd.FACTOR_BF16X3, dd.FACTOR_FP16X3])
def test_mixed_refinement_through_staged_solve(handle, factor):
    import torch

    n, nrhs, tol = 8200, 70, 1e-14
    print(f"case staged mixed n={n} nrhs={nrhs} factor={factor}", flush=True)
    a0, b0 = generate(handle, n, 33, 4e-4, n, nrhs)
    a, b = a0.clone(), b0.clone()
    A = dd.DistributedMatrix.from_torch(handle, a, n, n, 512, 512)
    B = dd.DistributedMatrix.from_torch(handle, b, n, nrhs, 512, 512)
    torch.cuda.synchronize()
    _, it, info = dd.pposv_mixed(A, B, "L", factor=factor, tol=tol)
    assert info == 0 and 1 <= it <= 30, (it, info)
    assert torch.equal(a, a0)  # A untouched when the refinement converges
    berr = backward_error(a0, b, b0, "L")
    print(f"  {it} corrections, backward error {berr:.2e}", flush=True)
    assert berr <= tol


def test_staged_wait_limit_reported(handle):
    import torch

    n, nrhs = 8192, 8
    print("case staged forced wait limit", flush=True)
    a0, b0 = generate(handle, n, 34, 0.5, n, nrhs)
    a = a0.clone()
    A = dd.DistributedMatrix.from_torch(handle, a, n, n, 512, 512)
    torch.cuda.synchronize()
    _, info = dd.ppotrf(A, "L")
    assert info == 0
    b = b0.clone()
    B = dd.DistributedMatrix.from_torch(handle, b, n, nrhs, 512, 512)
    previous = dd.lib.ddla_c_set_wait_limit(1)
    try:
        assert previous == 1 << 23
        with pytest.raises(RuntimeError, match="wait limit"):
            dd.ppotrs(A, B, "L")
    finally:
        assert dd.lib.ddla_c_set_wait_limit(0) == 1
    b.copy_(b0)
    torch.cuda.synchronize()
    dd.ppotrs(A, B, "L")
    assert backward_error(a0, b, b0, "L") <= 0.05 * n * EPS
