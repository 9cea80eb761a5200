"""Writes tests/golden/lapack_cases.npz: LAPACK ?potrf / ?potrs results on small inputs
(real d/s and Hermitian z/c).

The expected outputs come from scipy's bundled LAPACK (OpenBLAS 0.3.29), the
library the reference's CPU handle calls at a 1 x 1 grid
(src/Cholesky/potrf.cpp:46-49 dpotrf_/spotrf_, src/BLAS/trsm.cpp:135-141).
Inputs are numpy-seeded; the unused triangle of every potrf input is NaN, as in
tests/test_backend_lapack.cpp:261-334, so a kernel that reads it fails.
Run: python tests/golden/make_golden.py
"""
import os

import numpy as np
from scipy.linalg import lapack

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lapack_cases.npz")


def main():
    rng = np.random.RandomState(20261202)
    cases = {}
    for dtype, tag in ((np.float64, "d"), (np.float32, "s")):
        potrf = lapack.dpotrf if tag == "d" else lapack.spotrf
        potrs = lapack.dpotrs if tag == "d" else lapack.spotrs
        for n, pad in ((1, 0), (2, 3), (5, 0), (17, 3), (33, 0), (64, 2), (97, 0)):
            for uplo in ("L", "U"):
                m = rng.standard_normal((n, n))
                a = (m @ m.T / n + np.eye(n)).astype(dtype)
                lda = n + pad
                full = np.full((lda, n), np.nan, dtype=dtype, order="F")
                tri = np.tril if uplo == "L" else np.triu
                full[:n, :n] = np.where(tri(np.ones((n, n), bool)), a, np.nan)
                c, info = potrf(a, lower=(uplo == "L"), clean=0)
                assert info == 0
                factor = tri(c)
                b = rng.standard_normal((n, 3)).astype(dtype)
                x, info2 = potrs(factor, b, lower=(uplo == "L"))
                assert info2 == 0
                key = f"{tag}_{uplo}_{n}_{pad}"
                cases[f"potrf_in_{key}"] = full
                cases[f"potrf_out_{key}"] = factor.astype(dtype)
                cases[f"potrs_b_{key}"] = b
                cases[f"potrs_x_{key}"] = x.astype(dtype)
    # Hermitian cases (zpotrf / cpotrf / zpotrs / cpotrs, src/Cholesky/potrf.cpp:51-59), from a
    # generator of their own so that the real cases above keep their values.
    crng = np.random.RandomState(20261203)
    for dtype, tag in ((np.complex128, "z"), (np.complex64, "c")):
        potrf = lapack.zpotrf if tag == "z" else lapack.cpotrf
        potrs = lapack.zpotrs if tag == "z" else lapack.cpotrs
        for n, pad in ((1, 0), (2, 3), (5, 0), (17, 3), (33, 0), (64, 2), (97, 0)):
            for uplo in ("L", "U"):
                m = crng.standard_normal((n, n)) + 1j * crng.standard_normal((n, n))
                a = (m @ m.conj().T / n + np.eye(n)).astype(dtype)
                # A nonzero imaginary part on the diagonal, which zpotrf ignores.
                a[np.diag_indices(n)] += 1j * crng.standard_normal(n).astype(dtype)
                lda = n + pad
                full = np.full((lda, n), np.nan + 1j * np.nan, dtype=dtype, order="F")
                tri = np.tril if uplo == "L" else np.triu
                full[:n, :n] = np.where(tri(np.ones((n, n), bool)), a, np.nan + 1j * np.nan)
                c, info = potrf(a, lower=(uplo == "L"), clean=0)
                assert info == 0
                factor = tri(c)
                b = (crng.standard_normal((n, 3)) + 1j * crng.standard_normal((n, 3))).astype(dtype)
                x, info2 = potrs(factor, b, lower=(uplo == "L"))
                assert info2 == 0
                key = f"{tag}_{uplo}_{n}_{pad}"
                cases[f"potrf_in_{key}"] = full
                cases[f"potrf_out_{key}"] = factor.astype(dtype)
                cases[f"potrs_b_{key}"] = b
                cases[f"potrs_x_{key}"] = x.astype(dtype)
    np.savez_compressed(OUT, **cases)
    print(f"wrote {OUT}: {len(cases)} arrays, {os.path.getsize(OUT)} bytes")


if __name__ == "__main__":
    main()
