/*
 * ddla_oracle.h -- CPU restatement of LibDDLA's Cholesky path, used ONLY as
 * test infrastructure (the checker in tests/, the smoke check and the CPU
 * baseline leg of bench.py).  Nothing in the shipped library links it.
 *
 * At a 1x1 process grid the reference delegates every routine on this path to
 * a vendor library (LAPACK ?potrf, BLAS ?trsm, rocThrust for the log-det
 * reduction); the functions below restate those algorithms in plain C so that
 * the GPU kernels can be compared with a transparent CPU computation.
 * Citations are into the reference tree (userchba/LibDDLA).
 *
 * Pinning (what fixes each function independently of this restatement; the
 * tests in tests/test_oracle.py run every one of these checks):
 *   philox4x64_10      numpy.random.Philox raw output (same counter/key)
 *   fill_normal        Box-Muller of pinned uniforms; moments; closed form
 *   potrf              scipy LAPACK dpotrf/spotrf golden fixtures, the
 *                      analytic factor of tests/test_ppotrf_grids.cpp:17-26,
 *                      n=1 sqrt (tests/test_api_grid_ppotrf.cpp:17-30),
 *                      info = n for a final -1 pivot
 *                      (python/tests/test_cholesky.py:79-92), reconstruction
 *   potrf_bottom_right the analytic factor of
 *                      tests/test_potrf_bottom_right.cpp:96-147 (U U^T /
 *                      L^T L formed from a known factor), scipy's Cholesky
 *                      of the order-reversed matrix, the failing row of a
 *                      planted negative pivot, reconstruction
 *   potrs              scipy LAPACK dpotrs fixtures, closed-form X_true
 *                      (tests/api_grid_test_common.h:490-512)
 *   diag_logdet        long-double sum of logs, the special cases of
 *                      tests/test_diag_logdet.cpp:76-111
 *   gemm               numpy matmul; zero-k / alpha / beta rules of
 *                      src/vbatched/gemm_template_kernel_vbatched.hpp:46-105
 *   zpotrf / cpotrf    scipy LAPACK zpotrf/cpotrf golden fixtures (with a
 *                      nonzero imaginary diagonal, which LAPACK ignores),
 *                      the closed-form factor triangular_l_value of
 *                      tests/api_grid_test_common.h:502-507, upper factor =
 *                      conjugate transpose of the lower one, pivot info
 *   zpotrs / cpotrs    scipy LAPACK zpotrs/cpotrs fixtures; closed forms
 *                      hpd_value / x_value (tests/api_grid_test_common.h:
 *                      490-512) on both sides
 *   z/c diag_logdet    50-digit decimal log|z| and atan2 sums for the cases
 *                      of tests/test_diag_logdet.cpp:78-110 (largest value,
 *                      subnormals, polar(1, 0.8 pi) x 3, 1025 x -1), numpy
 *                      log|z| / angle on random entries
 *   z/c gemm           numpy matmul for every N/T/C pair, one real operand
 *                      ('C' = 'T' there), the alpha / k / beta rules
 * No function is left "parity unpinned".
 */
#ifndef DDLA_ORACLE_H
#define DDLA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Philox4x64-10 block (Salmon et al., SC'11): out = philox(ctr, key). */
void oracle_philox4x64_10(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]);

/* The project's "seed s" convention: element e of stream `stream` is taken
 * from Philox block e/4 (counter {e/4, stream, 0, 0}, key {seed, 0}); words
 * (0,1) and (2,3) of a block each give one Box-Muller pair
 * u1 = ((w0 >> 11) + 1) * 2^-53 in (0,1], u2 = (w1 >> 11) * 2^-53,
 * z0 = sqrt(-2 ln u1) cos(2 pi u2), z1 = sqrt(-2 ln u1) sin(2 pi u2).
 * Fills out[0..count) with elements first..first+count-1. */
void oracle_fill_normal(uint64_t seed, uint64_t stream, uint64_t first, int64_t count, double* out);

/* LAPACK ?potrf (src/Cholesky/potrf.cpp:41-59 calls dpotrf_/spotrf_):
 * A = L L^T (uplo 'L') or U^T U ('U'); only the uplo triangle is read and
 * written; info = j+1 for the first non-positive (or NaN) pivot. */
void oracle_dpotrf(char uplo, int n, double* A, int lda, int* info);
void oracle_spotrf(char uplo, int n, float* A, int lda, int* info);

/* potrf_bottom_right (include/ddla/ddla.h:1072-1093, src/Cholesky/
 * potrf_bottom_right.cpp:221-273): A = U U^T (uplo 'U') or L^T L ('L'),
 * columns from the last to the first; only the uplo triangle is read and
 * written; info = j+1 for the first non-positive pivot, j descending. */
void oracle_dpotrf_bottom_right(char uplo, int n, double* A, int lda, int* info);
void oracle_spotrf_bottom_right(char uplo, int n, float* A, int lda, int* info);

/* ppotrs at 1x1 (src/Cholesky/ppotrs.cpp:31-62 -> ptrtrs.cpp:150-155 ->
 * BLAS ?trsm twice).  side 'L': B (n x nrhs) := A^-1 B; side 'R':
 * B (nrhs x n) := B A^-1, with A = L L^T / U^T U given by its factor. */
void oracle_dpotrs(char side, char uplo, int n, int nrhs, const double* F, int ldf, double* B, int ldb);
void oracle_spotrs(char side, char uplo, int n, int nrhs, const float* F, int ldf, float* B, int ldb);

/* diag_logdet (src/diag_logdet_impl.h:50-113, src/diag_logdet.cpp:42-55),
 * real types: *magnitude = sum log|d_i| accumulated in double,
 * *flags = 1 nonfinite | 2 zero | 4 odd number of negative entries. */
void oracle_ddiag_logdet(int n, const double* A, int lda, double* magnitude, int* flags);
void oracle_sdiag_logdet(int n, const float* A, int lda, double* magnitude, int* flags);

/* Naive GEMM with the vbatched kernel's rules
 * (src/vbatched/gemm_template_kernel_vbatched.hpp:46-105): k == 0 or
 * alpha == 0 only scales C; beta == 0 never reads C. ops 'N','T','C'. */
void oracle_dgemm(char ta, char tb, int m, int n, int k, double alpha, const double* A, int lda,
                  const double* B, int ldb, double beta, double* C, int ldc);
void oracle_sgemm(char ta, char tb, int m, int n, int k, float alpha, const float* A, int lda,
                  const float* B, int ldb, float beta, float* C, int ldc);

/* Complex (Hermitian) routines; matrices are interleaved (real, imaginary)
 * column-major arrays (std::complex<R> / C99 complex layout).
 *
 * Hermitian ?potrf (LAPACK zpotf2/cpotf2, src/Cholesky/potrf.cpp:51-59):
 * A = L L^H (uplo 'L') or U^H U ('U'); only the uplo triangle is read and
 * written, the imaginary part of the diagonal is not read; info = j+1 for
 * the first pivot whose real value is not positive (or is NaN). */
void oracle_zpotrf(char uplo, int n, void* A, int lda, int* info);
void oracle_cpotrf(char uplo, int n, void* A, int lda, int* info);

/* Hermitian ppotrs at 1x1: side 'L': B := A^-1 B; side 'R': B := B A^-1. */
void oracle_zpotrs(char side, char uplo, int n, int nrhs, const void* F, int ldf, void* B, int ldb);
void oracle_cpotrs(char side, char uplo, int n, int nrhs, const void* F, int ldf, void* B, int ldb);

/* diag_logdet for complex types (src/diag_logdet_impl.h:50-94): magnitude
 * sum of log|d_i|, phase sum of arg d_i over the entries with a nonzero
 * imaginary part, flags 1 nonfinite | 2 zero | 4 odd count of negative
 * real entries. */
void oracle_zdiag_logdet(int n, const void* A, int lda, double* magnitude, double* phase, int* flags);
void oracle_cdiag_logdet(int n, const void* A, int lda, double* magnitude, double* phase, int* flags);

/* Complex GEMM, also with one real operand (a_complex / b_complex 0: that
 * operand is a real array; 'C' on it is 'T'), the mixed gemmVbatched
 * overloads of include/ddla/ddla.h:1878-1960.  alpha, beta: {re, im}. */
void oracle_zgemm(char ta, char tb, int m, int n, int k, const double* alpha, const void* A, int a_complex,
                  int lda, const void* B, int b_complex, int ldb, const double* beta, void* C, int ldc);
void oracle_cgemm(char ta, char tb, int m, int n, int k, const float* alpha, const void* A, int a_complex,
                  int lda, const void* B, int b_complex, int ldb, const float* beta, void* C, int ldc);

#ifdef __cplusplus
}
#endif
#endif
