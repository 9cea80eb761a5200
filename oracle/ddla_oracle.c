/*
 * ddla_oracle.c -- CPU restatement of LibDDLA's Cholesky path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker; never linked into, or called
 * by, the shipped library.  See ddla_oracle.h for what pins each function.
 */
#include "ddla_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Philox4x64-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as   */
/* easy as 1, 2, 3", SC'11).  Multipliers and Weyl increments as published. */
/* ------------------------------------------------------------------------ */
static void mulhilo64(uint64_t a, uint64_t b, uint64_t* hi, uint64_t* lo)
{
    const unsigned __int128 p = (unsigned __int128)a * b;
    *hi = (uint64_t)(p >> 64);
    *lo = (uint64_t)p;
}

void oracle_philox4x64_10(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4])
{
    uint64_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint64_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += 0x9E3779B97F4A7C15ULL;
            k1 += 0xBB67AE8584CAA73BULL;
        }
        uint64_t hi0, lo0, hi1, lo1;
        mulhilo64(0xD2E7470EE14C6C93ULL, c0, &hi0, &lo0);
        mulhilo64(0xCA5A826395121157ULL, c2, &hi1, &lo1);
        const uint64_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void oracle_fill_normal(uint64_t seed, uint64_t stream, uint64_t first, int64_t count, double* out)
{
    const double two_pi = 6.283185307179586476925286766559;
    const double scale = 1.0 / 9007199254740992.0; /* 2^-53 */
    const uint64_t key[2] = {seed, 0};
    for (int64_t t = 0; t < count; ++t) {
        const uint64_t e = first + (uint64_t)t;
        const uint64_t ctr[4] = {e >> 2, stream, 0, 0};
        uint64_t w[4];
        oracle_philox4x64_10(ctr, key, w);
        const int word = (int)(e & 3), pair = word >> 1;
        const double u1 = (double)((w[2 * pair] >> 11) + 1) * scale;
        const double u2 = (double)(w[2 * pair + 1] >> 11) * scale;
        const double r = sqrt(-2.0 * log(u1));
        out[t] = (word & 1) ? r * sin(two_pi * u2) : r * cos(two_pi * u2);
    }
}

/* ------------------------------------------------------------------------ */
/* ?potrf: the unblocked Cholesky of LAPACK ?potf2 (the arithmetic that      */
/* LAPACK ?potrf performs, blocked or not, up to rounding order).            */
/* src/Cholesky/potrf.cpp:46-49 is the call this replaces.                   */
/* ------------------------------------------------------------------------ */
#define DEFINE_POTRF(NAME, T, SQRT)                                                     \
void NAME(char uplo, int n, T* A, int lda, int* info)                                   \
{                                                                                       \
    *info = 0;                                                                          \
    const int lower = (uplo == 'L' || uplo == 'l');                                     \
    for (int j = 0; j < n; ++j) {                                                       \
        T ajj = A[j + (size_t)j * lda];                                                 \
        for (int k = 0; k < j; ++k) {                                                   \
            const T v = lower ? A[j + (size_t)k * lda] : A[k + (size_t)j * lda];        \
            ajj -= v * v;                                                               \
        }                                                                               \
        if (!(ajj > 0)) { A[j + (size_t)j * lda] = ajj; *info = j + 1; return; }        \
        ajj = SQRT(ajj);                                                                \
        A[j + (size_t)j * lda] = ajj;                                                   \
        const T r = (T)1 / ajj;                                                         \
        if (lower) {                                                                    \
            for (int k = 0; k < j; ++k) {                                               \
                const T ljk = A[j + (size_t)k * lda];                                   \
                const T* col = A + (size_t)k * lda;                                     \
                T* dst = A + (size_t)j * lda;                                           \
                for (int i = j + 1; i < n; ++i) dst[i] -= col[i] * ljk;                 \
            }                                                                           \
            T* dst = A + (size_t)j * lda;                                               \
            for (int i = j + 1; i < n; ++i) dst[i] *= r;                                \
        } else {                                                                        \
            for (int c = j + 1; c < n; ++c) {                                           \
                T s = A[j + (size_t)c * lda];                                           \
                const T* uc = A + (size_t)c * lda;                                      \
                const T* uj = A + (size_t)j * lda;                                      \
                for (int k = 0; k < j; ++k) s -= uj[k] * uc[k];                         \
                A[j + (size_t)c * lda] = s * r;                                         \
            }                                                                           \
        }                                                                               \
    }                                                                                   \
}
DEFINE_POTRF(oracle_dpotrf, double, sqrt)
DEFINE_POTRF(oracle_spotrf, float, sqrtf)

/* ------------------------------------------------------------------------ */
/* potrf_bottom_right: the Cholesky recurrence run from the last column to   */
/* the first (src/Cholesky/potrf_bottom_right.cpp:221-273 does it in blocks  */
/* of 128 on reversed copies).  uplo 'U': A = U U^T,                         */
/*   U(j,j)^2 = A(j,j) - sum_{k>j} U(j,k)^2,                                 */
/*   U(i,j)   = (A(i,j) - sum_{k>j} U(i,k) U(j,k)) / U(j,j),  i < j;         */
/* uplo 'L': A = L^T L,                                                      */
/*   L(j,j)^2 = A(j,j) - sum_{k>j} L(k,j)^2,                                 */
/*   L(j,i)   = (A(j,i) - sum_{k>j} L(k,j) L(k,i)) / L(j,j),  i < j.         */
/* info = j+1 for the first non-positive (or NaN) pivot met, j descending.   */
/* ------------------------------------------------------------------------ */
#define DEFINE_POTRF_BR(NAME, T, SQRT)                                                  \
void NAME(char uplo, int n, T* A, int lda, int* info)                                   \
{                                                                                       \
    *info = 0;                                                                          \
    const int lower = (uplo == 'L' || uplo == 'l');                                     \
    for (int j = n - 1; j >= 0; --j) {                                                  \
        T ajj = A[j + (size_t)j * lda];                                                 \
        for (int k = j + 1; k < n; ++k) {                                               \
            const T v = lower ? A[k + (size_t)j * lda] : A[j + (size_t)k * lda];        \
            ajj -= v * v;                                                               \
        }                                                                               \
        if (!(ajj > 0)) { A[j + (size_t)j * lda] = ajj; *info = j + 1; return; }        \
        ajj = SQRT(ajj);                                                                \
        A[j + (size_t)j * lda] = ajj;                                                   \
        for (int i = 0; i < j; ++i) {                                                   \
            T s;                                                                        \
            if (lower) {                                                                \
                s = A[j + (size_t)i * lda];                                             \
                for (int k = j + 1; k < n; ++k)                                         \
                    s -= A[k + (size_t)j * lda] * A[k + (size_t)i * lda];               \
                A[j + (size_t)i * lda] = s / ajj;                                       \
            } else {                                                                    \
                s = A[i + (size_t)j * lda];                                             \
                for (int k = j + 1; k < n; ++k)                                         \
                    s -= A[i + (size_t)k * lda] * A[j + (size_t)k * lda];               \
                A[i + (size_t)j * lda] = s / ajj;                                       \
            }                                                                           \
        }                                                                               \
    }                                                                                   \
}
DEFINE_POTRF_BR(oracle_dpotrf_bottom_right, double, sqrt)
DEFINE_POTRF_BR(oracle_spotrf_bottom_right, float, sqrtf)

/* ------------------------------------------------------------------------ */
/* ?potrs: two triangular solves (src/Cholesky/ppotrs.cpp:44-61).  For side  */
/* 'R' the system X A = B is A X^T = B^T (A is symmetric), solved row by row.*/
/* ------------------------------------------------------------------------ */
#define DEFINE_POTRS(NAME, T)                                                           \
static void NAME##_vector(int lower, int n, const T* F, int ldf, T* x)                  \
{                                                                                       \
    /* A = L L^T: L y = b, then L^T x = y.  A = U^T U: U^T y = b, then U x = y. */      \
    for (int i = 0; i < n; ++i) {                                                       \
        T s = x[i];                                                                     \
        for (int k = 0; k < i; ++k)                                                     \
            s -= (lower ? F[i + (size_t)k * ldf] : F[k + (size_t)i * ldf]) * x[k];      \
        x[i] = s / F[i + (size_t)i * ldf];                                              \
    }                                                                                   \
    for (int i = n - 1; i >= 0; --i) {                                                  \
        T s = x[i];                                                                     \
        for (int k = i + 1; k < n; ++k)                                                 \
            s -= (lower ? F[k + (size_t)i * ldf] : F[i + (size_t)k * ldf]) * x[k];      \
        x[i] = s / F[i + (size_t)i * ldf];                                              \
    }                                                                                   \
}                                                                                       \
void NAME(char side, char uplo, int n, int nrhs, const T* F, int ldf, T* B, int ldb)    \
{                                                                                       \
    const int lower = (uplo == 'L' || uplo == 'l');                                     \
    const int left = (side == 'L' || side == 'l');                                      \
    if (n <= 0 || nrhs <= 0) return;                                                    \
    if (left) {                                                                         \
        for (int j = 0; j < nrhs; ++j) NAME##_vector(lower, n, F, ldf, B + (size_t)j * ldb); \
        return;                                                                         \
    }                                                                                   \
    T* row = (T*)malloc(sizeof(T) * (size_t)n);                                         \
    for (int r = 0; r < nrhs; ++r) {                                                    \
        for (int i = 0; i < n; ++i) row[i] = B[r + (size_t)i * ldb];                    \
        NAME##_vector(lower, n, F, ldf, row);                                           \
        for (int i = 0; i < n; ++i) B[r + (size_t)i * ldb] = row[i];                    \
    }                                                                                   \
    free(row);                                                                          \
}
DEFINE_POTRS(oracle_dpotrs, double)
DEFINE_POTRS(oracle_spotrs, float)

/* ------------------------------------------------------------------------ */
/* diag_logdet for real types (src/diag_logdet_impl.h:50-94 term, :42-48     */
/* join).  For a real entry x: nonfinite -> flag 1; 0 -> flag 2; otherwise   */
/* log|x| (log(hi) + 0.5 log1p(0)) and a parity bit for x < 0.               */
/* ------------------------------------------------------------------------ */
#define DEFINE_LOGDET(NAME, T)                                                          \
void NAME(int n, const T* A, int lda, double* magnitude, int* flags)                    \
{                                                                                       \
    double m = 0;                                                                       \
    int f = 0, parity = 0;                                                              \
    const double limit = 1.7976931348623157e308;                                        \
    for (int i = 0; i < n; ++i) {                                                       \
        const double x = (double)A[(size_t)i * lda + i];                                \
        if (!(x <= limit && x >= -limit)) { f |= 1; continue; }                         \
        const double ax = fabs(x);                                                      \
        if (ax == 0) { f |= 2; continue; }                                              \
        m += log(ax); /* log(hi) + 0.5 log1p((lo/hi)^2) with lo = 0 */                   \
        if (x < 0) parity ^= 1;                                                         \
    }                                                                                   \
    *magnitude = m;                                                                     \
    *flags = f | (parity ? 4 : 0);                                                      \
}
DEFINE_LOGDET(oracle_ddiag_logdet, double)
DEFINE_LOGDET(oracle_sdiag_logdet, float)

/* ------------------------------------------------------------------------ */
/* GEMM with the vbatched kernel's special cases.                            */
/* ------------------------------------------------------------------------ */
#define DEFINE_GEMM(NAME, T, ACC)                                                       \
void NAME(char ta, char tb, int m, int n, int k, T alpha, const T* A, int lda,          \
          const T* B, int ldb, T beta, T* C, int ldc)                                   \
{                                                                                       \
    const int na = (ta == 'N' || ta == 'n'), nb = (tb == 'N' || tb == 'n');             \
    for (int j = 0; j < n; ++j) {                                                       \
        for (int i = 0; i < m; ++i) {                                                   \
            ACC s = 0;                                                                  \
            if (alpha != (T)0)                                                          \
                for (int p = 0; p < k; ++p) {                                           \
                    const T a = na ? A[i + (size_t)p * lda] : A[p + (size_t)i * lda];   \
                    const T b = nb ? B[p + (size_t)j * ldb] : B[j + (size_t)p * ldb];   \
                    s += (ACC)a * (ACC)b;                                               \
                }                                                                       \
            const T prod = (T)(alpha * (T)s);                                           \
            T* c = C + i + (size_t)j * ldc;                                             \
            *c = (beta == (T)0) ? prod : (T)(prod + beta * *c);                         \
        }                                                                               \
    }                                                                                   \
}
DEFINE_GEMM(oracle_dgemm, double, double)
DEFINE_GEMM(oracle_sgemm, float, double)

/* ======================================================================== */
/* Complex (Hermitian) routines.  Matrices are interleaved (real, imaginary) */
/* column-major arrays, the layout of std::complex<R> and of C99 complex.    */
/* ======================================================================== */
#include <complex.h>

/* ------------------------------------------------------------------------ */
/* Hermitian ?potrf: the unblocked Cholesky of LAPACK zpotf2 / cpotf2, which */
/* the reference's CPU path calls for complex types (src/Cholesky/           */
/* potrf.cpp:51-59 cpotrf_/zpotrf_).  uplo 'L': A = L L^H,                   */
/*   L(j,j)^2 = Re A(j,j) - sum_{k<j} |L(j,k)|^2,                            */
/*   L(i,j)   = (A(i,j) - sum_{k<j} L(i,k) conj(L(j,k))) / L(j,j),  i > j;   */
/* uplo 'U': A = U^H U with U = L^H.  The imaginary part of a diagonal entry */
/* is never read; the factor's diagonal is real.                             */
/* ------------------------------------------------------------------------ */
#define DEFINE_HPOTRF(NAME, R, C, SQRT)                                                 \
void NAME(char uplo, int n, void* Av, int lda, int* info)                               \
{                                                                                       \
    C* A = (C*)Av;                                                                      \
    *info = 0;                                                                          \
    const int lower = (uplo == 'L' || uplo == 'l');                                     \
    for (int j = 0; j < n; ++j) {                                                       \
        R ajj = creal(A[j + (size_t)j * lda]);                                          \
        for (int k = 0; k < j; ++k) {                                                   \
            const C v = lower ? A[j + (size_t)k * lda] : A[k + (size_t)j * lda];        \
            ajj -= creal(v) * creal(v) + cimag(v) * cimag(v);                           \
        }                                                                               \
        if (!(ajj > 0)) { A[j + (size_t)j * lda] = ajj; *info = j + 1; return; }        \
        ajj = SQRT(ajj);                                                                \
        A[j + (size_t)j * lda] = ajj;                                                   \
        if (lower) {                                                                    \
            for (int i = j + 1; i < n; ++i) {                                           \
                C s = A[i + (size_t)j * lda];                                           \
                for (int k = 0; k < j; ++k)                                             \
                    s -= A[i + (size_t)k * lda] * conj(A[j + (size_t)k * lda]);         \
                A[i + (size_t)j * lda] = s / ajj;                                       \
            }                                                                           \
        } else {                                                                        \
            for (int c = j + 1; c < n; ++c) {                                           \
                C s = A[j + (size_t)c * lda];                                           \
                for (int k = 0; k < j; ++k)                                             \
                    s -= conj(A[k + (size_t)j * lda]) * A[k + (size_t)c * lda];         \
                A[j + (size_t)c * lda] = s / ajj;                                       \
            }                                                                           \
        }                                                                               \
    }                                                                                   \
}
DEFINE_HPOTRF(oracle_zpotrf, double, double complex, sqrt)
DEFINE_HPOTRF(oracle_cpotrf, float, float complex, sqrtf)

/* ------------------------------------------------------------------------ */
/* Hermitian ?potrs (src/Cholesky/ppotrs.cpp:44-61, two ?trsm).  With        */
/* L(i,k) = F(i,k) (uplo 'L') or conj(F(k,i)) ('U'): L y = b, then           */
/* L^H x = y.  Side 'R' (X A = B): A Hermitian gives A conj(x)^T =           */
/* conj(b)^T for every row, solved as a left system and conjugated back.     */
/* ------------------------------------------------------------------------ */
#define DEFINE_HPOTRS(NAME, C)                                                          \
static C NAME##_l(int lower, const C* F, int ldf, int i, int k)                         \
{                                                                                       \
    return lower ? F[i + (size_t)k * ldf] : conj(F[k + (size_t)i * ldf]);               \
}                                                                                       \
static void NAME##_vector(int lower, int n, const C* F, int ldf, C* x)                  \
{                                                                                       \
    for (int i = 0; i < n; ++i) {                                                       \
        C s = x[i];                                                                     \
        for (int k = 0; k < i; ++k) s -= NAME##_l(lower, F, ldf, i, k) * x[k];          \
        x[i] = s / NAME##_l(lower, F, ldf, i, i);                                       \
    }                                                                                   \
    for (int i = n - 1; i >= 0; --i) {                                                  \
        C s = x[i];                                                                     \
        for (int k = i + 1; k < n; ++k) s -= conj(NAME##_l(lower, F, ldf, k, i)) * x[k]; \
        x[i] = s / conj(NAME##_l(lower, F, ldf, i, i));                                 \
    }                                                                                   \
}                                                                                       \
void NAME(char side, char uplo, int n, int nrhs, const void* Fv, int ldf, void* Bv, int ldb) \
{                                                                                       \
    const C* F = (const C*)Fv;                                                          \
    C* B = (C*)Bv;                                                                      \
    const int lower = (uplo == 'L' || uplo == 'l');                                     \
    const int left = (side == 'L' || side == 'l');                                      \
    if (n <= 0 || nrhs <= 0) return;                                                    \
    if (left) {                                                                         \
        for (int j = 0; j < nrhs; ++j) NAME##_vector(lower, n, F, ldf, B + (size_t)j * ldb); \
        return;                                                                         \
    }                                                                                   \
    C* row = (C*)malloc(sizeof(C) * (size_t)n);                                         \
    for (int r = 0; r < nrhs; ++r) {                                                    \
        for (int i = 0; i < n; ++i) row[i] = conj(B[r + (size_t)i * ldb]);              \
        NAME##_vector(lower, n, F, ldf, row);                                           \
        for (int i = 0; i < n; ++i) B[r + (size_t)i * ldb] = conj(row[i]);              \
    }                                                                                   \
    free(row);                                                                          \
}
DEFINE_HPOTRS(oracle_zpotrs, double complex)
DEFINE_HPOTRS(oracle_cpotrs, float complex)

/* ------------------------------------------------------------------------ */
/* diag_logdet for complex types: the term of src/diag_logdet_impl.h:50-94   */
/* and the join of :42-48.  For d = x + iy: nonfinite part -> flag 1;        */
/* d = 0 -> flag 2; otherwise log|d| = log(hi) + 0.5 log1p((lo/hi)^2)        */
/* (hi, lo = the larger and smaller of |x|, |y|, so |d| never overflows),    */
/* and arg d = atan2(y, x) added to the phase, except for a real d < 0,      */
/* which flips the parity bit (its pi is added once, by finish_logdet).      */
/* ------------------------------------------------------------------------ */
#define DEFINE_HLOGDET(NAME, R)                                                         \
void NAME(int n, const void* Av, int lda, double* magnitude, double* phase, int* flags) \
{                                                                                       \
    const R* A = (const R*)Av;                                                          \
    double m = 0, p = 0;                                                                \
    int f = 0, parity = 0;                                                              \
    const double limit = 1.7976931348623157e308;                                        \
    for (int i = 0; i < n; ++i) {                                                       \
        const size_t at = 2 * ((size_t)i * lda + i);                                    \
        const double x = (double)A[at], y = (double)A[at + 1];                          \
        if (!(x <= limit && x >= -limit && y <= limit && y >= -limit)) { f |= 1; continue; } \
        const double ax = fabs(x), ay = fabs(y);                                        \
        const double hi = ax > ay ? ax : ay, lo = ax > ay ? ay : ax;                    \
        if (hi == 0) { f |= 2; continue; }                                              \
        const double ratio = lo / hi;                                                   \
        m += log(hi) + 0.5 * log1p(ratio * ratio);                                      \
        if (y == 0) { if (x < 0) parity ^= 1; }                                         \
        else p += atan2(y, x);                                                          \
    }                                                                                   \
    *magnitude = m;                                                                     \
    *phase = p;                                                                         \
    *flags = f | (parity ? 4 : 0);                                                      \
}
DEFINE_HLOGDET(oracle_zdiag_logdet, double)
DEFINE_HLOGDET(oracle_cdiag_logdet, float)

/* ------------------------------------------------------------------------ */
/* Complex GEMM, also with one real operand (the mixed real/complex          */
/* gemmVbatched overloads, include/ddla/ddla.h:1878-1960): a_complex /        */
/* b_complex 0 means that operand is a real R array, on which 'C' is 'T'.    */
/* Accumulated in double complex.  alpha, beta: (real, imaginary).  k == 0   */
/* or alpha == 0 only scales C; beta == 0 never reads C.                     */
/* ------------------------------------------------------------------------ */
#define DEFINE_ZGEMM(NAME, R, C)                                                        \
static double complex NAME##_load(const void* X, int is_complex, size_t at, int conjugate) \
{                                                                                       \
    if (!is_complex) return (double)((const R*)X)[at];                                  \
    const double complex v = (double complex)((const C*)X)[at];                         \
    return conjugate ? conj(v) : v;                                                     \
}                                                                                       \
void NAME(char ta, char tb, int m, int n, int k, const R* alpha, const void* A, int a_complex, \
          int lda, const void* B, int b_complex, int ldb, const R* beta, void* Cv, int ldc) \
{                                                                                       \
    C* Cm = (C*)Cv;                                                                     \
    const int na = (ta == 'N' || ta == 'n'), nb = (tb == 'N' || tb == 'n');             \
    const int ca = (ta == 'C' || ta == 'c'), cb = (tb == 'C' || tb == 'c');             \
    const C al = alpha[0] + alpha[1] * I, be = beta[0] + beta[1] * I;                   \
    const int alpha_zero = (alpha[0] == 0 && alpha[1] == 0);                            \
    const int beta_zero = (beta[0] == 0 && beta[1] == 0);                               \
    for (int j = 0; j < n; ++j) {                                                       \
        for (int i = 0; i < m; ++i) {                                                   \
            double complex s = 0;                                                       \
            if (!alpha_zero)                                                            \
                for (int p = 0; p < k; ++p) {                                           \
                    const size_t ia = na ? i + (size_t)p * lda : p + (size_t)i * lda;   \
                    const size_t ib = nb ? p + (size_t)j * ldb : j + (size_t)p * ldb;   \
                    s += NAME##_load(A, a_complex, ia, ca) * NAME##_load(B, b_complex, ib, cb); \
                }                                                                       \
            const C prod = al * (C)s;                                                   \
            C* c = Cm + i + (size_t)j * ldc;                                            \
            *c = beta_zero ? prod : prod + be * *c;                                     \
        }                                                                               \
    }                                                                                   \
}
DEFINE_ZGEMM(oracle_zgemm, double, double complex)
DEFINE_ZGEMM(oracle_cgemm, float, float complex)
