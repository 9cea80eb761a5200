// Complex kernels of the Hermitian Cholesky path (zkernels.hip): launch
// functions in ddla::kern, matrices in the interleaved layout of
// std::complex<R> (real part first), column-major with leading dimensions.
#pragma once
#include <hip/hip_runtime.h>

#include <complex>
#include <cstddef>

namespace ddla {
namespace kern {

// Operation on a GEMM operand: 0 'N', 1 'T', 2 'C' (conjugate transpose;
// on a real operand the same as 'T').
int op_code(char op);

// C = alpha op(A) op(B) + beta C with complex C, alpha, beta; A and B
// complex, or one of them real (a_complex / b_complex false: the operand is
// an R array).  tri 0: every element, 1: only i + diag >= j, 2: only
// i + diag <= j.  alpha == 0 or k == 0 scale C without reading A, B; beta == 0
// never reads C.  The product may overwrite op(A)'s storage when C is that
// same block and n <= 64 (each tile reads its rows of A before it writes).
template <class R>
void zgemm(hipStream_t s, bool a_complex, bool b_complex, char ta, char tb, int m, int n, int k,
           std::complex<R> alpha, const void* A, long long lda, const void* B, long long ldb, std::complex<R> beta,
           std::complex<R>* C, long long ldc, int tri = 0, long long diag = 0, const int* abort = nullptr);

// The same over a batch with per-member sizes and leading dimensions (device
// arrays), pointer arrays of the members, one launch; max_m / max_n bound the
// members' m and n.
template <class R>
void zgemm_vbatched(hipStream_t s, bool a_complex, bool b_complex, char ta, char tb, const int* m, const int* n,
                    const int* k, std::complex<R> alpha, const void* const* A, const int* lda, const void* const* B,
                    const int* ldb, std::complex<R> beta, std::complex<R>* const* C, const int* ldc, int batch,
                    int max_m, int max_n);

// Factors the order-w (w = min(64, n - k0)) Hermitian diagonal block at
// A(k0, k0) in place (lower triangle: A = L L^H), writes its inverse L^-1 to
// inv (64 x 64, column-major, zero above the diagonal) and, on a pivot whose
// real part is not positive (or NaN), sets *info = k0 + j + 1 (1-based) and
// leaves the block partly factored.  Does nothing if *info != 0 on entry.
template <class R>
void zpotf2(hipStream_t s, int n, std::complex<R>* A, long long lda, int k0, std::complex<R>* inv, int* info);

// Inverse of every 64 x 64 diagonal block of the order-n lower factor L (L
// view: L(i, j) = A[i + j lda], or conj(A[j + i lda]) when upper) into inv
// (block b at inv + 4096 b).
template <class R>
void ztrtri_diag_blocks(hipStream_t s, bool upper, int n, const std::complex<R>* A, long long lda,
                        std::complex<R>* inv);

// out (n x m) = in^H (in m x n).
template <class R>
void zconj_transpose(hipStream_t s, int m, int n, const std::complex<R>* in, long long ldi, std::complex<R>* out,
                     long long ldo);

// A(i, j) <-> conj(A(j, i)) for i > j, in place (order n): moves a Hermitian
// matrix's upper triangle to the lower one and back.
template <class R>
void zswap_triangles(hipStream_t s, int n, std::complex<R>* A, long long lda);

// Sum of log |d_i| and of arg d_i over the diagonal (deterministic order),
// with flags 1 nonfinite entry, 2 zero entry, 4 odd count of negative real
// entries; out[0] = magnitude, out[1] = phase sum (doubles), *(int*)(out + 2)
// = flags.  work: zlogdet_work_bytes().
std::size_t zlogdet_work_bytes();
template <class R>
void zdiag_logdet(hipStream_t s, int n, const std::complex<R>* A, long long lda, void* work, double* out);

} // namespace kern
} // namespace ddla
