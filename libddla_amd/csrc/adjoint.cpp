// Reverse-mode adjoints of the Cholesky path, built from the same MFMA GEMM
// and triangular-solve sweep kernels as the forward path.  There is no
// counterpart in the reference.  All take the factor of A = L L^T in the
// lower triangle (uplo 'L'; the upper triangle of the factor's storage is
// ignored) and return synchronized.
//
//   potrf_backward: Abar = sym( L^-T Phi(L^T Lbar) L^-1 ),  sym(S) = (S + S^T)/2,
//                   Phi(M) = tril(M) with the diagonal halved
//                   (Murray 2016, "Differentiation of the Cholesky decomposition",
//                   the symmetric-input convention of the blocked algorithm),
//                   evaluated as L^-T Psi L^-1 with the symmetric
//                   Psi = sym(Phi(M)): only tril(M) and the lower triangle
//                   of the (symmetric) result are formed;
//   potrs_backward: X = A^-1 B:  Bbar = A^-1 Xbar,
//                   Lbar -= tril(Bbar (L^T X)^T + X (L^T Bbar)^T);
//   logdet_backward: logdet A = 2 sum log L(i, i):  Lbar(i, i) += 2 g / L(i, i).
// The inverted 64 x 64 diagonal blocks the sweeps need are recomputed from L
// (one launch) rather than kept from the forward pass.
#include "internal.h"
#include "kernels.h"

#include <algorithm>

namespace ddla {

namespace {

constexpr int kIB = 64;

std::size_t inv_elems(int n) { return (std::size_t)((n + kIB - 1) / kIB) * kIB * kIB; }

void check_common(const DdlaHandle_t& h, const char* who, char uplo, int n, const void* L, int ldl)
{
    require_gpu(h, who);
    if (uplo != 'L' && uplo != 'l') throw std::invalid_argument(std::string(who) + ": the factor must be uplo 'L'");
    if (n < 0) throw std::invalid_argument(std::string(who) + ": negative order");
    if (n > 0 && (L == nullptr || ldl < n)) throw std::invalid_argument(std::string(who) + ": invalid factor");
}

void finish(const DdlaHandle_t& h, int* error, const char* who)
{
    int e = 0;
    DDLA_HIP(hipMemcpyAsync(&e, error, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    DDLA_HIP(hipStreamSynchronize(h->stream));
    if (e != 0) throw std::runtime_error(std::string(who) + ": a triangular-solve sweep exceeded its wait limit");
}

// M = tril(L^T Lbar) for Lz, Bz with zero strict upper triangles: the terms
// of k in [k0, k1) reach only rows and columns below k1, so the product runs
// in four k blocks of growing order (about n^3 / 2 flops instead of 2 n^3),
// the last block -- every row, beta 0 -- first.
void k_block(int n, int q, int& k0, int& k1)
{
    k0 = q * (n / 4) / kIB * kIB;
    k1 = q == 3 ? n : (q + 1) * (n / 4) / kIB * kIB;
}

double lower_product_flops(int n)
{
    double flops = 0;
    for (int q = 0, k0, k1; q < 4; ++q) {
        k_block(n, q, k0, k1);
        if (k1 > k0) flops += (double)k1 * (k1 + 1) * (k1 - k0);
    }
    return flops;
}

template <class T>
void lower_product(hipStream_t s, int n, const T* Lz, const T* Bz, T* M, const kern::Batch& batch = kern::Batch())
{
    for (int q = 3, k0, k1; q >= 0; --q) {
        k_block(n, q, k0, k1);
        if (k1 <= k0) continue;
        kern::gemm<T>(s, 'T', 'N', k1, k1, k1 - k0, T(1), Lz + k0, n, Bz + k0, n, q == 3 ? T(0) : T(1), M, n,
                      kern::kTriLower, 0, nullptr, false, batch);
    }
}

// tril(X) of L^T X = C (order n, in place on C): rows j.. of column j of X
// depend only on rows j.. of C(:, j) and the trailing factor from j, so with
// L = [L11 0; L21 L22] and the columns split alike,
//   X21 = L22^-T C21,  tril(X22) from (L22, C22),
//   C11 -= L21^T X21 (lower triangle),  tril(X11) from (L11, C11)
// -- n^3 / 3 flops instead of n^3.  Leaves of order <= kLowerLeaf take the
// full solve.
constexpr int kLowerLeaf = 1024;

template <class T>
void lower_solve(hipStream_t s, int n, const T* L, long long ldl, const T* dinv, T* C, long long ldc, void* sync,
                 int* error)
{
    if (n <= kLowerLeaf) {
        kern::trsm<T>(s, false, true, n, n, L, ldl, dinv, C, ldc, sync, error);
        return;
    }
    const int n1 = (n / 2 + kIB - 1) / kIB * kIB, n2 = n - n1;
    const T* L21 = L + n1;
    const T* L22 = L + n1 + (long long)n1 * ldl;
    const T* dinv22 = dinv + (long long)(n1 / kIB) * kIB * kIB;
    T* C21 = C + n1;
    kern::trsm<T>(s, false, true, n2, n1, L22, ldl, dinv22, C21, ldc, sync, error);
    lower_solve<T>(s, n2, L22, ldl, dinv22, C21 + (long long)n1 * ldc, ldc, sync, error);
    kern::gemm<T>(s, 'T', 'N', n1, n1, n2, T(-1), L21, ldl, C21, ldc, T(1), C, ldc, kern::kTriLower, 0);
    lower_solve<T>(s, n1, L, ldl, dinv, C, ldc, sync, error);
}

// X = L^-1 (lower triangular, order n, in place on X, which holds the
// identity on entry): with L = [L11 0; L21 L22] split at a multiple of 64,
//   X11 = L11^-1,  X21 = -L22^-1 (L21 X11),  X22 = L22^-1,
// X12 staying zero; leaves of order <= kLowerLeaf solve L X = I directly
// (forward substitution keeps the zeros above the diagonal).  About n^3 / 2
// flops.
template <class T>
void lower_inverse(hipStream_t s, int n, const T* L, long long ldl, const T* dinv, T* X, long long ldx, void* sync,
                   int* error)
{
    if (n <= kLowerLeaf) {
        kern::trsm<T>(s, false, false, n, n, L, ldl, dinv, X, ldx, sync, error);
        return;
    }
    const int n1 = (n / 2 + kIB - 1) / kIB * kIB, n2 = n - n1;
    const T* L21 = L + n1;
    const T* L22 = L + n1 + (long long)n1 * ldl;
    const T* dinv22 = dinv + (long long)(n1 / kIB) * kIB * kIB;
    T* X21 = X + n1;
    lower_inverse<T>(s, n1, L, ldl, dinv, X, ldx, sync, error);
    kern::gemm<T>(s, 'N', 'N', n2, n1, n1, T(-1), L21, ldl, X, ldx, T(0), X21, ldx);  // -L21 X11
    kern::trsm<T>(s, false, false, n2, n1, L22, ldl, dinv22, X21, ldx, sync, error);   // X21 = L22^-1 (-L21 X11)
    lower_inverse<T>(s, n2, L22, ldl, dinv22, X21 + (long long)n1 * ldx, ldx, sync, error);
}

} // namespace

// M = tril(X^T X) for lower-triangular X (order n, leading dimension n) of
// each problem of the batch (the batched Gaussian-process likelihood).
void lower_gram_batched(hipStream_t s, int n, const double* X, double* M, const kern::Batch& batch)
{
    lower_product<double>(s, n, X, X, M, batch);
}
double lower_gram_flops(int n) { return lower_product_flops(n); }

template <class T>
void potri(const DdlaHandle_t& h, char uplo, int n, const T* L, int ldl, T* Ainv, int lda)
{
    check_common(h, "potri", uplo, n, L, ldl);
    if (n == 0) return;
    if (!Ainv || lda < n) throw std::invalid_argument("potri: invalid Ainv");
    const std::size_t nn = (std::size_t)n * n;
    const std::size_t bytes = 2 * detail::padded_bytes<T>(nn) + detail::padded_bytes<T>(inv_elems(n)) +
                              kern::solve_work_bytes(n) + detail::padded_bytes<int>(1);
    char* w = static_cast<char*>(h->ensure_scratch(bytes));
    T* X = detail::take<T>(w, nn);
    T* M = detail::take<T>(w, nn);
    T* dinv = detail::take<T>(w, inv_elems(n));
    void* sync = w;
    w += kern::solve_work_bytes(n);
    int* error = detail::take<int>(w, 1);
    hipStream_t s = h->stream;
    fill_int(s, error, 0);
    kern::trtri_diag_blocks<T>(s, false, n, L, ldl, dinv);
    kern::set_identity<T>(s, n, X, n);
    {
        LaunchTimer timer(h, kClassSolve, (double)nn * n / 2.0, 0);
        lower_inverse<T>(s, n, L, ldl, dinv, X, n, sync, error);  // X = L^-1
    }
    {
        LaunchTimer timer(h, kClassOther, lower_product_flops(n));
        lower_product<T>(s, n, X, X, M);  // M = tril(X^T X) = tril(A^-1)
    }
    kern::mirror_lower<T>(s, n, M, n, T(1), Ainv, lda);
    finish(h, error, "potri");
}

template <class T>
void potrf_backward(const DdlaHandle_t& h, char uplo, int n, const T* L, int ldl, const T* Lbar, int ldlb, T* Abar,
                    int ldab)
{
    check_common(h, "potrf_backward", uplo, n, L, ldl);
    if (n == 0) return;
    if (!Lbar || !Abar || ldlb < n || ldab < n) throw std::invalid_argument("potrf_backward: invalid Lbar or Abar");
    const std::size_t nn = (std::size_t)n * n;
    const std::size_t bytes = 3 * detail::padded_bytes<T>(nn) + detail::padded_bytes<T>(inv_elems(n)) +
                              kern::solve_work_bytes(n) + detail::padded_bytes<int>(1);
    char* w = static_cast<char*>(h->ensure_scratch(bytes));
    T* Lz = detail::take<T>(w, nn);
    T* Bz = detail::take<T>(w, nn);
    T* M = detail::take<T>(w, nn);
    T* dinv = detail::take<T>(w, inv_elems(n));
    void* sync = w;
    w += kern::solve_work_bytes(n);
    int* error = detail::take<int>(w, 1);
    hipStream_t s = h->stream;
    fill_int(s, error, 0);
    kern::tril_copy<T>(s, n, L, ldl, Lz, n);
    kern::tril_copy<T>(s, n, Lbar, ldlb, Bz, n);
    {
        LaunchTimer timer(h, kClassOther, lower_product_flops(n));
        lower_product<T>(s, n, Lz, Bz, M);  // M = tril(L^T Lbar)
    }
    kern::mirror_lower<T>(s, n, M, n, T(0.5), Bz, n);  // Psi = sym(Phi(M)) = sym(tril(M)) / 2
    {
        LaunchTimer timer(h, kClassTrtri, (double)n * kIB * kIB / 3.0);
        kern::trtri_diag_blocks<T>(s, false, n, L, ldl, dinv);
    }
    {
        LaunchTimer timer(h, kClassSolve, 4.0 / 3.0 * (double)nn * n, 0);
        kern::trsm<T>(s, false, true, n, n, L, ldl, dinv, Bz, n, sync, error);  // W = L^-T Psi
        kern::transpose<T>(s, n, n, Bz, n, M, n);                              // W^T = Psi L^-1
        lower_solve<T>(s, n, L, ldl, dinv, M, n, sync, error);                 // tril(L^-T Psi L^-1)
    }
    kern::mirror_lower<T>(s, n, M, n, T(1), Abar, ldab);
    finish(h, error, "potrf_backward");
}

template <class T>
void potrs_backward(const DdlaHandle_t& h, char uplo, int n, int nrhs, const T* L, int ldl, const T* X, int ldx,
                    const T* Xbar, int ldxb, T* Bbar, int ldbb, T* Lbar, int ldlb)
{
    check_common(h, "potrs_backward", uplo, n, L, ldl);
    if (nrhs < 0) throw std::invalid_argument("potrs_backward: negative nrhs");
    if (n == 0 || nrhs == 0) return;
    if (!X || !Xbar || !Bbar || !Lbar || ldx < n || ldxb < n || ldbb < n || ldlb < n)
        throw std::invalid_argument("potrs_backward: invalid X, Xbar, Bbar or Lbar");
    const std::size_t nn = (std::size_t)n * n, nr = (std::size_t)n * nrhs;
    // Few columns: P = [X | Bbar | X], Q = L^T [X | Bbar] = [U | V] in one pass
    // over the lower triangle, and Lbar -= [Bbar | X] [U | V]^T in one update.
    // Otherwise U and V are GEMMs with a zero-padded copy of L.
    const bool few = 2 * nrhs <= kern::kTmvColumns;
    const std::size_t bytes =
        (few ? detail::padded_bytes<T>(3 * nr) + detail::padded_bytes<T>(2 * nr)
             : detail::padded_bytes<T>(nn) + 2 * detail::padded_bytes<T>(nr)) +
        detail::padded_bytes<T>(inv_elems(n)) + kern::solve_work_bytes(nrhs) + detail::padded_bytes<int>(1);
    char* w = static_cast<char*>(h->ensure_scratch(bytes));
    T *P = nullptr, *Q = nullptr, *Lz = nullptr, *U = nullptr, *V = nullptr;
    if (few) {
        P = detail::take<T>(w, 3 * nr);
        Q = detail::take<T>(w, 2 * nr);
    } else {
        Lz = detail::take<T>(w, nn);
        U = detail::take<T>(w, nr);
        V = detail::take<T>(w, nr);
    }
    T* dinv = detail::take<T>(w, inv_elems(n));
    void* sync = w;
    w += kern::solve_work_bytes(nrhs);
    int* error = detail::take<int>(w, 1);
    hipStream_t s = h->stream;
    fill_int(s, error, 0);
    kern::convert<T, T>(s, n, nrhs, Xbar, ldxb, Bbar, ldbb);
    kern::trtri_diag_blocks<T>(s, false, n, L, ldl, dinv);
    {
        LaunchTimer timer(h, kClassSolve, 2.0 * (double)nn * nrhs, 0);
        kern::potrs_sweeps<T>(s, false, n, nrhs, L, ldl, dinv, Bbar, ldbb, sync, error);  // Bbar = A^-1 Xbar
    }
    if (few) {
        kern::convert<T, T>(s, n, nrhs, X, ldx, P, n);
        kern::convert<T, T>(s, n, nrhs, Bbar, ldbb, P + nr, n);
        kern::convert<T, T>(s, n, nrhs, X, ldx, P + 2 * nr, n);
        kern::lower_tmv<T>(s, n, 2 * nrhs, L, ldl, P, n, Q, n);  // [U | V] = L^T [X | Bbar]
        kern::gemm<T>(s, 'N', 'T', n, n, 2 * nrhs, T(-1), P + nr, n, Q, n, T(1), Lbar, ldlb, kern::kTriLower, 0);
    } else {
        kern::tril_copy<T>(s, n, L, ldl, Lz, n);
        kern::gemm<T>(s, 'T', 'N', n, nrhs, n, T(1), Lz, n, X, ldx, T(0), U, n);     // U = L^T X
        kern::gemm<T>(s, 'T', 'N', n, nrhs, n, T(1), Lz, n, Bbar, ldbb, T(0), V, n);  // V = L^T Bbar
        // Lbar (lower) -= Bbar U^T + X V^T
        kern::gemm<T>(s, 'N', 'T', n, n, nrhs, T(-1), Bbar, ldbb, U, n, T(1), Lbar, ldlb, kern::kTriLower, 0);
        kern::gemm<T>(s, 'N', 'T', n, n, nrhs, T(-1), X, ldx, V, n, T(1), Lbar, ldlb, kern::kTriLower, 0);
    }
    finish(h, error, "potrs_backward");
}

template <class T>
void logdet_backward(const DdlaHandle_t& h, int n, const T* L, int ldl, double g, T* Lbar, int ldlb)
{
    check_common(h, "logdet_backward", 'L', n, L, ldl);
    if (n == 0) return;
    if (!Lbar || ldlb < n) throw std::invalid_argument("logdet_backward: invalid Lbar");
    kern::diag_adjoint<T>(h->stream, n, L, ldl, g, Lbar, ldlb);
    DDLA_HIP(hipStreamSynchronize(h->stream));
}

#define DDLA_ADJOINT_INSTANTIATE(T)                                                                              \
    template void potrf_backward<T>(const DdlaHandle_t&, char, int, const T*, int, const T*, int, T*, int);    \
    template void potrs_backward<T>(const DdlaHandle_t&, char, int, int, const T*, int, const T*, int,         \
                                    const T*, int, T*, int, T*, int);                                           \
    template void logdet_backward<T>(const DdlaHandle_t&, int, const T*, int, double, T*, int);                 \
    template void potri<T>(const DdlaHandle_t&, char, int, const T*, int, T*, int);
DDLA_ADJOINT_INSTANTIATE(float)
DDLA_ADJOINT_INSTANTIATE(double)
#undef DDLA_ADJOINT_INSTANTIATE

} // namespace ddla
