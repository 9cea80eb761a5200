// General triangular solves on one GPU: the local trsm and the distributed
// ptrtrs at a 1 x 1 grid, on the same solve kernels as ppotrs.
//
// Reference behaviour followed:
//   trsm    include/ddla/ddla.h:1620 (src/BLAS/trsm.cpp:105-168): B := alpha
//           op(A)^-1 B (side 'L', A m x m) or alpha B op(A)^-1 (side 'R', A
//           n x n); flags case-insensitive, a bad flag throws
//           std::runtime_error (src/blas_enums.h:12-30); alpha == 0 sets B to
//           zero without reading A (BLAS convention).
//   ptrtrs  include/ddla/ddla.h:96 (src/PBLAS/ptrtrs.cpp:23-57 checks, :141-155
//           run): at 1 x 1 one trsm with alpha = 1 on the descriptors' LLDs;
//           flags upper case, bad arguments throw std::invalid_argument.
// The arithmetic is this project's: every case becomes one sweep over the
// "L view" of the stored triangle (kern::trsm): with L the lower triangle, or
// L = U^T for an upper one,
//   side 'L':  op(A) = L or U^T -> L X = B (forward), op(A) = L^T or U -> L^T X = B (backward);
//   side 'R':  X op(A) = B is op(A)^T X^T = B^T, solved on a transposed copy;
//   diag 'U':  the diagonal blocks' inverses take ones on the diagonal (the
//              stored diagonal is never read).
#include "internal.h"
#include "kernels.h"

#include <stdexcept>
#include <string>

namespace ddla {

namespace {

constexpr int kIB = 64;

std::size_t inverse_elems(int n) { return (std::size_t)((n + kIB - 1) / kIB) * kIB * kIB; }

char upper_case(char c) { return (c >= 'a' && c <= 'z') ? char(c - 'a' + 'A') : c; }

// side, uplo, trans, diag already validated and upper case
template <class T>
void triangular_solve(const DdlaHandle_t& h, char side, char uplo, char trans, char diag, int m, int n, T alpha,
                      const T* A, long long lda, T* B, long long ldb, const char* who)
{
    if (m == 0 || n == 0) return;
    hipStream_t s = h->stream;
    if (alpha == T(0)) {
        kern::scale_matrix<T>(s, m, n, T(0), B, ldb);
        DDLA_HIP(hipStreamSynchronize(s));
        return;
    }
    const bool left = side == 'L';
    const int order = left ? m : n, nrhs = left ? n : m;
    const bool upper = uplo == 'U';
    // op(A) in the L view: L (lower, 'N') or U^T (upper, 'T'/'C') is a forward sweep
    const bool op_lower = (uplo == 'L') == (trans == 'N');
    // side 'R' solves op(A)^T, the other triangle's sweep
    const bool backward = left ? !op_lower : op_lower;
    const std::size_t bytes = detail::padded_bytes<T>(inverse_elems(order)) + kern::solve_work_bytes(nrhs) +
                              detail::padded_bytes<int>(1) +
                              (left ? 0 : detail::padded_bytes<T>((std::size_t)m * n));
    char* w = static_cast<char*>(h->ensure_scratch(bytes));
    T* dinv = detail::take<T>(w, inverse_elems(order));
    void* sync = w;
    w += kern::solve_work_bytes(nrhs);
    int* error = detail::take<int>(w, 1);
    fill_int(s, error, 0);
    kern::trtri_diag_blocks<T>(s, upper, order, A, lda, dinv, diag == 'U');
    if (alpha != T(1)) kern::scale_matrix<T>(s, m, n, alpha, B, ldb);
    {
        LaunchTimer timer(h, kClassSolve, (double)order * order * nrhs, 0);
        if (left) {
            kern::trsm<T>(s, upper, backward, order, nrhs, A, lda, dinv, B, ldb, sync, error);
        } else {
            T* bt = detail::take<T>(w, (std::size_t)m * n);
            kern::transpose<T>(s, m, n, B, ldb, bt, n);  // B^T: n x m
            kern::trsm<T>(s, upper, backward, order, nrhs, A, lda, dinv, bt, n, sync, error);
            kern::transpose<T>(s, n, m, bt, n, B, ldb);
        }
    }
    int host_error = 0;
    DDLA_HIP(hipMemcpyAsync(&host_error, error, sizeof(int), hipMemcpyDeviceToHost, s));
    DDLA_HIP(hipStreamSynchronize(s));
    if (host_error != 0) throw std::runtime_error(std::string(who) + ": a triangular-solve sweep exceeded its wait limit");
}

} // namespace

template <class T>
void trsm(const DdlaHandle_t& h, char side, char uplo, char trans, char diag, int m, int n, const T& alpha, T* A,
          int lda, T* B, int ldb)
{
    require_gpu(h, "trsm");
    side = upper_case(side);
    uplo = upper_case(uplo);
    trans = upper_case(trans);
    diag = upper_case(diag);
    if (side != 'L' && side != 'R') throw std::runtime_error("ddla: invalid side flag (expected 'L' or 'R')");
    if (uplo != 'L' && uplo != 'U') throw std::runtime_error("ddla: invalid uplo flag (expected 'L' or 'U')");
    if (trans != 'N' && trans != 'T' && trans != 'C')
        throw std::runtime_error("ddla: invalid trans flag (expected 'N', 'T' or 'C')");
    if (diag != 'N' && diag != 'U') throw std::runtime_error("ddla: invalid diag flag (expected 'N' or 'U')");
    if (m < 0 || n < 0) throw std::invalid_argument("trsm: negative dimension");
    const int order = side == 'L' ? m : n;
    if (lda < std::max(1, order) || ldb < std::max(1, m)) throw std::invalid_argument("trsm: leading dimension too small");
    if (m && n && (B == nullptr || (alpha != T(0) && A == nullptr))) throw std::invalid_argument("trsm: null matrix");
    triangular_solve<T>(h, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, "trsm");
}

template <class T>
void ptrtrs(const DdlaHandle_t& h, char side, char uplo, char trans, char diag, int m, int n, T* A,
            const int* descA, T* B, const int* descB)
{
    check_desc(descB, h);
    check_desc(descA, h);
    require_gpu(h, "ptrtrs");
    if (descA[DDLA_MB_] != descA[DDLA_NB_]) throw std::invalid_argument("ptrtrs: A needs square blocks");
    if (descA[DDLA_MB_] != descB[DDLA_MB_]) throw std::invalid_argument("ptrtrs: A and B need the same row block size");
    if (side != 'L' && side != 'R') throw std::invalid_argument("ptrtrs: side must be 'L' or 'R'");
    if (side == 'L' && descA[DDLA_RSRC_] != descB[DDLA_RSRC_])
        throw std::invalid_argument("ptrtrs: side 'L' needs B's rows distributed like A's (RSRC)");
    if (side == 'R' && (descA[DDLA_NB_] != descB[DDLA_NB_] || descA[DDLA_CSRC_] != descB[DDLA_CSRC_]))
        throw std::invalid_argument("ptrtrs: side 'R' needs B's columns distributed like A's (NB, CSRC)");
    if (uplo != 'L' && uplo != 'U') throw std::invalid_argument("ptrtrs: uplo must be 'L' or 'U'");
    if (diag != 'U' && diag != 'N') throw std::invalid_argument("ptrtrs: diag must be 'U' or 'N'");
    if (trans != 'N' && trans != 'T' && trans != 'C') throw std::invalid_argument("ptrtrs: trans must be 'N', 'T' or 'C'");
    const int order = side == 'L' ? m : n;
    if (m < 0 || n < 0 || order > descA[DDLA_M_] || order > descA[DDLA_N_] || m > descB[DDLA_M_] || n > descB[DDLA_N_])
        throw std::invalid_argument("ptrtrs: m x n exceeds the descriptors");
    if (m == 0 || n == 0) return;
    if (A == nullptr || B == nullptr) throw std::invalid_argument("ptrtrs: null local matrix");
    triangular_solve<T>(h, side, uplo, trans, diag, m, n, T(1), A, descA[DDLA_LLD_], B, descB[DDLA_LLD_], "ptrtrs");
}

#define DDLA_INSTANTIATE(T)                                                                                    \
    template void trsm<T>(const DdlaHandle_t&, char, char, char, char, int, int, const T&, T*, int, T*, int); \
    template void ptrtrs<T>(const DdlaHandle_t&, char, char, char, char, int, int, T*, const int*, T*, const int*);
DDLA_INSTANTIATE(float)
DDLA_INSTANTIATE(double)
#undef DDLA_INSTANTIATE

} // namespace ddla
