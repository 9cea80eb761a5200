// Reverse-mode adjoints of the Cholesky path, built from the same MFMA GEMM
// and triangular-solve sweep kernels as the forward path.  There is no
// counterpart in the reference.  All take the factor of A = L L^T in the
// lower triangle (uplo 'L'; the upper triangle of the factor's storage is
// ignored) and return synchronized.
//
//   potrf_backward: Abar = sym( L^-T Phi(L^T Lbar) L^-1 ),  sym(S) = (S + S^T)/2,
//                   Phi(M) = tril(M) with the diagonal halved
//                   (Murray 2016, "Differentiation of the Cholesky decomposition",
//                   the symmetric-input convention of the blocked algorithm);
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

} // namespace

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
        LaunchTimer timer(h, kClassOther, 2.0 * n * (double)nn);
        kern::gemm<T>(s, 'T', 'N', n, n, n, T(1), Lz, n, Bz, n, T(0), M, n);  // M = L^T Lbar
    }
    kern::phi<T>(s, n, M, n);
    {
        LaunchTimer timer(h, kClassTrtri, (double)n * kIB * kIB / 3.0);
        kern::trtri_diag_blocks<T>(s, false, n, L, ldl, dinv);
    }
    {
        LaunchTimer timer(h, kClassSolve, 2.0 * (double)nn * n, 0);
        kern::trsm_sweep<T>(s, false, true, n, n, L, ldl, dinv, M, n, sync, error);   // W = L^-T Phi
        kern::transpose<T>(s, n, n, M, n, Bz, n);                                      // W^T
        kern::trsm_sweep<T>(s, false, true, n, n, L, ldl, dinv, Bz, n, sync, error);  // S^T = L^-T W^T
    }
    kern::sym_half<T>(s, n, Bz, n, Abar, ldab);
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
    const std::size_t bytes = detail::padded_bytes<T>(nn) + 2 * detail::padded_bytes<T>(nr) +
                              detail::padded_bytes<T>(inv_elems(n)) + kern::solve_work_bytes(nrhs) +
                              detail::padded_bytes<int>(1);
    char* w = static_cast<char*>(h->ensure_scratch(bytes));
    T* Lz = detail::take<T>(w, nn);
    T* U = detail::take<T>(w, nr);
    T* V = detail::take<T>(w, nr);
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
    kern::tril_copy<T>(s, n, L, ldl, Lz, n);
    kern::gemm<T>(s, 'T', 'N', n, nrhs, n, T(1), Lz, n, X, ldx, T(0), U, n);     // U = L^T X
    kern::gemm<T>(s, 'T', 'N', n, nrhs, n, T(1), Lz, n, Bbar, ldbb, T(0), V, n);  // V = L^T Bbar
    // Lbar (lower) -= Bbar U^T + X V^T
    kern::gemm<T>(s, 'N', 'T', n, n, nrhs, T(-1), Bbar, ldbb, U, n, T(1), Lbar, ldlb, kern::kTriLower, 0);
    kern::gemm<T>(s, 'N', 'T', n, n, nrhs, T(-1), X, ldx, V, n, T(1), Lbar, ldlb, kern::kTriLower, 0);
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
    template void logdet_backward<T>(const DdlaHandle_t&, int, const T*, int, double, T*, int);
DDLA_ADJOINT_INSTANTIATE(float)
DDLA_ADJOINT_INSTANTIATE(double)
#undef DDLA_ADJOINT_INSTANTIATE

} // namespace ddla
