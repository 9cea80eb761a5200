// Hermitian Cholesky path for std::complex<float> / std::complex<double> on one
// GPU: the blocked factorization, the two triangular solves and the
// log-determinant behind ppotrf / ppotrs / pposv / potrf / diag_logdet, and
// the complex and mixed real/complex gemmVbatched(2s).
//
// Reference behaviour followed (1 x 1 grid), as for the real types:
//   ppotrf   src/Cholesky/ppotrf.cpp:21-54 (A = L L^H for uplo 'L', U^H U for 'U')
//   ppotrs   src/Cholesky/ppotrs.cpp:38-61 (side 'L' applies L then L^H; side 'R'
//            the reverse; trans 'N' and 'C' solve the same Hermitian system)
//   logdet   src/diag_logdet_impl.h:50-113 (principal phase in (-pi, pi])
//   vbatched include/ddla/ddla.h:1878-1960 (mixed overloads: 'C' on a real operand is 'T')
// The arithmetic is this project's: right-looking over 64-column blocks, each
// diagonal block factored and inverted in one workgroup (zpotf2), the panel
// formed as A21 D^H and the trailing matrix updated by a lower-masked
// A22 -= L21 L21^H, both on the complex GEMM tile of zkernels.hip.
#include "internal.h"
#include "kernels.h"
#include "zkernels.h"

#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>

namespace ddla {
namespace complex_impl {

namespace {

constexpr int kZB = 64;

std::size_t zdiag_inverse_elems(int n) { return (std::size_t)((n + kZB - 1) / kZB) * kZB * kZB; }

} // namespace

// Bytes of the factorization's device work: the current diagonal block's
// inverse.
template <class R>
std::size_t potrf_work_bytes()
{
    return detail::padded_bytes<std::complex<R>>((std::size_t)kZB * kZB);
}

// A = L L^H in the lower triangle (upper: U^H U, by moving the upper
// triangle down conjugated and back); *info (device, zero on entry) receives
// the first failed pivot.  Asynchronous on the handle's stream.
template <class R>
void potrf_blocked(DdlaStream* h, bool upper, int n, std::complex<R>* A, long long lda, int* info, void* work)
{
    using C = std::complex<R>;
    hipStream_t s = h->stream;
    if (upper) {
        {
            LaunchTimer timer(h, kClassOther, 0, 4.0 * n * (n - 1) * sizeof(R));
            kern::zswap_triangles<R>(s, n, A, lda);
        }
        potrf_blocked<R>(h, false, n, A, lda, info, work);
        LaunchTimer timer(h, kClassOther, 0, 4.0 * n * (n - 1) * sizeof(R));
        kern::zswap_triangles<R>(s, n, A, lda);
        return;
    }
    C* inv = static_cast<C*>(work);
    for (int k = 0; k < n; k += kZB) {
        const int kb = std::min(kZB, n - k), m2 = n - k - kb;
        C* akk = A + k + (long long)k * lda;
        {
            LaunchTimer timer(h, kClassPanel, 4.0 * kb * kb * kb / 3.0);
            kern::zpotf2<R>(s, n, A, lda, k, inv, info);
        }
        if (m2 <= 0) break;
        C* a21 = akk + kb;
        {
            // L21 = A21 L11^-H, in place (one tile column)
            LaunchTimer timer(h, kClassPanel, 8.0 * m2 * kb * kb);
            kern::zgemm<R>(s, true, true, 'N', 'C', m2, kb, kb, C(1), a21, lda, inv, kZB, C(0), a21, lda, 0, 0,
                           info);
        }
        LaunchTimer timer(h, kClassTrailingUpdate, 4.0 * m2 * (m2 + 1.0) * kb);
        kern::zgemm<R>(s, true, true, 'N', 'C', m2, m2, kb, C(-1), a21, lda, a21, lda, C(1), a21 + (long long)kb * lda,
                       lda, 1, 0, info);
    }
}

// Solves A X = B in place (B n x nrhs) with A = L L^H given by its L view
// (lower: L = A's lower triangle; upper: L = U^H of A's upper triangle), the
// inverses of its 64 x 64 diagonal blocks in dinv: forward L Y = B, then
// backward L^H X = Y, one block of 64 rows at a time (X_t = D_t B_t, then
// the remaining rows updated by one GEMM).
template <class R>
void solve_left(DdlaStream* h, bool upper, int n, int nrhs, const std::complex<R>* A, long long lda,
                std::complex<R>* B, long long ldb, const std::complex<R>* dinv)
{
    using C = std::complex<R>;
    hipStream_t s = h->stream;
    LaunchTimer timer(h, kClassSolve, 8.0 * n * (double)n * nrhs, 2.0 * n * n * sizeof(R));
    // block (r0.., c0..) of the L view and the operation that reads it
    auto lblock = [&](long long r0, long long c0, char op, const C*& p, char& stored_op) {
        if (!upper) {
            p = A + r0 + c0 * lda;
            stored_op = op;
        } else {  // L(r, c) = conj(A(c, r)): the stored block at (c0, r0), conjugate transposed
            p = A + c0 + r0 * lda;
            stored_op = op == 'N' ? 'C' : 'N';
        }
    };
    const int nb = (n + kZB - 1) / kZB;
    for (int t = 0; t < nb; ++t) {
        const long long r0 = (long long)t * kZB;
        const int w = std::min(kZB, n - (int)r0), below = n - (int)r0 - w;
        C* bt = B + r0;
        kern::zgemm<R>(s, true, true, 'N', 'N', w, nrhs, w, C(1), dinv + (long long)t * kZB * kZB, kZB, bt, ldb, C(0),
                       bt, ldb);
        if (below > 0) {
            const C* p;
            char op;
            lblock(r0 + w, r0, 'N', p, op);
            kern::zgemm<R>(s, true, true, op, 'N', below, nrhs, w, C(-1), p, lda, bt, ldb, C(1), bt + w, ldb);
        }
    }
    for (int t = nb - 1; t >= 0; --t) {
        const long long r0 = (long long)t * kZB;
        const int w = std::min(kZB, n - (int)r0);
        C* bt = B + r0;
        kern::zgemm<R>(s, true, true, 'C', 'N', w, nrhs, w, C(1), dinv + (long long)t * kZB * kZB, kZB, bt, ldb, C(0),
                       bt, ldb);
        if (r0 > 0) {
            // B(0:r0) -= L(t rows, 0:r0)^H X_t
            const C* p;
            char op;
            lblock(r0, 0, 'C', p, op);
            kern::zgemm<R>(s, true, true, op, 'N', (int)r0, nrhs, w, C(-1), p, lda, bt, ldb, C(1), B, ldb);
        }
    }
}

template <class R>
std::size_t ppotrs_bytes(char side, int n, int nrhs)
{
    if (n == 0 || nrhs == 0) return 0;
    std::size_t b = detail::padded_bytes<std::complex<R>>(zdiag_inverse_elems(n));
    if (side == 'R') b += detail::padded_bytes<std::complex<R>>((std::size_t)n * nrhs);
    return b;
}

template <class R>
std::size_t ppotrf_bytes(int n)
{
    if (n <= 0) return 0;
    return detail::padded_bytes<int>(1) + potrf_work_bytes<R>();
}

template <class R>
void ppotrf_run(const DdlaHandle_t& h, char uplo, int n, std::complex<R>* A, const int* d, int& info,
                void* work, std::size_t bytes)
{
    if (n == 0) {
        info = 0;
        return;
    }
    if (bytes < ppotrf_bytes<R>(n) || work == nullptr)
        throw std::invalid_argument("ppotrf: backend workspace smaller than the workspace query");
    char* w = static_cast<char*>(work);
    int* d_info = detail::take<int>(w, 1);
    fill_int(h->stream, d_info, 0);
    potrf_blocked<R>(h, uplo == 'U', n, A, d[DDLA_LLD_], d_info, w);
    DDLA_HIP(hipMemcpyAsync(&info, d_info, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    DDLA_HIP(hipStreamSynchronize(h->stream));
}

template <class R>
void ppotrs_run(const DdlaHandle_t& h, char side, char uplo, int n, int nrhs, const std::complex<R>* A,
                const int* da, std::complex<R>* B, const int* db, void* work, std::size_t bytes)
{
    using C = std::complex<R>;
    if (n == 0 || nrhs == 0) return;
    if (bytes < ppotrs_bytes<R>(side, n, nrhs) || work == nullptr)
        throw std::invalid_argument("ppotrs: backend workspace smaller than the workspace query");
    const bool upper = uplo == 'U';
    const long long lda = da[DDLA_LLD_], ldb = db[DDLA_LLD_];
    char* w = static_cast<char*>(work);
    C* dinv = detail::take<C>(w, zdiag_inverse_elems(n));
    {
        LaunchTimer timer(h, kClassTrtri, 4.0 * n * kZB * kZB / 3.0);
        kern::ztrtri_diag_blocks<R>(h->stream, upper, n, A, lda, dinv);
    }
    if (side == 'L') {
        solve_left<R>(h, upper, n, nrhs, A, lda, B, ldb, dinv);
    } else {
        // X A = B (B nrhs x n) with Hermitian A: A X^H = B^H
        C* bt = detail::take<C>(w, (std::size_t)n * nrhs);
        kern::zconj_transpose<R>(h->stream, nrhs, n, B, ldb, bt, n);
        solve_left<R>(h, upper, n, nrhs, A, lda, bt, n, dinv);
        kern::zconj_transpose<R>(h->stream, n, nrhs, bt, n, B, ldb);
    }
    DDLA_HIP(hipStreamSynchronize(h->stream));
}

template <class R>
void potrf(const DdlaHandle_t& h, bool upper, int n, std::complex<R>* A, int lda, int* info)
{
    void* work = h->ensure_scratch(potrf_work_bytes<R>());
    fill_int(h->stream, info, 0);
    potrf_blocked<R>(h, upper, n, A, lda, info, work);
}

// sum of log |d_i| and the principal phase, src/diag_logdet_impl.h:50-113
template <class R>
void diag_logdet(const DdlaHandle_t& h, int n, const std::complex<R>* a, int lda, std::complex<R>& result)
{
    if (n == 0) {
        result = {R(0), R(0)};
        return;
    }
    const std::size_t part = detail::padded_bytes<char>(kern::zlogdet_work_bytes());
    char* w = static_cast<char*>(h->ensure_scratch(part + detail::padded_bytes<double>(3)));
    double* out = reinterpret_cast<double*>(w + part);
    kern::zdiag_logdet<R>(h->stream, n, a, lda, w, out);
    double host[3];
    DDLA_HIP(hipMemcpyAsync(host, out, sizeof(host), hipMemcpyDeviceToHost, h->stream));
    DDLA_HIP(hipStreamSynchronize(h->stream));
    int flags;
    std::memcpy(&flags, &host[2], sizeof(int));
    if (flags & 1) throw std::domain_error("diag_logdet: nonfinite diagonal entry");
    if (flags & 2) {
        result = {-std::numeric_limits<R>::infinity(), R(0)};
        return;
    }
    const double pi = std::acos(-1.0);
    double phase = std::remainder(host[1] + ((flags >> 2) & 1) * pi, 2 * pi);
    if (phase <= -pi) phase = pi;
    result = {static_cast<R>(host[0]), static_cast<R>(phase)};
}

#define DDLA_CINSTANTIATE(R)                                                                                      \
    template std::size_t ppotrs_bytes<R>(char, int, int);                                                        \
    template std::size_t ppotrf_bytes<R>(int);                                                                   \
    template void ppotrf_run<R>(const DdlaHandle_t&, char, int, std::complex<R>*, const int*, int&, void*,       \
                                std::size_t);                                                                    \
    template void ppotrs_run<R>(const DdlaHandle_t&, char, char, int, int, const std::complex<R>*, const int*,   \
                                std::complex<R>*, const int*, void*, std::size_t);                               \
    template void potrf<R>(const DdlaHandle_t&, bool, int, std::complex<R>*, int, int*);                         \
    template void diag_logdet<R>(const DdlaHandle_t&, int, const std::complex<R>*, int, std::complex<R>&);
DDLA_CINSTANTIATE(float)
DDLA_CINSTANTIATE(double)
#undef DDLA_CINSTANTIATE

} // namespace complex_impl
} // namespace ddla
