// Cholesky factorization from the bottom-right corner on one GPU: uplo 'U'
// gives A = U U^T (U upper), uplo 'L' gives A = L^T L (L lower), with the other
// triangle neither read nor written.
//
// Reference behaviour followed:
//   potrf_bottom_right   include/ddla/ddla.h:1072-1093 (src/Cholesky/
//                        potrf_bottom_right.cpp:169-292): checks throw
//                        std::runtime_error; info = i > 0 when the trailing
//                        principal submatrix starting at row i (1-based) is not
//                        positive definite; ends synchronized.
//   ppotrf_bottom_right  include/ddla/ddla.h:1095-1121 (src/Cholesky/
//                        ppotrf_bottom_right.cpp:61-85): descriptor checks throw
//                        std::invalid_argument; at 1 x 1 the local routine on
//                        the descriptor's LLD.
// The arithmetic is this project's.  With J the order reversal, S = J A J has
// the ordinary factorization S = W W^T (W lower), and
//   uplo 'U': U = J W J, i.e. U(i, j) = W(n-1-i, n-1-j), and S's lower triangle
//             is A's upper one rotated the same way;
//   uplo 'L': L = J W^T J, i.e. L(i, j) = W(n-1-j, n-1-i), and S's lower
//             triangle is A's lower one under the same anti-transpose.
// So A's stored triangle is reversed into the lower triangle of a scratch copy
// (its upper triangle zeroed), the blocked lower factorization (MFMA trailing
// updates, look-ahead) runs on the copy, and the factor is reversed back into
// A's triangle: two passes over the triangle in HBM beside the n^3/3 flops,
// instead of the reference's sequence of 128-wide blocks.  A failed pivot k of
// the copy is pivot n - k + 1 of A.
#include "internal.h"
#include "kernels.h"

#include <stdexcept>
#include <string>

namespace ddla {

template <class T>
void potrf_blocked_dispatch(DdlaStream* h, bool upper, int n, T* A, long long lda, int* info, void* work);

namespace {

// leading dimension of the reversed copy: the order rounded up to 32 elements
long long reversed_ld(int n) { return ((long long)n + 31) / 32 * 32; }

char upper_char(char c) { return (c >= 'a' && c <= 'z') ? char(c - 'a' + 'A') : c; }

} // namespace

namespace lapack_impl {

void potrf_bottom_right_check(char uplo, int n, const void* A, int lda, const DdlaHandle_t& h)
{
    require_gpu(h, "potrf_bottom_right");
    if (uplo != 'U' && uplo != 'L') throw std::runtime_error("potrf_bottom_right: uplo must be 'U' or 'L'");
    if (n < 0) throw std::runtime_error("potrf_bottom_right: negative matrix size");
    if (n == 0) return;
    if (A == nullptr) throw std::runtime_error("potrf_bottom_right: null A");
    if (lda < n) throw std::runtime_error("potrf_bottom_right: lda too small");
}

// [ info word | the reversed copy (reversed_ld(n) x n) | panel work ]
template <class T>
void potrf_bottom_right_workspace_size(const DdlaHandle_t&, char, int n, T*, std::size_t* backend_bytes)
{
    *backend_bytes = 0;
    if (n <= 0) return;
    *backend_bytes = detail::padded_bytes<int>(1) + detail::padded_bytes<T>((std::size_t)reversed_ld(n) * n) +
                     detail::padded_bytes<char>(kern::panel_work_bytes<T>(n));
}

template <class T>
void potrf_bottom_right_run(const DdlaHandle_t& h, char uplo, int n, T* A, int lda, int& info, void* backend_work,
                            std::size_t backend_bytes)
{
    info = 0;
    if (n <= 0) return;
    std::size_t need = 0;
    potrf_bottom_right_workspace_size<T>(h, uplo, n, A, &need);
    if (backend_work == nullptr || backend_bytes < need)
        throw std::invalid_argument("potrf_bottom_right: backend workspace smaller than the workspace query");
    const bool upper = upper_char(uplo) == 'U';
    const long long ldw = reversed_ld(n);
    char* w = static_cast<char*>(backend_work);
    int* d_info = detail::take<int>(w, 1);
    T* rev = detail::take<T>(w, (std::size_t)ldw * n);
    void* panel = detail::take<char>(w, kern::panel_work_bytes<T>(n));
    hipStream_t s = h->stream;
    fill_int(s, d_info, 0);
    {
        LaunchTimer timer(h, kClassOther, 0, ((double)n * (n + 1) / 2 + (double)n * n) * sizeof(T));
        // uplo 'U' rotates, uplo 'L' anti-transposes, into the copy's lower triangle
        kern::reverse_triangle<T>(s, false, !upper, n, A, lda, rev, ldw, true);
    }
    potrf_blocked_dispatch<T>(h, false, n, rev, ldw, d_info, panel);
    {
        LaunchTimer timer(h, kClassOther, 0, (double)n * (n + 1) * sizeof(T));
        kern::reverse_triangle<T>(s, upper, !upper, n, rev, ldw, A, lda, false);
    }
    int rinfo = 0, error = 0;
    DDLA_HIP(hipMemcpyAsync(&rinfo, d_info, sizeof(int), hipMemcpyDeviceToHost, s));
    DDLA_HIP(hipMemcpyAsync(&error, kern::panel_error_word<T>(n, panel), sizeof(int), hipMemcpyDeviceToHost, s));
    DDLA_HIP(hipStreamSynchronize(s));
    if (error != 0) throw std::runtime_error("potrf_bottom_right: a panel factorization wait exceeded its limit");
    info = rinfo == 0 ? 0 : n - rinfo + 1;
}

} // namespace lapack_impl

namespace cholesky_impl {

void ppotrf_bottom_right_check(const DdlaHandle_t& h, char uplo, int n, const void* A, const int* d)
{
    check_desc(d, h);
    require_gpu(h, "ppotrf_bottom_right");
    if ((uplo != 'L' && uplo != 'U') || n < 0 || n > d[DDLA_M_] || n > d[DDLA_N_] || d[DDLA_MB_] != d[DDLA_NB_])
        throw std::invalid_argument("ppotrf_bottom_right: invalid triangle, dimensions, blocks or grid");
    if (n && !A) throw std::invalid_argument("ppotrf_bottom_right: null local matrix");
}

template <class T>
void ppotrf_bottom_right_workspace_size(const DdlaHandle_t& h, char uplo, int n, T* A, const int*,
                                        std::size_t* backend_bytes, std::size_t* host_bytes)
{
    *host_bytes = 0;
    lapack_impl::potrf_bottom_right_workspace_size<T>(h, uplo, n, A, backend_bytes);
}

template <class T>
void ppotrf_bottom_right_run(const DdlaHandle_t& h, char uplo, int n, T* A, const int* d, int& info,
                             void* backend_work, std::size_t backend_bytes, void*, std::size_t)
{
    lapack_impl::potrf_bottom_right_run<T>(h, uplo, n, A, d[DDLA_LLD_], info, backend_work, backend_bytes);
}

} // namespace cholesky_impl

template <class T>
void potrf_bottom_right(char uplo, int n, T* A, int lda, int& info, const DdlaHandle_t& h)
{
    lapack_impl::potrf_bottom_right_check(uplo, n, A, lda, h);
    info = 0;
    if (n == 0) return;
    std::size_t bytes = 0;
    lapack_impl::potrf_bottom_right_workspace_size<T>(h, uplo, n, A, &bytes);
    lapack_impl::potrf_bottom_right_run<T>(h, uplo, n, A, lda, info, h->ensure_scratch(bytes), bytes);
}

template <class T>
void ppotrf_bottom_right(const DdlaHandle_t& h, char uplo, int n, T* A, const int* descA, int& info)
{
    cholesky_impl::ppotrf_bottom_right_check(h, uplo, n, A, descA);
    info = 0;
    if (n == 0) return;
    std::size_t bytes = 0, host = 0;
    cholesky_impl::ppotrf_bottom_right_workspace_size<T>(h, uplo, n, A, descA, &bytes, &host);
    cholesky_impl::ppotrf_bottom_right_run<T>(h, uplo, n, A, descA, info, h->ensure_scratch(bytes), bytes, nullptr,
                                              0);
}

#define DDLA_INSTANTIATE(T)                                                                                    \
    template void potrf_bottom_right<T>(char, int, T*, int, int&, const DdlaHandle_t&);                        \
    template void ppotrf_bottom_right<T>(const DdlaHandle_t&, char, int, T*, const int*, int&);                \
    template void lapack_impl::potrf_bottom_right_workspace_size<T>(const DdlaHandle_t&, char, int, T*,        \
                                                                    std::size_t*);                             \
    template void lapack_impl::potrf_bottom_right_run<T>(const DdlaHandle_t&, char, int, T*, int, int&, void*, \
                                                         std::size_t);                                         \
    template void cholesky_impl::ppotrf_bottom_right_workspace_size<T>(const DdlaHandle_t&, char, int, T*,     \
                                                                       const int*, std::size_t*, std::size_t*); \
    template void cholesky_impl::ppotrf_bottom_right_run<T>(const DdlaHandle_t&, char, int, T*, const int*,    \
                                                            int&, void*, std::size_t, void*, std::size_t);
DDLA_INSTANTIATE(float)
DDLA_INSTANTIATE(double)
#undef DDLA_INSTANTIATE

} // namespace ddla
