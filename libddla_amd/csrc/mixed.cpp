// Mixed-precision pposv: the FP64 system A X = B is factored in FP32 on the
// matrix cores and solved to FP64 accuracy by iterative refinement.  There is
// no counterpart in the reference; the iteration, stopping test and
// fallback follow LAPACK dsposv: X0 from the FP32 factor, then
// R = B - A X (FP64, on the GPU), stop when every column satisfies
// ||r||_inf <= tol ||A||_inf ||x||_inf, otherwise X += FP32 solve of R; after
// max_iter corrections without convergence (or if the FP32 factorization
// fails) the plain FP64 pposv runs instead.  iter reports the number of
// corrections applied, or -3 (FP32 factorization failed) / -(max_iter + 1)
// (no convergence) when the fallback ran.  A is left unchanged unless the
// fallback ran (then it holds the FP64 factor, like pposv); B receives X.
#include "internal.h"
#include "kernels.h"

#include <cmath>
#include <type_traits>
#include <vector>

namespace ddla {

// Schedules shared with cholesky.cpp.
template <class T>
void potrf_blocked_dispatch(DdlaStream* h, bool upper, int n, T* A, long long lda, int* info, void* work);
void potrf_split_dispatch(DdlaStream* h, bool upper, int n, float* A, long long lda, int* info, void* work,
                          int mode, void* const* bufs);

namespace {

constexpr int kIB = 64;

std::size_t diag_inverse_elems(int n) { return (std::size_t)((n + kIB - 1) / kIB) * kIB * kIB; }

struct MixedLayout {
    std::size_t a32, norm, scalars, x, r, w, info, panel, dinv, sync, error, stage, split[3], total;
};

MixedLayout mixed_layout(int n, int nrhs)
{
    MixedLayout l{};
    std::size_t off = 0;
    auto put = [&](std::size_t bytes) {
        const std::size_t at = off;
        off += (bytes + 255) / 256 * 256;
        return at;
    };
    const std::size_t nn = (std::size_t)n * n, nr = (std::size_t)n * nrhs;
    l.a32 = put(nn * sizeof(float));
    l.norm = put(kern::sym_norm_work_bytes(n));
    l.scalars = put((1 + 2 * (std::size_t)nrhs) * sizeof(double));
    l.x = put(nr * sizeof(double));
    l.r = put(nr * sizeof(double));
    l.w = put(nr * sizeof(float));
    l.info = put(sizeof(int));
    l.panel = put(kern::panel_work_bytes<float>(n));
    l.dinv = put(diag_inverse_elems(n) * sizeof(float));
    l.sync = put(kern::solve_work_bytes(nrhs));
    l.error = put(sizeof(int));
    l.stage = put(kern::stage_work_bytes(n, nrhs, sizeof(float)));
    for (std::size_t& b : l.split) b = put(kern::split_buffer_bytes(n, 2));  // the largest mode (three planes)
    l.total = off;
    return l;
}

void check_mixed_options(int factor, double tol, int max_iter)
{
    if (factor < 0 || factor > 3)
        throw std::invalid_argument("pposv_mixed: factor must be 0 (FP32), 1 (bf16x3), 2 (bf16x6) or 3 (fp16x3)");
    if (!(tol > 0) || max_iter < 0) throw std::invalid_argument("pposv_mixed: tol must be > 0 and max_iter >= 0");
}

} // namespace

namespace cholesky_impl {

void pposv_mixed_check(const DdlaHandle_t& h, char uplo, int n, int nrhs, const void* A, int ia, int ja,
                       const int* descA, const void* B, int ib, int jb, const int* descB)
{
    pposv_check(h, 'L', uplo, 'N', n, nrhs, A, ia, ja, descA, B, ib, jb, descB);
}

template <class T>
void pposv_mixed_workspace_size(const DdlaHandle_t& h, char uplo, int n, int nrhs, const int* descA,
                                const int* descB, std::size_t* backend_bytes, std::size_t* host_bytes)
{
    static_assert(std::is_same<T, double>::value, "pposv_mixed solves FP64 systems");
    *backend_bytes = *host_bytes = 0;
    if (n <= 0) return;
    std::size_t fb = 0, fh = 0;
    pposv_workspace_size<T>(h, 'L', uplo, n, nrhs, static_cast<T*>(nullptr), descA, descB, &fb, &fh);
    *backend_bytes = std::max(mixed_layout(n, nrhs).total, fb);
    *host_bytes = fh;
}

template <class T>
void pposv_mixed_run(const DdlaHandle_t& h, char uplo, int n, int nrhs, T* A, const int* descA, T* B,
                     const int* descB, int factor, double tol, int max_iter, int& iter, int& info,
                     void* backend_work, std::size_t backend_bytes)
{
    check_mixed_options(factor, tol, max_iter);
    iter = 0;
    info = 0;
    if (n == 0) return;
    std::size_t need = 0, host = 0;
    pposv_mixed_workspace_size<T>(h, uplo, n, nrhs, descA, descB, &need, &host);
    if (backend_bytes < need || backend_work == nullptr)
        throw std::invalid_argument("pposv_mixed: backend workspace smaller than the workspace query");
    const MixedLayout l = mixed_layout(n, nrhs);
    char* base = static_cast<char*>(backend_work);
    float* a32 = reinterpret_cast<float*>(base + l.a32);
    double* scalars = reinterpret_cast<double*>(base + l.scalars);  // [anorm | rnorm[nrhs] | xnorm[nrhs]]
    double* x = reinterpret_cast<double*>(base + l.x);
    double* r = reinterpret_cast<double*>(base + l.r);
    float* w = reinterpret_cast<float*>(base + l.w);
    int* info32 = reinterpret_cast<int*>(base + l.info);
    float* dinv = reinterpret_cast<float*>(base + l.dinv);
    int* error = reinterpret_cast<int*>(base + l.error);
    const bool upper = uplo == 'U';
    const long long lda = descA[DDLA_LLD_], ldb = descB[DDLA_LLD_];
    hipStream_t s = h->stream;

    // FP32 copy of the stored triangle and ||A||_inf
    {
        LaunchTimer timer(h, kClassOther, 0, (double)n * n * (sizeof(T) + sizeof(float)) / 2);
        kern::convert_sym_with_norm<float, double>(s, n, !upper, A, lda, a32, n, base + l.norm, scalars);
    }
    fill_int(s, info32, 0);
    if (factor == 0) {
        potrf_blocked_dispatch<float>(h, upper, n, a32, n, info32, base + l.panel);
    } else {
        void* const bufs[3] = {base + l.split[0], base + l.split[1], base + l.split[2]};
        potrf_split_dispatch(h, upper, n, a32, n, info32, base + l.panel, factor, bufs);
    }
    int host_info = 0, panel_error = 0;
    DDLA_HIP(hipMemcpyAsync(&host_info, info32, sizeof(int), hipMemcpyDeviceToHost, s));
    DDLA_HIP(hipMemcpyAsync(&panel_error, kern::panel_error_word<float>(n, base + l.panel), sizeof(int),
                            hipMemcpyDeviceToHost, s));
    DDLA_HIP(hipStreamSynchronize(s));
    if (panel_error != 0) throw std::runtime_error("pposv_mixed: a panel factorization wait exceeded its limit");

    bool converged = false;
    if (host_info == 0) {
        fill_int(s, error, 0);
        {
            LaunchTimer timer(h, kClassTrtri, (double)n * kIB * kIB / 3.0);
            kern::trtri_diag_blocks<float>(s, upper, n, a32, n, dinv);
        }
        // every correction solves with the same factor: the staged solve's
        // blocks are prepared once
        const bool staged = kern::stage_solve(n, nrhs);
        if (staged) {
            LaunchTimer timer(h, kClassTrtri, 2.0 * n * 512.0 * 512.0 + (double)n * 512.0 * 512.0 / 3.0);
            kern::stage_prepare<float>(s, upper, n, a32, n, dinv, base + l.stage);
        }
        auto solve32 = [&]() {
            LaunchTimer timer(h, kClassSolve, 2.0 * n * n * nrhs, (double)n * n * sizeof(float));
            if (staged) kern::stage_sweeps<float>(s, upper, n, nrhs, a32, n, w, n, base + l.stage, error);
            else kern::potrs_sweeps<float>(s, upper, n, nrhs, a32, n, dinv, w, n, base + l.sync, error);
        };
        kern::convert<float, double>(s, n, nrhs, B, ldb, w, n);
        solve32();
        kern::convert<double, float>(s, n, nrhs, w, n, x, n);
        std::vector<double> host(1 + 2 * (std::size_t)nrhs);
        for (int it = 0;; ++it) {
            {
                LaunchTimer timer(h, kClassResidual, 2.0 * n * n * nrhs, (double)n * n * sizeof(T));
                kern::symm_residual<double>(s, !upper, n, nrhs, A, lda, x, n, B, ldb, r, n);
            }
            kern::column_max<double>(s, n, nrhs, r, n, scalars + 1);
            kern::column_max<double>(s, n, nrhs, x, n, scalars + 1 + nrhs);
            DDLA_HIP(hipMemcpyAsync(host.data(), scalars, host.size() * sizeof(double), hipMemcpyDeviceToHost, s));
            DDLA_HIP(hipStreamSynchronize(s));
            const double anorm = host[0];
            bool ok = std::isfinite(anorm);
            for (int j = 0; j < nrhs && ok; ++j) {
                const double rn = host[1 + j], xn = host[1 + nrhs + j];
                ok = std::isfinite(rn) && std::isfinite(xn) && rn <= tol * anorm * xn;
            }
            if (ok) {
                converged = true;
                iter = it;
                break;
            }
            if (it == max_iter) break;
            kern::convert<float, double>(s, n, nrhs, r, n, w, n);
            solve32();
            kern::add_converted<float>(s, n, nrhs, w, n, x, n);
        }
        int host_error = 0;
        DDLA_HIP(hipMemcpyAsync(&host_error, error, sizeof(int), hipMemcpyDeviceToHost, s));
        DDLA_HIP(hipStreamSynchronize(s));
        if (host_error != 0)
            throw std::runtime_error("pposv_mixed: a triangular-solve sweep exceeded its wait limit");
        if (converged) kern::convert<double, double>(s, n, nrhs, x, n, B, ldb);
    }
    if (!converged) {
        iter = host_info != 0 ? -3 : -(max_iter + 1);
        pposv_run<T>(h, 'L', uplo, 'N', n, nrhs, A, 1, 1, descA, B, 1, 1, descB, info, backend_work, backend_bytes,
                     nullptr, 0);
    }
    DDLA_HIP(hipStreamSynchronize(s));
}

template void pposv_mixed_workspace_size<double>(const DdlaHandle_t&, char, int, int, const int*, const int*,
                                                 std::size_t*, std::size_t*);
template void pposv_mixed_run<double>(const DdlaHandle_t&, char, int, int, double*, const int*, double*, const int*,
                                      int, double, int, int&, int&, void*, std::size_t);

} // namespace cholesky_impl

template <class T>
void pposv_mixed(const DdlaHandle_t& h, char uplo, int n, int nrhs, T* A, int ia, int ja, const int* descA, T* B,
                 int ib, int jb, const int* descB, int factor, double tol, int max_iter, int& iter, int& info)
{
    cholesky_impl::pposv_mixed_check(h, uplo, n, nrhs, A, ia, ja, descA, B, ib, jb, descB);
    check_mixed_options(factor, tol, max_iter);
    if (n == 0) {
        iter = 0;
        info = 0;
        return;
    }
    std::size_t bb = 0, hb = 0;
    cholesky_impl::pposv_mixed_workspace_size<T>(h, uplo, n, nrhs, descA, descB, &bb, &hb);
    void* work = h->ensure_scratch(bb);
    cholesky_impl::pposv_mixed_run<T>(h, uplo, n, nrhs, A, descA, B, descB, factor, tol, max_iter, iter, info, work,
                                      bb);
}

template void pposv_mixed<double>(const DdlaHandle_t&, char, int, int, double*, int, int, const int*, double*, int,
                                  int, const int*, int, double, int, int&, int&);

} // namespace ddla
