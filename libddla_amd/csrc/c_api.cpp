// C ABI over the C++ API (declarations and reference citations in
// include/ddla_mi355x.h).  Same conventions as the reference's generated
// shim, fortran/ddla_c_api.cpp:43-98: integer statuses, a C++ exception
// becomes status 6 with its text kept per thread, the handle travels as void*.
#include "internal.h"
#include "kernels.h"

#include <ddla_mi355x.h>

#include <complex>
#include <exception>
#include <string>

namespace {

thread_local std::string g_last_error;

void record_error(const char* what)
{
    try {
        g_last_error = what ? what : "";
    } catch (...) {
        g_last_error.clear();
    }
}

template <typename F>
int guarded(F&& body)
{
    try {
        return body();
    } catch (const std::exception& e) {
        record_error(e.what());
        return DDLA_C_STATUS_EXCEPTION;
    } catch (...) {
        record_error("unknown C++ exception in a LibDDLA C API call");
        return DDLA_C_STATUS_EXCEPTION;
    }
}

// A computational call: a null handle is status 1; an unavailable backend 4.
template <typename F>
int run(void* handle, F&& body)
{
    if (handle == nullptr) return DDLA_C_STATUS_INVALID_HANDLE;
    return guarded([&]() -> int {
        const ddla::DdlaHandle_t h = static_cast<ddla::DdlaHandle_t>(handle);
        if (!ddla::ddlaBackendAvailable(ddla::ddlaGetBackend(h))) return DDLA_C_STATUS_BACKEND_UNAVAILABLE;
        return body(h);
    });
}

inline ddla::DdlaHandle_t H(void* p) { return static_cast<ddla::DdlaHandle_t>(p); }

// A scalar passed by address: T itself for the real types, a (real,
// imaginary) pair behind a void pointer for the complex ones.
template <class T, class S>
T scalar(const S* p)
{
    return *static_cast<const T*>(static_cast<const void*>(p));
}
inline int S(ddla::ddlaStatus_t s) { return static_cast<int>(s); }

} // namespace

extern "C" {

const char* ddla_c_error_string(int status)
{
    if (status == DDLA_C_STATUS_EXCEPTION && !g_last_error.empty()) return g_last_error.c_str();
    if (status >= 0 && status <= 5) return ddla::ddlaGetErrorString(static_cast<ddla::ddlaStatus_t>(status));
    return "unknown LibDDLA C API status";
}

int ddla_c_init_backend(void** handle, int backend)
{
    return guarded([&]() -> int {
        if (handle == nullptr) return DDLA_C_STATUS_INVALID_VALUE;
        if (backend < 0 || backend > 1) return DDLA_C_STATUS_INVALID_VALUE;
        const ddla::DdlaBackend be = backend == 0 ? ddla::DdlaBackend::CPU : ddla::DdlaBackend::GPU;
        if (!ddla::ddlaBackendAvailable(be)) return DDLA_C_STATUS_BACKEND_UNAVAILABLE;
        ddla::DdlaHandle_t h = nullptr;
        const ddla::ddlaStatus_t st = ddla::ddlaInit(h, be);
        if (st != ddla::ddlaStatus_t::DDLA_STATUS_SUCCESS) return S(st);
        *handle = h;
        return DDLA_C_STATUS_SUCCESS;
    });
}

int ddla_c_set(void* handle, int /*f_comm*/, int nprows, int npcols, char major)
{
    if (handle == nullptr) return DDLA_C_STATUS_INVALID_HANDLE;
    return guarded([&]() -> int { return S(ddla::ddlaSet(H(handle), nprows, npcols, major)); });
}

int ddla_c_destroy(void** handle)
{
    return guarded([&]() -> int {
        if (handle == nullptr || *handle == nullptr) return DDLA_C_STATUS_SUCCESS;
        ddla::DdlaHandle_t h = H(*handle);
        const ddla::ddlaStatus_t st = ddla::ddlaDestroy(h);
        *handle = h;
        return S(st);
    });
}

int ddla_c_backend_available(int backend)
{
    if (backend < 0 || backend > 1) return 0;
    return ddla::ddlaBackendAvailable(backend == 0 ? ddla::DdlaBackend::CPU : ddla::DdlaBackend::GPU) ? 1 : 0;
}

int ddla_c_get_backend(void* handle)
{
    if (handle == nullptr) return -DDLA_C_STATUS_INVALID_HANDLE;
    return static_cast<int>(ddla::ddlaGetBackend(H(handle)));
}

int ddla_c_grid_dims(void* handle, int* nprows, int* npcols)
{
    if (handle == nullptr) return DDLA_C_STATUS_INVALID_HANDLE;
    if (nprows == nullptr || npcols == nullptr) return DDLA_C_STATUS_INVALID_VALUE;
    ddla::ddlaGetGridDims(H(handle), *nprows, *npcols);
    return DDLA_C_STATUS_SUCCESS;
}

int ddla_c_grid_coords(void* handle, int* myprow, int* mypcol)
{
    if (handle == nullptr) return DDLA_C_STATUS_INVALID_HANDLE;
    if (myprow == nullptr || mypcol == nullptr) return DDLA_C_STATUS_INVALID_VALUE;
    ddla::ddlaGetGridCoords(H(handle), *myprow, *mypcol);
    return DDLA_C_STATUS_SUCCESS;
}

int ddla_c_rank(void* handle) { return handle == nullptr ? -1 : ddla::ddlaGetRank(H(handle)); }
int ddla_c_size(void* handle) { return handle == nullptr ? -1 : ddla::ddlaGetSize(H(handle)); }

int ddla_c_descinit(int* desc, void* handle, int m, int n, int mb, int nb, int irsrc, int icsrc, int lld)
{
    if (handle == nullptr) return DDLA_C_STATUS_INVALID_HANDLE;
    return S(ddla::ddlaDescInit(desc, H(handle), m, n, mb, nb, irsrc, icsrc, lld));
}

int ddla_c_numloc(int n, int nb, int iproc, int srcproc, int nprocs)
{
    return ddla::num_loc(n, nb, iproc, srcproc, nprocs);
}

int ddla_c_malloc(void* handle, void** ptr, long long bytes)
{
    if (handle == nullptr) return DDLA_C_STATUS_INVALID_HANDLE;
    if (bytes < 0) return DDLA_C_STATUS_INVALID_VALUE;
    return S(ddla::ddlaMalloc(H(handle), ptr, static_cast<std::size_t>(bytes)));
}

int ddla_c_free(void* handle, void* ptr)
{
    if (handle == nullptr) return DDLA_C_STATUS_INVALID_HANDLE;
    return S(ddla::ddlaFree(H(handle), ptr));
}

int ddla_c_memcpy(void* handle, void* dst, const void* src, long long bytes)
{
    if (handle == nullptr) return DDLA_C_STATUS_INVALID_HANDLE;
    if (bytes < 0) return DDLA_C_STATUS_INVALID_VALUE;
    return S(ddla::ddlaMemcpy(H(handle), dst, src, static_cast<std::size_t>(bytes)));
}

int ddla_c_memcpy2d(void* handle, void* dst, long long dpitch, const void* src, long long spitch, long long width,
                    long long height)
{
    if (handle == nullptr) return DDLA_C_STATUS_INVALID_HANDLE;
    if (dpitch < 0 || spitch < 0 || width < 0 || height < 0) return DDLA_C_STATUS_INVALID_VALUE;
    return S(ddla::ddlaMemcpy2D(H(handle), dst, (std::size_t)dpitch, src, (std::size_t)spitch, (std::size_t)width,
                                (std::size_t)height));
}

int ddla_c_set_communication(void* handle, int communication)
{
    return guarded([&]() -> int {
        return S(ddla::ddlaSetCommunication(H(handle), static_cast<ddla::DdlaCommunication>(communication)));
    });
}

int ddla_c_get_communication(void* handle, int* communication)
{
    if (handle == nullptr) return DDLA_C_STATUS_INVALID_HANDLE;
    if (communication == nullptr) return DDLA_C_STATUS_INVALID_VALUE;
    return guarded([&]() -> int {
        *communication = static_cast<int>(ddla::ddlaGetCommunication(H(handle)));
        return DDLA_C_STATUS_SUCCESS;
    });
}

int ddla_c_communication_available(void* handle, int communication)
{
    return ddla::ddlaCommunicationAvailable(H(handle), static_cast<ddla::DdlaCommunication>(communication)) ? 1 : 0;
}

int ddla_c_set_wait_limit(int spins) { return ddla::kern::set_wait_limit(spins); }
int ddla_c_set_panel_width(int cols) { return ddla::set_panel_width(cols); }

int ddla_c_potrf_schedule(int n, int panel_width, int* steps, int capacity, int* count)
{
    return guarded([&]() -> int {
        if (count == nullptr || capacity < 0 || (capacity > 0 && steps == nullptr)) return DDLA_C_STATUS_INVALID_VALUE;
        *count = ddla::potrf_schedule_steps(n, panel_width, steps, capacity);
        return DDLA_C_STATUS_SUCCESS;
    });
}

// One process: the communicator helpers describe a world of one rank.
int ddla_c_default_comm() { return 0; }
int ddla_c_comm_rank(int) { return 0; }
int ddla_c_comm_size(int) { return 1; }
void ddla_c_finalize() {}
int ddla_c_bcast_dims(int, int*, int*, int*, int) { return 0; }

int ddla_c_memset(void* handle, void* ptr, int value, long long bytes)
{
    if (handle == nullptr) return DDLA_C_STATUS_INVALID_HANDLE;
    if (bytes < 0) return DDLA_C_STATUS_INVALID_VALUE;
    return S(ddla::ddlaMemset(H(handle), ptr, value, static_cast<std::size_t>(bytes)));
}

int ddla_c_synchronize(void* handle)
{
    if (handle == nullptr) return DDLA_C_STATUS_INVALID_HANDLE;
    return S(ddla::ddlaSynchronize(H(handle)));
}

int ddla_c_set_stream(void* handle, void* stream)
{
    if (handle == nullptr) return DDLA_C_STATUS_INVALID_HANDLE;
    return guarded([&]() -> int { return S(ddla::ddlaSetStream(H(handle), stream)); });
}

void* ddla_c_get_stream(void* handle)
{
    if (handle == nullptr) return nullptr;
    return ddla::ddlaGetStream(H(handle));
}

// ---- drivers ---------------------------------------------------------------
#define DDLA_C_DRIVERS(P, T, R, S)                                                                                  \
    int ddla_##P##_ppotrf(void* handle, char uplo, int n, void* A, int ia, int ja, const int* descA, int* info) \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            if (info == nullptr) return DDLA_C_STATUS_INVALID_VALUE;                                           \
            ddla::ppotrf<T>(h, uplo, n, static_cast<T*>(A), ia, ja, descA, *info);                             \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }                                                                                                          \
    int ddla_##P##_ppotrs(void* handle, char side, char uplo, char trans, int n, int nrhs, const void* A,       \
                          const int* descA, void* B, const int* descB)                                         \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            ddla::ppotrs<T>(h, side, uplo, trans, n, nrhs, static_cast<const T*>(A), descA, static_cast<T*>(B), \
                            descB);                                                                            \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }                                                                                                          \
    int ddla_##P##_pposv(void* handle, char side, char uplo, char trans, int n, int nrhs, void* A, int ia,      \
                         int ja, const int* descA, void* B, int ib, int jb, const int* descB, int* info)       \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            if (info == nullptr) return DDLA_C_STATUS_INVALID_VALUE;                                           \
            ddla::pposv<T>(h, side, uplo, trans, n, nrhs, static_cast<T*>(A), ia, ja, descA,                   \
                           static_cast<T*>(B), ib, jb, descB, *info);                                          \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }                                                                                                          \
    int ddla_##P##_ppotrf_workspace_size(void* handle, char uplo, int n, const int* descA,                     \
                                         long long* backend_bytes, long long* host_bytes)                      \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            if (!backend_bytes || !host_bytes) return DDLA_C_STATUS_INVALID_VALUE;                             \
            std::size_t b = 0, hb = 0;                                                                         \
            ddla::cholesky_impl::ppotrf_workspace_size<T>(h, uplo, n, static_cast<T*>(nullptr), descA, &b,     \
                                                          &hb);                                                \
            *backend_bytes = (long long)b;                                                                     \
            *host_bytes = (long long)hb;                                                                       \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }                                                                                                          \
    int ddla_##P##_ppotrf_run(void* handle, char uplo, int n, void* A, int ia, int ja, const int* descA,       \
                              int* info, void* bw, long long bb, void* hw, long long hb)                       \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            if (info == nullptr || bb < 0 || hb < 0) return DDLA_C_STATUS_INVALID_VALUE;                       \
            ddla::cholesky_impl::ppotrf_run<T>(h, uplo, n, static_cast<T*>(A), ia, ja, descA, *info, bw,       \
                                               (std::size_t)bb, hw, (std::size_t)hb);                          \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }                                                                                                          \
    int ddla_##P##_ppotrs_workspace_size(void* handle, char side, int n, int nrhs, const int* descA,           \
                                         const int* descB, long long* backend_bytes, long long* host_bytes)    \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            if (!backend_bytes || !host_bytes) return DDLA_C_STATUS_INVALID_VALUE;                             \
            std::size_t b = 0, hb = 0;                                                                         \
            ddla::cholesky_impl::ppotrs_workspace_size<T>(h, side, n, nrhs, descA, descB, &b, &hb);            \
            *backend_bytes = (long long)b;                                                                     \
            *host_bytes = (long long)hb;                                                                       \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }                                                                                                          \
    int ddla_##P##_ppotrs_run(void* handle, char side, char uplo, char trans, int n, int nrhs, const void* A,  \
                              const int* descA, void* B, const int* descB, void* bw, long long bb, void* hw,   \
                              long long hb)                                                                    \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            if (bb < 0 || hb < 0) return DDLA_C_STATUS_INVALID_VALUE;                                          \
            ddla::cholesky_impl::ppotrs_run<T>(h, side, uplo, trans, n, nrhs, static_cast<const T*>(A), descA, \
                                               static_cast<T*>(B), descB, bw, (std::size_t)bb, hw,             \
                                               (std::size_t)hb);                                               \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }                                                                                                          \
    int ddla_##P##_pposv_workspace_size(void* handle, char side, char uplo, int n, int nrhs, const int* descA, \
                                        const int* descB, long long* backend_bytes, long long* host_bytes)     \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            if (!backend_bytes || !host_bytes) return DDLA_C_STATUS_INVALID_VALUE;                             \
            std::size_t b = 0, hb = 0;                                                                         \
            ddla::cholesky_impl::pposv_workspace_size<T>(h, side, uplo, n, nrhs, static_cast<T*>(nullptr),     \
                                                         descA, descB, &b, &hb);                               \
            *backend_bytes = (long long)b;                                                                     \
            *host_bytes = (long long)hb;                                                                       \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }                                                                                                          \
    int ddla_##P##_pposv_run(void* handle, char side, char uplo, char trans, int n, int nrhs, void* A, int ia, \
                             int ja, const int* descA, void* B, int ib, int jb, const int* descB, int* info,   \
                             void* bw, long long bb, void* hw, long long hb)                                   \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            if (info == nullptr || bb < 0 || hb < 0) return DDLA_C_STATUS_INVALID_VALUE;                       \
            ddla::cholesky_impl::pposv_run<T>(h, side, uplo, trans, n, nrhs, static_cast<T*>(A), ia, ja,       \
                                              descA, static_cast<T*>(B), ib, jb, descB, *info, bw,             \
                                              (std::size_t)bb, hw, (std::size_t)hb);                           \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }                                                                                                          \
    int ddla_##P##_potrf(void* handle, char uplo, int n, void* A, int lda, int* d_info)                        \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            ddla::potrf<T>(h, uplo, n, static_cast<T*>(A), lda, d_info);                                       \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }                                                                                                          \
    int ddla_##P##_diag_logdet(void* handle, int n, const void* A, int lda, R* result)                         \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            if (result == nullptr) return DDLA_C_STATUS_INVALID_VALUE;                                         \
            std::complex<R> z;                                                                                 \
            ddla::diag_logdet<T>(h, n, static_cast<const T*>(A), lda, z);                                      \
            result[0] = z.real();                                                                              \
            result[1] = z.imag();                                                                              \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }                                                                                                          \
    int ddla_##P##_pdiag_logdet(void* handle, int n, const void* A, const int* descA, R* result)               \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            if (result == nullptr) return DDLA_C_STATUS_INVALID_VALUE;                                         \
            std::complex<R> z;                                                                                 \
            ddla::pdiag_logdet<T>(h, n, static_cast<const T*>(A), descA, z);                                   \
            result[0] = z.real();                                                                              \
            result[1] = z.imag();                                                                              \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }                                                                                                          \
    int ddla_##P##_gemm_vbatched(void* handle, char ta, char tb, int* d_m, int* d_n, int* d_k, const S* alpha, \
                                 const void* const* d_A, int* d_lda, const void* const* d_B, int* d_ldb,       \
                                 const S* beta, void** d_C, int* d_ldc, int batch_count)                       \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            if (alpha == nullptr || beta == nullptr) return DDLA_C_STATUS_INVALID_VALUE;                       \
            ddla::gemmVbatched<T>(h, ta, tb, d_m, d_n, d_k, scalar<T>(alpha), reinterpret_cast<const T* const*>(d_A),    \
                                  d_lda, reinterpret_cast<const T* const*>(d_B), d_ldb, scalar<T>(beta),                 \
                                  reinterpret_cast<T**>(d_C), d_ldc, batch_count);                             \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }                                                                                                          \
    int ddla_##P##_gemm_vbatched2s(                                                                            \
        void* handle, char ta0, char tb0, int* d_m0, int* d_n0, int* d_k0, const S* alpha0,                    \
        const void* const* d_A0, int* d_lda0, const void* const* d_B0, int* d_ldb0, const S* beta0,            \
        void** d_C0, int* d_ldc0, char ta1, char tb1, int* d_m1, int* d_n1, int* d_k1, const S* alpha1,        \
        const void* const* d_AB1, int* d_lda1, int* d_ldb1, const S* beta1, void** d_C1, int* d_ldc1,          \
        int C0_left, int batch_count, const int* segment_sizes, int segment_count)                             \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            if (!alpha0 || !beta0 || !alpha1 || !beta1) return DDLA_C_STATUS_INVALID_VALUE;                    \
            ddla::gemmVbatched2s<T>(h, ta0, tb0, d_m0, d_n0, d_k0, scalar<T>(alpha0),                                    \
                                    reinterpret_cast<const T* const*>(d_A0), d_lda0,                           \
                                    reinterpret_cast<const T* const*>(d_B0), d_ldb0, scalar<T>(beta0),                   \
                                    reinterpret_cast<T**>(d_C0), d_ldc0, ta1, tb1, d_m1, d_n1, d_k1, scalar<T>(alpha1),  \
                                    reinterpret_cast<const T* const*>(d_AB1), d_lda1, d_ldb1, scalar<T>(beta1),          \
                                    reinterpret_cast<T**>(d_C1), d_ldc1, C0_left != 0, batch_count,            \
                                    segment_sizes, segment_count);                                             \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }

DDLA_C_DRIVERS(s, float, float, float)
DDLA_C_DRIVERS(d, double, double, double)
DDLA_C_DRIVERS(c, std::complex<float>, float, void)
DDLA_C_DRIVERS(z, std::complex<double>, double, void)
#undef DDLA_C_DRIVERS

#define DDLA_C_BOTTOM_RIGHT(P, T)                                                                                \
    int ddla_##P##_potrf_bottom_right(void* handle, char uplo, int n, void* A, int lda, int* info)             \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            if (info == nullptr) return DDLA_C_STATUS_INVALID_VALUE;                                           \
            ddla::potrf_bottom_right<T>(uplo, n, static_cast<T*>(A), lda, *info, h);                           \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }                                                                                                          \
    int ddla_##P##_ppotrf_bottom_right(void* handle, char uplo, int n, void* A, const int* descA, int* info)   \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            if (info == nullptr) return DDLA_C_STATUS_INVALID_VALUE;                                           \
            ddla::ppotrf_bottom_right<T>(h, uplo, n, static_cast<T*>(A), descA, *info);                        \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }

DDLA_C_BOTTOM_RIGHT(s, float)
DDLA_C_BOTTOM_RIGHT(d, double)
#undef DDLA_C_BOTTOM_RIGHT

// Mixed real/complex gemmVbatched(2s) (the C++ overloads of
// include/ddla/ddla.h:1878-1960; no C entry in the reference): real_operand 0
// makes A (stage 0: A_0) the real operand, 1 makes B (B_0) real; ab1_real
// makes the stage-1 external operand real.
#define DDLA_C_MIXED(P, R)                                                                                     \
    int ddla_##P##_gemm_vbatched_mixed(void* handle, int real_operand, char ta, char tb, int* d_m, int* d_n,   \
                                       int* d_k, const void* alpha, const void* const* d_A,          \
                                       int* d_lda, const void* const* d_B, int* d_ldb,                          \
                                       const void* beta, void** d_C, int* d_ldc, int batch_count)    \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            using C = std::complex<R>;                                                                         \
            if (alpha == nullptr || beta == nullptr || (real_operand != 0 && real_operand != 1))               \
                return DDLA_C_STATUS_INVALID_VALUE;                                                            \
            if (real_operand == 0)                                                                             \
                ddla::gemmVbatched<R>(h, ta, tb, d_m, d_n, d_k, scalar<C>(alpha), reinterpret_cast<const R* const*>(d_A),\
                                      d_lda, reinterpret_cast<const C* const*>(d_B), d_ldb, scalar<C>(beta),             \
                                      reinterpret_cast<C**>(d_C), d_ldc, batch_count);                         \
            else                                                                                               \
                ddla::gemmVbatched<R>(h, ta, tb, d_m, d_n, d_k, scalar<C>(alpha), reinterpret_cast<const C* const*>(d_A),\
                                      d_lda, reinterpret_cast<const R* const*>(d_B), d_ldb, scalar<C>(beta),             \
                                      reinterpret_cast<C**>(d_C), d_ldc, batch_count);                         \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }                                                                                                          \
    int ddla_##P##_gemm_vbatched2s_mixed(                                                                      \
        void* handle, int real_operand, int ab1_real, char ta0, char tb0, int* d_m0, int* d_n0, int* d_k0,     \
        const void* alpha0, const void* const* d_A0, int* d_lda0, const void* const* d_B0,          \
        int* d_ldb0, const void* beta0, void** d_C0, int* d_ldc0, char ta1, char tb1, int* d_m1,    \
        int* d_n1, int* d_k1, const void* alpha1, const void* const* d_AB1, int* d_lda1,            \
        int* d_ldb1, const void* beta1, void** d_C1, int* d_ldc1, int C0_left, int batch_count,     \
        const int* segment_sizes, int segment_count)                                                           \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            using C = std::complex<R>;                                                                         \
            if (!alpha0 || !beta0 || !alpha1 || !beta1 || (real_operand != 0 && real_operand != 1))            \
                return DDLA_C_STATUS_INVALID_VALUE;                                                            \
            auto cc = [](const void* const* p) { return reinterpret_cast<const C* const*>(p); };               \
            auto rr = [](const void* const* p) { return reinterpret_cast<const R* const*>(p); };               \
            C** c0 = reinterpret_cast<C**>(d_C0);                                                              \
            C** c1 = reinterpret_cast<C**>(d_C1);                                                              \
            if (real_operand == 0 && ab1_real)                                                                 \
                ddla::gemmVbatched2s<R>(h, ta0, tb0, d_m0, d_n0, d_k0, scalar<C>(alpha0), rr(d_A0), d_lda0, cc(d_B0),     \
                                        d_ldb0, scalar<C>(beta0), c0, d_ldc0, ta1, tb1, d_m1, d_n1, d_k1, scalar<C>(alpha1),        \
                                        rr(d_AB1), d_lda1, d_ldb1, scalar<C>(beta1), c1, d_ldc1, C0_left != 0,            \
                                        batch_count, segment_sizes, segment_count);                            \
            else if (real_operand == 1 && ab1_real)                                                            \
                ddla::gemmVbatched2s<R>(h, ta0, tb0, d_m0, d_n0, d_k0, scalar<C>(alpha0), cc(d_A0), d_lda0, rr(d_B0),     \
                                        d_ldb0, scalar<C>(beta0), c0, d_ldc0, ta1, tb1, d_m1, d_n1, d_k1, scalar<C>(alpha1),        \
                                        rr(d_AB1), d_lda1, d_ldb1, scalar<C>(beta1), c1, d_ldc1, C0_left != 0,            \
                                        batch_count, segment_sizes, segment_count);                            \
            else if (real_operand == 0)                                                                        \
                ddla::gemmVbatched2s<R>(h, ta0, tb0, d_m0, d_n0, d_k0, scalar<C>(alpha0), rr(d_A0), d_lda0, cc(d_B0),     \
                                        d_ldb0, scalar<C>(beta0), c0, d_ldc0, ta1, tb1, d_m1, d_n1, d_k1, scalar<C>(alpha1),        \
                                        cc(d_AB1), d_lda1, d_ldb1, scalar<C>(beta1), c1, d_ldc1, C0_left != 0,            \
                                        batch_count, segment_sizes, segment_count);                            \
            else                                                                                               \
                ddla::gemmVbatched2s<R>(h, ta0, tb0, d_m0, d_n0, d_k0, scalar<C>(alpha0), cc(d_A0), d_lda0, rr(d_B0),     \
                                        d_ldb0, scalar<C>(beta0), c0, d_ldc0, ta1, tb1, d_m1, d_n1, d_k1, scalar<C>(alpha1),        \
                                        cc(d_AB1), d_lda1, d_ldb1, scalar<C>(beta1), c1, d_ldc1, C0_left != 0,            \
                                        batch_count, segment_sizes, segment_count);                            \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }

DDLA_C_MIXED(c, float)
DDLA_C_MIXED(z, double)
#undef DDLA_C_MIXED

// ---- triangular solves -----------------------------------------------------------
#define DDLA_C_TRIANGULAR(P, T)                                                                                \
    int ddla_##P##_ptrtrs(void* handle, char side, char uplo, char trans, char diag, int m, int n, void* A,   \
                          const int* descA, void* B, const int* descB)                                         \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            ddla::ptrtrs<T>(h, side, uplo, trans, diag, m, n, static_cast<T*>(A), descA, static_cast<T*>(B),   \
                            descB);                                                                            \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }                                                                                                          \
    int ddla_##P##_trsm(void* handle, char side, char uplo, char trans, char diag, int m, int n,              \
                        const T* alpha, void* A, int lda, void* B, int ldb)                                    \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            if (alpha == nullptr) return DDLA_C_STATUS_INVALID_VALUE;                                          \
            ddla::trsm<T>(h, side, uplo, trans, diag, m, n, *alpha, static_cast<T*>(A), lda, static_cast<T*>(B), \
                          ldb);                                                                                \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }
DDLA_C_TRIANGULAR(s, float)
DDLA_C_TRIANGULAR(d, double)
#undef DDLA_C_TRIANGULAR

// ---- adjoints ------------------------------------------------------------------
#define DDLA_C_ADJOINTS(P, T)                                                                                  \
    int ddla_##P##_potrf_backward(void* handle, char uplo, int n, const void* L, int ldl, const void* Lbar,     \
                                  int ldlb, void* Abar, int ldab)                                              \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            ddla::potrf_backward<T>(h, uplo, n, static_cast<const T*>(L), ldl, static_cast<const T*>(Lbar),    \
                                    ldlb, static_cast<T*>(Abar), ldab);                                        \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }                                                                                                          \
    int ddla_##P##_potrs_backward(void* handle, char uplo, int n, int nrhs, const void* L, int ldl,            \
                                  const void* X, int ldx, const void* Xbar, int ldxb, void* Bbar, int ldbb,    \
                                  void* Lbar, int ldlb)                                                        \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            ddla::potrs_backward<T>(h, uplo, n, nrhs, static_cast<const T*>(L), ldl, static_cast<const T*>(X), \
                                    ldx, static_cast<const T*>(Xbar), ldxb, static_cast<T*>(Bbar), ldbb,       \
                                    static_cast<T*>(Lbar), ldlb);                                              \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }                                                                                                          \
    int ddla_##P##_logdet_backward(void* handle, int n, const void* L, int ldl, double g, void* Lbar,          \
                                   int ldlb)                                                                   \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            ddla::logdet_backward<T>(h, n, static_cast<const T*>(L), ldl, g, static_cast<T*>(Lbar), ldlb);     \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }                                                                                                          \
    int ddla_##P##_potri(void* handle, char uplo, int n, const void* L, int ldl, void* Ainv, int lda)          \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            ddla::potri<T>(h, uplo, n, static_cast<const T*>(L), ldl, static_cast<T*>(Ainv), lda);             \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }
DDLA_C_ADJOINTS(s, float)
DDLA_C_ADJOINTS(d, double)
#undef DDLA_C_ADJOINTS

int ddla_d_gp_batched_workspace_size(int n, long long* bytes_per_problem)
{
    return guarded([&]() -> int {
        if (bytes_per_problem == nullptr) return DDLA_C_STATUS_INVALID_VALUE;
        *bytes_per_problem = (long long)ddla::gpBatchedWorkspaceSize(n);
        return DDLA_C_STATUS_SUCCESS;
    });
}

int ddla_d_gp_batched_forward(void* handle, int batch, int n, const void* K, long long k_stride, const void* y,
                              long long y_stride, void* work, void* loss, int* info)
{
    return run(handle, [&](ddla::DdlaHandle_t h) -> int {
        ddla::gpBatchedForward(h, batch, n, static_cast<const double*>(K), k_stride, static_cast<const double*>(y),
                               y_stride, work, static_cast<double*>(loss), info);
        return DDLA_C_STATUS_SUCCESS;
    });
}

int ddla_d_gp_batched_backward(void* handle, int batch, int n, void* work, const void* g, void* Kbar,
                               long long kbar_stride, void* ybar, long long ybar_stride)
{
    return run(handle, [&](ddla::DdlaHandle_t h) -> int {
        ddla::gpBatchedBackward(h, batch, n, work, static_cast<const double*>(g), static_cast<double*>(Kbar),
                                kbar_stride, static_cast<double*>(ybar), ybar_stride);
        return DDLA_C_STATUS_SUCCESS;
    });
}

// ---- batched SPD solves --------------------------------------------------------
#define DDLA_C_BATCHED(P, T)                                                                                   \
    int ddla_##P##_pposv_vbatched(void* handle, char uplo, const int* d_n, void* const* d_A, const int* d_lda, \
                                  int nrhs, void* const* d_B, const int* d_ldb, int* d_info, double* d_logdet, \
                                  int batch_count, int max_n)                                                  \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            ddla::pposvVbatched<T>(h, uplo, d_n, reinterpret_cast<T* const*>(d_A), d_lda, nrhs,                \
                                   reinterpret_cast<T* const*>(d_B), d_ldb, d_info, d_logdet, batch_count,     \
                                   max_n);                                                                     \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }                                                                                                          \
    int ddla_##P##_pposv_vbatched_workspace_size(void* handle, int batch_count, int max_n,                     \
                                                 long long* backend_bytes)                                     \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            if (backend_bytes == nullptr) return DDLA_C_STATUS_INVALID_VALUE;                                  \
            if (batch_count < 0 || max_n < 0 || max_n > 1024) return DDLA_C_STATUS_INVALID_VALUE;              \
            std::size_t b = 0;                                                                                 \
            ddla::cholesky_impl::pposv_vbatched_workspace_size<T>(h, batch_count, max_n, &b);                  \
            *backend_bytes = (long long)b;                                                                     \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }                                                                                                          \
    int ddla_##P##_pposv_vbatched_run(void* handle, char uplo, const int* d_n, void* const* d_A,               \
                                      const int* d_lda, int nrhs, void* const* d_B, const int* d_ldb,          \
                                      int* d_info, double* d_logdet, int batch_count, int max_n,               \
                                      void* backend_work, long long backend_bytes)                             \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            if (backend_bytes < 0) return DDLA_C_STATUS_INVALID_VALUE;                                         \
            ddla::cholesky_impl::pposv_vbatched_check(h, uplo, batch_count, nrhs, max_n, d_n, d_A, d_lda, d_B, \
                                                      d_ldb, d_info, d_logdet);                                \
            ddla::cholesky_impl::pposv_vbatched_run<T>(h, uplo, d_n, reinterpret_cast<T* const*>(d_A), d_lda,  \
                                                       nrhs, reinterpret_cast<T* const*>(d_B), d_ldb, d_info,  \
                                                       d_logdet, batch_count, max_n, backend_work,             \
                                                       (std::size_t)backend_bytes);                            \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }                                                                                                          \
    int ddla_##P##_pposv_vbatched_tune(void* handle, char uplo, const int* d_n, void* const* d_A,              \
                                       const int* d_lda, int nrhs, void* const* d_B, const int* d_ldb,         \
                                       int batch_count, int max_n, int* occupancy, int* ordered)               \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            int occ = 0;                                                                                       \
            bool ord = false;                                                                                  \
            ddla::pposvVbatchedTune<T>(h, uplo, d_n, reinterpret_cast<T* const*>(d_A), d_lda, nrhs,            \
                                       reinterpret_cast<T* const*>(d_B), d_ldb, batch_count, max_n, &occ,      \
                                       &ord);                                                                  \
            if (occupancy) *occupancy = occ;                                                                   \
            if (ordered) *ordered = ord ? 1 : 0;                                                               \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }                                                                                                          \
    int ddla_##P##_pposv_vbatched_graph(void* handle, char uplo, const int* d_n, void* const* d_A,             \
                                        const int* d_lda, int nrhs, void* const* d_B, const int* d_ldb,        \
                                        int* d_info, double* d_logdet, int batch_count, int max_n,             \
                                        void** graph)                                                          \
    {                                                                                                          \
        return run(handle, [&](ddla::DdlaHandle_t h) -> int {                                                  \
            if (graph == nullptr) return DDLA_C_STATUS_INVALID_VALUE;                                          \
            *graph = ddla::pposvVbatchedGraph<T>(h, uplo, d_n, reinterpret_cast<T* const*>(d_A), d_lda, nrhs,  \
                                                 reinterpret_cast<T* const*>(d_B), d_ldb, d_info, d_logdet,    \
                                                 batch_count, max_n);                                          \
            return DDLA_C_STATUS_SUCCESS;                                                                      \
        });                                                                                                    \
    }
DDLA_C_BATCHED(s, float)
DDLA_C_BATCHED(d, double)
#undef DDLA_C_BATCHED

int ddla_c_graph_launch(void* graph, void* handle)
{
    return run(handle, [&](ddla::DdlaHandle_t h) -> int {
        if (graph == nullptr) return DDLA_C_STATUS_INVALID_VALUE;
        ddla::batchedGraphLaunch(graph, h);
        return DDLA_C_STATUS_SUCCESS;
    });
}

int ddla_c_vbatched_schedule(void* handle, int element_bytes, int* tuned, int* occupancy, int* ordered,
                             int* split_blocks)
{
    return run(handle, [&](ddla::DdlaHandle_t h) -> int {
        if (element_bytes != 4 && element_bytes != 8) return DDLA_C_STATUS_INVALID_VALUE;
        const auto& t = h->batched_tuning[element_bytes == 8 ? 1 : 0];
        if (tuned) *tuned = t.tuned ? 1 : 0;
        if (occupancy) *occupancy = t.occupancy;
        if (ordered) *ordered = t.ordered ? 1 : 0;
        if (split_blocks) *split_blocks = t.split_blocks;
        return DDLA_C_STATUS_SUCCESS;
    });
}

int ddla_c_vbatched_set_schedule(void* handle, int element_bytes, int occupancy, int ordered, int split_blocks)
{
    return run(handle, [&](ddla::DdlaHandle_t h) -> int {
        if ((element_bytes != 4 && element_bytes != 8) || (occupancy != 1 && occupancy != 2) || split_blocks < 1 ||
            split_blocks > ddla::kern::kBatchedMaxBlocks)
            return DDLA_C_STATUS_INVALID_VALUE;
        h->batched_tuning[element_bytes == 8 ? 1 : 0] =
            ddla::DdlaStream::BatchedTuning{true, occupancy, ordered != 0, split_blocks, false, 1};
        return DDLA_C_STATUS_SUCCESS;
    });
}

int ddla_c_vbatched_kernel(void* handle, int element_bytes, int* kernel)
{
    return run(handle, [&](ddla::DdlaHandle_t h) -> int {
        if (element_bytes != 4 && element_bytes != 8) return DDLA_C_STATUS_INVALID_VALUE;
        const auto& t = h->batched_tuning[element_bytes == 8 ? 1 : 0];
        if (kernel) *kernel = t.left_looking ? 1 : (t.step_blocks == 2 ? 2 : 0);
        return DDLA_C_STATUS_SUCCESS;
    });
}

int ddla_c_vbatched_set_kernel(void* handle, int element_bytes, int kernel)
{
    return run(handle, [&](ddla::DdlaHandle_t h) -> int {
        if ((element_bytes != 4 && element_bytes != 8) || kernel < 0 || kernel > 2) return DDLA_C_STATUS_INVALID_VALUE;
        auto& t = h->batched_tuning[element_bytes == 8 ? 1 : 0];
        if (!t.tuned) t = ddla::DdlaStream::BatchedTuning{true, 1, true, ddla::kern::kBatchedMaxBlocks, false, 1};
        t.left_looking = kernel == 1;
        t.step_blocks = kernel == 2 ? 2 : 1;
        return DDLA_C_STATUS_SUCCESS;
    });
}

int ddla_c_graph_destroy(void* graph)
{
    return guarded([&]() -> int {
        ddla::batchedGraphDestroy(graph);
        return DDLA_C_STATUS_SUCCESS;
    });
}

// ---- FP64 pposv on the static scratch (allocates nothing) ---------------------
int ddla_d_pposv_noalloc(void* handle, char side, char uplo, char trans, int n, int nrhs, void* A, int ia, int ja,
                         const int* descA, void* B, int ib, int jb, const int* descB, int* info)
{
    return run(handle, [&](ddla::DdlaHandle_t h) -> int {
        if (info == nullptr) return DDLA_C_STATUS_INVALID_VALUE;
        ddla::pposv_noalloc(h, side, uplo, trans, n, nrhs, static_cast<double*>(A), ia, ja, descA,
                            static_cast<double*>(B), ib, jb, descB, *info);
        return DDLA_C_STATUS_SUCCESS;
    });
}

// ---- mixed-precision pposv -----------------------------------------------------
int ddla_d_pposv_mixed(void* handle, char uplo, int n, int nrhs, void* A, int ia, int ja, const int* descA, void* B,
                       int ib, int jb, const int* descB, int factor_precision, double tol, int max_iter, int* iter,
                       int* info)
{
    return run(handle, [&](ddla::DdlaHandle_t h) -> int {
        if (iter == nullptr || info == nullptr) return DDLA_C_STATUS_INVALID_VALUE;
        ddla::pposv_mixed<double>(h, uplo, n, nrhs, static_cast<double*>(A), ia, ja, descA, static_cast<double*>(B),
                                  ib, jb, descB, factor_precision, tol, max_iter, *iter, *info);
        return DDLA_C_STATUS_SUCCESS;
    });
}

int ddla_d_pposv_mixed_workspace_size(void* handle, char uplo, int n, int nrhs, const int* descA, const int* descB,
                                      long long* backend_bytes, long long* host_bytes)
{
    return run(handle, [&](ddla::DdlaHandle_t h) -> int {
        if (!backend_bytes || !host_bytes) return DDLA_C_STATUS_INVALID_VALUE;
        std::size_t b = 0, hb = 0;
        ddla::cholesky_impl::pposv_mixed_workspace_size<double>(h, uplo, n, nrhs, descA, descB, &b, &hb);
        *backend_bytes = (long long)b;
        *host_bytes = (long long)hb;
        return DDLA_C_STATUS_SUCCESS;
    });
}

int ddla_d_pposv_mixed_run(void* handle, char uplo, int n, int nrhs, void* A, const int* descA, void* B,
                           const int* descB, int factor_precision, double tol, int max_iter, int* iter, int* info,
                           void* backend_work, long long backend_bytes)
{
    return run(handle, [&](ddla::DdlaHandle_t h) -> int {
        if (iter == nullptr || info == nullptr || backend_bytes < 0) return DDLA_C_STATUS_INVALID_VALUE;
        ddla::cholesky_impl::pposv_mixed_run<double>(h, uplo, n, nrhs, static_cast<double*>(A), descA,
                                                     static_cast<double*>(B), descB, factor_precision, tol, max_iter,
                                                     *iter, *info, backend_work, (std::size_t)backend_bytes);
        return DDLA_C_STATUS_SUCCESS;
    });
}

// ---- harness helpers ---------------------------------------------------------
int ddla_c_fill_normal(void* handle, unsigned long long seed, unsigned long long stream, unsigned long long first,
                       long long count, double* d_out)
{
    return run(handle, [&](ddla::DdlaHandle_t h) -> int {
        ddla::require_gpu(h, "fill_normal");
        if (count < 0 || (count > 0 && d_out == nullptr)) return DDLA_C_STATUS_INVALID_VALUE;
        ddla::kern::fill_normal(h->stream, seed, stream, first, count, d_out);
        return DDLA_C_STATUS_SUCCESS;
    });
}

int ddla_d_spd_generate(void* handle, int n, unsigned long long seed, double delta, double* A, int lda, double* work)
{
    return run(handle, [&](ddla::DdlaHandle_t h) -> int {
        ddla::require_gpu(h, "spd_generate");
        if (n < 0 || lda < std::max(1, n) || (n > 0 && (A == nullptr || work == nullptr)))
            return DDLA_C_STATUS_INVALID_VALUE;
        if (n == 0) return DDLA_C_STATUS_SUCCESS;
        ddla::kern::fill_normal(h->stream, seed, 0, 0, (long long)n * n, work);
        ddla::kern::gemm<double>(h->stream, 'N', 'T', n, n, n, 1.0 / n, work, n, work, n, 0.0, A, lda);
        ddla::kern::add_diagonal<double>(h->stream, n, A, lda, delta);
        return DDLA_C_STATUS_SUCCESS;
    });
}

int ddla_c_spd_generate_batch(void* handle, int batch, const int* n, unsigned long long seed, double delta, int fp32,
                              void* const* A, const int* lda, int nrhs, void* const* B, const int* ldb, double* work)
{
    return run(handle, [&](ddla::DdlaHandle_t h) -> int {
        ddla::require_gpu(h, "spd_generate_batch");
        if (batch < 0 || nrhs < 0 || (batch > 0 && (!n || !A || !lda || !work || (nrhs > 0 && (!B || !ldb)))))
            return DDLA_C_STATUS_INVALID_VALUE;
        for (int i = 0; i < batch; ++i) {
            const int m = n[i];
            if (m < 0 || lda[i] < std::max(1, m) || (nrhs > 0 && ldb[i] < std::max(1, m)))
                return DDLA_C_STATUS_INVALID_VALUE;
            if (m == 0) continue;
            const long long mm = (long long)m * m;
            double* g = work;
            double* a64 = fp32 ? work + mm : static_cast<double*>(A[i]);
            const int ld64 = fp32 ? m : lda[i];
            ddla::kern::fill_normal(h->stream, seed, 2ull * i, 0, mm, g);
            ddla::kern::gemm<double>(h->stream, 'N', 'T', m, m, m, 1.0 / m, g, m, g, m, 0.0, a64, ld64);
            ddla::kern::add_diagonal<double>(h->stream, m, a64, ld64, delta);
            if (fp32) ddla::kern::convert<float, double>(h->stream, m, m, a64, ld64, static_cast<float*>(A[i]), lda[i]);
            if (nrhs > 0) {
                double* b64 = fp32 ? g : static_cast<double*>(B[i]);
                const int ldb64 = fp32 ? m : ldb[i];
                if (ldb64 == m || fp32) {
                    ddla::kern::fill_normal(h->stream, seed, 2ull * i + 1, 0, (long long)m * nrhs, b64);
                } else {
                    ddla::kern::fill_normal(h->stream, seed, 2ull * i + 1, 0, (long long)m * nrhs, g);
                    ddla::kern::convert<double, double>(h->stream, m, nrhs, g, m, b64, ldb64);
                }
                if (fp32)
                    ddla::kern::convert<float, double>(h->stream, m, nrhs, b64, m, static_cast<float*>(B[i]), ldb[i]);
            }
        }
        return DDLA_C_STATUS_SUCCESS;
    });
}

// ---- profiler ------------------------------------------------------------------
int ddla_c_profile(void* handle, int enable)
{
    if (handle == nullptr) return DDLA_C_STATUS_INVALID_HANDLE;
    return guarded([&]() -> int {
        ddla::DdlaHandle_t h = H(handle);
        if (h->stream != nullptr || h->initialized) DDLA_HIP(hipStreamSynchronize(h->stream));
        h->release_events();
        h->profiling = enable != 0;
        return DDLA_C_STATUS_SUCCESS;
    });
}

int ddla_c_profile_read(void* handle, int cls, int* count, double* ms, double* flops, double* bytes)
{
    if (handle == nullptr) return DDLA_C_STATUS_INVALID_HANDLE;
    if (count == nullptr || ms == nullptr || flops == nullptr || bytes == nullptr) return DDLA_C_STATUS_INVALID_VALUE;
    return guarded([&]() -> int {
        ddla::DdlaHandle_t h = H(handle);
        DDLA_HIP(hipStreamSynchronize(h->stream));
        int c = 0;
        double t = 0, f = 0, b = 0;
        for (const auto& r : h->records) {
            if (r.cls != cls) continue;
            float e = 0;
            DDLA_HIP(hipEventElapsedTime(&e, r.start, r.stop));
            ++c;
            t += e;
            f += r.flops;
            b += r.bytes;
        }
        *count = c;
        *ms = t;
        *flops = f;
        *bytes = b;
        return DDLA_C_STATUS_SUCCESS;
    });
}

} // extern "C"
