// ddla/ddla.h -- C++ API of the MI355X-native LibDDLA Cholesky path.
//
// Drop-in for the reference's public templates on this path (same names,
// argument meaning and error behaviour), for a 1x1 process grid on one GPU:
//   handle lifecycle        include/ddla/ddla_handle_t.h:51-67
//   memory helpers          include/ddla/ddla_handle_t.h:130-284
//   descriptor, index maps  include/ddla/ddla_desc.h:10-135
//   ddla::vector            include/ddla/ddla_vector.h:27-110 (ddla/ddla_vector.h,
//                           included at the end of this header)
//   ppotrf / ppotrs / pposv include/ddla/ddla.h:1065, 1149, 1184
//   gemmVbatched(2s)        include/ddla/ddla.h:1836, 1857
//   potrf                   include/ddla/ddla.h:2160
//   potrf_bottom_right(p)   include/ddla/ddla.h:1090, 1115
//   diag_logdet             include/ddla/ddla.h:2901
//   check / size / run      src/Cholesky/cholesky_workspace.h:133-249
// Differences: the library itself needs no MPI.  ddlaSet takes the grid
// shape and accepts 1 x 1; a caller that compiles with MPI (DDLA_USE_MPI
// defined, or <mpi.h> on the include path) also gets the reference's
// communicator overloads, inline here, which accept a one-rank communicator.
// Only the GPU backend exists (ddlaInit with the CPU backend throws, as the
// reference does for a backend that is not built).
#ifndef DDLA_MI355X_DDLA_H
#define DDLA_MI355X_DDLA_H

#include <algorithm>
#include <complex>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <type_traits>

#if !defined(DDLA_USE_MPI) && !defined(DDLA_NO_MPI) && defined(__has_include)
#if __has_include(<mpi.h>)
#define DDLA_USE_MPI 1
#endif
#endif
#if defined(DDLA_USE_MPI)
#include <mpi.h>
#endif

namespace ddla {

enum class DdlaBackend { CPU, GPU };

// Communication mode of a handle (include/ddla/ddla_handle_t.h:22-26).  A 1 x 1
// grid exchanges nothing, so every mode is accepted and only recorded.
enum class DdlaCommunication { MPI = 0, HostStagedMPI = 1, CCL = 2 };

enum class ddlaStatus_t {
    DDLA_STATUS_SUCCESS = 0,
    DDLA_STATUS_INVALID_HANDLE,
    DDLA_STATUS_INVALID_VALUE,
    DDLA_STATUS_ALLOC_FAILED,
    DDLA_STATUS_BACKEND_UNAVAILABLE,
    DDLA_STATUS_INTERNAL_ERROR
};

class DdlaStream;
using DdlaHandle_t = DdlaStream*;

// Lifecycle.  ddlaSet builds the 1 x 1 grid and the device resources (a
// compute stream on the current device); nprows * npcols must be 1 (values
// <= 0 select the default 1 x 1 grid).
ddlaStatus_t ddlaInit(DdlaHandle_t& handle, DdlaBackend requested_backend);
#if defined(DDLA_USE_MPI)
ddlaStatus_t ddlaSet(DdlaHandle_t handle, int nprows, int npcols, char major = 'R');
// The reference's signatures (include/ddla/ddla_handle_t.h:58-64).  The caller
// owns MPI initialization; the communicator must hold exactly one rank (this
// build runs one process on one GPU), else std::invalid_argument, before any
// device work.
inline ddlaStatus_t ddlaSet(DdlaHandle_t handle, const MPI_Comm& comm, int nprows, int npcols, char major = 'R')
{
    int size = 0;
    if (MPI_Comm_size(comm, &size) != MPI_SUCCESS) return ddlaStatus_t::DDLA_STATUS_INVALID_VALUE;
    if (size != 1)
        throw std::invalid_argument("ddlaSet: this build runs one process on one GPU; the communicator must hold one rank");
    return ddlaSet(handle, nprows, npcols, major);
}
inline ddlaStatus_t ddlaSet(DdlaHandle_t handle, const MPI_Comm& comm = MPI_COMM_WORLD, char major = 'R')
{
    return ddlaSet(handle, comm, 1, 1, major);
}
#else
ddlaStatus_t ddlaSet(DdlaHandle_t handle, int nprows = 1, int npcols = 1, char major = 'R');
#endif
ddlaStatus_t ddlaDestroy(DdlaHandle_t& handle);
bool ddlaBackendAvailable(DdlaBackend backend);
DdlaBackend ddlaGetBackend(const DdlaHandle_t& handle);

// Memory and ordering on the handle's queue.
ddlaStatus_t ddlaMalloc(const DdlaHandle_t& handle, void** ptr, std::size_t bytes);
template <typename T>
inline ddlaStatus_t ddlaMalloc(const DdlaHandle_t& handle, T** ptr, std::size_t bytes)
{
    return ddlaMalloc(handle, reinterpret_cast<void**>(ptr), bytes);
}
ddlaStatus_t ddlaFree(const DdlaHandle_t& handle, void* ptr);
ddlaStatus_t ddlaMemcpy(const DdlaHandle_t& handle, void* dst, const void* src, std::size_t bytes);
ddlaStatus_t ddlaMemset(const DdlaHandle_t& handle, void* ptr, int value, std::size_t bytes);
ddlaStatus_t ddlaSynchronize(const DdlaHandle_t& handle);
const char* ddlaGetErrorString(ddlaStatus_t status);

// Further helpers of the reference's handle API (include/ddla/ddla_handle_t.h:
// 146-155, 205-280, 328-334), for the GPU handle:
//   ddlaMallocHost / ddlaFreeHost  pinned host memory (zero bytes: *ptr = null;
//                                  freeing null succeeds)
//   ddlaMemsetTri / ddlaMemcpyTri  one triangle of a column-major m x n matrix
//                                  (uplo 'U': row <= column, 'L': row >= column,
//                                  either case; strict without the diagonal);
//                                  memset repeats the low byte of value; device
//                                  memory takes one kernel on the handle's
//                                  stream, host/device copies go column segment
//                                  by segment, host-to-host copies are
//                                  immediate.  An empty selection succeeds
//                                  untouched; bad uplo, extents, leading
//                                  dimensions or null pointers are
//                                  DDLA_STATUS_INVALID_VALUE, and so is host
//                                  memory for ddlaMemsetTri.
//   ddlaDeviceSynchronize, events (record on / wait on / synchronize with the
//   handle's stream), ddlaGetLastError, ddlaSetDevice, ddlaGetDeviceCount,
//   ddlaMemGetInfo, and the rank <-> grid coordinate maps of the handle's
//   grid order ('R' or 'C').
using ddlaEvent_t = void*;
ddlaStatus_t ddlaMallocHost(const DdlaHandle_t& handle, void** ptr, std::size_t bytes);
template <typename T>
inline ddlaStatus_t ddlaMallocHost(const DdlaHandle_t& handle, T** ptr, std::size_t bytes)
{
    return ddlaMallocHost(handle, reinterpret_cast<void**>(ptr), bytes);
}
ddlaStatus_t ddlaFreeHost(const DdlaHandle_t& handle, void* ptr);
namespace detail {
ddlaStatus_t memset_tri(const DdlaHandle_t& handle, char uplo, int m, int n, void* ptr, int lda, int value,
                        std::size_t elem, bool include_diagonal);
ddlaStatus_t memcpy_tri(const DdlaHandle_t& handle, char uplo, int m, int n, void* dst, int ldd, const void* src,
                        int lds, std::size_t elem, bool include_diagonal);
} // namespace detail
template <class T>
ddlaStatus_t ddlaMemsetTri(const DdlaHandle_t& handle, char uplo, int m, int n, T* ptr, int lda, int value = 0,
                           bool include_diagonal = true)
{
    static_assert(std::is_trivially_copyable<T>::value, "triangular memory requires trivially copyable elements");
    return detail::memset_tri(handle, uplo, m, n, ptr, lda, value, sizeof(T), include_diagonal);
}
template <class T>
ddlaStatus_t ddlaMemcpyTri(const DdlaHandle_t& handle, char uplo, int m, int n, T* dst, int ldd, const T* src,
                           int lds, bool include_diagonal = true)
{
    static_assert(std::is_trivially_copyable<T>::value, "triangular memory requires trivially copyable elements");
    return detail::memcpy_tri(handle, uplo, m, n, dst, ldd, src, lds, sizeof(T), include_diagonal);
}
ddlaStatus_t ddlaDeviceSynchronize(const DdlaHandle_t& handle);
ddlaStatus_t ddlaEventCreate(const DdlaHandle_t& handle, ddlaEvent_t* event);
ddlaStatus_t ddlaEventDestroy(const DdlaHandle_t& handle, ddlaEvent_t event);
ddlaStatus_t ddlaEventRecord(const DdlaHandle_t& handle, ddlaEvent_t event);
ddlaStatus_t ddlaEventWait(const DdlaHandle_t& handle, ddlaEvent_t event);
ddlaStatus_t ddlaEventSynchronize(const DdlaHandle_t& handle, ddlaEvent_t event);
ddlaStatus_t ddlaGetLastError(const DdlaHandle_t& handle);
ddlaStatus_t ddlaSetDevice(const DdlaHandle_t& handle);
ddlaStatus_t ddlaGetDeviceCount(const DdlaHandle_t& handle, int* count);
ddlaStatus_t ddlaMemGetInfo(const DdlaHandle_t& handle, std::size_t* free_bytes, std::size_t* total_bytes);
void ddlaRankToRc(const DdlaHandle_t& handle, int rank, int& row, int& col);
int ddlaRcToRank(const DdlaHandle_t& handle, int row, int col);
// Replace the compute stream (the handle owns a non-null replacement; null
// selects the default stream, which is never destroyed).
// Rectangle copy of height rows of width bytes between pitched buffers
// (include/ddla/ddla_handle_t.h:180-197): direction from where each rectangle
// starts, host-to-host copies finish before returning; an empty rectangle
// succeeds without touching either pointer; null pointers or a pitch below
// width are DDLA_STATUS_INVALID_VALUE.
ddlaStatus_t ddlaMemcpy2D(const DdlaHandle_t& handle, void* dst, std::size_t dpitch, const void* src,
                          std::size_t spitch, std::size_t width, std::size_t height);
// Communication selection (include/ddla/ddla_handle_t.h:80-99): a valid mode
// is recorded (default CCL, the reference's GPU default); an invalid one is
// DDLA_STATUS_INVALID_VALUE and keeps the previous mode; null handles are
// DDLA_STATUS_INVALID_HANDLE for the setter, std::runtime_error for the
// getter and false for the query.
ddlaStatus_t ddlaSetCommunication(const DdlaHandle_t& handle, DdlaCommunication mode);
DdlaCommunication ddlaGetCommunication(const DdlaHandle_t& handle);
bool ddlaCommunicationAvailable(const DdlaHandle_t& handle, DdlaCommunication mode);
ddlaStatus_t ddlaSetStream(const DdlaHandle_t& handle, void* stream);
void* ddlaGetStream(const DdlaHandle_t& handle);

int ddlaGetRank(const DdlaHandle_t& handle);
int ddlaGetSize(const DdlaHandle_t& handle);
void ddlaGetGridCoords(const DdlaHandle_t& handle, int& myprow, int& mypcol);
void ddlaGetGridDims(const DdlaHandle_t& handle, int& nprows, int& npcols);

inline void DDLA_CHECK(ddlaStatus_t status, const char* file = __builtin_FILE(), int line = __builtin_LINE())
{
    if (status != ddlaStatus_t::DDLA_STATUS_SUCCESS)
        throw std::runtime_error(std::string("ddla error ") + std::to_string(static_cast<int>(status)) + " (" +
                                 ddlaGetErrorString(status) + ") at " + file + ":" + std::to_string(line));
}

// ---------------------------------------------------------------------------
// ScaLAPACK int[9] descriptor (include/ddla/ddla_desc.h).
// ---------------------------------------------------------------------------
constexpr int DDLA_DTYPE_ = 0;
constexpr int DDLA_CTXT_ = 1;
constexpr int DDLA_M_ = 2;
constexpr int DDLA_N_ = 3;
constexpr int DDLA_MB_ = 4;
constexpr int DDLA_NB_ = 5;
constexpr int DDLA_RSRC_ = 6;
constexpr int DDLA_CSRC_ = 7;
constexpr int DDLA_LLD_ = 8;
constexpr int DDLA_DLEN_ = 9;
constexpr int DDLA_BLOCK_CYCLIC_2D = 1;

// Number of rows (columns) of an n-long block-cyclic dimension owned by
// process iproc (the ScaLAPACK NUMROC count).
inline int num_loc(int n, int nb, int iproc, int srcproc, int nprocs)
{
    const int mydist = (nprocs + iproc - srcproc) % nprocs;
    const int nblocks = n / nb;
    int count = (nblocks / nprocs) * nb;
    const int extra = nblocks % nprocs;
    if (mydist < extra) count += nb;
    else if (mydist == extra) count += n % nb;
    return count;
}

// Global <-> local index maps of the block-cyclic distribution (0-based; the
// ScaLAPACK tools INDXG2P, INDXG2L and INDXL2G; the reference's
// include/ddla/ddla_desc.h:10-24).  g: global index, l: local index, nb:
// block size, src: process holding the first block, p: process count.
inline int indxg2p(int g, int nb, int src, int p) { return (src + g / nb) % p; }
inline int indxg2l(int g, int nb, int p) { return (g / (nb * p)) * nb + g % nb; }
inline int indxl2g(int l, int nb, int iproc, int src, int p)
{
    const int dist = (p + iproc - src) % p;  // position of iproc in the cycle
    return ((l / nb) * p + dist) * nb + l % nb;
}

// DESCINIT semantics: lld is raised to max(1, LOCr(m)); lld < 0 derives it.
ddlaStatus_t ddlaDescInit(int* desc, const DdlaHandle_t& handle, int m, int n, int mb, int nb, int irsrc, int icsrc,
                          int lld = -1);

// Validation of a descriptor against the handle's grid (src/ddla_desc.cpp:9-60).
void check_desc(const int* desc, const DdlaHandle_t& handle);

// Descriptor-level index maps on the handle's grid (the reference's
// include/ddla/ddla_desc.h:143-191): local row / column of a global one, or -1
// when this process does not own it (or it lies beyond M / N); global row /
// column of a local one.  A null handle gives -1.
namespace desc_detail {
inline bool grid(const DdlaHandle_t& h, int& nprows, int& npcols, int& myprow, int& mypcol)
{
    if (h == nullptr) return false;
    ddlaGetGridDims(h, nprows, npcols);
    ddlaGetGridCoords(h, myprow, mypcol);
    return nprows > 0 && npcols > 0;
}
} // namespace desc_detail
inline int indx_g2l_r(const int* desc, const DdlaHandle_t& handle, int gindx)
{
    int pr = 0, pc = 0, r = -1, c = -1;
    if (!desc_detail::grid(handle, pr, pc, r, c) || gindx >= desc[DDLA_M_] ||
        r != indxg2p(gindx, desc[DDLA_MB_], desc[DDLA_RSRC_], pr))
        return -1;
    return indxg2l(gindx, desc[DDLA_MB_], pr);
}
inline int indx_g2l_c(const int* desc, const DdlaHandle_t& handle, int gindx)
{
    int pr = 0, pc = 0, r = -1, c = -1;
    if (!desc_detail::grid(handle, pr, pc, r, c) || gindx >= desc[DDLA_N_] ||
        c != indxg2p(gindx, desc[DDLA_NB_], desc[DDLA_CSRC_], pc))
        return -1;
    return indxg2l(gindx, desc[DDLA_NB_], pc);
}
inline int indx_l2g_r(const int* desc, const DdlaHandle_t& handle, int lindx)
{
    int pr = 0, pc = 0, r = -1, c = -1;
    if (!desc_detail::grid(handle, pr, pc, r, c)) return -1;
    return indxl2g(lindx, desc[DDLA_MB_], r, desc[DDLA_RSRC_], pr);
}
inline int indx_l2g_c(const int* desc, const DdlaHandle_t& handle, int lindx)
{
    int pr = 0, pc = 0, r = -1, c = -1;
    if (!desc_detail::grid(handle, pr, pc, r, c)) return -1;
    return indxl2g(lindx, desc[DDLA_NB_], c, desc[DDLA_CSRC_], pc);
}

template <class T> struct real_type { using type = T; };
template <class R> struct real_type<std::complex<R>> { using type = R; };

// ---------------------------------------------------------------------------
// Cholesky path.  float and double are instantiated.  The factorizations
// (ppotrf, pposv, potrf) return the strict triangle opposite uplo as they found
// it (the reference's docs/hip-dtk.md:3-7 leaves it unspecified); with uplo 'U'
// it is working storage during the call.
// ---------------------------------------------------------------------------
template <typename T>
void ppotrf(const DdlaHandle_t& handle, char uplo, int n, T* A, int ia, int ja, const int* descA, int& info);

template <typename T>
void ppotrs(const DdlaHandle_t& handle, char side, char uplo, char trans, int n, int nrhs, const T* A,
            const int* descA, T* B, const int* descB);

template <typename T>
void pposv(const DdlaHandle_t& handle, char side, char uplo, char trans, int n, int nrhs, T* A, int ia, int ja,
           const int* descA, T* B, int ib, int jb, const int* descB, int& info);

// Triangular solves: B := op(A)^-1 B (side 'L') or B op(A)^-1 (side 'R'),
// A triangular (uplo), trans 'N' / 'T' / 'C', diag 'N' / 'U' (unit, the
// stored diagonal not read); ptrtrs on descriptors at a 1 x 1 grid (upper-case
// flags), trsm on leading dimensions with alpha (case-insensitive flags;
// alpha == 0 sets B to zero without reading A).  Both return synchronized.
template <typename T>
void ptrtrs(const DdlaHandle_t& handle, char side, char uplo, char trans, char diag, int m, int n, T* A,
            const int* descA, T* B, const int* descB);
template <typename T>
void trsm(const DdlaHandle_t& handle, char side, char uplo, char trans, char diag, int m, int n, const T& alpha,
          T* A, int lda, T* B, int ldb);

// Local factorization with a device-resident info word; asynchronous.  *info
// receives 0, the 1-based index of the first failed pivot, or -1000 when a
// wait inside the factorization exceeded its spin limit (factor not valid).
template <typename T>
void potrf(const DdlaHandle_t& handle, char uplo, int n, T* A, int lda, int* info);

// Cholesky from the bottom-right corner (the reference's include/ddla/ddla.h:1090,
// 1115): uplo 'U' gives A = U U^T in the upper triangle, 'L' gives A = L^T L in
// the lower one; the other triangle is neither read nor written.  info = i > 0
// when the trailing principal submatrix from row i (1-based) is not positive
// definite (A then holds a partial result).  Synchronized.
template <typename T>
void potrf_bottom_right(char uplo, int n, T* A, int lda, int& info, const DdlaHandle_t& handle);
template <typename T>
void ppotrf_bottom_right(const DdlaHandle_t& handle, char uplo, int n, T* A, const int* descA, int& info);

namespace lapack_impl {
// potrf_bottom_right in three components (src/lapack_impl.h:137-155 of the reference).
void potrf_bottom_right_check(char uplo, int n, const void* A, int lda, const DdlaHandle_t& handle);
template <class T>
void potrf_bottom_right_workspace_size(const DdlaHandle_t& handle, char uplo, int n, T* A,
                                       std::size_t* backend_bytes);
template <class T>
void potrf_bottom_right_run(const DdlaHandle_t& handle, char uplo, int n, T* A, int lda, int& info,
                            void* backend_work, std::size_t backend_bytes);
} // namespace lapack_impl

template <typename T>
void diag_logdet(const DdlaHandle_t& h, int n, const T* A, int lda, std::complex<typename real_type<T>::type>& result);
// pdiag_logdet (the reference's include/ddla/ddla.h:2914): diag_logdet of the
// leading n x n block of a descriptor-described matrix, at a 1 x 1 grid.
template <typename T>
void pdiag_logdet(const DdlaHandle_t& h, int n, const T* A, const int* descA,
                  std::complex<typename real_type<T>::type>& result);

template <typename T>
void gemmVbatched(const DdlaHandle_t& handle, char transA, char transB, int* d_m, int* d_n, int* d_k, T alpha,
                  const T* const* d_A_array, int* d_lda, const T* const* d_B_array, int* d_ldb, T beta,
                  T** d_C_array, int* d_ldc, int batch_count);

// Two dependent variable-size stages sharing a temporary (reference
// include/ddla/ddla.h:1857-1877): per host segment, in order, stage 0 writes
// C0 (beta_0 must be zero), then stage 1 computes C1 = alpha_1 op(C0) op(AB1)
// + beta_1 C1 (C0_left) or alpha_1 op(AB1) op(C0) + beta_1 C1.
template <typename T>
void gemmVbatched2s(const DdlaHandle_t& handle, char transA_0, char transB_0, int* d_m_0, int* d_n_0, int* d_k_0,
                    T alpha_0, const T* const* d_A_array_0, int* d_lda_0, const T* const* d_B_array_0, int* d_ldb_0,
                    T beta_0, T** d_C_array_0, int* d_ldc_0, char transA_1, char transB_1, int* d_m_1, int* d_n_1,
                    int* d_k_1, T alpha_1, const T* const* d_AB_array_1, int* d_lda_1, int* d_ldb_1, T beta_1,
                    T** d_C_array_1, int* d_ldc_1, bool C0_left, int batch_count, const int* segment_sizes,
                    int segment_count);

// Mixed real/complex overloads (the reference's include/ddla/ddla.h:1878-1960):
// one operand real, the output and scalars complex; 'C' on a real operand is
// 'T'.  Two-stage forms: stage 0 with the typed operands, the complex
// intermediate C0, and a real or complex external stage-1 operand.
template <typename R>
void gemmVbatched(const DdlaHandle_t& handle, char transA, char transB, int* m, int* n, int* k, std::complex<R> alpha,
                  const R* const* A, int* lda, const std::complex<R>* const* B, int* ldb, std::complex<R> beta,
                  std::complex<R>** C, int* ldc, int batch_count);
template <typename R>
void gemmVbatched(const DdlaHandle_t& handle, char transA, char transB, int* m, int* n, int* k, std::complex<R> alpha,
                  const std::complex<R>* const* A, int* lda, const R* const* B, int* ldb, std::complex<R> beta,
                  std::complex<R>** C, int* ldc, int batch_count);
template <typename R>
void gemmVbatched2s(const DdlaHandle_t& handle, char transA_0, char transB_0, int* m_0, int* n_0, int* k_0,
                    std::complex<R> alpha_0, const R* const* A_0, int* lda_0, const std::complex<R>* const* B_0,
                    int* ldb_0, std::complex<R> beta_0, std::complex<R>** C_0, int* ldc_0, char transA_1,
                    char transB_1, int* m_1, int* n_1, int* k_1, std::complex<R> alpha_1, const R* const* AB_1,
                    int* lda_1, int* ldb_1, std::complex<R> beta_1, std::complex<R>** C_1, int* ldc_1, bool C0_left,
                    int batch_count, const int* segment_sizes, int segment_count);
template <typename R>
void gemmVbatched2s(const DdlaHandle_t& handle, char transA_0, char transB_0, int* m_0, int* n_0, int* k_0,
                    std::complex<R> alpha_0, const std::complex<R>* const* A_0, int* lda_0, const R* const* B_0,
                    int* ldb_0, std::complex<R> beta_0, std::complex<R>** C_0, int* ldc_0, char transA_1,
                    char transB_1, int* m_1, int* n_1, int* k_1, std::complex<R> alpha_1, const R* const* AB_1,
                    int* lda_1, int* ldb_1, std::complex<R> beta_1, std::complex<R>** C_1, int* ldc_1, bool C0_left,
                    int batch_count, const int* segment_sizes, int segment_count);
template <typename R>
void gemmVbatched2s(const DdlaHandle_t& handle, char transA_0, char transB_0, int* m_0, int* n_0, int* k_0,
                    std::complex<R> alpha_0, const R* const* A_0, int* lda_0, const std::complex<R>* const* B_0,
                    int* ldb_0, std::complex<R> beta_0, std::complex<R>** C_0, int* ldc_0, char transA_1,
                    char transB_1, int* m_1, int* n_1, int* k_1, std::complex<R> alpha_1,
                    const std::complex<R>* const* AB_1, int* lda_1, int* ldb_1, std::complex<R> beta_1,
                    std::complex<R>** C_1, int* ldc_1, bool C0_left, int batch_count, const int* segment_sizes,
                    int segment_count);
template <typename R>
void gemmVbatched2s(const DdlaHandle_t& handle, char transA_0, char transB_0, int* m_0, int* n_0, int* k_0,
                    std::complex<R> alpha_0, const std::complex<R>* const* A_0, int* lda_0, const R* const* B_0,
                    int* ldb_0, std::complex<R> beta_0, std::complex<R>** C_0, int* ldc_0, char transA_1,
                    char transB_1, int* m_1, int* n_1, int* k_1, std::complex<R> alpha_1,
                    const std::complex<R>* const* AB_1, int* lda_1, int* ldb_1, std::complex<R> beta_1,
                    std::complex<R>** C_1, int* ldc_1, bool C0_left, int batch_count, const int* segment_sizes,
                    int segment_count);

// Mixed-precision pposv (no counterpart in the reference; LAPACK dsposv
// semantics): factor in FP32 on the matrix cores -- trailing updates in FP32
// (factor = 0) or as split bfloat16 products accumulated in FP32, bf16x3
// (factor = 1, ~2^-16 relative) or bf16x6 (factor = 2, ~FP32) -- refine in FP64
// until every column has ||r||_inf <= tol ||A||_inf ||x||_inf, at most
// max_iter corrections; otherwise (or if the FP32 factor fails) run pposv.
// iter = corrections applied, -3 (FP32 factor failed) or -(max_iter + 1)
// (no convergence, fallback ran).  Side 'L', trans 'N'.
template <typename T>
void pposv_mixed(const DdlaHandle_t& handle, char uplo, int n, int nrhs, T* A, int ia, int ja, const int* descA,
                 T* B, int ib, int jb, const int* descB, int factor, double tol, int max_iter, int& iter, int& info);

// Batched SPD solves of uneven order (no counterpart in the reference): for
// every i < batch_count, independently, factor A_i (order d_n[i] <= max_n <=
// 1024, triangle uplo, leading dimension d_lda[i]) in place, solve A_i X_i =
// B_i (d_n[i] x nrhs, d_ldb[i]) in place when the factorization succeeds, and
// write d_logdet[i] = log det A_i = 2 sum log L_ii.  d_info[i] = 0, the
// 1-based failed pivot (B_i untouched, logdet NaN), or -1 for an invalid
// order or leading dimension.  All arrays are device resident; the call is
// asynchronous on the handle's stream (fused potrf + potrs + logdet in one
// persistent launch).
template <typename T>
void pposvVbatched(const DdlaHandle_t& handle, char uplo, const int* d_n, T* const* d_A, const int* d_lda, int nrhs,
                   T* const* d_B, const int* d_ldb, int* d_info, double* d_logdet, int batch_count, int max_n);
// Times the batched schedules on copies of these inputs (left unchanged) and
// keeps the fastest for later calls on this handle; synchronizes.
template <typename T>
void pposvVbatchedTune(const DdlaHandle_t& handle, char uplo, const int* d_n, T* const* d_A, const int* d_lda,
                       int nrhs, T* const* d_B, const int* d_ldb, int batch_count, int max_n, int* occupancy,
                       bool* ordered);
// Captures one batched call (tuning first if the handle has not been tuned)
// into an instantiated hipGraph that owns its workspace; replay it with
// batchedGraphLaunch on the handle's stream, free it with batchedGraphDestroy.
template <typename T>
void* pposvVbatchedGraph(const DdlaHandle_t& handle, char uplo, const int* d_n, T* const* d_A, const int* d_lda,
                         int nrhs, T* const* d_B, const int* d_ldb, int* d_info, double* d_logdet, int batch_count,
                         int max_n);
void batchedGraphLaunch(void* graph, const DdlaHandle_t& handle);
void batchedGraphDestroy(void* graph);

// Reverse-mode adjoints (no counterpart in the reference), for the factor
// of A = L L^T held in the lower triangle (uplo 'L'); synchronized.
//   potrf_backward:  Abar = sym(L^-T Phi(L^T Lbar) L^-1) (symmetric, full)
//   potrs_backward:  X = A^-1 B: Bbar = A^-1 Xbar, Lbar -= tril(Bbar (L^T X)^T + X (L^T Bbar)^T)
//   logdet_backward: logdet A = 2 sum log L_ii: Lbar_ii += 2 g / L_ii
template <typename T>
void potrf_backward(const DdlaHandle_t& handle, char uplo, int n, const T* L, int ldl, const T* Lbar, int ldlb,
                    T* Abar, int ldab);
template <typename T>
void potrs_backward(const DdlaHandle_t& handle, char uplo, int n, int nrhs, const T* L, int ldl, const T* X, int ldx,
                    const T* Xbar, int ldxb, T* Bbar, int ldbb, T* Lbar, int ldlb);
template <typename T>
void logdet_backward(const DdlaHandle_t& handle, int n, const T* L, int ldl, double g, T* Lbar, int ldlb);
// A^-1 (full symmetric, n x n at Ainv) from the factor of A = L L^T in the
// lower triangle (LAPACK ?potri, which returns one triangle); synchronized.
template <typename T>
void potri(const DdlaHandle_t& handle, char uplo, int n, const T* L, int ldl, T* Ainv, int lda);

// Batched Gaussian-process negative log-likelihood (no counterpart in the
// reference): for `batch` independent problems of order n (FP64, device
// memory), loss_b = y_b^T K_b^-1 y_b / 2 + log det K_b / 2 from the lower
// triangle of K_b (column-major, leading dimension n, problem b at
// K + b * k_stride), y_b at y + b * y_stride.  Every launch serves all
// problems at once.  work holds batch * gpBatchedWorkspaceSize(n) bytes
// (256-byte aligned) and carries the factor, L^-1 and L^-1 y from the
// forward call to the backward call.  The forward call writes loss[b] and
// info[b] (device; info as for potrf, the loss is meaningless where it is
// nonzero) and returns synchronized; the backward call writes, for the
// upstream gradients g[b] (device), Kbar_b = g_b (K_b^-1 - a a^T) / 2 (full
// symmetric, a = K_b^-1 y_b, at Kbar + b * kbar_stride) and, when ybar is not
// null, ybar_b = g_b a, asynchronously on the handle's stream.
std::size_t gpBatchedWorkspaceSize(int n);
void gpBatchedForward(const DdlaHandle_t& handle, int batch, int n, const double* K, long long k_stride,
                      const double* y, long long y_stride, void* work, double* loss, int* info);
void gpBatchedBackward(const DdlaHandle_t& handle, int batch, int n, void* work, const double* g, double* Kbar,
                       long long kbar_stride, double* ybar, long long ybar_stride);

namespace cholesky_impl {
// Three components per routine: <r>_check (validation), <r>_workspace_size
// (backend and host bytes; no allocation) and <r>_run on caller scratch (no
// validation, no allocation).
void ppotrf_check(const DdlaHandle_t& h, char uplo, int n, const void* A, int ia, int ja, const int* descA);
template <class T>
void ppotrf_workspace_size(const DdlaHandle_t& h, char uplo, int n, T* A, const int* descA,
                           std::size_t* backend_bytes, std::size_t* host_bytes);
template <class T>
void ppotrf_run(const DdlaHandle_t& h, char uplo, int n, T* A, int ia, int ja, const int* descA, int& info,
                void* backend_work, std::size_t backend_bytes, void* host_work, std::size_t host_bytes);

void ppotrs_check(const DdlaHandle_t& h, char side, char uplo, char trans, int n, int nrhs, const void* A,
                  const int* descA, const void* B, const int* descB);
template <class T>
void ppotrs_workspace_size(const DdlaHandle_t& h, char side, int n, int nrhs, const int* descA, const int* descB,
                           std::size_t* backend_bytes, std::size_t* host_bytes);
template <class T>
void ppotrs_run(const DdlaHandle_t& h, char side, char uplo, char trans, int n, int nrhs, const T* A,
                const int* descA, T* B, const int* descB, void* backend_work, std::size_t backend_bytes,
                void* host_work, std::size_t host_bytes);

void pposv_check(const DdlaHandle_t& h, char side, char uplo, char trans, int n, int nrhs, const void* A, int ia,
                 int ja, const int* descA, const void* B, int ib, int jb, const int* descB);
template <class T>
void pposv_workspace_size(const DdlaHandle_t& h, char side, char uplo, int n, int nrhs, T* A, const int* descA,
                          const int* descB, std::size_t* backend_bytes, std::size_t* host_bytes);
template <class T>
void pposv_run(const DdlaHandle_t& h, char side, char uplo, char trans, int n, int nrhs, T* A, int ia, int ja,
               const int* descA, T* B, int ib, int jb, const int* descB, int& info, void* backend_work,
               std::size_t backend_bytes, void* host_work, std::size_t host_bytes);
void pposv_mixed_check(const DdlaHandle_t& h, char uplo, int n, int nrhs, const void* A, int ia, int ja,
                       const int* descA, const void* B, int ib, int jb, const int* descB);
template <class T>
void pposv_mixed_workspace_size(const DdlaHandle_t& h, char uplo, int n, int nrhs, const int* descA,
                                const int* descB, std::size_t* backend_bytes, std::size_t* host_bytes);
template <class T>
void pposv_mixed_run(const DdlaHandle_t& h, char uplo, int n, int nrhs, T* A, const int* descA, T* B,
                     const int* descB, int factor, double tol, int max_iter, int& iter, int& info,
                     void* backend_work, std::size_t backend_bytes);
void pposv_vbatched_check(const DdlaHandle_t& h, char uplo, int batch, int nrhs, int max_n, const void* d_n,
                          const void* d_A, const void* d_lda, const void* d_B, const void* d_ldb, const void* d_info,
                          const void* d_logdet);
template <class T>
void pposv_vbatched_workspace_size(const DdlaHandle_t& h, int batch, int max_n, std::size_t* backend_bytes);
// No allocation and no synchronization: may be captured into a graph.
template <class T>
void pposv_vbatched_run(const DdlaHandle_t& h, char uplo, const int* d_n, T* const* d_A, const int* d_lda, int nrhs,
                        T* const* d_B, const int* d_ldb, int* d_info, double* d_logdet, int batch, int max_n,
                        void* backend_work, std::size_t backend_bytes);
void ppotrf_bottom_right_check(const DdlaHandle_t& h, char uplo, int n, const void* A, const int* descA);
template <class T>
void ppotrf_bottom_right_workspace_size(const DdlaHandle_t& h, char uplo, int n, T* A, const int* descA,
                                        std::size_t* backend_bytes, std::size_t* host_bytes);
template <class T>
void ppotrf_bottom_right_run(const DdlaHandle_t& h, char uplo, int n, T* A, const int* descA, int& info,
                             void* backend_work, std::size_t backend_bytes, void* host_work, std::size_t host_bytes);
} // namespace cholesky_impl

} // namespace ddla

#include "ddla_vector.h"

#endif
