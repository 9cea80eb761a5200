// Private runtime state of a handle (the reference's DdlaStream,
// src/ddla_stream.h:18-128, reduced to one GPU and a 1 x 1 grid).
#pragma once
#include <hip/hip_runtime.h>

#include <ddla/ddla.h>

#include <complex>
#include <cstddef>
#include <type_traits>
#include <stdexcept>
#include <string>
#include <vector>

namespace ddla {

#define DDLA_HIP(call)                                                                                         \
    do {                                                                                                       \
        const hipError_t ddla_e_ = (call);                                                                     \
        if (ddla_e_ != hipSuccess)                                                                             \
            throw std::runtime_error(std::string(#call) + " failed: " + hipGetErrorString(ddla_e_));           \
    } while (0)

// Kernel classes timed by the optional per-launch event profiler.
enum KernelClass {
    kClassTrailingUpdate = 0,  // syrk-shaped trailing update of potrf (MFMA)
    kClassPanel = 1,           // persistent panel factorization (diagonal blocks, inverses, rows below)
    kClassSolve = 2,           // potrs triangular-solve sweeps (MFMA)
    kClassTrtri = 3,           // inversion of the diagonal blocks for potrs
    kClassResidual = 4,        // FP64 residual R = B - A X of the refinement (MFMA)
    kClassOther = 5,
    kClassLookahead = 6,       // look-ahead updates of the next panels' columns (MFMA, beside the trailing update)
    kClassCount = 7
};

class DdlaStream {
public:
    DdlaBackend backend = DdlaBackend::GPU;
    bool initialized = false;
    int device = 0;
    hipStream_t stream = nullptr;
    bool owns_stream = false;
    DdlaCommunication communication = DdlaCommunication::CCL;  // recorded only: a 1 x 1 grid exchanges nothing
    // Streams of the look-ahead factorization (non-blocking, default
    // priority): `side` factors panels, `look` updates the next two panels'
    // columns; the three events order them against `stream`.
    hipStream_t side = nullptr, look = nullptr;
    hipEvent_t ev_look = nullptr, ev_panel = nullptr, ev_rest = nullptr;
    int nprows_ = 0, npcols_ = 0, myprow_ = 0, mypcol_ = 0, nprocs = 0;
    char major = 'R';
    int num_cus = 0;  // compute units of the device (set by ddlaSet)

    // Batched-solve schedule chosen by the autotuner, per element size:
    // workgroups per CU (kernel variant and grid) and largest-first order.
    struct BatchedTuning {
        bool tuned = false;
        int occupancy = 1;
        bool ordered = true;
        int split_blocks = 16;  // members of more blocks take the blocked path (16: none)
        bool left_looking = false;  // every member on the left-looking persistent kernel
        int step_blocks = 1;        // 64-column block columns per step of the blocked path (1 or 2)
    };
    BatchedTuning batched_tuning[2];  // [0] float, [1] double

    // Cached device scratch for the public entry points (grown on demand).
    void* scratch = nullptr;
    std::size_t scratch_bytes = 0;
    void* ensure_scratch(std::size_t bytes);

    // Per-launch event profiler (off by default).
    struct Record {
        int cls;
        hipEvent_t start, stop;
        double flops;
        double bytes;
    };
    bool profiling = false;
    std::vector<Record> records;
    std::vector<hipEvent_t> event_pool;
    std::size_t events_used = 0;
    hipEvent_t next_event();
    void release_events();
};

// Brackets one kernel launch with events when the handle profiles.
class LaunchTimer {
public:
    LaunchTimer(DdlaStream* h, int cls, double flops, double bytes = 0, hipStream_t s = nullptr);
    ~LaunchTimer();
    LaunchTimer(const LaunchTimer&) = delete;
    LaunchTimer& operator=(const LaunchTimer&) = delete;

private:
    DdlaStream* h_;
    int cls_;
    double flops_, bytes_;
    hipStream_t s_;
    hipEvent_t start_ = nullptr;
};

void require_gpu(const DdlaHandle_t& h, const char* who);

namespace detail {
// Bytes take<T>(cursor, count) consumes: count elements padded to 256 bytes
// (src/plan_util.h:28-43).
template <typename T>
constexpr std::size_t padded_bytes(std::size_t count)
{
    return (count * sizeof(T) + 255) / 256 * 256;
}
template <typename T>
inline T* take(char*& cursor, std::size_t count)
{
    T* p = reinterpret_cast<T*>(cursor);
    cursor += padded_bytes<T>(count);
    return p;
}
} // namespace detail

// Launch helpers shared by the Cholesky schedules.
void fill_int(hipStream_t s, int* p, int value);

// Panel width (the trailing update's depth) of the blocked factorization,
// process-wide: cols <= 0 restores the built-in rule (768 columns for FP64
// from order 16384, else 512); other values are rounded down to a multiple of
// 64 within [64, 768] (the split-product path caps it at 512).  Returns the
// previous override (0 = none).
int set_panel_width(int cols);

// The stream schedule of the blocked factorization: the host issues these
// steps in order.  op kReset / kFactor (a = panel) / kUpdate (panel a applied
// to columns [b, c), rows from b) launch on `stream` (kMain = the handle's
// stream, kSide, kLook); kRecord / kWait record event a on, or make `stream`
// wait for, one of the three events.
namespace sched {
enum Op { kReset = 0, kFactor = 1, kUpdate = 2, kRecord = 3, kWait = 4 };
enum Stream { kMain = 0, kSide = 1, kLook = 2 };
enum Event { kEvPanel = 0, kEvLook = 1, kEvRest = 2 };
struct Step {
    int op, stream, a, b, c;
};
std::vector<Step> potrf_schedule(int n, int NB);
} // namespace sched
// Writes min(count, capacity) steps of the schedule for order n and panel
// width panel_width (<= 0: the FP64 rule) as (op, stream, a, b, c) ints;
// returns the step count.
int potrf_schedule_steps(int n, int panel_width, int* steps, int capacity);

template <class T> struct is_complex : std::false_type {};
template <class R> struct is_complex<std::complex<R>> : std::true_type {};

// Hermitian path of the complex types (complex.cpp); arguments already checked.
namespace complex_impl {
template <class R>
std::size_t ppotrf_bytes(int n);
template <class R>
std::size_t ppotrs_bytes(char side, int n, int nrhs);
template <class R>
void ppotrf_run(const DdlaHandle_t& h, char uplo, int n, std::complex<R>* A, const int* descA, int& info, void* work,
                std::size_t bytes);
template <class R>
void ppotrs_run(const DdlaHandle_t& h, char side, char uplo, int n, int nrhs, const std::complex<R>* A,
                const int* descA, std::complex<R>* B, const int* descB, void* work, std::size_t bytes);
template <class R>
void potrf(const DdlaHandle_t& h, bool upper, int n, std::complex<R>* A, int lda, int* info);
template <class R>
void diag_logdet(const DdlaHandle_t& h, int n, const std::complex<R>* A, int lda, std::complex<R>& result);
} // namespace complex_impl

// pposv with the 64-row sweeps whatever the order (no staged-solve work, so
// the workspace stays at about 512 bytes per row): its backend bytes and its
// run (ia = ja = ib = jb = 1, arguments already checked; cholesky.cpp).
template <class T>
std::size_t pposv_unstaged_bytes(char side, int n, int nrhs);
template <class T>
void pposv_run_unstaged(const DdlaHandle_t& h, char side, char uplo, int n, int nrhs, T* A, const int* descA, T* B,
                        const int* descB, int& info, void* backend_work, std::size_t backend_bytes);

// FP64 pposv on the library's static device scratch (noalloc.cpp): no
// hipMalloc / hipFree; throws std::invalid_argument when the call needs more
// scratch than the static array holds.  Returns synchronized.
void pposv_noalloc(const DdlaHandle_t& h, char side, char uplo, char trans, int n, int nrhs, double* A, int ia, int ja,
                   const int* descA, double* B, int ib, int jb, const int* descB, int& info);

} // namespace ddla
