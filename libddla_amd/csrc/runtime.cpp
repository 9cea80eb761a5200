// Handle lifecycle, memory helpers, descriptor and the launch profiler.
// Follows the reference's runtime for a single GPU and a 1 x 1 grid:
// src/ddla_handle_t.cpp:94-200 (init / set / destroy), :322-770 (memory),
// src/ddla_desc.cpp:9-60 (check_desc), include/ddla/ddla_desc.h:102-135.
#include "internal.h"
#include "kernels.h"

#include <cstring>
#include <limits>

namespace ddla {

namespace {
thread_local std::string g_status_text;

bool is_device_pointer(const void* p)
{
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

__global__ void fill_int_kernel(int* p, int value) { *p = value; }
} // namespace

void fill_int(hipStream_t s, int* p, int value)
{
    fill_int_kernel<<<1, 1, 0, s>>>(p, value);
    DDLA_HIP(hipGetLastError());
}

void require_gpu(const DdlaHandle_t& h, const char* who)
{
    if (h == nullptr) throw std::runtime_error(std::string(who) + ": null handle");
    if (!h->initialized) throw std::runtime_error(std::string(who) + ": handle has no process grid (call ddlaSet)");
    if (h->backend != DdlaBackend::GPU) throw std::runtime_error(std::string(who) + ": GPU handle required");
}

void* DdlaStream::ensure_scratch(std::size_t bytes)
{
    if (bytes <= scratch_bytes) return scratch;
    if (scratch != nullptr) {
        DDLA_HIP(hipStreamSynchronize(stream));
        DDLA_HIP(hipFree(scratch));
        scratch = nullptr;
        scratch_bytes = 0;
    }
    DDLA_HIP(hipMalloc(&scratch, bytes));
    scratch_bytes = bytes;
    return scratch;
}

hipEvent_t DdlaStream::next_event()
{
    if (events_used == event_pool.size()) {
        hipEvent_t e;
        DDLA_HIP(hipEventCreate(&e));
        event_pool.push_back(e);
    }
    return event_pool[events_used++];
}

void DdlaStream::release_events()
{
    records.clear();
    events_used = 0;
}

LaunchTimer::LaunchTimer(DdlaStream* h, int cls, double flops, double bytes, hipStream_t s)
    : h_(h), cls_(cls), flops_(flops), bytes_(bytes), s_(s != nullptr ? s : (h != nullptr ? h->stream : nullptr))
{
    if (h_ != nullptr && h_->profiling) {
        start_ = h_->next_event();
        DDLA_HIP(hipEventRecord(start_, s_));
    }
}

LaunchTimer::~LaunchTimer()
{
    if (start_ == nullptr) return;
    hipEvent_t stop = h_->next_event();
    if (hipEventRecord(stop, s_) == hipSuccess)
        h_->records.push_back({cls_, start_, stop, flops_, bytes_});
}

bool ddlaBackendAvailable(DdlaBackend backend) { return backend == DdlaBackend::GPU; }

ddlaStatus_t ddlaInit(DdlaHandle_t& handle, DdlaBackend requested_backend)
{
    if (!ddlaBackendAvailable(requested_backend))
        throw std::runtime_error("ddlaInit: requested backend is not available in this build (GPU only)");
    handle = new DdlaStream;
    handle->backend = requested_backend;
    return ddlaStatus_t::DDLA_STATUS_SUCCESS;
}

ddlaStatus_t ddlaSet(DdlaHandle_t handle, int nprows, int npcols, char major)
{
    if (handle == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE;
    if (nprows <= 0 || npcols <= 0) nprows = npcols = 1;
    if (nprows * npcols != 1)
        throw std::invalid_argument("ddlaSet: this build runs one process on one GPU; the grid must be 1 x 1");
    if (major != 'R' && major != 'C' && major != 'r' && major != 'c')
        return ddlaStatus_t::DDLA_STATUS_INVALID_VALUE;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return ddlaStatus_t::DDLA_STATUS_BACKEND_UNAVAILABLE;
    }
    if (!handle->initialized) {
        DDLA_HIP(hipGetDevice(&handle->device));
        DDLA_HIP(hipDeviceGetAttribute(&handle->num_cus, hipDeviceAttributeMultiprocessorCount, handle->device));
        DDLA_HIP(hipStreamCreate(&handle->stream));
        handle->owns_stream = true;
        DDLA_HIP(hipStreamCreateWithFlags(&handle->side, hipStreamNonBlocking));
        DDLA_HIP(hipStreamCreateWithFlags(&handle->look, hipStreamNonBlocking));
        DDLA_HIP(hipEventCreateWithFlags(&handle->ev_look, hipEventDisableTiming));
        DDLA_HIP(hipEventCreateWithFlags(&handle->ev_panel, hipEventDisableTiming));
        DDLA_HIP(hipEventCreateWithFlags(&handle->ev_rest, hipEventDisableTiming));
    }
    handle->nprows_ = handle->npcols_ = handle->nprocs = 1;
    handle->myprow_ = handle->mypcol_ = 0;
    handle->major = (major == 'c' || major == 'C') ? 'C' : 'R';
    handle->initialized = true;
    return ddlaStatus_t::DDLA_STATUS_SUCCESS;
}

ddlaStatus_t ddlaDestroy(DdlaHandle_t& handle)
{
    if (handle == nullptr) return ddlaStatus_t::DDLA_STATUS_SUCCESS;
    if (handle->initialized) {
        (void)hipStreamSynchronize(handle->stream);
        if (handle->scratch != nullptr) (void)hipFree(handle->scratch);
        for (hipEvent_t e : handle->event_pool) (void)hipEventDestroy(e);
        for (hipStream_t st : {handle->side, handle->look}) {
            if (st == nullptr) continue;
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
        }
        for (hipEvent_t e : {handle->ev_look, handle->ev_panel, handle->ev_rest})
            if (e != nullptr) (void)hipEventDestroy(e);
        if (handle->owns_stream && handle->stream != nullptr) (void)hipStreamDestroy(handle->stream);
    }
    delete handle;
    handle = nullptr;
    return ddlaStatus_t::DDLA_STATUS_SUCCESS;
}

DdlaBackend ddlaGetBackend(const DdlaHandle_t& handle)
{
    if (handle == nullptr) throw std::runtime_error("ddlaGetBackend: null handle");
    return handle->backend;
}

static ddlaStatus_t internal_error(hipError_t e)
{
    g_status_text = hipGetErrorString(e);
    (void)hipGetLastError();
    return ddlaStatus_t::DDLA_STATUS_INTERNAL_ERROR;
}

ddlaStatus_t ddlaMalloc(const DdlaHandle_t& handle, void** ptr, std::size_t bytes)
{
    if (handle == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE;
    if (ptr == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_VALUE;
    if (bytes == 0) {
        *ptr = nullptr;
        return ddlaStatus_t::DDLA_STATUS_SUCCESS;
    }
    const hipError_t e = hipMalloc(ptr, bytes);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        return ddlaStatus_t::DDLA_STATUS_ALLOC_FAILED;
    }
    if (e != hipSuccess) return internal_error(e);
    return ddlaStatus_t::DDLA_STATUS_SUCCESS;
}

ddlaStatus_t ddlaFree(const DdlaHandle_t& handle, void* ptr)
{
    if (handle == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE;
    if (ptr == nullptr) return ddlaStatus_t::DDLA_STATUS_SUCCESS;
    if (handle->stream != nullptr) (void)hipStreamSynchronize(handle->stream);
    const hipError_t e = hipFree(ptr);
    if (e != hipSuccess) return internal_error(e);
    return ddlaStatus_t::DDLA_STATUS_SUCCESS;
}

ddlaStatus_t ddlaMemcpy(const DdlaHandle_t& handle, void* dst, const void* src, std::size_t bytes)
{
    if (handle == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE;
    if (bytes == 0) return ddlaStatus_t::DDLA_STATUS_SUCCESS;
    if (dst == nullptr || src == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_VALUE;
    const bool dd = is_device_pointer(dst), sd = is_device_pointer(src);
    if (!dd && !sd) {
        std::memcpy(dst, src, bytes);
        return ddlaStatus_t::DDLA_STATUS_SUCCESS;
    }
    const hipMemcpyKind kind = sd ? (dd ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost) : hipMemcpyHostToDevice;
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, handle->stream);
    if (e != hipSuccess) return internal_error(e);
    return ddlaStatus_t::DDLA_STATUS_SUCCESS;
}

ddlaStatus_t ddlaMemcpy2D(const DdlaHandle_t& handle, void* dst, std::size_t dpitch, const void* src,
                          std::size_t spitch, std::size_t width, std::size_t height)
{
    if (handle == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE;
    if (width == 0 || height == 0) return ddlaStatus_t::DDLA_STATUS_SUCCESS;
    if (dst == nullptr || src == nullptr || dpitch < width || spitch < width)
        return ddlaStatus_t::DDLA_STATUS_INVALID_VALUE;
    const bool dd = is_device_pointer(dst), sd = is_device_pointer(src);
    if (!dd && !sd) {
        for (std::size_t r = 0; r < height; ++r)
            std::memcpy(static_cast<char*>(dst) + r * dpitch, static_cast<const char*>(src) + r * spitch, width);
        return ddlaStatus_t::DDLA_STATUS_SUCCESS;
    }
    const hipMemcpyKind kind = sd ? (dd ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost) : hipMemcpyHostToDevice;
    const hipError_t e = hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, kind, handle->stream);
    if (e != hipSuccess) return internal_error(e);
    return ddlaStatus_t::DDLA_STATUS_SUCCESS;
}

namespace {
bool valid_communication(DdlaCommunication mode)
{
    return mode == DdlaCommunication::MPI || mode == DdlaCommunication::HostStagedMPI ||
           mode == DdlaCommunication::CCL;
}
} // namespace

ddlaStatus_t ddlaSetCommunication(const DdlaHandle_t& handle, DdlaCommunication mode)
{
    if (handle == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE;
    if (!valid_communication(mode)) return ddlaStatus_t::DDLA_STATUS_INVALID_VALUE;
    handle->communication = mode;
    return ddlaStatus_t::DDLA_STATUS_SUCCESS;
}

DdlaCommunication ddlaGetCommunication(const DdlaHandle_t& handle)
{
    if (handle == nullptr) throw std::runtime_error("ddlaGetCommunication: handle is null");
    return handle->communication;
}

bool ddlaCommunicationAvailable(const DdlaHandle_t& handle, DdlaCommunication mode)
{
    return handle != nullptr && valid_communication(mode);
}

ddlaStatus_t ddlaMemset(const DdlaHandle_t& handle, void* ptr, int value, std::size_t bytes)
{
    if (handle == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE;
    if (bytes == 0) return ddlaStatus_t::DDLA_STATUS_SUCCESS;
    if (ptr == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_VALUE;
    const hipError_t e = hipMemsetAsync(ptr, value, bytes, handle->stream);
    if (e != hipSuccess) return internal_error(e);
    return ddlaStatus_t::DDLA_STATUS_SUCCESS;
}

ddlaStatus_t ddlaSynchronize(const DdlaHandle_t& handle)
{
    if (handle == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE;
    const hipError_t e = hipStreamSynchronize(handle->stream);
    if (e != hipSuccess) return internal_error(e);
    return ddlaStatus_t::DDLA_STATUS_SUCCESS;
}

const char* ddlaGetErrorString(ddlaStatus_t status)
{
    const char* name = "DDLA_STATUS_UNKNOWN";
    switch (status) {
    case ddlaStatus_t::DDLA_STATUS_SUCCESS: name = "DDLA_STATUS_SUCCESS"; break;
    case ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE: name = "DDLA_STATUS_INVALID_HANDLE"; break;
    case ddlaStatus_t::DDLA_STATUS_INVALID_VALUE: name = "DDLA_STATUS_INVALID_VALUE"; break;
    case ddlaStatus_t::DDLA_STATUS_ALLOC_FAILED: name = "DDLA_STATUS_ALLOC_FAILED"; break;
    case ddlaStatus_t::DDLA_STATUS_BACKEND_UNAVAILABLE: name = "DDLA_STATUS_BACKEND_UNAVAILABLE"; break;
    case ddlaStatus_t::DDLA_STATUS_INTERNAL_ERROR: name = "DDLA_STATUS_INTERNAL_ERROR"; break;
    }
    thread_local std::string text;
    text = name;
    if (status == ddlaStatus_t::DDLA_STATUS_INTERNAL_ERROR && !g_status_text.empty()) text += ": " + g_status_text;
    return text.c_str();
}

ddlaStatus_t ddlaSetStream(const DdlaHandle_t& handle, void* stream)
{
    if (handle == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE;
    if (!handle->initialized) return ddlaStatus_t::DDLA_STATUS_INVALID_VALUE;
    hipStream_t next = static_cast<hipStream_t>(stream);
    if (next == handle->stream) return ddlaStatus_t::DDLA_STATUS_SUCCESS;
    // The look-ahead streams are the handle's own: taking one as the compute
    // stream would destroy it twice and serialize the look-ahead.
    if (next != nullptr && (next == handle->side || next == handle->look))
        return ddlaStatus_t::DDLA_STATUS_INVALID_VALUE;
    DDLA_HIP(hipStreamSynchronize(handle->stream));
    if (handle->owns_stream && handle->stream != nullptr) DDLA_HIP(hipStreamDestroy(handle->stream));
    handle->stream = next;
    handle->owns_stream = next != nullptr;
    return ddlaStatus_t::DDLA_STATUS_SUCCESS;
}

void* ddlaGetStream(const DdlaHandle_t& handle)
{
    if (handle == nullptr) throw std::runtime_error("ddlaGetStream: null handle");
    return static_cast<void*>(handle->stream);
}

int ddlaGetRank(const DdlaHandle_t& handle) { return handle == nullptr ? -1 : 0; }
int ddlaGetSize(const DdlaHandle_t& handle) { return handle == nullptr ? -1 : (handle->initialized ? 1 : 0); }

void ddlaGetGridCoords(const DdlaHandle_t& handle, int& myprow, int& mypcol)
{
    if (handle == nullptr) throw std::runtime_error("ddlaGetGridCoords: null handle");
    myprow = handle->myprow_;
    mypcol = handle->mypcol_;
}

void ddlaGetGridDims(const DdlaHandle_t& handle, int& nprows, int& npcols)
{
    if (handle == nullptr) throw std::runtime_error("ddlaGetGridDims: null handle");
    nprows = handle->nprows_;
    npcols = handle->npcols_;
}

ddlaStatus_t ddlaMallocHost(const DdlaHandle_t& handle, void** ptr, std::size_t bytes)
{
    if (handle == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE;
    if (ptr == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_VALUE;
    *ptr = nullptr;
    if (bytes == 0) return ddlaStatus_t::DDLA_STATUS_SUCCESS;
    if (hipHostMalloc(ptr, bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        *ptr = nullptr;
        return ddlaStatus_t::DDLA_STATUS_ALLOC_FAILED;
    }
    return ddlaStatus_t::DDLA_STATUS_SUCCESS;
}

ddlaStatus_t ddlaFreeHost(const DdlaHandle_t& handle, void* ptr)
{
    if (handle == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE;
    if (ptr == nullptr) return ddlaStatus_t::DDLA_STATUS_SUCCESS;
    const hipError_t e = hipHostFree(ptr);
    if (e != hipSuccess) return internal_error(e);
    return ddlaStatus_t::DDLA_STATUS_SUCCESS;
}

namespace {
// Validation shared by the triangular helpers; *empty when nothing is
// selected (then no pointer or leading dimension is looked at).
ddlaStatus_t check_tri(char uplo, int m, int n, std::size_t elem, bool diag, bool& upper, bool& empty)
{
    const char u = (uplo >= 'a' && uplo <= 'z') ? char(uplo - 'a' + 'A') : uplo;
    if ((u != 'U' && u != 'L') || m < 0 || n < 0 || elem == 0) return ddlaStatus_t::DDLA_STATUS_INVALID_VALUE;
    upper = u == 'U';
    // strict upper needs a second column, strict lower a second row
    empty = m == 0 || n == 0 || (!diag && (upper ? n < 2 : m < 2));
    return ddlaStatus_t::DDLA_STATUS_SUCCESS;
}

bool tri_span_ok(int m, int n, int ld, std::size_t elem)
{
    if (ld < std::max(1, m)) return false;
    const std::size_t limit = std::numeric_limits<std::size_t>::max() / elem;
    return (std::size_t)ld <= limit / (std::size_t)n;  // ld * n elements fit in size_t bytes
}

// rows [lo, hi) of column j selected by the triangle
void tri_rows(bool upper, bool diag, int m, int j, int& lo, int& hi)
{
    if (upper) {
        lo = 0;
        hi = std::min(m, diag ? j + 1 : j);
    } else {
        lo = std::min(m, diag ? j : j + 1);
        hi = m;
    }
}
} // namespace

namespace detail {

ddlaStatus_t memset_tri(const DdlaHandle_t& handle, char uplo, int m, int n, void* ptr, int lda, int value,
                        std::size_t elem, bool include_diagonal)
{
    if (handle == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE;
    bool upper = false, empty = false;
    const ddlaStatus_t st = check_tri(uplo, m, n, elem, include_diagonal, upper, empty);
    if (st != ddlaStatus_t::DDLA_STATUS_SUCCESS || empty) return st;
    if (ptr == nullptr || !tri_span_ok(m, n, lda, elem) || !is_device_pointer(ptr))
        return ddlaStatus_t::DDLA_STATUS_INVALID_VALUE;
    try {
        kern::memset_tri_bytes(handle->stream, upper, include_diagonal, m, n, ptr, lda, elem,
                               (unsigned char)(value & 0xff));
    } catch (const std::exception&) {
        return ddlaStatus_t::DDLA_STATUS_INTERNAL_ERROR;
    }
    return ddlaStatus_t::DDLA_STATUS_SUCCESS;
}

ddlaStatus_t memcpy_tri(const DdlaHandle_t& handle, char uplo, int m, int n, void* dst, int ldd, const void* src,
                        int lds, std::size_t elem, bool include_diagonal)
{
    if (handle == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE;
    bool upper = false, empty = false;
    const ddlaStatus_t st = check_tri(uplo, m, n, elem, include_diagonal, upper, empty);
    if (st != ddlaStatus_t::DDLA_STATUS_SUCCESS || empty) return st;
    if (dst == nullptr || src == nullptr || !tri_span_ok(m, n, ldd, elem) || !tri_span_ok(m, n, lds, elem))
        return ddlaStatus_t::DDLA_STATUS_INVALID_VALUE;
    if (dst == src && ldd == lds) return ddlaStatus_t::DDLA_STATUS_SUCCESS;  // self-copy
    const bool dd = is_device_pointer(dst), sd = is_device_pointer(src);
    auto* d = static_cast<unsigned char*>(dst);
    const auto* s = static_cast<const unsigned char*>(src);
    if (dd && sd) {
        try {
            kern::memcpy_tri_bytes(handle->stream, upper, include_diagonal, m, n, dst, ldd, src, lds, elem);
        } catch (const std::exception&) {
            return ddlaStatus_t::DDLA_STATUS_INTERNAL_ERROR;
        }
        return ddlaStatus_t::DDLA_STATUS_SUCCESS;
    }
    const hipMemcpyKind kind = sd ? hipMemcpyDeviceToHost : hipMemcpyHostToDevice;
    for (int j = 0; j < n; ++j) {
        int lo = 0, hi = 0;
        tri_rows(upper, include_diagonal, m, j, lo, hi);
        if (hi <= lo) continue;
        unsigned char* to = d + ((std::size_t)j * ldd + lo) * elem;
        const unsigned char* from = s + ((std::size_t)j * lds + lo) * elem;
        const std::size_t bytes = (std::size_t)(hi - lo) * elem;
        if (!dd && !sd) {
            std::memcpy(to, from, bytes);
        } else {
            const hipError_t e = hipMemcpyAsync(to, from, bytes, kind, handle->stream);
            if (e != hipSuccess) return internal_error(e);
        }
    }
    return ddlaStatus_t::DDLA_STATUS_SUCCESS;
}

} // namespace detail

ddlaStatus_t ddlaDeviceSynchronize(const DdlaHandle_t& handle)
{
    if (handle == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE;
    hipError_t e = hipSetDevice(handle->device);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) return internal_error(e);
    return ddlaStatus_t::DDLA_STATUS_SUCCESS;
}

ddlaStatus_t ddlaEventCreate(const DdlaHandle_t& handle, ddlaEvent_t* event)
{
    if (handle == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE;
    if (event == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_VALUE;
    hipEvent_t ev = nullptr;
    const hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e != hipSuccess) return internal_error(e);
    *event = ev;
    return ddlaStatus_t::DDLA_STATUS_SUCCESS;
}

ddlaStatus_t ddlaEventDestroy(const DdlaHandle_t& handle, ddlaEvent_t event)
{
    if (handle == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE;
    if (event == nullptr) return ddlaStatus_t::DDLA_STATUS_SUCCESS;
    const hipError_t e = hipEventDestroy(static_cast<hipEvent_t>(event));
    if (e != hipSuccess) return internal_error(e);
    return ddlaStatus_t::DDLA_STATUS_SUCCESS;
}

ddlaStatus_t ddlaEventRecord(const DdlaHandle_t& handle, ddlaEvent_t event)
{
    if (handle == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE;
    if (event == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_VALUE;
    const hipError_t e = hipEventRecord(static_cast<hipEvent_t>(event), handle->stream);
    if (e != hipSuccess) return internal_error(e);
    return ddlaStatus_t::DDLA_STATUS_SUCCESS;
}

ddlaStatus_t ddlaEventWait(const DdlaHandle_t& handle, ddlaEvent_t event)
{
    if (handle == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE;
    if (event == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_VALUE;
    const hipError_t e = hipStreamWaitEvent(handle->stream, static_cast<hipEvent_t>(event), 0);
    if (e != hipSuccess) return internal_error(e);
    return ddlaStatus_t::DDLA_STATUS_SUCCESS;
}

ddlaStatus_t ddlaEventSynchronize(const DdlaHandle_t& handle, ddlaEvent_t event)
{
    if (handle == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE;
    if (event == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_VALUE;
    const hipError_t e = hipEventSynchronize(static_cast<hipEvent_t>(event));
    if (e != hipSuccess) return internal_error(e);
    return ddlaStatus_t::DDLA_STATUS_SUCCESS;
}

ddlaStatus_t ddlaGetLastError(const DdlaHandle_t& handle)
{
    if (handle == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE;
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return internal_error(e);
    return ddlaStatus_t::DDLA_STATUS_SUCCESS;
}

ddlaStatus_t ddlaSetDevice(const DdlaHandle_t& handle)
{
    if (handle == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE;
    const hipError_t e = hipSetDevice(handle->device);
    if (e != hipSuccess) return internal_error(e);
    return ddlaStatus_t::DDLA_STATUS_SUCCESS;
}

ddlaStatus_t ddlaGetDeviceCount(const DdlaHandle_t& handle, int* count)
{
    if (handle == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE;
    if (count == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_VALUE;
    const hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) return internal_error(e);
    return ddlaStatus_t::DDLA_STATUS_SUCCESS;
}

ddlaStatus_t ddlaMemGetInfo(const DdlaHandle_t& handle, std::size_t* free_bytes, std::size_t* total_bytes)
{
    if (handle == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE;
    if (free_bytes == nullptr || total_bytes == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_VALUE;
    hipError_t e = hipSetDevice(handle->device);
    if (e == hipSuccess) e = hipMemGetInfo(free_bytes, total_bytes);
    if (e != hipSuccess) return internal_error(e);
    return ddlaStatus_t::DDLA_STATUS_SUCCESS;
}

// src/ddla_stream.h:114-125 of the reference; -1 for a null handle or one
// without a grid.
void ddlaRankToRc(const DdlaHandle_t& handle, int rank, int& row, int& col)
{
    row = col = -1;
    if (handle == nullptr || handle->nprows_ <= 0 || handle->npcols_ <= 0) return;
    if (handle->major == 'R') {
        row = rank / handle->npcols_;
        col = rank % handle->npcols_;
    } else {
        row = rank % handle->nprows_;
        col = rank / handle->nprows_;
    }
}

int ddlaRcToRank(const DdlaHandle_t& handle, int row, int col)
{
    if (handle == nullptr || handle->nprows_ <= 0 || handle->npcols_ <= 0) return -1;
    return handle->major == 'R' ? row * handle->npcols_ + col : col * handle->nprows_ + row;
}

ddlaStatus_t ddlaDescInit(int* desc, const DdlaHandle_t& handle, int m, int n, int mb, int nb, int irsrc, int icsrc,
                          int lld)
{
    if (desc == nullptr) return ddlaStatus_t::DDLA_STATUS_INVALID_VALUE;
    if (handle == nullptr || !handle->initialized || handle->nprows_ <= 0 || handle->npcols_ <= 0)
        return ddlaStatus_t::DDLA_STATUS_INVALID_HANDLE;
    if (m < 0 || n < 0 || mb <= 0 || nb <= 0) return ddlaStatus_t::DDLA_STATUS_INVALID_VALUE;
    if (irsrc < 0 || irsrc >= handle->nprows_ || icsrc < 0 || icsrc >= handle->npcols_)
        return ddlaStatus_t::DDLA_STATUS_INVALID_VALUE;
    const int lld_min = std::max(1, num_loc(m, mb, handle->myprow_, irsrc, handle->nprows_));
    desc[DDLA_DTYPE_] = DDLA_BLOCK_CYCLIC_2D;
    desc[DDLA_CTXT_] = 0;
    desc[DDLA_M_] = m;
    desc[DDLA_N_] = n;
    desc[DDLA_MB_] = mb;
    desc[DDLA_NB_] = nb;
    desc[DDLA_RSRC_] = irsrc;
    desc[DDLA_CSRC_] = icsrc;
    desc[DDLA_LLD_] = std::max(lld, lld_min);
    return ddlaStatus_t::DDLA_STATUS_SUCCESS;
}

void check_desc(const int* desc, const DdlaHandle_t& handle)
{
    if (desc == nullptr) throw std::invalid_argument("ddla: null ScaLAPACK descriptor");
    if (handle == nullptr) throw std::invalid_argument("ddla: null handle for ScaLAPACK descriptor");
    if (desc[DDLA_DTYPE_] != DDLA_BLOCK_CYCLIC_2D)
        throw std::invalid_argument("ddla: unsupported descriptor type " + std::to_string(desc[DDLA_DTYPE_]) +
                                    "; only dense block-cyclic (DTYPE_ == 1) is supported");
    if (!handle->initialized || handle->nprows_ <= 0 || handle->npcols_ <= 0)
        throw std::invalid_argument("ddla: handle carries no process grid");
    if (desc[DDLA_M_] < 0 || desc[DDLA_N_] < 0)
        throw std::invalid_argument("ddla: descriptor M/N must be non-negative, got M=" +
                                    std::to_string(desc[DDLA_M_]) + " N=" + std::to_string(desc[DDLA_N_]));
    if (desc[DDLA_MB_] <= 0 || desc[DDLA_NB_] <= 0)
        throw std::invalid_argument("ddla: descriptor MB/NB must be positive, got MB=" +
                                    std::to_string(desc[DDLA_MB_]) + " NB=" + std::to_string(desc[DDLA_NB_]));
    if (desc[DDLA_RSRC_] < 0 || desc[DDLA_RSRC_] >= handle->nprows_ || desc[DDLA_CSRC_] < 0 ||
        desc[DDLA_CSRC_] >= handle->npcols_)
        throw std::invalid_argument("ddla: descriptor RSRC/CSRC outside the process grid");
    const int required =
        std::max(1, num_loc(desc[DDLA_M_], desc[DDLA_MB_], handle->myprow_, desc[DDLA_RSRC_], handle->nprows_));
    if (desc[DDLA_LLD_] < required)
        throw std::invalid_argument("ddla: ScaLAPACK descriptor LLD_A=" + std::to_string(desc[DDLA_LLD_]) +
                                    " is smaller than the required max(1, LOCr(M_A))=" + std::to_string(required));
}

} // namespace ddla
