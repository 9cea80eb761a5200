// FP64 pposv that allocates nothing: the scratch of the factorization and the
// solves lives in a file-scope device array instead of the handle's scratch
// (which hipMalloc grows on demand).  For callers that cannot allocate device
// memory: code under stream capture, tight serving loops, harnesses that
// forbid hipMalloc.  No counterpart in the reference.
//
// The array, `scratch`, holds kNoAllocScratchBytes per GPU for the life of
// the process.  Host code never uses its name as an address (the name is a
// host-side stand-in): the first call on a device resolves the device address
// with hipGetSymbolAddress, with that device current, and keeps it per
// device.  The kernels clear what must start at zero (the info word, the
// panel counters, the sweep counters and the error words) on every call, so
// the contents left by the previous call do not matter.  Calls that overlap
// on one GPU (two handles, two host threads) share the array and race.
#include "internal.h"
#include "kernels.h"

#include <array>
#include <mutex>
#include <string>

namespace {

constexpr std::size_t kNoAllocScratchBytes = std::size_t(64) << 20;  // 64 MiB
constexpr int kMaxDevices = 64;

static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

std::mutex g_mu;
std::array<void*, kMaxDevices> g_scratch_ptr{};  // device address of `scratch`, per device

// Device address of this device's copy of `scratch` (resolved once, with the
// device current).
void* scratch_address(int device)
{
    if (device < 0 || device >= kMaxDevices) throw std::runtime_error("pposv_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_scratch_ptr[device] != nullptr) return g_scratch_ptr[device];
    int current = 0;
    DDLA_HIP(hipGetDevice(&current));
    if (current != device) DDLA_HIP(hipSetDevice(device));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (current != device) DDLA_HIP(hipSetDevice(current));
    if (e != hipSuccess || ptr == nullptr)
        throw std::runtime_error(std::string("pposv_noalloc: hipGetSymbolAddress(scratch) failed: ") +
                                 hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
    return ptr;
}

} // namespace

namespace ddla {

void pposv_noalloc(const DdlaHandle_t& h, char side, char uplo, char trans, int n, int nrhs, double* A, int ia, int ja,
                   const int* descA, double* B, int ib, int jb, const int* descB, int& info)
{
    cholesky_impl::pposv_check(h, side, uplo, trans, n, nrhs, A, ia, ja, descA, B, ib, jb, descB);
    if (n == 0) {
        info = 0;
        return;
    }
    // The 64-row sweeps keep the workspace at about 512 bytes per row; the
    // staged solve's prepared blocks (12 KB per row in FP64) would not fit.
    const std::size_t bytes = pposv_unstaged_bytes<double>(side, n, nrhs);
    if (bytes > kNoAllocScratchBytes)
        throw std::invalid_argument("pposv_noalloc: this order and right-hand-side count need " +
                                    std::to_string(bytes) + " bytes of scratch, more than the " +
                                    std::to_string(kNoAllocScratchBytes) +
                                    " bytes of the static scratch (use pposv, which allocates)");
    void* work = scratch_address(h->device);
    pposv_run_unstaged<double>(h, side, uplo, n, nrhs, A, descA, B, descB, info, work, bytes);
    DDLA_HIP(hipStreamSynchronize(h->stream));
}

} // namespace ddla
