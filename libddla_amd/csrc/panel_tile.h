// One block column (panel) of the blocked Cholesky factorization in a single
// persistent launch (device code, included by kernels.hip).
//
// The panel is columns [k, k + kb) (kb <= 768) of the L view, rows k .. n-1,
// cut into 64 x 64 blocks: P = ceil(kb / 64) block columns and R = ceil((n -
// k) / 64) block rows.  Workgroup t (t from a ticket counter, so tickets are
// handed out in dispatch order) owns block row t and factors it left-looking:
// for each block column j it accumulates
//     acc = A(t, j) - sum_{i < j} L(t, i) L(j, i)^T
// and, once the inverse D_j = L(j, j)^-1 is published, writes
//     L(t, j) = acc D_j^T.
// Block rows t < P also own diagonal block t: they keep
//     A(t, t) - sum_{j < t} L(t, j) L(t, j)^T
// in registers, updated as soon as each L(t, j) is formed, and factor and
// invert it in LDS (potf2_tile.h) when its turn comes.  Progress of block row
// t < P is published as progress[t] = number of its block columns done (t + 1
// includes D_t), so the chain from D_j to D_{j+1} is one 64^3 product, one
// 64^3 rank update and the diagonal factorization.  A workgroup waits only on
// block rows with smaller tickets, which are running, so the panel always
// progresses; each wait is still bounded (spin limit -> error word, and the
// panel is abandoned).  A failed pivot sets *info and a failure flag that
// releases every waiter.
// Publication: every wavefront's stores acknowledged, a barrier, then one
// thread's agent-scope release fence (L2 write-back, waited for) and the
// counter store; readers poll relaxed and read blocks published by other
// workgroups with agent-scope atomic loads (never served from a stale L1 or
// L2 line).  A workgroup reads back the blocks it wrote itself with plain
// loads after a barrier (same-workgroup visibility).
#pragma once
#include <hip/hip_runtime.h>

#include "gemm_tile.h"
#include "potf2_tile.h"

namespace ddla {
namespace kern {
namespace panel_detail {

using gemm_detail::Mfma;
using potf2_detail::kB;
using potf2_detail::kL;

constexpr int kMaxCols = 12;  // block columns per panel (panel width <= 768)
constexpr int kSyncInts = 16; // per panel: [0] ticket, [1 + t] progress of block row t, [1 + kMaxCols] failure
constexpr int kPer = kB * kB / 256;

template <class T>
struct PanelArgs {
    T* A;
    long long lda;
    int upper;
    int n, k, kb;
    int* info;       // device info: the panel is skipped if nonzero on entry; receives k + 64 t + j (1-based),
                     // or kInfoWaitLimit when a wait gave up
    T* inv;          // kMaxCols inverses D_j, 64 x 64 column-major each
    int* sync;       // this panel's kSyncInts counters, zero on entry
    int* error;      // set to 1 when a wait exceeds spin_limit
    int spin_limit;
    long long bstride;  // grid y = independent problems: problem b's A, info, inv, sync and error sit
                        // b * bstride bytes after problem 0's
};

template <class T, bool COHERENT>
__device__ inline T ld(const T* p)
{
    if (COHERENT) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return *p;
}

// buf[c * kL + r] = L(i0 + r, j0 + c) for r < rows, c < cols, zero elsewhere.
// Consecutive threads read consecutive addresses (r fastest for uplo L, c for U).
template <class T, bool COHERENT>
__device__ inline void load_lblock(const PanelArgs<T>& a, long long i0, long long j0, int rows, int cols, T* buf)
{
    T v[kPer];
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
        const int idx = threadIdx.x + 256 * e, fast = idx & (kB - 1), slow = idx >> 6;
        const int r = a.upper ? slow : fast, c = a.upper ? fast : slow;
        v[e] = T(0);
        if (r < rows && c < cols) {
            const T* p = a.upper ? a.A + (j0 + c) + (i0 + r) * a.lda : a.A + (i0 + r) + (j0 + c) * a.lda;
            v[e] = ld<T, COHERENT>(p);
        }
    }
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
        const int idx = threadIdx.x + 256 * e, fast = idx & (kB - 1), slow = idx >> 6;
        const int r = a.upper ? slow : fast, c = a.upper ? fast : slow;
        buf[c * kL + r] = v[e];
    }
}

// L(i0 + r, j0 + c) = buf[c * kL + r] for r < rows, c < cols.
template <class T>
__device__ inline void store_lblock(const PanelArgs<T>& a, long long i0, long long j0, int rows, int cols,
                                    const T* buf)
{
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
        const int idx = threadIdx.x + 256 * e, fast = idx & (kB - 1), slow = idx >> 6;
        const int r = a.upper ? slow : fast, c = a.upper ? fast : slow;
        if (r < rows && c < cols) {
            T* p = a.upper ? a.A + (j0 + c) + (i0 + r) * a.lda : a.A + (i0 + r) + (j0 + c) * a.lda;
            *p = buf[c * kL + r];
        }
    }
}

// buf[c * kL + r] = D_j(r, c) (published inverse, column-major 64 x 64).
template <class T>
__device__ inline void load_inverse(const T* d, T* buf)
{
    T v[kPer];
#pragma unroll
    for (int e = 0; e < kPer; ++e) v[e] = ld<T, true>(d + threadIdx.x + 256 * e);
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
        const int idx = threadIdx.x + 256 * e;
        buf[(idx >> 6) * kL + (idx & (kB - 1))] = v[e];
    }
}

// Wavefront w owns rows 32 (w / 2) .. +31 and columns 32 (w % 2) .. +31 of a
// 64 x 64 product as 2 x 2 MFMA tiles: acc[i][jj] element e is
// (32 (w / 2) + 16 i + row(lane, e), 32 (w % 2) + 16 jj + lane % 16).
template <class T>
struct Acc64 {
    typename Mfma<T>::acc_t v[2][2];
    __device__ void zero()
    {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int jj = 0; jj < 2; ++jj) v[i][jj] = typename Mfma<T>::acc_t{0, 0, 0, 0};
    }
    // from / to a buffer laid out buf[col * kL + row]
    __device__ void load(const T* buf)
    {
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int ro = 32 * (wave >> 1), co = 32 * (wave & 1);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int jj = 0; jj < 2; ++jj)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    v[i][jj][e] = buf[(co + 16 * jj + (lane & 15)) * kL + ro + 16 * i + Mfma<T>::row(lane, e)];
    }
    __device__ void store(T* buf) const
    {
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int ro = 32 * (wave >> 1), co = 32 * (wave & 1);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int jj = 0; jj < 2; ++jj)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    buf[(co + 16 * jj + (lane & 15)) * kL + ro + 16 * i + Mfma<T>::row(lane, e)] = v[i][jj][e];
    }
    // += sign * As Bs^T with As(row, k) = As[k * kL + row], Bs(col, k) = Bs[k * kL + col]
    __device__ void mma_nt(const T* As, const T* Bs, T sign)
    {
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
        const int ro = 32 * (wave >> 1), co = 32 * (wave & 1);
#pragma unroll 4
        for (int ks = 0; ks < kB; ks += 4) {
            const int kk = (ks + lk) * kL;
            const T a0 = sign * As[kk + ro + li], a1 = sign * As[kk + ro + 16 + li];
            const T b0 = Bs[kk + co + li], b1 = Bs[kk + co + 16 + li];
            v[0][0] = Mfma<T>::mma(a0, b0, v[0][0]);
            v[0][1] = Mfma<T>::mma(a0, b1, v[0][1]);
            v[1][0] = Mfma<T>::mma(a1, b0, v[1][0]);
            v[1][1] = Mfma<T>::mma(a1, b1, v[1][1]);
        }
    }
};

template <class T>
__global__ __launch_bounds__(256) void panel_kernel(PanelArgs<T> a)
{
    __shared__ T S[kB * kL], X[kB * kL], rdiag[kB];
    __shared__ int s_ticket, s_flag;
    const int tid = threadIdx.x;
    if (gridDim.y > 1) {
        const long long off = (long long)blockIdx.y * a.bstride;
        a.A = reinterpret_cast<T*>(reinterpret_cast<char*>(a.A) + off);
        a.info = reinterpret_cast<int*>(reinterpret_cast<char*>(a.info) + off);
        a.inv = reinterpret_cast<T*>(reinterpret_cast<char*>(a.inv) + off);
        a.sync = reinterpret_cast<int*>(reinterpret_cast<char*>(a.sync) + off);
        a.error = reinterpret_cast<int*>(reinterpret_cast<char*>(a.error) + off);
    }
    int* const progress = a.sync + 1;
    int* const failed = a.sync + 1 + kMaxCols;

    if (tid == 0) {
        s_ticket = atomicAdd(a.sync, 1);
        s_flag = __hip_atomic_load(a.info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (s_flag != 0) return;  // an earlier panel (or this one) failed
    const int t = s_ticket;
    const int P = (a.kb + kB - 1) / kB;
    const long long r0 = (long long)a.k + (long long)kB * t;
    const int h = (int)min((long long)kB, (long long)a.n - r0);

    // Waits until progress[j] >= need.  Uniform result: 1 ready, 0 the panel
    // failed, -1 spin limit (error word set).  Every call is followed by a
    // barrier before the next one, so s_flag is read by all threads before
    // thread 0 writes it again.
    auto wait_progress = [&](int j, int need) -> int {
        if (tid == 0) {
            int st = 1, spins = 0;
            while (__hip_atomic_load(progress + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
                if (__hip_atomic_load(failed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                    st = 0;
                    break;
                }
                if (++spins > a.spin_limit) {
                    st = -1;
                    atomicExch(a.error, 1);
                    // the factor is not valid: the device info says so too
                    // (unless a failed pivot was recorded first), and every
                    // later panel and update of this factorization is skipped
                    atomicCAS(a.info, 0, kInfoWaitLimit);
                    atomicExch(failed, 1);
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
            s_flag = st;
        }
        __syncthreads();
        return s_flag;
    };
    // every wavefront's stores acknowledged, then one L2 write-back (release
    // fence), waited for explicitly before the counter moves (the compiler
    // emits no wait between the write-back and a following store)
    auto publish = [&](int value) {
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (tid == 0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __builtin_amdgcn_s_waitcnt(0);
            __hip_atomic_store(progress + t, value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    };

    const bool diag_row = t < P;
    const int ncols = diag_row ? t : P;  // off-diagonal block columns of this block row
    Acc64<T> dacc;                       // A(t, t) - sum_j L(t, j) L(t, j)^T
    if (diag_row) {
        load_lblock<T, false>(a, r0, r0, h, h, S);
        __syncthreads();
        dacc.load(S);
        __syncthreads();
    }
    for (int j = 0; j < ncols; ++j) {
        const long long c0 = (long long)a.k + (long long)kB * j;
        const int wj = min(kB, a.kb - kB * j);
        Acc64<T> acc;
        load_lblock<T, false>(a, r0, c0, h, wj, S);
        __syncthreads();
        acc.load(S);
        for (int i = 0; i < j; ++i) {
            const long long ci = (long long)a.k + (long long)kB * i;
            const int wi = min(kB, a.kb - kB * i);
            if (wait_progress(j, i + 1) <= 0) return;
            load_lblock<T, false>(a, r0, ci, h, wi, S); // L(t, i), written by this workgroup
            load_lblock<T, true>(a, c0, ci, wj, wi, X); // L(j, i), published by block row j
            __syncthreads();
            acc.mma_nt(S, X, T(-1));
            __syncthreads();
        }
        if (wait_progress(j, j + 1) <= 0) return;
        acc.store(S);
        load_inverse<T>(a.inv + (long long)j * kB * kB, X);
        __syncthreads();
        Acc64<T> l;
        l.zero();
        l.mma_nt(S, X, T(1));  // L(t, j) = acc D_j^T
        __syncthreads();
        l.store(S);
        __syncthreads();
        store_lblock<T>(a, r0, c0, h, wj, S);
        if (diag_row) {
            publish(j + 1);
            dacc.mma_nt(S, S, T(-1));
        }
        __syncthreads();
    }
    if (!diag_row) return;

    // diagonal block t (order h); the identity beyond h
    dacc.store(S);
    __syncthreads();
    if (tid >= h && tid < kB) S[tid * kL + tid] = T(1);
    __syncthreads();
    const int fail = potf2_detail::factor_invert64<T>(S, X, rdiag, &s_flag);
    potf2_detail::store_factor<T>(a.upper != 0, h, S, a.A + r0 + r0 * a.lda, a.lda);
    if (fail != 0) {
        if (tid == 0) {
            atomicExch(a.info, (int)r0 + fail);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            atomicExch(failed, 1);
        }
        return;
    }
    potf2_detail::store_inverse<T>(X, a.inv + (long long)t * kB * kB);
    publish(t + 1);
}

} // namespace panel_detail
} // namespace kern
} // namespace ddla
