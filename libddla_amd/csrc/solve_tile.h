// Triangular solve with the Cholesky factor in one launch per direction
// (device code, included by kernels.hip).
//
// Forward (L Y = B) and backward (L^T X = Y) substitution over 64-row blocks.
// Each workgroup takes a ticket t from a per-column-slab counter and solves
// block t of the sweep (block t from the top for the forward sweep, from the
// bottom for the backward one).  Blocks are published in ticket order through
// a per-slab counter `done`: block t accumulates -L(t, j) X_j (forward) or
// -L(j, t)^T X_j (backward) for every published block j on the matrix cores,
// then applies the inverse of its diagonal block, writes X_t in place of B_t
// and sets done = t + 1.  A workgroup waits only on tickets handed out
// before its own, i.e. to workgroups that are already running, so the sweep
// always progresses; every wait is nevertheless bounded by a spin limit, after
// which the workgroup flags an error and carries on (the host reports it).
// Publication uses agent-scope release / acquire fences, so blocks written on
// one XCD are seen on the others.
#pragma once
#include <hip/hip_runtime.h>

#include "gemm_tile.h"

namespace ddla {
namespace kern {
namespace solve_detail {

using gemm_detail::Mfma;

template <class T>
struct SolveArgs {
    int n, nrhs, nblk;
    const T* L;   // the factor, L view: L(i, j) = upper ? L[j + i*ldl] : L[i + j*ldl]
    long long ldl;
    int upper;
    const T* dinv;  // inverse of every 64 x 64 diagonal block of L (column-major, 64 x 64 each)
    T* B;
    long long ldb;
    int* sync;   // per 64-column slab s: sync[2s] ticket counter, sync[2s+1] published blocks; zero on entry
    int* error;  // set to 1 if a wait exceeds spin_limit
    int spin_limit;
};

constexpr int kSB = 64;       // block rows and right-hand-side columns per workgroup
constexpr int kSL = kSB + 1;  // LDS leading dimension

// Ls[k][r] := block(r, k), rows r < rows, k < cols of the L view starting at
// (i0, j0); transposed: block(r, k) = L(j0 + k, i0 + r).
template <class T>
__device__ inline void load_factor_block(const SolveArgs<T>& a, long long i0, long long j0, int rows, int cols,
                                         bool transposed, T (*Ls)[kSL])
{
    // element (r, k) of the block is L(i, j) with (i, j) = (i0 + r, j0 + k), or (j0 + k, i0 + r) transposed
    // stored at L[i + j*ldl] (lower) or L[j + i*ldl] (upper); choose the thread order so that
    // consecutive threads touch consecutive addresses.
    const bool rows_contiguous = (!transposed && !a.upper) || (transposed && a.upper);
    for (int idx = threadIdx.x; idx < kSB * kSB; idx += 256) {
        const int r = rows_contiguous ? (idx & (kSB - 1)) : (idx >> 6);
        const int k = rows_contiguous ? (idx >> 6) : (idx & (kSB - 1));
        T v = T(0);
        if (r < rows && k < cols) {
            const long long i = transposed ? j0 + k : i0 + r;
            const long long j = transposed ? i0 + r : j0 + k;
            v = a.upper ? a.L[j + i * a.ldl] : a.L[i + j * a.ldl];
        }
        Ls[k][r] = v;
    }
}

template <class T, bool BACKWARD>
__global__ __launch_bounds__(256) void solve_sweep_kernel(SolveArgs<T> a)
{
    using Op = Mfma<T>;
    using acc_t = typename Op::acc_t;
    __shared__ T Ls[kSB][kSL];  // [k][row] of the factor block (or of the diagonal inverse)
    __shared__ T Xs[kSB][kSL];  // [k][column] of a published solution block (or of the accumulator)
    __shared__ int s_ticket, s_avail;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int slab = blockIdx.y, c0 = slab * kSB, nc = min(kSB, a.nrhs - c0);
    int* ticket_ctr = a.sync + 2 * slab;
    int* done_ctr = a.sync + 2 * slab + 1;

    if (tid == 0) s_ticket = atomicAdd(ticket_ctr, 1);
    __syncthreads();
    const int t = s_ticket;
    const int bt = BACKWARD ? a.nblk - 1 - t : t;
    const long long r0 = (long long)bt * kSB;
    const int w = (int)min((long long)kSB, a.n - r0);

    // accumulator: wavefront `wave` owns right-hand-side columns [16 wave, 16 wave + 16)
    // and the 4 row tiles of the block.
    acc_t acc[4];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = rt * 16 + Op::row(lane, r), col = wave * 16 + (lane & 15);
            acc[rt][r] = (row < w && col < nc) ? a.B[(r0 + row) + (long long)(c0 + col) * a.ldb] : T(0);
        }

    int processed = 0;
    bool failed = false;
    while (processed < t) {
        if (tid == 0) {
            // relaxed polling (no cache invalidation per poll); the acquire
            // fence below orders the reads of the published blocks
            int d = __hip_atomic_load(done_ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int spins = 0;
            while (d <= processed) {
                if (++spins > a.spin_limit) {
                    d = -1;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
                d = __hip_atomic_load(done_ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            s_avail = d;
        }
        __syncthreads();
        const int avail = s_avail;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        __syncthreads();
        if (avail < 0) {
            failed = true;
            break;
        }
        const int upto = min(avail, t);
        for (int q = processed; q < upto; ++q) {
            const int bj = BACKWARD ? a.nblk - 1 - q : q;
            const long long j0 = (long long)bj * kSB;
            const int wj = (int)min((long long)kSB, a.n - j0);
            // block(r, k) = L(r0 + r, j0 + k) forward, L(j0 + k, r0 + r) backward; r < w, k < wj
            load_factor_block(a, r0, j0, w, wj, BACKWARD, Ls);
            for (int idx = tid; idx < kSB * kSB; idx += 256) {
                const int k = idx & (kSB - 1), col = idx >> 6;
                Xs[k][col] = (k < wj && col < nc) ? a.B[(j0 + k) + (long long)(c0 + col) * a.ldb] : T(0);
            }
            __syncthreads();
#pragma unroll 4
            for (int ks = 0; ks < kSB; ks += 4) {
                const int kk = ks + (lane >> 4);
                const T b = Xs[kk][wave * 16 + (lane & 15)];
#pragma unroll
                for (int rt = 0; rt < 4; ++rt) acc[rt] = Op::mma(-Ls[kk][rt * 16 + (lane & 15)], b, acc[rt]);
            }
            __syncthreads();
        }
        processed = upto;
    }
    if (failed && tid == 0) atomicExch(a.error, 1);

    // X_t = Dinv_bt * acc (forward) or Dinv_bt^T * acc (backward)
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int r = 0; r < 4; ++r) Xs[rt * 16 + Op::row(lane, r)][wave * 16 + (lane & 15)] = acc[rt][r];
    const T* d = a.dinv + (long long)bt * kSB * kSB;
    for (int idx = tid; idx < kSB * kSB; idx += 256) {
        const int row = idx & (kSB - 1), k = idx >> 6;
        // Ls[k][row] = D(row, k), D = Dinv (forward) or Dinv^T (backward)
        Ls[k][row] = BACKWARD ? d[k + row * kSB] : d[row + k * kSB];
    }
    __syncthreads();
    acc_t x[4];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) x[rt] = acc_t{0, 0, 0, 0};
#pragma unroll 4
    for (int ks = 0; ks < kSB; ks += 4) {
        const int kk = ks + (lane >> 4);
        const T b = Xs[kk][wave * 16 + (lane & 15)];
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) x[rt] = Op::mma(Ls[kk][rt * 16 + (lane & 15)], b, x[rt]);
    }
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = rt * 16 + Op::row(lane, r), col = wave * 16 + (lane & 15);
            if (row < w && col < nc) a.B[(r0 + row) + (long long)(c0 + col) * a.ldb] = x[rt][r];
        }
    // publish block t: every wavefront's stores are released before the counter moves
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (tid == 0) __hip_atomic_store(done_ctr, t + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void zero_ints_kernel(int* p, int count)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) p[i] = 0;
}

} // namespace solve_detail
} // namespace kern
} // namespace ddla
