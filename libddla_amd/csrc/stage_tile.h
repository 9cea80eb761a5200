// Triangular solve with the Cholesky factor in 512-row stages (device code,
// included by kernels.hip).
//
// The 64-row sweep of solve_tile.h hands every 64-row block of the solution
// from one workgroup to the next: 512 cross-CU hand-overs per sweep at order
// 32768, each about 4 us.  Here the rows are grouped in stages of S = 512 (8
// blocks of 64) and the hand-over chain runs over stages.  For stage s (in
// sweep order; the backward sweep takes the stages from the bottom and the
// transposed factor):
//
//   acc'_s = B_s - sum_{q <= s-2} Op(s, q) X_q
//   X_s    = E_s acc'_s - M_s X_{s-1}
//
// with E_s the inverse of the stage's S x S diagonal block of Op and
// M_s = E_s Op(s, s-1), both formed once per factor (stage_inverse_kernel,
// stage_product_kernel).  One workgroup per 64-row block (tickets in sweep
// order, as in solve_tile.h) accumulates its part of acc' as the stages are
// published, then forms its rows of E_s acc'_s from the stage's other blocks
// and, once stage s - 1 is published, subtracts its rows of M_s X_{s-1}: one
// 64 x 512 x nrhs product after the hand-over, per 512 rows instead of per 64.
//
// Publication and reading of published blocks follow solve_tile.h: the
// stores are acknowledged, the workgroup meets at a barrier, one thread
// releases at agent scope (L2 write-back, waited for) and then sets a flag or
// adds to the stage's counter; readers poll (relaxed) and read the published
// rows with agent-scope atomic loads.  Every wait is on a smaller ticket and
// bounded by the spin limit (then the error word is set and the workgroup
// carries on).
#pragma once
#include <hip/hip_runtime.h>

#include "gemm_tile.h"
#include "solve_tile.h"

namespace ddla {
namespace kern {
namespace stage_detail {

using gemm_detail::Mfma;
using solve_detail::block_mma;
using solve_detail::kPer;
using solve_detail::kSB;  // 64
using solve_detail::kSL;
using solve_detail::kThreads;  // 512

constexpr int kS = 512;            // stage rows
constexpr int kBlocks = kS / kSB;  // 64-row blocks per stage

// Ints of one right-hand-side slab's counters: ticket, stage_done[nst], acc_flag[nblk].
__host__ __device__ inline int sync_ints(int nst, int nblk) { return (1 + nst + nblk + 63) / 64 * 64; }

template <class T>
struct StageArgs {
    int n, nrhs, nst, nblk;
    const T* L;  // the factor, L view: L(i, j) = upper ? L[j + i*ldl] : L[i + j*ldl]
    long long ldl;
    int upper;
    const T* dinv;  // inverses of the 64 x 64 diagonal blocks (trtri_kernel)
    T* E;           // nst blocks of S x S (column-major, ld S): inverse of each stage's diagonal block
    T* M;           // nst blocks of S x S: M of sweep step s >= 1 (block s)
    T* B;
    long long ldb;
    T* acc;     // per slab an n x 64 block (ld n): acc' of the stages
    int* sync;  // per slab sync_ints(nst, nblk) ints, zero on entry
    int* error;
    int spin_limit;
};

template <class T>
__device__ inline T lview(const StageArgs<T>& a, long long i, long long j)
{
    return a.upper ? a.L[j + i * a.ldl] : a.L[i + j * a.ldl];
}

// Releases this workgroup's stores and adds one to *ctr (see the file comment).
__device__ inline void publish_one(int* ctr)
{
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __builtin_amdgcn_s_waitcnt(0);
        __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Waits until *ctr >= target (thread 0 polls, relaxed); false if the spin
// limit was hit.  Ends with a barrier.
__device__ inline bool wait_count(const int* ctr, int target, int spin_limit, int* s_flag)
{
    if (threadIdx.x == 0) {
        int d = __hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int spins = 0;
        while (d < target) {
            if (++spins > spin_limit) {
                d = -1;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
            d = __hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        *s_flag = d >= target ? 1 : 0;
    }
    __syncthreads();
    return *s_flag != 0;
}

template <class T>
__device__ inline T coherent_load(const T* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Raw element (r, k) of the block at `base` (a uniform pointer) of a
// column-major matrix with leading dimension ld, clamped to rows x cols:
// (r, k) is read at base + min(r, rows - 1) + min(k, cols - 1) * ld, with a
// 32-bit element offset (the host keeps 64 ld below 2^31).
template <class T>
__device__ inline T block_raw(const T* base, int ld, int r, int k, int rows, int cols)
{
    const unsigned off = (unsigned)min(r, rows - 1) + (unsigned)min(k, cols - 1) * (unsigned)ld;
    return base[off];
}
template <class T>
__device__ inline T block_raw_coherent(const T* base, int ld, int r, int k, int rows, int cols)
{
    const unsigned off = (unsigned)min(r, rows - 1) + (unsigned)min(k, cols - 1) * (unsigned)ld;
    return coherent_load(base + off);
}

// ---------------------------------------------------------------------------
// E_T, one workgroup per (stage T, 64-column block j): column block j of the
// inverse of the stage's diagonal block (L view, always lower), by block
// forward substitution: E_jj = D_j, E_ij = -D_i sum_{j <= l < i} L_il E_lj.
// Everything outside the stage's order (a short last stage) is zero.
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(kThreads) void stage_inverse_kernel(StageArgs<T> a)
{
    using Op = Mfma<T>;
    using acc_t = typename Op::acc_t;
    __shared__ T Ls[kSB][kSL];
    __shared__ T Xs[kSB][kSL];
    const int st = blockIdx.x, j = blockIdx.y, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col = (wave & 3) * 16 + (lane & 15), rt0 = (wave >> 2) * 2;
    const long long r0 = (long long)st * kS;
    const int m = (int)min((long long)kS, a.n - r0);
    const int nb = (m + kSB - 1) / kSB;
    T* E = a.E + (long long)st * kS * kS;
    auto width = [&](int b) { return min(kSB, m - b * kSB); };
    // rows of blocks above j and columns of blocks past the order are zero
    for (int i = 0; i < kBlocks; ++i) {
        if (j < nb && i >= j && i < nb) continue;
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            const int idx = tid + kThreads * e, r = idx & (kSB - 1), c = idx >> 6;
            E[(i * kSB + r) + (long long)(j * kSB + c) * kS] = T(0);
        }
    }
    if (j >= nb) return;
    const int wj = width(j);
    // E_jj = D_j (masked to the block's order)
    {
        const T* d = a.dinv + ((r0 + (long long)j * kSB) / kSB) * kSB * kSB;
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            const int idx = tid + kThreads * e, r = idx & (kSB - 1), c = idx >> 6;
            E[(j * kSB + r) + (long long)(j * kSB + c) * kS] = (r < wj && c < wj) ? d[r + c * kSB] : T(0);
        }
    }
    __syncthreads();
    for (int i = j + 1; i < nb; ++i) {
        const int wi = width(i);
        acc_t acc[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) acc[q] = acc_t{0, 0, 0, 0};
        for (int l = j; l < i; ++l) {
            const int wl = width(l);
            // Ls[k][r] = L(stage row 64i + r, stage column 64l + k); Xs[k][c] = E(64l + k, 64j + c)
#pragma unroll
            for (int e = 0; e < kPer; ++e) {
                const int idx = tid + kThreads * e;
                const int fast = idx & (kSB - 1), slow = idx >> 6;
                const int r = a.upper ? slow : fast, k = a.upper ? fast : slow;
                Ls[k][r] = (r < wi && k < wl) ? lview(a, r0 + i * kSB + r, r0 + l * kSB + k) : T(0);
                Xs[fast][slow] = E[(l * kSB + fast) + (long long)(j * kSB + slow) * kS];
            }
            __syncthreads();
            block_mma<T, false>(acc, Ls, Xs, T(1));
            __syncthreads();
        }
        // E_ij = -D_i acc: Xs[k][c] = acc(k, c), Ls[k][r] = D_i(r, k)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r) Xs[(rt0 + q) * 16 + Op::row(lane, r)][col] = acc[q][r];
        {
            const T* d = a.dinv + ((r0 + (long long)i * kSB) / kSB) * kSB * kSB;
#pragma unroll
            for (int e = 0; e < kPer; ++e) {
                const int idx = tid + kThreads * e, r = idx & (kSB - 1), k = idx >> 6;
                Ls[k][r] = (r < wi && k < wi) ? d[r + k * kSB] : T(0);
            }
        }
        __syncthreads();
        acc_t z[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) z[q] = acc_t{0, 0, 0, 0};
        block_mma<T, false>(z, Ls, Xs, T(-1));
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = (rt0 + q) * 16 + Op::row(lane, r);
                E[(i * kSB + row) + (long long)(j * kSB + col) * kS] = (row < wi && col < wj) ? z[q][r] : T(0);
            }
        __syncthreads();  // the block is read back (as E_lj) by the next rows
    }
}

// ---------------------------------------------------------------------------
// M of sweep step s >= 1, one workgroup per (s, 64-row block, 64-column block):
// forward (stage T = s)       M = E_T   L(T, T-1)
// backward (stage T = nst-1-s) M = E_T^T L(T+1, T)^T
// ---------------------------------------------------------------------------
template <class T, bool BACKWARD>
__global__ __launch_bounds__(kThreads) void stage_product_kernel(StageArgs<T> a)
{
    using Op = Mfma<T>;
    using acc_t = typename Op::acc_t;
    __shared__ T As[kSB][kSL];
    __shared__ T Bs[kSB][kSL];
    const int s = blockIdx.x + 1, ib = blockIdx.y, cb = blockIdx.z, tid = threadIdx.x, lane = tid & 63,
              wave = tid >> 6;
    const int col = (wave & 3) * 16 + (lane & 15), rt0 = (wave >> 2) * 2;
    const int st = BACKWARD ? a.nst - 1 - s : s, sp = BACKWARD ? st + 1 : st - 1;
    const long long r0 = (long long)st * kS, p0 = (long long)sp * kS;
    const int m = (int)min((long long)kS, a.n - r0), mp = (int)min((long long)kS, a.n - p0);
    const T* E = a.E + (long long)st * kS * kS;
    T* M = a.M + (long long)s * kS * kS;
    acc_t acc[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) acc[q] = acc_t{0, 0, 0, 0};
    const int nb = (m + kSB - 1) / kSB;
    // E_T is lower triangular, E_T^T upper: blocks lb <= ib, resp. lb >= ib
    const int lb0 = BACKWARD ? ib : 0, lb1 = BACKWARD ? nb : min(ib + 1, nb);
    for (int lb = lb0; lb < lb1; ++lb) {
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            const int idx = tid + kThreads * e, fast = idx & (kSB - 1), slow = idx >> 6;
            // As[k][r] = A(64 ib + r, 64 lb + k); A = E_T (contiguous along r) or E_T^T (along k)
            if (!BACKWARD) As[slow][fast] = E[(ib * kSB + fast) + (long long)(lb * kSB + slow) * kS];
            else As[fast][slow] = E[(lb * kSB + fast) + (long long)(ib * kSB + slow) * kS];
            // Bs[k][c] = Bm(64 lb + k, 64 cb + c); Bm = L(T, T-1) or L(T+1, T)^T, in L-view storage
            // order: lower: L(i, j) contiguous along i; upper: along j.
            const int l = lb * kSB, c = cb * kSB;
            int k, cc;
            bool k_fast;
            if (!BACKWARD) k_fast = !a.upper;  // L(r0 + k, p0 + cc): contiguous along k when lower
            else k_fast = a.upper;             // L(p0 + cc, r0 + k): contiguous along k when upper
            k = k_fast ? fast : slow;
            cc = k_fast ? slow : fast;
            T v = T(0);
            if (l + k < m && c + cc < mp)
                v = BACKWARD ? lview(a, p0 + c + cc, r0 + l + k) : lview(a, r0 + l + k, p0 + c + cc);
            Bs[k][cc] = v;
        }
        __syncthreads();
        block_mma<T, false>(acc, As, Bs, T(1));
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = (rt0 + q) * 16 + Op::row(lane, r);
            const int i = ib * kSB + row, c = cb * kSB + col;
            M[i + (long long)c * kS] = (i < m && c < mp) ? acc[q][r] : T(0);
        }
}

// ---------------------------------------------------------------------------
// acc += sign * sum_{q < nq} A_q R_q over 64 x 64 blocks whose operands can
// all be read now, software-pipelined: block q sits in register slot q % D
// while the next D - 1 are in flight; once q is in LDS its slot takes
// q + D.  Every load is unconditional (clamped addresses; past the last
// block the last one is read again and dropped) and the masks are applied
// when the values go to LDS, so that the compiler waits for each slot by
// count instead of draining every outstanding load.
//   a_pos(e, r, k)      place (r, k) of element e of this thread's share
//   a_load(q, e)        its raw value in A_q (clamped address)
//   a_dims(q, rows, cols) the extent of A_q (rows = 0 past the last block)
//   r_load(q, k, c)     raw element (k, c) of R_q (clamped address)
//   r_dims(q, rows, cols) the extent of R_q
// ---------------------------------------------------------------------------
template <class T, int D, class APos, class ALoad, class ADims, class RLoad, class RDims>
__device__ inline void known_mma(typename Mfma<T>::acc_t (&acc)[2], int nq, T sign, APos a_pos, ALoad a_load,
                                 ADims a_dims, RLoad r_load, RDims r_dims, T (*Ls)[kSL], T (*Xs)[kSL])
{
    if (nq <= 0) return;
    const int tid = threadIdx.x;
    T va0[kPer], vr0[kPer], va1[kPer], vr1[kPer];
    auto issue = [&](int q, T (&va)[kPer], T (&vr)[kPer]) {
        const int qc = min(q, nq - 1);
#pragma unroll
        for (int e = 0; e < kPer; ++e) va[e] = a_load(qc, e);
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            const int idx = tid + kThreads * e;
            vr[e] = r_load(qc, idx & (kSB - 1), idx >> 6);
        }
    };
    auto step = [&](int q, T (&va)[kPer], T (&vr)[kPer]) {
        int rows = 0, cols = 0, rrows = 0, rcols = 0;
        if (q < nq) {
            a_dims(q, rows, cols);
            r_dims(q, rrows, rcols);
        }
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            int r, k;
            a_pos(e, r, k);
            Ls[k][r] = (r < rows && k < cols) ? va[e] : T(0);
        }
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            const int idx = tid + kThreads * e, k = idx & (kSB - 1), c = idx >> 6;
            Xs[k][c] = (k < rrows && c < rcols) ? vr[e] : T(0);
        }
        __syncthreads();
        issue(q + D, va, vr);
        block_mma<T, false>(acc, Ls, Xs, sign);
        __syncthreads();
    };
    const int steps = (nq + D - 1) / D * D;
    issue(0, va0, vr0);
    if (D == 2) issue(1, va1, vr1);
    for (int q = 0; q < steps; q += D) {
        step(q, va0, vr0);
        if (D == 2) step(q + 1, va1, vr1);
    }
}

// Place (r, k) of element e of this thread's share of a 64 x 64 block, for a
// block contiguous along r (rows_fast) or along k.
__device__ inline void block_pos(bool rows_fast, int e, int& r, int& k)
{
    const int idx = threadIdx.x + kThreads * e;
    const int fast = idx & (kSB - 1), slow = idx >> 6;
    r = rows_fast ? fast : slow;
    k = rows_fast ? slow : fast;
}

// ---------------------------------------------------------------------------
// One sweep (forward: L Y = B; backward: L^T X = Y); grid (nblk, slabs).  A
// workgroup takes a ticket t and owns 64-row block t in sweep order (stage s,
// block b of the stage):
//   1. acc = B_b minus the blocks of the stages at least two steps back, as
//      they are published (as the 64-row sweep does);
//   2. publishes acc as its part of acc'_s (flag per block);
//   3. Y = sum over the stage's blocks b' <= b (forward; b' >= b backward,
//      all with smaller tickets) of E(b, b') acc'_b';
//   4. once stage s - 1 is published: X_b = Y - sum_k M(b, k) X_{s-1, k};
//   5. stores X_b and adds one to the stage's counter.
// Every wait is on smaller tickets.
// ---------------------------------------------------------------------------
template <class T, bool BACKWARD>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) void stage_sweep_kernel(
    StageArgs<T> a)
{
    using Op = Mfma<T>;
    using acc_t = typename Op::acc_t;
    // One register slot: a second (prefetch depth 2) does not fit the 128
    // registers of two workgroups per CU beside block_mma's operand preload.
    constexpr int D = 1;
    __shared__ T Ls[kSB][kSL];
    __shared__ T Xs[kSB][kSL];
    __shared__ int s_ticket, s_flag;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int slab = blockIdx.y, c0 = slab * kSB, nc = min(kSB, a.nrhs - c0);
    int* sync = a.sync + (long long)slab * sync_ints(a.nst, a.nblk);
    int* stage_done = sync + 1;          // per sweep stage: blocks published
    int* acc_flag = sync + 1 + a.nst;    // per global block: acc' published
    T* accb = a.acc + (long long)slab * a.n * kSB;

    if (tid == 0) s_ticket = atomicAdd(sync, 1);
    __syncthreads();
    const int t = s_ticket;
    const int bq = BACKWARD ? a.nblk - 1 - t : t;  // global 64-row block
    const int st = bq / kBlocks, b = bq % kBlocks, s = BACKWARD ? a.nst - 1 - st : st;
    const long long r0 = (long long)st * kS, i0 = (long long)bq * kSB;
    const int m = (int)min((long long)kS, a.n - r0), w = (int)min((long long)kSB, a.n - i0);
    const int nbs = (m + kSB - 1) / kSB;  // blocks of this stage
    auto stage_blocks = [&](int sq) {    // blocks of sweep stage sq
        const int sst = BACKWARD ? a.nst - 1 - sq : sq;
        return (int)min((long long)kBlocks, (a.n - (long long)sst * kS + kSB - 1) / kSB);
    };
    const int col = (wave & 3) * 16 + (lane & 15), rt0 = (wave >> 2) * 2;
    bool failed = false;

    acc_t acc[2];
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = (rt0 + q) * 16 + Op::row(lane, r);
            acc[q][r] = (row < w && col < nc) ? a.B[(i0 + row) + (long long)(c0 + col) * a.ldb] : T(0);
        }

    // Orientation of the factor's blocks: contiguous along the block's rows
    // (lower, forward; upper, backward) or along its columns.
    const bool l_rows_fast = BACKWARD == (a.upper != 0);
    // raw element of the factor's block at (i0.., j0..), rows < rows, columns < cols (clamped)
    // element (r, k) of Op's block at (fi0, fj0) is L(i, j) with (i, j) = (fi0 + r, fj0 + k) (forward) or
    // (fj0 + k, fi0 + r) (backward), stored at L[i + j ldl] (lower) or L[j + i ldl] (upper): at
    // base + r + k ldl when the block is contiguous along r, base + k + r ldl otherwise
    const int ldl = (int)a.ldl;
    auto factor_raw = [&](long long fi0, long long fj0, int rows, int cols, int e) -> T {
        int r, k;
        block_pos(l_rows_fast, e, r, k);
        const long long bi = BACKWARD ? fj0 : fi0, bj = BACKWARD ? fi0 : fj0;
        const T* base = a.upper ? a.L + bj + bi * a.ldl : a.L + bi + bj * a.ldl;
        return l_rows_fast ? block_raw(base, ldl, r, k, rows, cols) : block_raw(base, ldl, k, r, cols, rows);
    };

    // 1. blocks of the sweep stages <= s - 2, stage by stage in sweep order
    for (int sq = 0; sq + 2 <= s && !failed; ++sq) {
        if (!wait_count(stage_done + sq, stage_blocks(sq), a.spin_limit, &s_flag)) {
            failed = true;
            break;
        }
        const int sst = BACKWARD ? a.nst - 1 - sq : sq;
        const int nq = stage_blocks(sq);
        auto j0_of = [&](int q) { return (long long)sst * kS + (long long)q * kSB; };
        auto width = [&](int q) { return (int)min((long long)kSB, a.n - j0_of(q)); };
        known_mma<T, D>(
            acc, nq, T(-1), [&](int e, int& r, int& k) { block_pos(l_rows_fast, e, r, k); },
            [&](int q, int e) { return factor_raw(i0, j0_of(q), w, width(q), e); },
            [&](int q, int& rows, int& cols) { rows = w, cols = width(q); },
            [&](int q, int k, int c) {
                return block_raw_coherent(a.B + j0_of(q) + (long long)c0 * a.ldb, (int)a.ldb, k, c, width(q), nc);
            },
            [&](int q, int& rows, int& cols) { rows = width(q), cols = nc; }, Ls, Xs);
    }

    // 2. publish acc'_b.  From here on the workgroup is on the sweep's
    // critical path: its wavefronts take issue priority over the
    // accumulation of later blocks sharing the CU.
    __builtin_amdgcn_s_setprio(3);
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = (rt0 + q) * 16 + Op::row(lane, r);
            if (row < w && col < nc) accb[(i0 + row) + (long long)col * a.n] = acc[q][r];
        }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tid == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __builtin_amdgcn_s_waitcnt(0);
        __hip_atomic_store(acc_flag + bq, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }

    // 3. Y = sum_b' E(b, b') acc'_b' (forward: E lower, b' <= b; backward: E^T upper, b' >= b).
    // The stage's blocks publish their parts at about the same time: wait for
    // all of them, then run the products with every load in flight early.
    acc_t y[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) y[q] = acc_t{0, 0, 0, 0};
    {
        const T* E = a.E + (long long)st * kS * kS;
        const int nq = BACKWARD ? nbs - b : b + 1;
        const int bp0 = BACKWARD ? b : 0;
        if (!failed && tid == 0) {
            int ok = 1;
            for (int q = 0; q < nq && ok; ++q) {
                const int* f = acc_flag + (int)(r0 / kSB) + bp0 + q;
                int spins = 0;
                while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
                    if (++spins > a.spin_limit) {
                        ok = 0;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
            }
            s_flag = ok;
        }
        __syncthreads();
        if (!failed && s_flag == 0) failed = true;
        // forward: E(64b + r, 64b' + k), contiguous along r; backward: E^T(64b + r, 64b' + k) =
        // E(64b' + k, 64b + r), contiguous along k
        auto e_raw = [&](int q, int e) {
            const int bp = bp0 + q, wb = min(kSB, m - bp * kSB);
            int r, k;
            block_pos(!BACKWARD, e, r, k);
            return BACKWARD ? block_raw(E + bp * kSB + (long long)b * kSB * kS, kS, k, r, wb, w)
                            : block_raw(E + b * kSB + (long long)bp * kSB * kS, kS, r, k, w, wb);
        };
        if (!failed)
            known_mma<T, D>(
                y, nq, T(1), [&](int e, int& r, int& k) { block_pos(!BACKWARD, e, r, k); }, e_raw,
                [&](int q, int& rows, int& cols) { rows = w, cols = min(kSB, m - (bp0 + q) * kSB); },
                [&](int q, int k, int c) {
                    return block_raw_coherent(accb + r0 + (long long)(bp0 + q) * kSB, a.n, k, c,
                                              min(kSB, m - (bp0 + q) * kSB), nc);
                },
                [&](int q, int& rows, int& cols) { rows = min(kSB, m - (bp0 + q) * kSB), cols = nc; }, Ls, Xs);
    }

    // 4. Y -= M X_{s-1} once the previous stage of the sweep is published
    if (s >= 1 && !failed) {
        const int sp = BACKWARD ? st + 1 : st - 1;
        const long long p0 = (long long)sp * kS;
        const int mp = (int)min((long long)kS, a.n - p0);
        const T* M = a.M + (long long)s * kS * kS;
        const int nq = (mp + kSB - 1) / kSB;
        auto m_raw = [&](int q, int e) {
            int r, k;
            block_pos(true, e, r, k);
            return block_raw(M + b * kSB + (long long)q * kSB * kS, kS, r, k, w, min(kSB, mp - q * kSB));
        };
        if (!wait_count(stage_done + s - 1, stage_blocks(s - 1), a.spin_limit, &s_flag))
            failed = true;
        else
            known_mma<T, D>(
                y, nq, T(-1), [&](int e, int& r, int& k) { block_pos(true, e, r, k); }, m_raw,
                [&](int q, int& rows, int& cols) { rows = w, cols = min(kSB, mp - q * kSB); },
                [&](int q, int k, int c) {
                    return block_raw_coherent(a.B + p0 + (long long)q * kSB + (long long)c0 * a.ldb, (int)a.ldb, k, c,
                                              min(kSB, mp - q * kSB), nc);
                },
                [&](int q, int& rows, int& cols) { rows = min(kSB, mp - q * kSB), cols = nc; }, Ls, Xs);
    }
    if (failed && tid == 0) atomicExch(a.error, 1);

    // 5. X_b through LDS so that it leaves as 16-byte stores along the columns
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int r = 0; r < 4; ++r) Xs[col][(rt0 + q) * 16 + Op::row(lane, r)] = y[q][r];
    __syncthreads();
    {
        constexpr int VW = Mfma<T>::kVec;
        constexpr int NV = kSB * kSB / VW / kThreads;
#pragma unroll
        for (int e = 0; e < NV; ++e) {
            const int v = tid + kThreads * e, c = v / (kSB / VW), row = (v % (kSB / VW)) * VW;
            if (c >= nc) continue;
            T* dst = a.B + (i0 + row) + (long long)(c0 + c) * a.ldb;
            if (row + VW <= w && (((uintptr_t)dst) & 15) == 0) {
                typename Mfma<T>::vec_t x;
#pragma unroll
                for (int q = 0; q < VW; ++q) x[q] = Xs[c][row + q];
                *reinterpret_cast<typename Mfma<T>::vec_t*>(dst) = x;
            } else {
#pragma unroll
                for (int q = 0; q < VW; ++q)
                    if (row + q < w) dst[q] = Xs[c][row + q];
            }
        }
    }
    publish_one(stage_done + s);
}

} // namespace stage_detail
} // namespace kern
} // namespace ddla
