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
// Publication: every wavefront waits for its stores of the block to be
// acknowledged, the workgroup meets at a barrier, and one thread runs an
// agent-scope release fence (L2 write-back, waited for) and then stores the
// counter (one write-back per block: the barrier orders the other
// wavefronts' stores before that fence); readers
// poll the counter and read the published block with agent-scope atomic
// loads, which see the written-back data on every XCD without invalidating
// their L2.
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

constexpr int kThreads = 512;                  // 8 wavefronts per workgroup
constexpr int kPer = kSB * kSB / kThreads;     // elements of a 64 x 64 block per thread

// Ls[k][r] := block(r, k), rows r < rows, k < cols of the L view starting at
// (i0, j0); transposed: block(r, k) = L(j0 + k, i0 + r).  All loads of a
// thread are issued before any is stored to LDS.
template <class T>
__device__ inline void load_factor_block(const SolveArgs<T>& a, long long i0, long long j0, int rows, int cols,
                                         bool transposed, T (*Ls)[kSL])
{
    // element (r, k) is L(i, j), (i, j) = (i0 + r, j0 + k) or (j0 + k, i0 + r) transposed, stored at
    // L[i + j*ldl] (lower) or L[j + i*ldl] (upper); the thread order makes consecutive threads touch
    // consecutive addresses.
    const bool rows_contiguous = (!transposed && !a.upper) || (transposed && a.upper);
    T v[kPer];
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
        const int idx = threadIdx.x + kThreads * e;
        const int r = rows_contiguous ? (idx & (kSB - 1)) : (idx >> 6);
        const int k = rows_contiguous ? (idx >> 6) : (idx & (kSB - 1));
        v[e] = T(0);
        if (r < rows && k < cols) {
            const long long i = transposed ? j0 + k : i0 + r;
            const long long j = transposed ? i0 + r : j0 + k;
            v[e] = a.upper ? a.L[j + i * a.ldl] : a.L[i + j * a.ldl];
        }
    }
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
        const int idx = threadIdx.x + kThreads * e;
        const int r = rows_contiguous ? (idx & (kSB - 1)) : (idx >> 6);
        const int k = rows_contiguous ? (idx >> 6) : (idx & (kSB - 1));
        Ls[k][r] = v[e];
    }
}

// Xs[k][c] := B(j0 + k, c0 + c), k < rows, c < cols (zero elsewhere), for a
// block published by another workgroup: read with agent-scope atomic loads,
// which are served coherently across the XCDs' L2 caches, so no L2
// invalidation is needed (the writer released the block before publishing).
template <class T>
__device__ inline void load_rhs_block(const SolveArgs<T>& a, long long j0, int c0, int rows, int cols, T (*Xs)[kSL])
{
    T v[kPer];
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
        const int idx = threadIdx.x + kThreads * e, k = idx & (kSB - 1), c = idx >> 6;
        v[e] = (k < rows && c < cols) ? __hip_atomic_load(a.B + (j0 + k) + (long long)(c0 + c) * a.ldb,
                                                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                      : T(0);
    }
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
        const int idx = threadIdx.x + kThreads * e, k = idx & (kSB - 1), c = idx >> 6;
        Xs[k][c] = v[e];
    }
}

// Wavefront w owns column tile (w & 3) and row tiles 2(w >> 2) and 2(w >> 2) + 1
// of a 64 x 64 block.  acc[i] += sign * As(row, k) * Bs(k, col) over k < 64,
// with As given as As[k][row] and Bs as Bs[k][col] (BT: Bs(k, col) = Bs[col][k]).
template <class T, bool BT>
__device__ inline void block_mma(typename Mfma<T>::acc_t (&acc)[2], const T (*As)[kSL], const T (*Bs)[kSL], T sign)
{
    using Op = Mfma<T>;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = (wave & 3) * 16 + (lane & 15), rt0 = (wave >> 2) * 2;
    // The operands of KC k-steps (24 registers) are read from LDS before
    // their products, so that the reads overlap one another instead of each
    // pair of products waiting for its own.
    constexpr int KC = sizeof(T) == 4 ? 8 : 4;
#pragma unroll
    for (int k0 = 0; k0 < kSB / 4; k0 += KC) {
        T av[2][KC], bv[KC];
#pragma unroll
        for (int j = 0; j < KC; ++j) {
            const int kk = 4 * (k0 + j) + (lane >> 4);
            bv[j] = sign * (BT ? Bs[col][kk] : Bs[kk][col]);
#pragma unroll
            for (int i = 0; i < 2; ++i) av[i][j] = As[kk][(rt0 + i) * 16 + (lane & 15)];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < KC; ++j)
#pragma unroll
            for (int i = 0; i < 2; ++i) acc[i] = Op::mma(av[i][j], bv[j], acc[i]);
    }
}

// Element e of this thread's share of a factor block (as load_factor_block)
// and its place (r, k) in Ls[k][r].
template <class T>
__device__ inline T factor_element(const SolveArgs<T>& a, long long i0, long long j0, int rows, int cols,
                                   bool transposed, int e, int& r, int& k)
{
    const bool rows_contiguous = (!transposed && !a.upper) || (transposed && a.upper);
    const int idx = threadIdx.x + kThreads * e;
    r = rows_contiguous ? (idx & (kSB - 1)) : (idx >> 6);
    k = rows_contiguous ? (idx >> 6) : (idx & (kSB - 1));
    if (rows <= 0 || cols <= 0) return T(0);
    // Branch-free: outside the block the load is clamped to its last row /
    // column and the value dropped, so every lane issues the same loads and
    // the compiler can wait for them by count instead of draining all
    // outstanding loads (which would also wait for the prefetched block).
    const bool in = r < rows && k < cols;
    const int rc = min(r, rows - 1), kc = min(k, cols - 1);
    const long long i = transposed ? j0 + kc : i0 + rc;
    const long long j = transposed ? i0 + rc : j0 + kc;
    const T v = a.upper ? a.L[j + i * a.ldl] : a.L[i + j * a.ldl];
    return in ? v : T(0);
}

// Ds[k][row] := D(row, k) with D = Dinv of diagonal block bt (forward) or
// its transpose (backward).
template <class T, bool BACKWARD>
__device__ inline void load_diag_inverse(const SolveArgs<T>& a, int bt, T (*Ds)[kSL])
{
    const T* d = a.dinv + (long long)bt * kSB * kSB;
    T v[kPer];
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
        const int idx = threadIdx.x + kThreads * e, row = idx & (kSB - 1), k = idx >> 6;
        v[e] = BACKWARD ? d[k + row * kSB] : d[row + k * kSB];
    }
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
        const int idx = threadIdx.x + kThreads * e, row = idx & (kSB - 1), k = idx >> 6;
        Ds[k][row] = v[e];
    }
}

// M = D L(t, t-1) is formed (into its own LDS block) before the accumulation
// loop; the loop multiplies block q from LDS while the factor block q + 1
// (and X_{q+1}, when it is already published) are loaded into registers.
// After the loop, Z = D acc is formed, so that once X_{t-1} is published one
// product is left, M X_{t-1}.
template <class T, bool BACKWARD, int WAVES_PER_EU = 2>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(WAVES_PER_EU))) void solve_sweep_kernel(
    SolveArgs<T> a)
{
    using Op = Mfma<T>;
    using acc_t = typename Op::acc_t;
    __shared__ T Ls[kSB][kSL];  // [k][row] of a factor block; later D
    __shared__ T Xs[kSB][kSL];  // [k][column] of a published solution block; D at the start; the accumulator
    __shared__ T Ms[kSB][kSL];  // [k][row] of M = D L(t, t-1)
    __shared__ int s_ticket, s_avail;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int slab = blockIdx.y, c0 = slab * kSB, nc = min(kSB, a.nrhs - c0);
    int* ticket_ctr = a.sync + 2 * slab;
    int* done_ctr = a.sync + 2 * slab + 1;
    const int col = (wave & 3) * 16 + (lane & 15), rt0 = (wave >> 2) * 2;

    if (tid == 0) s_ticket = atomicAdd(ticket_ctr, 1);
    __syncthreads();
    const int t = s_ticket;
    const int bt = BACKWARD ? a.nblk - 1 - t : t;
    const long long r0 = (long long)bt * kSB;
    const int w = (int)min((long long)kSB, a.n - r0);
    auto block_of = [&](int q, long long& j0, int& wj) {
        const int bj = BACKWARD ? a.nblk - 1 - q : q;
        j0 = (long long)bj * kSB;
        wj = (int)min((long long)kSB, a.n - j0);
    };

    acc_t acc[2], m[2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = (rt0 + i) * 16 + Op::row(lane, r);
            acc[i][r] = (row < w && col < nc) ? a.B[(r0 + row) + (long long)(c0 + col) * a.ldb] : T(0);
        }
    // M(row, c) = sum_k D(row, k) L(t, t-1)(k, c), with D in Xs and Ls[c][k] = L(t, t-1)(k, c)
    if (t >= 1) {
        long long j0;
        int wj;
        block_of(t - 1, j0, wj);
        load_diag_inverse<T, BACKWARD>(a, bt, Xs);
        load_factor_block(a, r0, j0, w, wj, BACKWARD, Ls);
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 2; ++i) m[i] = acc_t{0, 0, 0, 0};
        block_mma<T, true>(m, Xs, Ls, T(1));
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) Ms[col][(rt0 + i) * 16 + Op::row(lane, r)] = m[i][r];  // Ms[k][row] = M(row, k)
        __syncthreads();
    }

    int avail = 0;
    bool failed = false;
    // Returns false if the spin limit was hit.  Afterwards blocks [0, avail) are published and visible.
    auto wait_for = [&](int q) -> bool {
        if (q < avail) return true;
        if (tid == 0) {
            // relaxed polling; the published blocks are then read with coherent loads
            int d = __hip_atomic_load(done_ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int spins = 0;
            while (d <= q) {
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
        avail = s_avail;
        return avail >= 0;
    };

    // every published block but the last: acc -= L(t, q) X_q, pipelined
    T vl[kPer], vx[kPer];
    auto load_l = [&](int q) {
        long long j0;
        int wj;
        block_of(q, j0, wj);
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            int r, k;
            vl[e] = factor_element(a, r0, j0, w, wj, BACKWARD, e, r, k);
        }
    };
    auto load_x = [&](int q) {
        long long j0;
        int wj;
        block_of(q, j0, wj);
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            const int idx = tid + kThreads * e, k = idx & (kSB - 1), c = idx >> 6;
            vx[e] = (k < wj && c < nc) ? __hip_atomic_load(a.B + (j0 + k) + (long long)(c0 + c) * a.ldb,
                                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                       : T(0);
        }
    };
    const int nq = t - 1;  // blocks q < t - 1
    if (nq > 0) {
        load_l(0);
        if (wait_for(0)) load_x(0);
        else failed = true;
    }
    for (int q = 0; q < nq && !failed; ++q) {
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            int r, k;
            factor_element(a, 0, 0, 0, 0, BACKWARD, e, r, k);  // (r, k) of element e only
            Ls[k][r] = vl[e];
        }
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            const int idx = tid + kThreads * e, k = idx & (kSB - 1), c = idx >> 6;
            Xs[k][c] = vx[e];
        }
        __syncthreads();
        const bool more = q + 1 < nq;
        const bool x_ready = more && q + 1 < avail;
        if (more) load_l(q + 1);
        if (x_ready) load_x(q + 1);
        block_mma<T, false>(acc, Ls, Xs, T(-1));
        __syncthreads();
        if (more && !x_ready) {
            if (wait_for(q + 1)) load_x(q + 1);
            else failed = true;
        }
    }

    // X_t = D (acc - L(t, t-1) X_{t-1}) = Z - M X_{t-1}, Z = D acc: only M X_{t-1}
    // is left once X_{t-1} is published.
    load_diag_inverse<T, BACKWARD>(a, bt, Ls);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) Xs[(rt0 + i) * 16 + Op::row(lane, r)][col] = acc[i][r];
    __syncthreads();
    acc_t z[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) z[i] = acc_t{0, 0, 0, 0};
    block_mma<T, false>(z, Ls, Xs, T(1));  // Z(row, col) = sum_k D(row, k) acc(k, col)
    __syncthreads();  // Xs is overwritten next (wait_for returns without a barrier when X_{t-1} is already known)
    if (t >= 1 && !failed) {
        if (!wait_for(t - 1)) {
            failed = true;
        } else {
            long long j0;
            int wj;
            block_of(t - 1, j0, wj);
            load_rhs_block(a, j0, c0, wj, nc, Xs);
            __syncthreads();
            block_mma<T, false>(z, Ms, Xs, T(-1));
        }
    }
    if (failed && tid == 0) atomicExch(a.error, 1);
    // X_t through LDS (Xs[row][col]) so that it leaves as 16-byte stores along
    // the columns
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) Xs[col][(rt0 + i) * 16 + Op::row(lane, r)] = z[i][r];
    __syncthreads();
    {
        constexpr int VW = Mfma<T>::kVec;              // elements per 16 bytes
        constexpr int NV = kSB * kSB / VW / kThreads;  // vectors per thread
#pragma unroll
        for (int e = 0; e < NV; ++e) {
            const int v = tid + kThreads * e, c = v / (kSB / VW), row = (v % (kSB / VW)) * VW;
            if (c >= nc) continue;
            T* dst = a.B + (r0 + row) + (long long)(c0 + c) * a.ldb;
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
    // publish block t: each wavefront's stores are acknowledged before the
    // barrier; one thread then writes the L2 back (the release fence), waits
    // for the write-back explicitly, and only then moves the counter.  (The
    // compiler emits no wait between the write-back and a following store,
    // so without it the counter could become visible before the block.)
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tid == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __builtin_amdgcn_s_waitcnt(0);
        __hip_atomic_store(done_ctr, t + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---------------------------------------------------------------------------
// Residual R = B - A X of a symmetric A given by one stored triangle
// (lower: A(i, k) = S[i + k*lds] for i >= k, S[k + i*lds] otherwise; upper:
// the mirror).  One 512-thread workgroup per 64-row block of R and 64-column
// slab: it walks the 64-column blocks k of A, staging A(i-block, k-block) and
// X(k-block) in LDS, and accumulates -A X on the matrix cores.  Off-diagonal
// blocks are read entirely from the stored triangle or entirely from its
// mirror (coalesced either way); diagonal blocks element by element.
// ---------------------------------------------------------------------------
template <class T>
struct SymmArgs {
    int n, nrhs, nblk;
    const T* S;
    long long lds;
    int lower;
    const T* X;
    long long ldx;
    const T* B;
    long long ldb;
    T* R;
    long long ldr;
};

// Element e of this thread's share of block (i0.., k0..) of A (rows < rows,
// columns < cols, zero elsewhere) and its place in Ls[k][r].  direct: element
// (r, k) is S[(i0 + r) + (k0 + k) * lds] (contiguous along r); otherwise
// (mirrored) S[(k0 + k) + (i0 + r) * lds] (contiguous along k); diagonal
// blocks pick per element.
template <class T>
__device__ inline T sym_element(const SymmArgs<T>& a, long long i0, long long k0, int rows, int cols, bool diagonal,
                                bool direct, int e, int& r, int& k)
{
    const bool rows_contiguous = diagonal || direct;
    const int idx = threadIdx.x + kThreads * e;
    r = rows_contiguous ? (idx & (kSB - 1)) : (idx >> 6);
    k = rows_contiguous ? (idx >> 6) : (idx & (kSB - 1));
    if (r >= rows || k >= cols) return T(0);
    const bool d = diagonal ? (a.lower ? r >= k : r <= k) : direct;
    const long long i = i0 + r, j = k0 + k;
    return d ? a.S[i + j * a.lds] : a.S[j + i * a.lds];
}

// One workgroup per 64-row block of R and 64-column slab walks the 64-column
// blocks of A; the loads of block kb + 1 (A and X, into registers) are in
// flight while block kb is multiplied from LDS.
template <class T>
__global__ __launch_bounds__(kThreads) void symm_residual_kernel(SymmArgs<T> a)
{
    using Op = Mfma<T>;
    using acc_t = typename Op::acc_t;
    __shared__ T Ls[kSB][kSL];
    __shared__ T Xs[kSB][kSL];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ib = blockIdx.x, slab = blockIdx.y, c0 = slab * kSB, nc = min(kSB, a.nrhs - c0);
    const int col = (wave & 3) * 16 + (lane & 15), rt0 = (wave >> 2) * 2;
    const long long i0 = (long long)ib * kSB;
    const int w = (int)min((long long)kSB, a.n - i0);
    acc_t acc[2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = (rt0 + i) * 16 + Op::row(lane, r);
            acc[i][r] = (row < w && col < nc) ? a.B[(i0 + row) + (long long)(c0 + col) * a.ldb] : T(0);
        }
    T va[kPer], vx[kPer];
    auto load = [&](int kb) {
        const long long k0 = (long long)kb * kSB;
        const int wk = (int)min((long long)kSB, a.n - k0);
        const bool diagonal = kb == ib, direct = a.lower ? kb < ib : kb > ib;
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            int r, k;
            va[e] = sym_element(a, i0, k0, w, wk, diagonal, direct, e, r, k);
        }
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            const int idx = threadIdx.x + kThreads * e, k = idx & (kSB - 1), c = idx >> 6;
            vx[e] = (k < wk && c < nc) ? a.X[(k0 + k) + (long long)(c0 + c) * a.ldx] : T(0);
        }
    };
    auto store = [&](int kb) {
        const bool diagonal = kb == ib, direct = a.lower ? kb < ib : kb > ib;
        const bool rows_contiguous = diagonal || direct;
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            const int idx = threadIdx.x + kThreads * e;
            const int r = rows_contiguous ? (idx & (kSB - 1)) : (idx >> 6);
            const int k = rows_contiguous ? (idx >> 6) : (idx & (kSB - 1));
            Ls[k][r] = va[e];
        }
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            const int idx = threadIdx.x + kThreads * e, k = idx & (kSB - 1), c = idx >> 6;
            Xs[k][c] = vx[e];
        }
    };
    load(0);
    for (int kb = 0; kb < a.nblk; ++kb) {
        store(kb);
        __syncthreads();
        if (kb + 1 < a.nblk) load(kb + 1);
        block_mma<T, false>(acc, Ls, Xs, T(-1));
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = (rt0 + i) * 16 + Op::row(lane, r);
            if (row < w && col < nc) a.R[(i0 + row) + (long long)(c0 + col) * a.ldr] = acc[i][r];
        }
}

__global__ void zero_ints_kernel(int* p, int count)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) p[i] = 0;
}

} // namespace solve_detail
} // namespace kern
} // namespace ddla
