// Batched Cholesky solve of many independent SPD systems of uneven order in
// one persistent launch (device code, included by kernels.hip).
//
// A grid of resident workgroups pulls matrix indices from an atomic queue, in
// the order given by `order` (largest order first, from order_by_blocks), so
// the long factorizations start first and the short ones fill the tail.  For
// matrix i (order n, stored triangle uplo), one 256-thread workgroup:
//   * factors it blocked right-looking on 64 x 64 blocks: the diagonal block
//     in LDS (potf2_tile.h), the blocks below it as A(r, j) D_j^T and the
//     trailing lower triangle A(r, c) -= L(r, j) L(c, j)^T on the matrix
//     cores, all in place in global memory (L2-resident at these orders);
//     D_j = L(j, j)^-1 goes to the workgroup's slot of xinv;
//   * if every pivot is positive, solves A X = B in place (forward sweep
//     y_j = D_j (b_j - sum_{i<j} L(j, i) y_i), backward sweep
//     x_j = D_j^T (y_j - sum_{r>j} L(r, j)^T x_r)) in 16-column slabs;
//   * writes logdet[i] = 2 sum_k log L(k, k) (double, fixed summation order)
//     and info[i] = 0; on a failed pivot p, info[i] = p (1-based), B is left
//     untouched and logdet[i] = NaN.  An order above 64 * max_blocks, a
//     negative order or a leading dimension below max(1, n) gives info[i] =
//     -1, logdet[i] = NaN and touches nothing.
#pragma once
#include <hip/hip_runtime.h>

#include "gemm_tile.h"
#include "potf2_tile.h"

namespace ddla {
namespace kern {
namespace batched_detail {

using gemm_detail::Mfma;
using potf2_detail::kB;
using potf2_detail::kL;

constexpr int kMaxBlocks = 16;  // orders up to 1024
constexpr int kThreads = 256;
constexpr int kPer = kB * kB / kThreads;
constexpr int kSlab = 16;       // right-hand-side columns per solve pass

template <class T>
struct BatchArgs {
    int batch, nrhs, upper, max_blocks;
    const int* n;
    T* const* A;
    const int* lda;
    T* const* B;
    const int* ldb;
    int* info;
    double* logdet;
    const int* order;  // processing order (a permutation of [0, batch))
    int* queue;        // work counter, zero on entry
    T* xinv;           // gridDim.x slots of max_blocks 64 x 64 inverses
    const int* first;  // null, or the device count of leading order[] entries handled elsewhere
};

struct View {
    long long lda;
    int upper;
};

// buf[c * kL + r] = L(i0 + r, j0 + c) (L view of a stored triangle), zero outside rows x cols.
template <class T>
__device__ inline void load_view_block(const T* A, View v, long long i0, long long j0, int rows, int cols, T* buf)
{
    T x[kPer];
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
        const int idx = threadIdx.x + kThreads * e, fast = idx & (kB - 1), slow = idx >> 6;
        const int r = v.upper ? slow : fast, c = v.upper ? fast : slow;
        x[e] = T(0);
        if (r < rows && c < cols)
            x[e] = v.upper ? A[(j0 + c) + (i0 + r) * v.lda] : A[(i0 + r) + (j0 + c) * v.lda];
    }
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
        const int idx = threadIdx.x + kThreads * e, fast = idx & (kB - 1), slow = idx >> 6;
        const int r = v.upper ? slow : fast, c = v.upper ? fast : slow;
        buf[c * kL + r] = x[e];
    }
}

// L(i0 + r, j0 + c) = buf[c * kL + r] for r < rows, c < cols (and r >= c when lower_only).
template <class T>
__device__ inline void store_view_block(T* A, View v, long long i0, long long j0, int rows, int cols,
                                        bool lower_only, const T* buf)
{
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
        const int idx = threadIdx.x + kThreads * e, fast = idx & (kB - 1), slow = idx >> 6;
        const int r = v.upper ? slow : fast, c = v.upper ? fast : slow;
        if (r < rows && c < cols && (!lower_only || r >= c)) {
            if (v.upper) A[(j0 + c) + (i0 + r) * v.lda] = buf[c * kL + r];
            else A[(i0 + r) + (j0 + c) * v.lda] = buf[c * kL + r];
        }
    }
}

// 64 x 64 accumulator tile of the workgroup (layout as in panel_tile.h).
template <class T>
struct Tile {
    typename Mfma<T>::acc_t v[2][2];
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
    __device__ void zero()
    {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int jj = 0; jj < 2; ++jj) v[i][jj] = typename Mfma<T>::acc_t{0, 0, 0, 0};
    }
    // += sign * As Bs^T, As(row, k) = As[k * kL + row], Bs(col, k) = Bs[k * kL + col]
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

// out(r, c) = sum_k M(r, k) V(k, c) for r < rows, k < cols, c < nc (nc <=
// kSlab): thread (r, q) = (tid % 64, tid / 64) sums k in [16q, 16q + 16)
// with M(r, k) = m[r * rs + k * ks] read from global memory (coalesced along
// r when rs = 1) and V(k, c) = V[c * kB + k] from LDS, then the four partial
// sums are added in a fixed order.  part holds kSlab x 4 x kB elements of LDS
// scratch; thread (r, q) receives column c = q + 4 s in out[s].
template <class T>
__device__ inline void block_apply(const T* m, long long rs, long long ks, int rows, int cols, const T* V, int nc,
                                   T* part, T (&out)[kSlab / 4])
{
    const int r = threadIdx.x & 63, q = threadIdx.x >> 6;
    T mv[16];
#pragma unroll
    for (int k = 0; k < 16; ++k)
        mv[k] = (r < rows && 16 * q + k < cols) ? m[r * rs + (long long)(16 * q + k) * ks] : T(0);
    for (int c = 0; c < nc; ++c) {
        T s = T(0);
#pragma unroll
        for (int k = 0; k < 16; ++k) s = fma(mv[k], V[c * kB + 16 * q + k], s);
        part[(c * 4 + q) * kB + r] = s;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kSlab / 4; ++s) {
        const int c = q + 4 * s;
        out[s] = c < nc ? ((part[(c * 4 + 0) * kB + r] + part[(c * 4 + 1) * kB + r]) +
                           (part[(c * 4 + 2) * kB + r] + part[(c * 4 + 3) * kB + r]))
                        : T(0);
    }
    __syncthreads();
}

// Block column j of the order-n member A (L view v, nb blocks): factors the
// diagonal block (potf2_tile.h), writes D_j = L(j, j)^-1 with exact zeros
// above its diagonal to dj (64 x 64, column-major), and (below) forms the
// blocks below as L(r, j) = A(r, j) D_j^T on the matrix cores, all in place.  Returns 0, or
// the 1-based global index of a failed pivot (then only L(j, j) is written).
// logpart (thread 0) receives sum log L(k, k) over the block (fixed-order
// butterfly) when with_log.  S, X: kB * kL elements of LDS each; rdiag: kB;
// s_flag: one int.  Without the blocks below, X keeps its scratch above the
// diagonal (dj still gets the zeros) and no barrier follows the stores: the
// caller synchronizes before it reads dj or reuses S and X.
template <class T, bool OUT_OF_LINE = false>
__device__ inline int factor_block_column(T* A, View v, int n, int nb, int j, T* dj, T* S, T* X, T* rdiag,
                                          int* s_flag, double& logpart, bool below = true, bool with_log = true)
{
    const int tid = threadIdx.x;
    const long long j0 = (long long)kB * j, lda = v.lda;
    const int wj = min(kB, n - (int)j0);
    T* Ajj = A + j0 + j0 * lda;
    potf2_detail::load_block<T>(v.upper != 0, wj, Ajj, lda, S);
    __syncthreads();
    const int f = OUT_OF_LINE ? potf2_detail::factor_invert64_call<T>(S, X, rdiag, s_flag)
                              : potf2_detail::factor_invert64<T>(S, X, rdiag, s_flag);
    potf2_detail::store_factor<T>(v.upper != 0, wj, S, Ajj, lda);
    if (f != 0) return (int)j0 + f;
    // log of the diagonal: wave 0, fixed-order butterfly
    if (with_log && tid < 64) {
        double l = tid < wj ? log((double)S[tid * kL + tid]) : 0.0;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
        if (tid == 0) logpart = l;
    }
    if (!below) {
        // X is complete (invert_lower ends with a barrier); dj gets the zeros from store_inverse
        potf2_detail::store_inverse<T>(X, dj);
        return 0;
    }
    // D_j with exact zeros above the diagonal: kept in X for the panel, saved for the solves
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
        const int idx = tid + kThreads * e, r = idx & (kB - 1), c = idx >> 6;
        if (r < c) X[c * kL + r] = T(0);
    }
    __syncthreads();
    potf2_detail::store_inverse<T>(X, dj);
    // blocks below: L(r, j) = A(r, j) D_j^T
    for (int r = j + 1; r < nb; ++r) {
        const long long r0 = (long long)kB * r;
        const int wr = min(kB, n - (int)r0);
        load_view_block<T>(A, v, r0, j0, wr, wj, S);
        __syncthreads();
        Tile<T> l;
        l.zero();
        l.mma_nt(S, X, T(1));
        __syncthreads();
        l.store(S);
        __syncthreads();
        store_view_block<T>(A, v, r0, j0, wr, wj, false, S);
        __syncthreads();
    }
    return 0;
}

// Forward then backward substitution for member A (L view v, order n, nb
// blocks) with its inverted diagonal blocks D_j at dinv + j * 64 * 64, in
// place on Bm (n x nrhs, leading dimension ldb), 16 columns at a time.
// part / V: LDS scratch (kSlab * 4 * kB and kSlab * kB elements).
template <class T>
__device__ inline void solve_member(const T* A, View v, int n, int nb, T* Bm, long long ldb, int nrhs, const T* dinv,
                                    T* part, T* V)
{
    const int tid = threadIdx.x;
    const long long lda = v.lda;
    for (int c0 = 0; c0 < nrhs; c0 += kSlab) {
        const int nc = min(kSlab, nrhs - c0);
        // forward: y_j = D_j (b_j - sum_{i<j} L(j, i) y_i), y written over b
        for (int j = 0; j < nb; ++j) {
            const long long j0 = (long long)kB * j;
            const int wj = min(kB, n - (int)j0);
            T acc[kSlab / 4];
            {
                const int r = tid & 63, q = tid >> 6;
#pragma unroll
                for (int s = 0; s < kSlab / 4; ++s) {
                    const int c = q + 4 * s;
                    acc[s] = (r < wj && c < nc) ? Bm[(j0 + r) + (c0 + c) * ldb] : T(0);
                }
            }
            for (int i = 0; i < j; ++i) {
                const long long i0 = (long long)kB * i;
                // V = y_i
                const int wi = min(kB, n - (int)i0);
                for (int e = tid; e < kB * kSlab; e += kThreads) {
                    const int k = e & (kB - 1), c = e >> 6;
                    V[c * kB + k] = (c < nc && k < wi) ? Bm[(i0 + k) + (c0 + c) * ldb] : T(0);
                }
                __syncthreads();
                T t[kSlab / 4];
                // M = L(j, i): element (r, k) at (j0 + r, i0 + k) of the L view
                const T* m = v.upper ? A + i0 + j0 * lda : A + j0 + i0 * lda;
                block_apply<T>(m, v.upper ? lda : 1, v.upper ? 1 : lda, wj, wi, V, nc, part, t);
#pragma unroll
                for (int s = 0; s < kSlab / 4; ++s) acc[s] -= t[s];
            }
            // V = acc, then y_j = D_j V
            {
                const int r = tid & 63, q = tid >> 6;
#pragma unroll
                for (int s = 0; s < kSlab / 4; ++s) {
                    const int c = q + 4 * s;
                    if (c < kSlab) V[c * kB + r] = acc[s];
                }
            }
            __syncthreads();
            T y[kSlab / 4];
            block_apply<T>(dinv + j0 * kB, 1, kB, kB, kB, V, nc, part, y);
            {
                const int r = tid & 63, q = tid >> 6;
#pragma unroll
                for (int s = 0; s < kSlab / 4; ++s) {
                    const int c = q + 4 * s;
                    if (r < wj && c < nc) Bm[(j0 + r) + (c0 + c) * ldb] = y[s];
                }
            }
            __syncthreads();
        }
        // backward: x_j = D_j^T (y_j - sum_{r>j} L(r, j)^T x_r)
        for (int j = nb - 1; j >= 0; --j) {
            const long long j0 = (long long)kB * j;
            const int wj = min(kB, n - (int)j0);
            T acc[kSlab / 4];
            {
                const int r = tid & 63, q = tid >> 6;
#pragma unroll
                for (int s = 0; s < kSlab / 4; ++s) {
                    const int c = q + 4 * s;
                    acc[s] = (r < wj && c < nc) ? Bm[(j0 + r) + (c0 + c) * ldb] : T(0);
                }
            }
            for (int i = j + 1; i < nb; ++i) {
                const long long i0 = (long long)kB * i;
                const int wi = min(kB, n - (int)i0);
                for (int e = tid; e < kB * kSlab; e += kThreads) {
                    const int k = e & (kB - 1), c = e >> 6;
                    V[c * kB + k] = (c < nc && k < wi) ? Bm[(i0 + k) + (c0 + c) * ldb] : T(0);
                }
                __syncthreads();
                T t[kSlab / 4];
                // M = L(i, j)^T: element (r, k) = L(i0 + k, j0 + r)
                const T* m = v.upper ? A + j0 + i0 * lda : A + i0 + j0 * lda;
                block_apply<T>(m, v.upper ? 1 : lda, v.upper ? lda : 1, wj, wi, V, nc, part, t);
#pragma unroll
                for (int s = 0; s < kSlab / 4; ++s) acc[s] -= t[s];
            }
            {
                const int r = tid & 63, q = tid >> 6;
#pragma unroll
                for (int s = 0; s < kSlab / 4; ++s) {
                    const int c = q + 4 * s;
                    if (c < kSlab) V[c * kB + r] = acc[s];
                }
            }
            __syncthreads();
            T x[kSlab / 4];
            // D_j^T: element (r, k) = D_j(k, r) = slot[j][k + 64 r]
            block_apply<T>(dinv + j0 * kB, kB, 1, kB, kB, V, nc, part, x);
            {
                const int r = tid & 63, q = tid >> 6;
#pragma unroll
                for (int s = 0; s < kSlab / 4; ++s) {
                    const int c = q + 4 * s;
                    if (r < wj && c < nc) Bm[(j0 + r) + (c0 + c) * ldb] = x[s];
                }
            }
            __syncthreads();
        }
    }
}

// OCC = resident workgroups per CU the register allocation is sized for (1:
// no spills; 2: two workgroups per CU share each SIMD, at the cost of a few
// spilled registers); the autotuner picks one.
template <class T, int OCC>
__global__ __launch_bounds__(kThreads, OCC) void batched_pposv_kernel(BatchArgs<T> a)
{
    __shared__ T S[kB * kL], X[kB * kL], rdiag[kB];
    __shared__ int s_item, s_flag;
    static_assert(kSlab * 4 * kB <= kB * kL && kSlab * kB <= kB * kL, "solve scratch must fit the factor buffers");
    T* const part = S;  // the solves reuse the factorization's buffers
    T* const V = X;
    const int tid = threadIdx.x;
    T* const slot = a.xinv + (long long)blockIdx.x * a.max_blocks * kB * kB;

    for (;;) {
        if (tid == 0) {
            const int q = atomicAdd(a.queue, 1) + (a.first != nullptr ? *a.first : 0);
            s_item = q < a.batch ? a.order[q] : -1;
        }
        __syncthreads();
        const int item = s_item;
        __syncthreads();
        if (item < 0) return;

        const int n = a.n[item];
        const long long lda = a.lda[item], ldb = a.ldb[item];
        const int nb = (n + kB - 1) / kB;
        if (n < 0 || n > a.max_blocks * kB || lda < (n > 1 ? n : 1) || (a.nrhs > 0 && ldb < (n > 1 ? n : 1))) {
            if (tid == 0) {
                a.info[item] = -1;
                a.logdet[item] = __builtin_nan("");
            }
            continue;
        }
        T* A = a.A[item];
        const View v{lda, a.upper};
        int fail = 0;
        double logsum = 0.0;  // meaningful in thread 0

        for (int j = 0; j < nb && fail == 0; ++j) {
            const long long j0 = (long long)kB * j;
            const int wj = min(kB, n - (int)j0);
            double l = 0.0;
            // two workgroups per CU leave too few registers to inline the factorization
            const int f = factor_block_column<T, (OCC > 1)>(A, v, n, nb, j, slot + (long long)j * kB * kB, S, X,
                                                            rdiag, &s_flag, l);
            if (f != 0) {
                fail = f;
                break;
            }
            logsum += l;
            // trailing update: A(r, c) -= L(r, j) L(c, j)^T, j < c <= r
            for (int c = j + 1; c < nb; ++c) {
                const long long c0 = (long long)kB * c;
                const int wc = min(kB, n - (int)c0);
                __syncthreads();
                load_view_block<T>(A, v, c0, j0, wc, wj, X);  // L(c, j)
                for (int r = c; r < nb; ++r) {
                    const long long r0 = (long long)kB * r;
                    const int wr = min(kB, n - (int)r0);
                    load_view_block<T>(A, v, r0, c0, wr, wc, S);  // A(r, c)
                    __syncthreads();
                    Tile<T> acc;
                    acc.load(S);
                    __syncthreads();
                    load_view_block<T>(A, v, r0, j0, wr, wj, S);  // L(r, j)
                    __syncthreads();
                    acc.mma_nt(S, X, T(-1));
                    __syncthreads();
                    acc.store(S);
                    __syncthreads();
                    store_view_block<T>(A, v, r0, c0, wr, wc, r == c, S);
                    __syncthreads();
                }
            }
        }
        __syncthreads();
        if (fail != 0) {
            if (tid == 0) {
                a.info[item] = fail;
                a.logdet[item] = __builtin_nan("");
            }
            continue;
        }

        solve_member<T>(A, v, n, nb, a.B[item], ldb, a.nrhs, slot, part, V);
        if (tid == 0) {
            a.info[item] = 0;
            a.logdet[item] = 2.0 * logsum;
        }
    }
}

// ---------------------------------------------------------------------------
// Blocked path for the larger members: the members with more than tb blocks
// (the leading count entries of order[], which is sorted by block count) are
// factored in steps of one or two 64-column block columns: per block column
// j, blocked_panel_kernel factors every such member's diagonal block (and
// keeps D_j) and one gemmVbatched launch forms all their panels L(r, j) =
// A(r, j) D_j^T in place; inside a two-column step the first block column
// then updates only the second, and the step's last block column applies the
// whole step's panel to the trailing matrices in one more launch (k = 64 or
// 128; arrays from blocked_prep_kernel), so a two-column step reads and
// writes the trailing matrices half as often; then blocked_solve_kernel
// solves them.  Each member's D_j is kept at
// dinv + (b * max_blocks + j) * 64 * 64 (b = its position in order[]) and its
// per-block log sums at logparts[item * kMaxBlocks + j].
// ---------------------------------------------------------------------------
template <class T>
struct BlockedArgs {
    int batch, upper, max_blocks, nrhs;
    const int* n;
    T* const* A;
    const int* lda;
    const int* ldb;
    int* info;
    const int* order;
    const int* count;
    T* dinv;
    double* logparts;
};

// *count = the number of valid members with more than tb blocks (one 256-thread workgroup).
__global__ __launch_bounds__(256) void count_large_kernel(int batch, int max_blocks, const int* n, int tb, int* count)
{
    __shared__ int part[256];
    int c = 0;
    for (int i = threadIdx.x; i < batch; i += 256) {
        const int m = n[i];
        if (m >= 0 && m <= max_blocks * kB && (m + kB - 1) / kB > tb) ++c;
    }
    part[threadIdx.x] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *count = part[0];
}

// Diagonal block j of member order[b] for every b < *count (one workgroup each).
// Members are taken as valid (count_large_kernel counts only valid orders);
// the leading dimensions are checked at j = 0 (info = -1, nothing touched,
// as on the fused path).  info = 0 is set at j = 0; a member whose info is
// nonzero is skipped.
template <class T>
__global__ __launch_bounds__(kThreads) void blocked_panel_kernel(BlockedArgs<T> a, int j)
{
    __shared__ T S[kB * kL], X[kB * kL], rdiag[kB];
    __shared__ int s_flag;
    const int b = blockIdx.x;
    if (b >= *a.count) return;
    const int item = a.order[b];
    const int n = a.n[item];
    const int nb = (n + kB - 1) / kB;
    if (j >= nb) return;
    const long long lda = a.lda[item];
    if (j == 0) {
        if (lda < (n > 1 ? n : 1) || (a.nrhs > 0 && a.ldb[item] < (n > 1 ? n : 1))) {
            if (threadIdx.x == 0) a.info[item] = -1;
            return;
        }
        if (threadIdx.x == 0) a.info[item] = 0;
    } else if (a.info[item] != 0) {
        return;
    }
    double lp = 0.0;
    const int f = factor_block_column<T>(a.A[item], View{lda, a.upper}, n, nb, j,
                                         a.dinv + ((long long)b * a.max_blocks + j) * kB * kB, S, X, rdiag, &s_flag, lp,
                                         false);
    if (threadIdx.x == 0) {
        if (f != 0) a.info[item] = f;
        else a.logparts[(long long)item * kMaxBlocks + j] = lp;
    }
}

// Arrays of block column j's gemmVbatched launches, in a step of block
// columns J .. (J <= j): for b < *count with a successful member of order >
// 64 (j + 1): m[b] = n - 64 (j + 1) (rows below the diagonal block), k[b] =
// 64, ld[b] = lda, ld64[b] = 64, P[b] = the stored address of L-view element
// (64 (j + 1), 64 j) (block column j's panel), Q[b] that of (64 (j + 1),
// 64 J) (the step's panel, kq[b] = 64 (j + 1 - J) columns), C[b] that of
// (64 (j + 1), 64 (j + 1)) (the trailing matrix), D[b] = D_j.  The panel
// launch is m x 64 (lower: L = A D_j^T) or 64 x m (upper: the stored panel
// U = D_j U) -- pm[b], pn[b]; the update of block column j + 1 alone is
// m x min(64, m) (lower) or its mirror (upper) -- cm[b], cn[b].  Every size
// is 0 (no work) otherwise.
template <class T>
struct StepArrays {
    int *m, *k, *kq, *ld, *ld64, *pm, *pn, *cm, *cn;
    T **P, **Q, **C, **D;
};

template <class T>
__global__ void blocked_prep_kernel(BlockedArgs<T> a, int j, int J, StepArrays<T> s)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= a.batch) return;
    int mm = 0;
    if (b < *a.count) {
        const int item = a.order[b];
        const int n = a.n[item];
        const long long r0 = (long long)kB * (j + 1), j0 = (long long)kB * j, q0 = (long long)kB * J;
        if (n > r0 && a.info[item] == 0) {
            const long long lda = a.lda[item];
            T* A = a.A[item];
            mm = n - (int)r0;
            s.ld[b] = (int)lda;
            s.P[b] = a.upper ? A + j0 + r0 * lda : A + r0 + j0 * lda;
            s.Q[b] = a.upper ? A + q0 + r0 * lda : A + r0 + q0 * lda;
            s.C[b] = A + r0 + r0 * lda;
            s.D[b] = a.dinv + ((long long)b * a.max_blocks + j) * kB * kB;
        }
    }
    const int c64 = mm < kB ? mm : kB;
    s.m[b] = mm;
    s.pm[b] = a.upper ? (mm > 0 ? kB : 0) : mm;
    s.pn[b] = a.upper ? mm : (mm > 0 ? kB : 0);
    s.cm[b] = a.upper ? c64 : mm;
    s.cn[b] = a.upper ? mm : c64;
    s.k[b] = kB;
    s.kq[b] = kB * (j + 1 - J);
    s.ld64[b] = kB;
}

// ---------------------------------------------------------------------------
// Left-looking persistent kernel: one 256-thread workgroup per member, every
// member in the same launch.  Block column j (64 columns) is formed once:
//   C = A(j0:n, j) - L(j0:n, 0:j0) L(j, 0:j0)^T      (one MFMA product, the
//       gemm_tile engine over 128 x 64 tiles, lower triangle of the diagonal
//       block masked; the accumulators stay in registers across all of k),
//   L(j, j), D_j = L(j, j)^-1 from the diagonal block (potf2_tile.h),
//   L(j0 + 64:n, j) = C(j0 + 64:n, j) D_j^T            (in place, MFMA).
// So each block column is written twice in all, where the right-looking
// kernels read and write the whole trailing matrix at every step.  The
// solve then runs as column sweeps (solve_member_ll) and the log-determinant
// is one fixed-order sum over the factor's diagonal per member.  Uplo 'U'
// runs the same steps on the stored upper triangle (transposed products).
// ---------------------------------------------------------------------------

// LDS of the left-looking kernel, in elements of T: the larger of the
// products' double-buffered slabs (128 x 64 tiles), the diagonal block's S,
// X and rdiag, and the solve's kLLCols-column right-hand-side buffer.
constexpr int kLLCols = 4;
template <class T>
constexpr int ll_lds_elems()
{
    constexpr int g = gemm_detail::lds_elems<128, 64, 16>();
    constexpr int f = 2 * kB * kL + kB;
    constexpr int s = kMaxBlocks * kB * kLLCols;
    return g > f ? (g > s ? g : s) : (f > s ? f : s);
}

// The value of p (v) in the workgroup's first lane, held in scalar registers.
template <class P>
__device__ inline P* uniform_ptr(P* p)
{
    const unsigned long long u = reinterpret_cast<unsigned long long>(p);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
    return reinterpret_cast<P*>(((unsigned long long)hi << 32) | lo);
}

__device__ inline long long uniform_ll(long long v)
{
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)((unsigned long long)v >> 32));
    return (long long)(((unsigned long long)hi << 32) | lo);
}

// One product of the left-looking kernel, tile after tile: BN == 64 runs
// down the rows (128 x 64 tiles), BM == 64 along the columns (64 x 128
// tiles).  Out of line, with its arguments made uniform again: inlined into
// the persistent loop, the address arithmetic of every product would be
// hoisted out of that loop and stay live throughout (the kernel then spilled
// about 400 registers).  OCC: the calling kernel's workgroups per CU, so that
// each instantiation is reached from one kernel and compiled to its register
// budget.
template <class T, int OCC, int BM, int BN, bool TA, bool TB>
__device__ __attribute__((noinline)) void ll_product(gemm_detail::GemmProblem<T> q, T* smem)
{
    gemm_detail::GemmProblem<T> p;
    p.m = __builtin_amdgcn_readfirstlane(q.m);
    p.n = __builtin_amdgcn_readfirstlane(q.n);
    p.k = __builtin_amdgcn_readfirstlane(q.k);
    p.alpha = q.alpha;
    p.beta = q.beta;
    p.A = uniform_ptr(q.A);
    p.B = uniform_ptr(q.B);
    p.C = uniform_ptr(q.C);
    p.lda = uniform_ll(q.lda);
    p.ldb = uniform_ll(q.ldb);
    p.ldc = uniform_ll(q.ldc);
    p.tri = __builtin_amdgcn_readfirstlane(q.tri);
    p.diag = uniform_ll(q.diag);
    p.vec_ab = __builtin_amdgcn_readfirstlane(q.vec_ab);
    static_assert(BM == 64 || BN == 64, "one tile dimension spans the block column");
    if (BN == 64) {
        for (int m0 = 0; m0 < p.m; m0 += BM) gemm_detail::gemm_tile<T, BM, BN, 16, TA, TB, 2, false>(p, m0, 0, smem);
    } else {
        for (int n0 = 0; n0 < p.n; n0 += BN) gemm_detail::gemm_tile<T, BM, BN, 16, TA, TB, 2, false>(p, 0, n0, smem);
    }
}

// Solve A X = B in place on B (n x nrhs, leading dimension ldb) from the
// member's factor (L view v) and its diagonal-block inverses D_j at dinv +
// j * 64 * 64 (column-major, zeros above the diagonal), NC right-hand-side
// columns per pass held in LDS (y, n x NC).  Forward: for each block j,
// y_j = D_j y_j, then y_i -= L(i, j) y_j for every row i below; backward: for
// j from the last block, x_j = D_j^T y_j, then y_i -= L(j, i)^T x_j for every
// row i above.  So each sweep takes one step per block, and every access to
// L runs along a column (lower rows i: consecutive lanes read consecutive
// rows) or along a row of L (each lane reads one contiguous piece).
template <class T, int NC, int OCC>
__device__ __attribute__((noinline)) void solve_member_ll(const T* A, View v, int n, int nb, T* Bm, long long ldb,
                                                          int nrhs, const T* dinv, T* y)
{
    const int tid = threadIdx.x;
    A = uniform_ptr(A);
    Bm = uniform_ptr(Bm);
    dinv = uniform_ptr(dinv);
    v.lda = uniform_ll(v.lda);
    v.upper = __builtin_amdgcn_readfirstlane(v.upper);
    n = __builtin_amdgcn_readfirstlane(n);
    nb = __builtin_amdgcn_readfirstlane(nb);
    ldb = uniform_ll(ldb);
    nrhs = __builtin_amdgcn_readfirstlane(nrhs);
    const long long lda = v.lda;
    for (int c0 = 0; c0 < nrhs; c0 += NC) {
        const int nc = min(NC, nrhs - c0);
        for (int e = tid; e < n * NC; e += kThreads) {
            const int i = e / NC, c = e % NC;
            y[e] = c < nc ? Bm[i + (long long)(c0 + c) * ldb] : T(0);
        }
        __syncthreads();
        // Lane layout of a step: wavefront w, lane l handles row 16 w + l / 4 of a
        // 64-row pass and the quarter k in [16 (l % 4), 16 (l % 4) + 16) of the
        // 64-long dot products, so each lane has 16 independent loads in flight;
        // the four quarters meet by two fixed-order cross-lane sums.
        const int lrow = 16 * (tid >> 6) + ((tid & 63) >> 2), kq = 16 * (tid & 3);
        // sum over the four lanes of a row, in a fixed order
        auto quarter_sum = [&](T x) {
            x += __shfl_xor(x, 1);
            return x + __shfl_xor(x, 2);
        };
        // z(r) = sum_k D(r, k) y_j(k) (or D(k, r)) for the block's rows r < wj, written over y_j
        auto diag_product = [&](const T* D, int j0, int wj, bool transposed) {
            T dv[16];
#pragma unroll
            for (int t = 0; t < 16; ++t) dv[t] = transposed ? D[kq + t + kB * lrow] : D[lrow + kB * (kq + t)];
            T z[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                T acc = T(0);
#pragma unroll
                for (int t = 0; t < 16; ++t) acc = fma(dv[t], kq + t < wj ? y[(j0 + kq + t) * NC + c] : T(0), acc);
                z[c] = quarter_sum(acc);
            }
            __syncthreads();
            if (lrow < wj && (tid & 3) == 0) {
#pragma unroll
                for (int c = 0; c < NC; ++c)
                    if (c < nc) y[(j0 + lrow) * NC + c] = z[c];
            }
            __syncthreads();
        };
        // y_i -= sum_k L(row i, k) y_j(k) for rows i0 <= i < i1 (li(i): the row's
        // first element, ks: the stride along it)
        auto rows_update = [&](int i0, int i1, int j0, int wj, bool below) {
            for (int base = i0; base < i1; base += 64) {
                const int i = base + lrow;
                const bool row = i < i1;
                const T* li = below ? (v.upper ? A + j0 + (long long)i * lda : A + i + (long long)j0 * lda)
                                    : (v.upper ? A + i + (long long)j0 * lda : A + j0 + (long long)i * lda);
                const long long ks = below ? (v.upper ? 1 : lda) : (v.upper ? lda : 1);
                T lv[16];
#pragma unroll
                for (int t = 0; t < 16; ++t) lv[t] = row && kq + t < wj ? li[(kq + t) * ks] : T(0);
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    T acc = T(0);
#pragma unroll
                    for (int t = 0; t < 16; ++t) acc = fma(lv[t], kq + t < wj ? y[(j0 + kq + t) * NC + c] : T(0), acc);
                    acc = quarter_sum(acc);
                    if (row && (tid & 3) == 0 && c < nc) y[i * NC + c] -= acc;
                }
            }
            __syncthreads();
        };
        for (int j = 0; j < nb; ++j) {
            const int j0 = kB * j, wj = min(kB, n - j0);
            diag_product(dinv + (long long)j * kB * kB, j0, wj, false);
            rows_update(j0 + wj, n, j0, wj, true);
        }
        for (int j = nb - 1; j >= 0; --j) {
            const int j0 = kB * j, wj = min(kB, n - j0);
            diag_product(dinv + (long long)j * kB * kB, j0, wj, true);
            rows_update(0, j0, j0, wj, false);  // L(j0 + k, i) for the rows i above the block
        }
        for (int e = tid; e < n * NC; e += kThreads) {
            const int i = e / NC, c = e % NC;
            if (c < nc) Bm[i + (long long)(c0 + c) * ldb] = y[e];
        }
        __syncthreads();
    }
}

// Block column j of the member (L view v, order n): the left-looking product
// C = A(j0:n, j) - L(j0:n, 0:j0) L(j, 0:j0)^T in place (j > 0), the diagonal
// block (factor_block_column: L(j, j), D_j to dj), then the panel
// L(j0 + 64:n, j) = C D_j^T in place.  Returns 0 or the failed pivot
// (1-based).  smem: ll_lds_elems<T>() elements.
template <class T, int OCC>
__device__ inline int ll_block_column(T* A, View v, int n, int nb, int j, T* dj, T* smem, int* s_flag)
{
    using gemm_detail::GemmProblem;
    const long long lda = v.lda;
    const int j0 = kB * j, wj = min(kB, n - j0), m = n - j0;
    constexpr int VW = Mfma<T>::kVec;
    const bool aligned = ((uintptr_t)A % 16 == 0) && lda % VW == 0;
    if (j > 0) {
        GemmProblem<T> p;
        p.k = j0;
        p.alpha = T(-1);
        p.beta = T(1);
        p.lda = p.ldb = p.ldc = lda;
        p.diag = 0;
        p.vec_ab = aligned;
        if (!v.upper) {
            // C_L(i, jj) -= sum_k L(j0 + i, k) L(j0 + jj, k), i >= jj
            p.m = m;
            p.n = wj;
            p.A = A + j0;
            p.B = A + j0;
            p.C = A + j0 + (long long)j0 * lda;
            p.tri = gemm_detail::kTriLowerD;
            ll_product<T, OCC, 128, 64, false, true>(p, smem);
        } else {
            // stored U(j0 + jj, j0 + i) -= sum_k U(k, j0 + jj) U(k, j0 + i), jj <= i
            p.m = wj;
            p.n = m;
            p.A = A + (long long)j0 * lda;
            p.B = A + (long long)j0 * lda;
            p.C = A + j0 + (long long)j0 * lda;
            p.tri = gemm_detail::kTriUpperD;
            ll_product<T, OCC, 64, 128, true, false>(p, smem);
        }
        __syncthreads();
    }
    T* S = smem;
    T* X = smem + kB * kL;
    T* rdiag = smem + 2 * kB * kL;
    double unused = 0.0;
    const int f = factor_block_column<T, (OCC > 1)>(A, v, n, nb, j, dj, S, X, rdiag, s_flag, unused, false, false);
    if (f != 0) return f;
    __syncthreads();
    if (m > kB) {
        GemmProblem<T> p;
        p.k = kB;
        p.alpha = T(1);
        p.beta = T(0);
        p.diag = 0;
        p.tri = gemm_detail::kTriNoneD;
        p.vec_ab = aligned;
        if (!v.upper) {
            // L(j0 + 64 + i, j0 + jj) = sum_k C(j0 + 64 + i, j0 + k) D_j(jj, k)
            p.m = m - kB;
            p.n = kB;
            p.A = A + j0 + kB + (long long)j0 * lda;
            p.lda = lda;
            p.B = dj;
            p.ldb = kB;
            p.C = A + j0 + kB + (long long)j0 * lda;
            p.ldc = lda;
            ll_product<T, OCC, 128, 64, false, true>(p, smem);
        } else {
            // stored U(j0 + jj, j0 + 64 + i) = sum_k D_j(jj, k) U(j0 + k, j0 + 64 + i)
            p.m = kB;
            p.n = m - kB;
            p.A = dj;
            p.lda = kB;
            p.B = A + j0 + (long long)(j0 + kB) * lda;
            p.ldb = lda;
            p.C = A + j0 + (long long)(j0 + kB) * lda;
            p.ldc = lda;
            ll_product<T, OCC, 64, 128, false, false>(p, smem);
        }
        __syncthreads();
    }
    return 0;
}

// OCC = resident workgroups per CU the register allocation is sized for.
template <class T, int OCC>
__global__ __launch_bounds__(kThreads, OCC) void batched_ll_kernel(BatchArgs<T> a)
{
    __shared__ T smem[ll_lds_elems<T>()];
    __shared__ int s_item, s_flag;
    __shared__ double s_log[kThreads / 64];
    const int tid = threadIdx.x;
    T* const slot = a.xinv + (long long)blockIdx.x * a.max_blocks * kB * kB;

    for (;;) {
        if (tid == 0) {
            const int q = atomicAdd(a.queue, 1);
            s_item = q < a.batch ? a.order[q] : -1;
        }
        __syncthreads();
        // uniform over the workgroup: every per-member value below then lives
        // in scalar registers (read back from LDS it would be per lane)
        const int item = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (item < 0) return;

        const int n = __builtin_amdgcn_readfirstlane(a.n[item]);
        const long long lda = __builtin_amdgcn_readfirstlane(a.lda[item]);
        const long long ldb = a.nrhs > 0 ? __builtin_amdgcn_readfirstlane(a.ldb[item]) : 1;
        const int nb = (n + kB - 1) / kB;
        if (n < 0 || n > a.max_blocks * kB || lda < (n > 1 ? n : 1) || (a.nrhs > 0 && ldb < (n > 1 ? n : 1))) {
            if (tid == 0) {
                a.info[item] = -1;
                a.logdet[item] = __builtin_nan("");
            }
            continue;
        }
        T* A = uniform_ptr(a.A[item]);
        const View v{lda, a.upper};
        int fail = 0;
        for (int j = 0; j < nb; ++j) {
            fail = ll_block_column<T, OCC>(A, v, n, nb, j, slot + (long long)j * kB * kB, smem, &s_flag);
            if (fail != 0) break;
        }
        __syncthreads();
        if (fail != 0) {
            if (tid == 0) {
                a.info[item] = fail;
                a.logdet[item] = __builtin_nan("");
            }
            continue;
        }
        // sum log L(k, k) once per member, in a fixed order: thread t takes
        // k = t, t + 256, ..., then a butterfly per wavefront and the four
        // wavefronts' sums in order (the factor's stores are visible to the
        // workgroup after the barrier above)
        {
            double l = 0.0;
            for (int k = tid; k < n; k += kThreads) l += log((double)A[k + (long long)k * lda]);
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
            if ((tid & 63) == 0) s_log[tid >> 6] = l;
        }
        __syncthreads();
        const double logsum = (s_log[0] + s_log[1]) + (s_log[2] + s_log[3]);  // used by thread 0
        if (a.nrhs == 1) solve_member_ll<T, 1, OCC>(A, v, n, nb, uniform_ptr(a.B[item]), ldb, 1, slot, smem);
        else if (a.nrhs > 1)
            solve_member_ll<T, kLLCols, OCC>(A, v, n, nb, uniform_ptr(a.B[item]), ldb, a.nrhs, slot, smem);
        if (tid == 0) {
            a.info[item] = 0;
            a.logdet[item] = 2.0 * logsum;
        }
    }
}

// Solves and log-determinants of the members factored on the blocked path.
template <class T>
// (four workgroups per CU: the solve is latency-bound, and its out-of-line
// instance for this kernel is compiled to that register budget)
__global__ __launch_bounds__(kThreads, 4) void blocked_solve_kernel(BlockedArgs<T> a, T* const* B, double* logdet)
{
    __shared__ T part[kSlab * 4 * kB];
    static_assert(kSlab * 4 * kB >= kMaxBlocks * kB * kLLCols, "the solve's right-hand sides must fit");
    const int b = blockIdx.x;
    if (b >= *a.count) return;
    const int item = a.order[b];
    const int n = a.n[item];
    const int nb = (n + kB - 1) / kB;
    if (a.info[item] != 0) {
        if (threadIdx.x == 0) logdet[item] = __builtin_nan("");
        return;
    }
    const View v{a.lda[item], a.upper};
    const T* dinv = a.dinv + (long long)b * a.max_blocks * kB * kB;
    if (a.nrhs == 1) solve_member_ll<T, 1, 4>(a.A[item], v, n, nb, B[item], a.ldb[item], 1, dinv, part);
    else if (a.nrhs > 1) solve_member_ll<T, kLLCols, 4>(a.A[item], v, n, nb, B[item], a.ldb[item], a.nrhs, dinv, part);
    if (threadIdx.x == 0) {
        double sum = 0.0;
        for (int jj = 0; jj < nb; ++jj) sum += a.logparts[(long long)item * kMaxBlocks + jj];
        logdet[item] = 2.0 * sum;
    }
}

// order[] = the matrix indices sorted by block count, largest first, stable
// (one 256-thread workgroup: per-thread bucket counts over a contiguous index
// range, an exclusive scan in bucket-major order, then a scatter in index
// order).  Orders outside [0, 64 * max_blocks] go last.
constexpr int kOrderThreads = 256;
__global__ __launch_bounds__(kOrderThreads) void order_by_blocks_kernel(int batch, int max_blocks, const int* n,
                                                                        int* order)
{
    constexpr int kBuckets = kMaxBlocks + 2;
    __shared__ int cnt[kBuckets][kOrderThreads];
    __shared__ int total[kBuckets];
    const int tid = threadIdx.x;
    const int per = (batch + kOrderThreads - 1) / kOrderThreads;
    const int lo = min(batch, tid * per), hi = min(batch, lo + per);
    auto bucket = [&](int i) {
        const int m = n[i];
        if (m < 0 || m > max_blocks * kB) return kBuckets - 1;
        return kMaxBlocks - (m + kB - 1) / kB;  // 16 blocks -> 0, an empty matrix -> 16
    };
    for (int b = 0; b < kBuckets; ++b) cnt[b][tid] = 0;
    for (int i = lo; i < hi; ++i) ++cnt[bucket(i)][tid];
    __syncthreads();
    if (tid < kBuckets) {
        int s = 0;
        for (int t = 0; t < kOrderThreads; ++t) {
            const int c = cnt[tid][t];
            cnt[tid][t] = s;
            s += c;
        }
        total[tid] = s;
    }
    __syncthreads();
    if (tid < kBuckets) {
        int base = 0;
        for (int b = 0; b < tid; ++b) base += total[b];
        for (int t = 0; t < kOrderThreads; ++t) cnt[tid][t] += base;
    }
    __syncthreads();
    for (int i = lo; i < hi; ++i) order[cnt[bucket(i)][tid]++] = i;
}

} // namespace batched_detail
} // namespace kern
} // namespace ddla
