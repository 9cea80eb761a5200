// Cholesky factorization and triangular inverse of one 64 x 64 diagonal block
// held in LDS, for a 256-thread workgroup on gfx950 (device code, included by
// kernels.hip).
//
// The block S is column-major in LDS with column stride kL (S[c * kL + r] =
// A(r, c)) and only its lower triangle is referenced.  The factorization is
// blocked in 16-column strips:
//   * one wavefront factors a strip and solves every row below it in one
//     pass: lane i holds row 16b + i of the strip in registers, and at column
//     step j the pivot and the 15 - j multipliers of the columns to its right
//     are broadcast with v_readlane, so a strip needs no barrier and no LDS
//     traffic until it is written back;
//   * all four wavefronts then apply the strip to the trailing lower triangle
//     on the matrix cores (16 x 16 tiles, v_mfma_*_16x16x4).
// The inverse X = L^-1 is formed blockwise: the four 16 x 16 diagonal
// inverses by forward substitution (one wavefront each, in parallel), then
// X(2p+1, 2p) = -D(2p+1) L(2p+1, 2p) D(2p) and finally X21 = -X22 L21 X11 on
// 32-row halves, all products on the matrix cores.  Pivots use the hardware
// reciprocal square root refined by Newton steps (1 / L(j, j) and
// L(j, j) = d / sqrt(d) to within an ulp or two of the correctly rounded
// values).
#pragma once
#include <hip/hip_runtime.h>

#include "gemm_tile.h"

namespace ddla {
namespace kern {
namespace potf2_detail {

using gemm_detail::Mfma;

constexpr int kB = 64;  // block order
constexpr int kL = 65;  // LDS column stride, in elements
constexpr int kW = 16;  // strip width

__device__ inline double lane_value(double x, int l)
{
    const long long b = __double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ inline float lane_value(float x, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

// 1 / sqrt(d) for d > 0: hardware estimate plus Newton steps
// y <- y + y (1/2 - (d/2) y^2).
__device__ inline double rsqrt_refined(double d)
{
    double y = __builtin_amdgcn_rsq(d);
    const double h = 0.5 * d;
    double e = fma(-h * y, y, 0.5);
    y = fma(y, e, y);
    e = fma(-h * y, y, 0.5);
    return fma(y, e, y);
}

__device__ inline float rsqrt_refined(float d)
{
    float y = __builtin_amdgcn_rsqf(d);
    const float h = 0.5f * d;
    const float e = fmaf(-h * y, y, 0.5f);
    return fmaf(y, e, y);
}

// Factors strip b (columns 16b .. 16b+15, rows 16b .. 63) of the lower
// triangle of S in one wavefront and writes it back; rdiag[16b + j] receives
// 1 / L(16b + j, 16b + j).  Returns 0, or the 1-based block index of the
// first pivot that is not positive (or is NaN); the columns from that pivot on
// then hold no meaningful values.  The 16 column steps are straight-line code
// with no branch (the failure is recorded, not acted on, and rdiag is stored
// once at the end), so the compiler overlaps one column's pivot chain with the
// previous column's remaining updates: this halved the strip's time.
template <class T>
__device__ __attribute__((always_inline)) inline int factor_strip(T* S, T* rdiag, int b)
{
    const int lane = threadIdx.x & 63;
    const int r0 = kW * b, row = r0 + lane;
    const bool act = row < kB;
    T v[kW];
#pragma unroll
    for (int c = 0; c < kW; ++c) v[c] = act ? S[(r0 + c) * kL + row] : T(0);
    int fail = 0;
    T my_y = T(0);  // lane j keeps 1 / L(r0 + j, r0 + j)
#pragma unroll
    for (int j = 0; j < kW; ++j) {
        const T d = lane_value(v[j], j);
        fail = (fail == 0 && !(d > T(0))) ? r0 + j + 1 : fail;
        const T y = rsqrt_refined(d);
        v[j] = lane == j ? d * y : v[j] * y;
        my_y = lane == j ? y : my_y;
#pragma unroll
        for (int c = j + 1; c < kW; ++c) v[c] = fma(-v[j], lane_value(v[j], c), v[c]);
    }
    if (lane < kW) rdiag[r0 + lane] = my_y;
#pragma unroll
    for (int c = 0; c < kW; ++c)
        if (act && lane >= c) S[(r0 + c) * kL + row] = v[c];
    return fail;
}

// Trailing update after strip b: S(16r.., 16c..) -= L(16r.., strip) L(16c.., strip)^T
// for the 16 x 16 tiles b < c <= r < 4, spread over the four wavefronts.
template <class T>
__device__ __attribute__((always_inline)) inline void update_after_strip(T* S, int b)
{
    using Op = Mfma<T>;
    using acc_t = typename Op::acc_t;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    int q = 0;
    for (int c = b + 1; c < 4; ++c) {
        for (int r = c; r < 4; ++r, ++q) {
            if ((q & 3) != wave) continue;
            acc_t acc;
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[e] = S[(kW * c + li) * kL + kW * r + Op::row(lane, e)];
#pragma unroll
            for (int ks = 0; ks < kW; ks += 4) {
                const int kk = kW * b + ks + lk;
                acc = Op::mma(-S[kk * kL + kW * r + li], S[kk * kL + kW * c + li], acc);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) S[(kW * c + li) * kL + kW * r + Op::row(lane, e)] = acc[e];
        }
    }
}

// X = L^-1 for the lower triangle of S, rdiag[i] = 1 / L(i, i).  The lower
// triangle of X is complete on return; its strict upper triangle holds
// scratch (callers write zeros there when they store X).
template <class T>
__device__ __attribute__((always_inline)) inline void invert_lower(const T* S, const T* rdiag, T* X)
{
    using Op = Mfma<T>;
    using acc_t = typename Op::acc_t;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    // diagonal 16 x 16 inverses; in wavefront p, lane c < 16 owns column c of D(p)
    {
        const int o = kW * wave;
        T x[kW];
#pragma unroll
        for (int i = 0; i < kW; ++i) {
            T t = (i == lane) ? T(1) : T(0);
#pragma unroll
            for (int k = 0; k < i; ++k) t = fma(-S[(o + k) * kL + o + i], x[k], t);
            x[i] = t * rdiag[o + i];
        }
        if (lane < kW) {
#pragma unroll
            for (int i = 0; i < kW; ++i) X[(o + lane) * kL + o + i] = x[i];
        }
    }
    __syncthreads();
    // X(r, c) = -D(r) L(r, c) D(c) for (r, c) = (1, 0), (3, 2); T = L(r, c) D(c)
    // is staged in the unused block X(c, r)
    const int r2 = 2 * wave + 1, c2 = 2 * wave;
    if (wave < 2) {
        acc_t acc = {0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < kW; ks += 4) {
            const int kk = ks + lk;
            acc = Op::mma(S[(kW * c2 + kk) * kL + kW * r2 + li], X[(kW * c2 + li) * kL + kW * c2 + kk], acc);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) X[(kW * r2 + li) * kL + kW * c2 + Op::row(lane, e)] = acc[e];
    }
    __syncthreads();
    if (wave < 2) {
        acc_t acc = {0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < kW; ks += 4) {
            const int kk = ks + lk;
            acc = Op::mma(-X[(kW * r2 + kk) * kL + kW * r2 + li], X[(kW * r2 + li) * kL + kW * c2 + kk], acc);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) X[(kW * c2 + li) * kL + kW * r2 + Op::row(lane, e)] = acc[e];
    }
    __syncthreads();
    // X21 = -X22 (L21 X11) on 32-row halves; wavefront w owns 16 x 16 tile
    // (a, bc) = (w / 2, w % 2); T = L21 X11 is staged in the unused X[0:32, 32:64]
    const int a = wave >> 1, bc = wave & 1;
    {
        acc_t acc = {0, 0, 0, 0};
        for (int kb = bc; kb < 2; ++kb) {
#pragma unroll
            for (int ks = 0; ks < kW; ks += 4) {
                const int kk = kW * kb + ks + lk;
                acc = Op::mma(S[kk * kL + 32 + kW * a + li], X[(kW * bc + li) * kL + kk], acc);
            }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) X[(32 + kW * bc + li) * kL + kW * a + Op::row(lane, e)] = acc[e];
    }
    __syncthreads();
    {
        acc_t acc = {0, 0, 0, 0};
        for (int kb = 0; kb <= a; ++kb) {
#pragma unroll
            for (int ks = 0; ks < kW; ks += 4) {
                const int kk = kW * kb + ks + lk;
                acc = Op::mma(-X[(32 + kk) * kL + 32 + kW * a + li], X[(32 + kW * bc + li) * kL + kk], acc);
            }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) X[(kW * bc + li) * kL + 32 + kW * a + Op::row(lane, e)] = acc[e];
    }
    __syncthreads();
}

// Factorization of the lower triangle of S in place and, when it succeeds,
// X = L^-1.  Returns 0 or the 1-based index of the first failed pivot
// (uniform over the workgroup).  Every thread of the workgroup must call it.
// Inlined with the four strips unrolled: as an out-of-line call with a
// generic strip loop it took about 21 us instead of 16 us (FP64).
template <class T>
__device__ __attribute__((always_inline)) inline int factor_invert64(T* S, T* X, T* rdiag, int* s_flag)
{
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        if (wave == 0) {
            const int f = factor_strip<T>(S, rdiag, b);
            if ((threadIdx.x & 63) == 0) *s_flag = f;
        }
        __syncthreads();
        const int f = *s_flag;
        if (f != 0) return f;
        if (b < 3) {
            update_after_strip<T>(S, b);
            __syncthreads();
        }
    }
    invert_lower<T>(S, rdiag, X);
    return 0;
}

// The same as an out-of-line function, for kernels whose register budget
// cannot hold it inlined (it would spill there).
template <class T>
__device__ __attribute__((noinline)) int factor_invert64_call(T* S, T* X, T* rdiag, int* s_flag)
{
    return factor_invert64<T>(S, X, rdiag, s_flag);
}

// Copies the referenced triangle of the w x w diagonal block at A (uplo
// L: A[r + c * lda], r >= c; uplo U: A[c + r * lda]) into the lower triangle
// of S, zeros above it and the identity beyond order w.
template <class T>
__device__ void load_block(bool upper, int w, const T* A, long long lda, T* S)
{
#pragma unroll 4
    for (int e = 0; e < kB * kB / 256; ++e) {
        const int idx = threadIdx.x + 256 * e;
        const int fast = idx & (kB - 1), slow = idx >> 6;
        const int r = upper ? slow : fast, c = upper ? fast : slow;
        T v = T(0);
        if (r < w && c < w) {
            if (r >= c) v = upper ? A[c + (long long)r * lda] : A[r + (long long)c * lda];
        } else if (r == c) {
            v = T(1);
        }
        S[c * kL + r] = v;
    }
}

// Writes the lower triangle of S (order w) back to the referenced triangle at A.
template <class T>
__device__ void store_factor(bool upper, int w, const T* S, T* A, long long lda)
{
#pragma unroll 4
    for (int e = 0; e < kB * kB / 256; ++e) {
        const int idx = threadIdx.x + 256 * e;
        const int fast = idx & (kB - 1), slow = idx >> 6;
        const int r = upper ? slow : fast, c = upper ? fast : slow;
        if (r < w && c < w && r >= c) {
            if (upper) A[c + (long long)r * lda] = S[c * kL + r];
            else A[r + (long long)c * lda] = S[c * kL + r];
        }
    }
}

// Stores X as a full 64 x 64 column-major block (leading dimension 64) with
// exact zeros above the diagonal.
template <class T>
__device__ void store_inverse(const T* X, T* inv)
{
#pragma unroll 4
    for (int e = 0; e < kB * kB / 256; ++e) {
        const int idx = threadIdx.x + 256 * e;
        const int r = idx & (kB - 1), c = idx >> 6;
        inv[idx] = r >= c ? X[c * kL + r] : T(0);
    }
}

} // namespace potf2_detail
} // namespace kern
} // namespace ddla
