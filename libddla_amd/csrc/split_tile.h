// BF16 split-product trailing update for the FP32 factorization (device
// code, included by kernels.hip).  There is no counterpart in the reference.
//
// An FP32 value x is written as the sum of bfloat16 planes
//   x0 = bf16(x),  x1 = bf16(x - x0),  x2 = bf16(x - x0 - x1)
// (round to nearest even; each difference is exact in FP32, and three planes
// hold all 24 significand bits).  A product then expands into plane products
// that v_mfma_f32_16x16x32_bf16 forms exactly and accumulates in FP32
// (Henry, Tang & Heinecke 2019, "Leveraging the bfloat16 artificial
// intelligence datatype for higher-precision computations"):
//   bf16x3: x0 y0 + x0 y1 + x1 y0                     (two planes, ~2^-16 relative)
//   bf16x6: ... + x0 y2 + x1 y1 + x2 y0               (three planes, ~FP32)
// The products are issued smallest first.
//
// The panel just factored is split once per step (split_panel_kernel) into a
// fragment-ordered buffer: 16 rows x 32 columns of one plane are 1 KB, and
// lane l of a wavefront holds 16 contiguous bytes, row l % 16, columns
// 8 (l / 16) .. + 7 -- exactly the operand of the MFMA instruction, for
// either operand (A[l%16][8(l/16)+q] and B[8(l/16)+q][l%16], checked on
// gfx950).  The update kernel (split_syrk_kernel) then moves whole 1 KB
// fragments from HBM through LDS into registers with 16-byte accesses and
// no bank conflicts, with no splitting in its inner loop.
#pragma once
#include <hip/hip_runtime.h>

#include "gemm_tile.h"

namespace ddla {
namespace kern {
namespace split_detail {

typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u4_t __attribute__((ext_vector_type(4)));  // 16 bytes: 8 bf16 of one lane

constexpr int kFragRows = 16;   // rows of one fragment
constexpr int kSlab = 32;       // columns (k) of one fragment
constexpr int kFragU4 = 64;     // 16-byte units per fragment (one per lane)
constexpr int kTile = 128;      // rows and columns of C per workgroup
constexpr int kTileFrags = kTile / kFragRows;

// Fragment-ordered split of an m x kb block (rows padded to `rows`, columns
// to nks * 32, zero filled): 16-byte unit
//   ((rb * nks + s) * NP + p) * 64 + lane
// holds plane p of rows rb * 16 + lane % 16, columns s * 32 + 8 (lane / 16) .. + 7.
__device__ __forceinline__ long long frag_index(int rb, int s, int p, int nks, int np)
{
    return (((long long)rb * nks + s) * np + p) * kFragU4;
}

__device__ __forceinline__ u4_t pack8(const __bf16 (&v)[8])
{
    bf8_t t;
#pragma unroll
    for (int q = 0; q < 8; ++q) t[q] = v[q];
    return __builtin_bit_cast(u4_t, t);
}

// Splits rows [0, m) and columns [0, kb) of the L view X (element (i, j) at
// X[i + j * ldx], or X[j + i * ldx] when upper) into NP planes.  Grid:
// (rows / 256, 4 * nks), 256 threads; thread = one row, 8 columns.
template <int NP>
__global__ __launch_bounds__(256) void split_panel_kernel(int m, int kb, int rows, int nks, const float* X,
                                                          long long ldx, int upper, u4_t* out)
{
    const int row = blockIdx.x * 256 + threadIdx.x;
    if (row >= rows) return;
    const int g = blockIdx.y;  // 8-column group
    const int s = g >> 2, kq = g & 3;
    float x[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int col = g * 8 + q;
        x[q] = (row < m && col < kb) ? (upper ? X[col + (long long)row * ldx] : X[row + (long long)col * ldx]) : 0.0f;
    }
    __bf16 pl[NP][8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        float r = x[q];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const __bf16 h = (__bf16)r;
            pl[p][q] = h;
            r -= (float)h;
        }
    }
    const int rb = row >> 4, lane = (row & 15) + 16 * kq;
#pragma unroll
    for (int p = 0; p < NP; ++p) out[frag_index(rb, s, p, nks, NP) + lane] = pack8(pl[p]);
}

// C (L view, element (i, j) at C[i + j * ldc], or C[j + i * ldc] when
// upper) minus= A B^T over the m x n block, where the rows of A and of B are
// split buffers (row offsets already applied, at least ceil(m / 128) * 128
// and ceil(n / 128) * 128 rows).  tri: write only i + diag >= j.
struct SplitProblem {
    int m, n, nks;
    const u4_t* A;
    const u4_t* B;
    float* C;
    long long ldc;
    int tri;
    long long diag;
};

template <int NP, int NBUF>
constexpr int split_lds_u4()
{
    return NBUF * 2 * kTileFrags * NP * kFragU4;
}

// One 128 x 128 tile per workgroup: 4 wavefronts (2 x 2), each 64 x 64 = 4 x 4
// MFMA tiles.  Per 32-column slab the workgroup stages 8 row fragments of A
// and 8 of B (NP planes each) in LDS; the next slab's fragments wait in
// registers (NBUF = 1) or in the second LDS buffer (NBUF = 2).  The
// accumulators hold D = B A^T for the lower layout (lane l % 16 runs along i,
// the contiguous dimension of C) and D = A B^T for the upper layout.
template <int NPROD, bool UPPER, int NBUF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void split_syrk_kernel(SplitProblem p, gemm_detail::TileOrder o,
                                                            const int* abort)
{
    static_assert(NPROD == 3 || NPROD == 6, "bf16x3 or bf16x6");
    constexpr int NP = NPROD == 6 ? 3 : 2;
    constexpr int FR = kTileFrags * NP;  // fragments of one operand per slab
    constexpr int LPT = 2 * FR / 4;      // fragments each wavefront loads per slab
    if (abort != nullptr && *abort != 0) return;
    int ti, tj;
    if (!gemm_detail::map_tile(o, blockIdx.x, ti, tj)) return;
    const int m0 = ti * kTile, n0 = tj * kTile;
    if (p.tri && (long long)(m0 + kTile - 1) + p.diag < n0) return;

    __shared__ u4_t lds[split_lds_u4<NP, NBUF>()];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;

    // global base of this tile's fragments: fragment f < FR of A is row block
    // f / NP, plane f % NP; f >= FR addresses B the same way
    const u4_t* gA = p.A + (long long)ti * kTileFrags * p.nks * NP * kFragU4;
    const u4_t* gB = p.B + (long long)tj * kTileFrags * p.nks * NP * kFragU4;
    auto gaddr = [&](int f, int s) -> const u4_t* {
        const bool isb = f >= FR;
        const int ff = isb ? f - FR : f;
        const int rb = ff / NP, pl = ff % NP;
        return (isb ? gB : gA) + frag_index(rb, s, pl, p.nks, NP) + lane;
    };
    u4_t reg[LPT];
    auto load = [&](int s) {
#pragma unroll
        for (int e = 0; e < LPT; ++e) reg[e] = *gaddr(wave + 4 * e, s);
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int e = 0; e < LPT; ++e) lds[(buf * 2 * FR + wave + 4 * e) * kFragU4 + lane] = reg[e];
    };

    f4_t acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = f4_t{0, 0, 0, 0};

    auto compute = [&](int buf) {
        const u4_t* la = lds + buf * 2 * FR * kFragU4;
        const u4_t* lb = la + FR * kFragU4;
        // every fragment of the slab is read up front (the reads then overlap
        // the products instead of waiting one by one)
        bf8_t bfr[4][NP], afr4[4][NP];
#pragma unroll
        for (int y = 0; y < 4; ++y)
#pragma unroll
            for (int q = 0; q < NP; ++q)
                bfr[y][q] = __builtin_bit_cast(bf8_t, lb[((wn * 4 + y) * NP + q) * kFragU4 + lane]);
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int q = 0; q < NP; ++q)
                afr4[x][q] = __builtin_bit_cast(bf8_t, la[((wm * 4 + x) * NP + q) * kFragU4 + lane]);
        __builtin_amdgcn_sched_barrier(0);  // keep the reads ahead of the products
#pragma unroll
        for (int x = 0; x < 4; ++x) {
            const bf8_t* afr = afr4[x];
            // products smallest first: (2,0) (1,1) (0,2) for bf16x6, then (1,0) (0,1) (0,0)
            auto mma = [&](int y, const bf8_t& a, const bf8_t& b) {
                if (UPPER) acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[x][y], 0, 0, 0);
                else acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b, a, acc[x][y], 0, 0, 0);
            };
            if constexpr (NPROD == 6) {
#pragma unroll
                for (int y = 0; y < 4; ++y) mma(y, afr[2], bfr[y][0]);
#pragma unroll
                for (int y = 0; y < 4; ++y) mma(y, afr[1], bfr[y][1]);
#pragma unroll
                for (int y = 0; y < 4; ++y) mma(y, afr[0], bfr[y][2]);
            }
#pragma unroll
            for (int y = 0; y < 4; ++y) mma(y, afr[1], bfr[y][0]);
#pragma unroll
            for (int y = 0; y < 4; ++y) mma(y, afr[0], bfr[y][1]);
#pragma unroll
            for (int y = 0; y < 4; ++y) mma(y, afr[0], bfr[y][0]);
        }
    };

    const int nks = p.nks;
    if (NBUF == 2) {
        load(0);
        store(0);
        __syncthreads();
        for (int s = 0; s < nks; ++s) {
            if (s + 1 < nks) load(s + 1);
            compute(s & 1);
            if (s + 1 < nks) store((s & 1) ^ 1);
            __syncthreads();
        }
    } else {
        load(0);
        store(0);
        __syncthreads();
        for (int s = 0; s < nks; ++s) {
            if (s + 1 < nks) load(s + 1);
            compute(0);
            __syncthreads();
            if (s + 1 < nks) {
                store(0);
                __syncthreads();
            }
        }
    }

    // C := C - acc.  Interior tiles (in bounds and wholly on the written side
    // of the diagonal) issue all 64 loads of a lane before the first store (a
    // store between them would have to complete before the next load, since
    // the compiler cannot tell the addresses apart), addressed as one uniform
    // tile base plus 32-bit lane offsets (the host keeps 128 * ldc below
    // 2^31).  Edge and diagonal tiles take the element-wise masked path.
    const bool full = m0 + kTile <= p.m && n0 + kTile <= p.n;
    const bool interior = full && (!p.tri || (long long)m0 + p.diag >= n0 + kTile - 1);
    const int rr0 = 4 * (lane >> 4), cc = lane & 15;
    if (interior) {
        const int ldc = (int)p.ldc;
        float* Ct = UPPER ? p.C + n0 + (long long)m0 * p.ldc : p.C + m0 + (long long)n0 * p.ldc;
        // element (x, y, r): row i = wm*64 + 16x + (UPPER ? rr0 + r : cc), column j = wn*64 + 16y + (UPPER ? cc : rr0 + r)
        const int base = UPPER ? (wm * 64 + rr0) * ldc + wn * 64 + cc : wm * 64 + cc + (wn * 64 + rr0) * ldc;
        auto off = [&](int x, int y, int r) {
            return UPPER ? base + (16 * x + r) * ldc + 16 * y : base + 16 * x + (16 * y + r) * ldc;
        };
        float cv[4][4][4];
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y)
#pragma unroll
                for (int r = 0; r < 4; ++r) cv[x][y][r] = Ct[off(x, y, r)];
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y)
#pragma unroll
                for (int r = 0; r < 4; ++r) Ct[off(x, y, r)] = cv[x][y][r] - acc[x][y][r];
        return;
    }
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = m0 + (wm * 4 + x) * 16 + (UPPER ? rr0 + r : cc);
                const int j = n0 + (wn * 4 + y) * 16 + (UPPER ? cc : rr0 + r);
                if (!full && (i >= p.m || j >= p.n)) continue;
                if (p.tri && (long long)i + p.diag < j) continue;
                float* c = UPPER ? p.C + j + (long long)i * p.ldc : p.C + i + (long long)j * p.ldc;
                *c -= acc[x][y][r];
            }
}

} // namespace split_detail
} // namespace kern
} // namespace ddla
