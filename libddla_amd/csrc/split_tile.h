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
// fp16x3 uses half-precision planes instead: 11 significand bits each, so
// the dropped product x1 y1 is ~2^-22 of x y (bf16x3 drops ~2^-16), at the
// cost of bf16x3 (v_mfma_f32_16x16x32_f16 runs at the bfloat16 rate, and an
// 11 x 11-bit product is exact in FP32).  Half precision's range is narrow,
// so every row i of the panel is first scaled by 2^e_i, e_i chosen to bring
// the row's largest magnitude into [2^14, 2^15) (split_rowexp_kernel); the
// update then takes acc * 2^-(e_i + e_j), exact, before subtracting it.  Entries
// far below their row's largest lose low bits to half-precision subnormals,
// which bounds the error by ~2^-22 of that row's largest magnitude.
//
// The panel just factored is split once per step (split_panel_kernel) into a
// fragment-ordered buffer: 16 rows x 32 columns of one plane are 1 KB, and
// lane l of a wavefront holds 16 contiguous bytes, row l % 16, columns
// 8 (l / 16) .. + 7 -- exactly the operand of the MFMA instruction, for
// either operand (A[l%16][8(l/16)+q] and B[8(l/16)+q][l%16], checked on
// gfx950).  The update kernel (split_syrk_kernel) then moves whole 1 KB
// fragments from HBM into registers (through LDS for the shared operand)
// with 16-byte accesses and no bank conflicts, with no splitting in its
// inner loop.
#pragma once
#include <hip/hip_runtime.h>

#include "gemm_tile.h"

#include <type_traits>

namespace ddla {
namespace kern {
namespace split_detail {

typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));
typedef _Float16 h8_t __attribute__((ext_vector_type(8)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u4_t __attribute__((ext_vector_type(4)));  // 16 bytes: 8 bf16 of one lane

constexpr int kFragRows = 16;   // rows of one fragment
constexpr int kSlab = 32;       // columns (k) of one fragment
constexpr int kFragU4 = 64;     // 16-byte units per fragment (one per lane)

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

__device__ __forceinline__ u4_t pack8(const _Float16 (&v)[8])
{
    h8_t t;
#pragma unroll
    for (int q = 0; q < 8; ++q) t[q] = v[q];
    return __builtin_bit_cast(u4_t, t);
}

// Split modes (the mixed solver's factor codes): 1 bf16x3, 2 bf16x6, 3 fp16x3.
constexpr int kSplitBf16x3 = 1, kSplitBf16x6 = 2, kSplitFp16x3 = 3;
__host__ __device__ constexpr int split_planes(int mode) { return mode == kSplitBf16x6 ? 3 : 2; }

// Per-row exponents of the fp16x3 split: exps[r] = e with the largest
// |X(r, c)|, c < kb, times 2^e in [2^14, 2^15); 0 for a zero or nonfinite
// largest magnitude and for the padding rows m <= r < rows.  A workgroup
// takes 64 rows; thread (r, g) scans column quarter g of row r (consecutive
// rows, so every column's loads coalesce), and the four partial maxima meet
// in LDS.  The maximum does not depend on the grouping (fmaxf drops NaNs
// either way), so the exponents are those of a plain row scan; one thread
// per row walking all columns took ~270 us per 512-column panel at order
// 32768, four times the panel's own factorization.
__global__ __launch_bounds__(256) void split_rowexp_kernel(int m, int kb, int rows, const float* X, long long ldx,
                                                           int upper, int* exps)
{
    __shared__ float part[4][64];
    const int r = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int row = blockIdx.x * 64 + r;
    float big = 0.0f;
    if (row < m) {
        const int q = (kb + 3) / 4, c0 = g * q, c1 = min(kb, c0 + q);
#pragma unroll 8
        for (int c = c0; c < c1; ++c)
            big = fmaxf(big, fabsf(upper ? X[c + (long long)row * ldx] : X[row + (long long)c * ldx]));
    }
    part[g][r] = big;
    __syncthreads();
    if (g != 0 || row >= rows) return;
    big = fmaxf(fmaxf(part[0][r], part[1][r]), fmaxf(part[2][r], part[3][r]));
    int e = 0;
    if (big > 0.0f && big <= 3.402823466e38f) {
        int q;
        (void)frexpf(big, &q);  // big = f 2^q, f in [0.5, 1)
        e = 15 - q;
    }
    exps[row] = e;
}

// Splits rows [0, m) and columns [0, kb) of the L view X (element (i, j) at
// X[i + j * ldx], or X[j + i * ldx] when upper) into NP planes.  Grid:
// (rows / 256, 4 * nks), 256 threads; thread = one row, 8 columns.
// F16 (fp16x3): the row is first scaled by 2^exps[row] and split into two
// half-precision planes.
template <int NP, bool F16>
__global__ __launch_bounds__(256) void split_panel_kernel(int m, int kb, int rows, int nks, const float* X,
                                                          long long ldx, int upper, const int* exps, u4_t* out)
{
    const int row = blockIdx.x * 256 + threadIdx.x;
    if (row >= rows) return;
    const int g = blockIdx.y;  // 8-column group
    const int s = g >> 2, kq = g & 3;
    const int e = F16 ? exps[row] : 0;
    float x[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int col = g * 8 + q;
        x[q] = (row < m && col < kb) ? (upper ? X[col + (long long)row * ldx] : X[row + (long long)col * ldx]) : 0.0f;
        if (F16) x[q] = ldexpf(x[q], e);  // exact: a power-of-two scale into [0, 2^15)
    }
    using H = typename std::conditional<F16, _Float16, __bf16>::type;
    H pl[NP][8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        float r = x[q];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const H h = (H)r;
            pl[p][q] = h;
            r -= (float)h;
        }
    }
    const int rb = row >> 4, lane = (row & 15) + 16 * kq;
#pragma unroll
    for (int p = 0; p < NP; ++p) out[frag_index(rb, s, p, nks, NP) + lane] = pack8(pl[p]);
}

// C (L view, element (i, j) at C[i + j * ldc]) minus= A B^T over the m x n
// block, where the rows of A and of B are
// split buffers (row offsets already applied, at least ceil(m / 256) * 256
// and ceil(n / 128) * 128 rows).  tri: write only i + diag >= j.
struct SplitProblem {
    int m, n, nks;
    const u4_t* A;
    const u4_t* B;
    float* C;
    long long ldc;
    int tri;
    long long diag;
    const int* ea;  // fp16x3: row exponents of A's rows and of B's rows (offsets applied)
    const int* eb;
};

constexpr int kRows = 256;  // rows (A) of C per workgroup
constexpr int kCols = 128;  // columns (B) of C per workgroup

// Tile schedule of one launch.  lower (square lower-triangular C, diag 0):
// the tiles (I, tj) with 256 I + 255 >= 128 tj, i.e. tj <= 2 I + 1, in
// groups of 8 tile columns, within a group by I, then tj; otherwise every
// (I, tj) of the nI x nt grid, in groups of 8 tile columns likewise.
// Workgroup b takes tile (b % 8) * per_xcd + b / 8, so XCD x (which runs
// the workgroups b = x mod 8) sweeps a contiguous run of tiles.
struct SplitOrder {
    int nI, nt;  // 256-row blocks, 128-column blocks
    int total, per_xcd;
    int lower;
};

// Tiles of group g (columns 8g .. min(8g + 8, nt) - 1) in the lower schedule:
// row I holds the columns tj <= 2 I + 1, so rows 4g .. 4g + 3 hold 2, 4, 6, 8
// columns (fewer in a short last group) and later rows the whole group.
__host__ __device__ inline long long split_group_tiles(int g, int nI, int nt)
{
    const int j0 = 8 * g, j1 = min(j0 + 8, nt), ifirst = j0 / 2;
    long long c = 0;
    for (int I = ifirst; I < nI && I < ifirst + 4; ++I) c += max(0, min(j1, 2 * I + 2) - j0);
    if (ifirst + 4 < nI) c += (long long)(nI - ifirst - 4) * (j1 - j0);
    return c;
}

__device__ inline bool split_tile_of(const SplitOrder& o, int t, int& I, int& tj)
{
    if (t >= o.total) return false;
    if (!o.lower) {
        const int per_group = 8 * o.nI;
        const int g = t / per_group, j0 = 8 * g, geff = min(8, o.nt - j0), r = t - g * per_group;
        I = r / geff;
        tj = j0 + r % geff;
        return true;
    }
    const int ng = (o.nt + 7) / 8;
    long long before = 0;
    int g = 0;
    for (; g < ng; ++g) {
        const long long c = split_group_tiles(g, o.nI, o.nt);
        if (t < before + c) break;
        before += c;
    }
    int r = (int)(t - before);
    const int j0 = 8 * g, j1 = min(j0 + 8, o.nt), ifirst = j0 / 2;
    for (I = ifirst; I < o.nI && I < ifirst + 4; ++I) {
        const int cnt = max(0, min(j1, 2 * I + 2) - j0);
        if (r < cnt) {
            tj = j0 + r;
            return true;
        }
        r -= cnt;
    }
    const int w = j1 - j0;
    I = ifirst + 4 + r / w;
    tj = j0 + r % w;
    return I < o.nI;
}

// One 256 x 128 tile per workgroup of 4 wavefronts; wavefront w owns rows
// 64w .. 64w + 63 and all 128 columns (4 x 8 MFMA tiles).  Its A fragments
// (4 row blocks x NP planes per 32-column slab) are private, so they go from
// HBM straight into registers; the B fragments (8 row blocks x NP planes) are
// shared by the four wavefronts and staged once per slab in LDS (two
// buffers), each wavefront loading a quarter of them.  Per MFMA this reads
// 0.17 (bf16x3) / 0.125 (bf16x6) 1 KB fragments from LDS, against 0.33 / 0.25
// with four 64 x 64 wavefront blocks.  A's next slab is loaded right after
// the slab's products (a second register set for loading it a slab ahead
// measured no faster and leaves no registers for bf16x6).  The accumulators hold
// D = B A^T (lane l % 16 runs along i, the contiguous dimension of C).
// Every accumulator takes its products in the same order, smallest first,
// slab after slab.
template <int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void split_syrk_kernel(
    SplitProblem p, SplitOrder o, const int* abort)
{
    static_assert(MODE == kSplitBf16x3 || MODE == kSplitBf16x6 || MODE == kSplitFp16x3, "split mode");
    constexpr int NPROD = MODE == kSplitBf16x6 ? 6 : 3;
    constexpr bool F16 = MODE == kSplitFp16x3;
    constexpr int NP = split_planes(MODE);
    constexpr int BF = 8 * NP;   // B fragments per slab
    constexpr int BPW = BF / 4;  // B fragments each wavefront stages per slab
    if (abort != nullptr && *abort != 0) return;
    const int b = blockIdx.x;
    int I, tj;
    if (!split_tile_of(o, (b & 7) * o.per_xcd + (b >> 3), I, tj)) return;
    const int m0 = I * kRows, n0 = tj * kCols;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int mw = m0 + 64 * wave;  // this wavefront's first row
    // a wavefront wholly above the written triangle or below the last row
    // computes nothing (it still stages its share of B)
    const bool active = mw < p.m && (!p.tri || (long long)(mw + 63) + p.diag >= n0);
    __shared__ u4_t lds[2 * BF * kFragU4];
    const int nks = p.nks;
    const int rbs = nks * NP * kFragU4;  // 16-byte units per 16-row block
    // Fragment loads through buffer instructions: the slab's origin (per
    // wavefront for A) goes into the resource, the fragment's place into the
    // scalar offset and the lane's 16 bytes into the one vector offset that
    // every load shares, so no load needs vector address arithmetic.
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    const char* aw = reinterpret_cast<const char*>(p.A + (long long)((m0 + 64 * wv) / 16) * rbs);
    const char* bw = reinterpret_cast<const char*>(p.B + (long long)(n0 / 16) * rbs);
    const unsigned loff = (unsigned)lane * 16u;
    const int rbb = rbs * 16;                   // bytes per 16-row block
    constexpr int kFragBytes = kFragU4 * 16;    // one fragment
    u4_t ra[4][NP];
    u4_t rb[BPW];
    auto load_a = [&](int s) {
        const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<char*>(aw + (long long)s * NP * kFragBytes), 0, 0x7fffffff, 0x00020000);
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int q = 0; q < NP; ++q)
                ra[x][q] = __builtin_bit_cast(u4_t, __builtin_amdgcn_raw_buffer_load_b128(r, loff, x * rbb + q * kFragBytes, 0));
    };
    auto load_b = [&](int s) {
        const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<char*>(bw + (long long)s * NP * kFragBytes), 0, 0x7fffffff, 0x00020000);
#pragma unroll
        for (int e = 0; e < BPW; ++e) {
            const int f = wv + 4 * e, rbk = f / NP, q = f % NP;
            rb[e] = __builtin_bit_cast(u4_t, __builtin_amdgcn_raw_buffer_load_b128(r, loff, rbk * rbb + q * kFragBytes, 0));
        }
    };
    auto store_b = [&](int buf) {
#pragma unroll
        for (int e = 0; e < BPW; ++e) lds[(buf * BF + wave + 4 * e) * kFragU4 + lane] = rb[e];
    };
    f4_t acc[4][8];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 8; ++y) acc[x][y] = f4_t{0, 0, 0, 0};
    auto compute = [&](int buf) {
        const u4_t* lb = lds + buf * BF * kFragU4 + lane;
#pragma unroll
        for (int y = 0; y < 8; ++y) {
            u4_t bq[NP];
#pragma unroll
            for (int q = 0; q < NP; ++q) bq[q] = lb[(y * NP + q) * kFragU4];
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                auto A = [&](int q) { return ra[x][q]; };
                auto mma = [&](const u4_t& a, const u4_t& bb) {
                    if constexpr (F16)
                        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h8_t, bb),
                                                                           __builtin_bit_cast(h8_t, a), acc[x][y], 0, 0, 0);
                    else
                        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8_t, bb),
                                                                            __builtin_bit_cast(bf8_t, a), acc[x][y], 0, 0, 0);
                };
                // products smallest first: (2,0) (1,1) (0,2) for bf16x6, then (1,0) (0,1) (0,0)
                if constexpr (NPROD == 6) {
                    mma(A(2), bq[0]);
                    mma(A(1), bq[1]);
                    mma(A(0), bq[2]);
                }
                mma(A(1), bq[0]);
                mma(A(0), bq[1]);
                mma(A(0), bq[0]);
            }
        }
    };
    load_a(0);
    load_b(0);
    store_b(0);
    __syncthreads();
    for (int s = 0; s < nks; ++s) {
        // the next slab (clamped: the last slab reloads itself, unused)
        const int sn = min(s + 1, nks - 1);
        load_b(sn);
        if (active) compute(s & 1);
        load_a(sn);
        if (s + 1 < nks) store_b((s & 1) ^ 1);
        __syncthreads();
    }
    // fp16x3: the products of rows scaled by 2^e_i and 2^e_j are unscaled
    // exactly; the tile's 256 row and 128 column exponents are staged in LDS
    // once (clamped indices: rows and columns past the end are never written)
    __shared__ int se[F16 ? kRows + kCols : 1];
    if constexpr (F16) {
        for (int t = tid; t < kRows + kCols; t += 256)
            se[t] = t < kRows ? p.ea[min(m0 + t, p.m - 1)] : p.eb[min(n0 + t - kRows, p.n - 1)];
        __syncthreads();
    }
    if (!active) return;
    const int rr0 = 4 * (lane >> 4), cc = lane & 15;
    int ex[4] = {0, 0, 0, 0};
    if constexpr (F16) {
#pragma unroll
        for (int x = 0; x < 4; ++x) ex[x] = se[64 * wave + 16 * x + cc];
    }
    auto update = [&](int x, int y, int r) -> float {
        if constexpr (F16) return ldexpf(acc[x][y][r], -(ex[x] + se[kRows + 16 * y + rr0 + r]));
        else return acc[x][y][r];
    };

    // C := C - acc.  A wavefront whose 64 x 128 block is in bounds and wholly
    // on the written side of the diagonal reads 64 columns of C at a time
    // (all 64 loads of a lane before its stores: a store between them would
    // have to complete before the next load, since the compiler cannot tell
    // the addresses apart), addressed as one uniform base plus 32-bit lane
    // offsets (the host keeps 128 * ldc below 2^31); other blocks take the
    // element-wise masked path.
    const bool interior =
        mw + 64 <= p.m && n0 + kCols <= p.n && (!p.tri || (long long)mw + p.diag >= n0 + kCols - 1);
    if (interior) {
        const int ldc = (int)p.ldc;
        float* Ct = p.C + mw + (long long)n0 * p.ldc;
        // element (x, y, r): row i = 16x + cc, column j = 16y + rr0 + r
        const int base = cc + rr0 * ldc;
        auto off = [&](int x, int y, int r) { return base + 16 * x + (16 * y + r) * ldc; };
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            float cv[4][4][4];
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y)
#pragma unroll
                    for (int r = 0; r < 4; ++r) cv[x][y][r] = Ct[off(x, 4 * h + y, r)];
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y)
#pragma unroll
                    for (int r = 0; r < 4; ++r) Ct[off(x, 4 * h + y, r)] = cv[x][y][r] - update(x, 4 * h + y, r);
        }
        return;
    }
    // The same in rounds of 16 elements per lane, each masked: all reads of a
    // round before its writes (element by element, each read would also wait
    // for the previous element's write, one counter covering both).
    auto live = [&](int i, int j) { return i < p.m && j < p.n && (!p.tri || (long long)i + p.diag >= j); };
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int i = mw + 16 * x + cc;
            float cv[4][4];
#pragma unroll
            for (int y = 0; y < 4; ++y)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int j = n0 + 16 * (4 * h + y) + rr0 + r;
                    cv[y][r] = live(i, j) ? p.C[i + (long long)j * p.ldc] : 0.0f;
                }
#pragma unroll
            for (int y = 0; y < 4; ++y)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int j = n0 + 16 * (4 * h + y) + rr0 + r;
                    if (live(i, j)) p.C[i + (long long)j * p.ldc] = cv[y][r] - update(x, 4 * h + y, r);
                }
        }
}

} // namespace split_detail
} // namespace kern
} // namespace ddla
