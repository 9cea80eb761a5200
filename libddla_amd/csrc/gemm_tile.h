// MFMA GEMM tile for gfx950 (device code, included by kernels.hip).
//
// C := alpha * op(A) * op(B) + beta * C on 64-wide wavefronts:
//   * a 256-thread workgroup (4 wavefronts, 2 x 2) owns a BM x BN tile of C,
//     each wavefront a (BM/2) x (BN/2) block of 16 x 16 MFMA tiles
//     (v_mfma_f64_16x16x4_f64 / v_mfma_f32_16x16x4_f32);
//   * op(A) and op(B) stream through two alternating LDS slabs, BK deep; the
//     global loads of slab t+1 are in flight while slab t is multiplied;
//   * interior tiles with 16-byte aligned operands load whole 16-byte vectors
//     along the contiguous dimension without bounds checks;
//   * the 1-D grid is remapped so that each of the 8 XCDs (workgroup b runs on
//     XCD b % 8) sweeps a contiguous run of tiles in column groups, keeping the
//     operand rows its concurrent tiles share resident in that XCD's L2; a
//     lower-triangular C (syrk) is enumerated directly, with no empty
//     workgroups.
#pragma once
#include <hip/hip_runtime.h>

namespace ddla {
namespace kern {
namespace gemm_detail {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef double d2_t __attribute__((ext_vector_type(2)));

// Operand layout of both 16x16x4 instructions: lane l supplies A[l%16][l/16]
// and B[l/16][l%16].  The accumulator layouts differ (checked on gfx950).
template <class T> struct Mfma;
template <> struct Mfma<double> {
    using acc_t = d4_t;
    using vec_t = d2_t;  // 16 bytes
    static constexpr int kVec = 2;
    __device__ static acc_t mma(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // lane l, accumulator r holds D[l/16 + 4r][l%16]
    __device__ static int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct Mfma<float> {
    using acc_t = f4_t;
    using vec_t = f4_t;  // 16 bytes
    static constexpr int kVec = 4;
    __device__ static acc_t mma(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // lane l, accumulator r holds D[4(l/16) + r][l%16]
    __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

enum { kTriNoneD = 0, kTriLowerD = 1, kTriUpperD = 2 };

template <class T>
struct GemmProblem {
    int m, n, k;
    T alpha, beta;
    const T* A;
    long long lda;
    const T* B;
    long long ldb;
    T* C;
    long long ldc;
    int tri;
    long long diag;
    int vec_ab;  // A, B and their leading dimensions allow 16-byte loads along the contiguous dimension
    // launches over independent problems (grid z = problem * inner + sub-block):
    // problem b moves every pointer by b * bstride bytes, sub-block i by
    // i * istride bytes more (the abort word only by the problem's offset)
    int inner = 1;
    long long bstride = 0, istride = 0;
};

// Moves p's operands, result and the abort word to grid z's problem and
// sub-block (nothing for a single-problem launch).
template <class T>
__device__ inline void batch_offset(GemmProblem<T>& p, const int*& abort)
{
    if (gridDim.z == 1) return;
    const long long z = blockIdx.z, prob = z / p.inner, sub = z - prob * p.inner;
    const long long off = prob * p.bstride + sub * p.istride;
    p.A = reinterpret_cast<const T*>(reinterpret_cast<const char*>(p.A) + off);
    p.B = reinterpret_cast<const T*>(reinterpret_cast<const char*>(p.B) + off);
    p.C = reinterpret_cast<T*>(reinterpret_cast<char*>(p.C) + off);
    if (abort != nullptr) abort = reinterpret_cast<const int*>(reinterpret_cast<const char*>(abort) + prob * p.bstride);
}

// Tile schedule of one launch (see map_tile).
struct TileOrder {
    int ntm, ntn;    // tiles along m and n
    int total;       // tiles to compute
    int per_xcd;     // ceil(total / 8)
    int group;       // tile columns per group
    int lower_enum;  // enumerate only the lower triangle (ntm == ntn, square tiles, diag 0)
};

// Workgroup b -> tile (ti, tj).  Tile t = (b % 8) * per_xcd + b / 8, so XCD x
// owns tiles [x * per_xcd, (x + 1) * per_xcd).  Tiles are ordered in groups of
// `group` tile columns; within a group by tile row, then column.
__device__ inline bool tile_of(const TileOrder& o, int t, int& ti, int& tj);

__device__ inline bool map_tile(const TileOrder& o, int b, int& ti, int& tj)
{
    return tile_of(o, (b & 7) * o.per_xcd + (b >> 3), ti, tj);
}

// Tile t (< o.total) of the schedule of map_tile.
__device__ inline bool tile_of(const TileOrder& o, int t, int& ti, int& tj)
{
    if (t >= o.total) return false;
    const int G = o.group;
    if (!o.lower_enum) {
        const int per_group = G * o.ntm;
        const int g = t / per_group;
        const int j0 = g * G;
        const int geff = min(G, o.ntn - j0);
        const int r = t - g * per_group;
        ti = r / geff;
        tj = j0 + r % geff;
        return true;
    }
    // Lower triangle, groups of G tile columns.  Group g (full) holds
    // G (nt - gG) - G (G - 1) / 2 tiles; S(g) = tiles before group g.
    const long long nt = o.ntm;
    auto S = [&](long long g) { return g * (G * nt - (long long)G * (G - 1) / 2) - (long long)G * G * g * (g - 1) / 2; };
    const int nfull = (int)(nt / G);
    int g;
    if (t >= S(nfull)) {
        g = nfull;
    } else {
        int lo = 0, hi = nfull - 1;  // S(lo) <= t < S(hi + 1)
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (S(mid) <= t) lo = mid;
            else hi = mid - 1;
        }
        g = lo;
    }
    const int j0 = g * G;
    const int geff = min((long long)G, nt - j0);
    const int o2 = t - (int)S(g);
    const int tri = geff * (geff - 1) / 2;
    if (o2 < tri) {
        int r = (int)((sqrtf(8.0f * o2 + 1.0f) - 1.0f) * 0.5f);
        while (r * (r + 1) / 2 > o2) --r;
        while ((r + 1) * (r + 2) / 2 <= o2) ++r;
        ti = j0 + r;
        tj = j0 + (o2 - r * (r + 1) / 2);
    } else {
        const int q = o2 - tri;
        ti = j0 + geff - 1 + q / geff;
        tj = j0 + q % geff;
    }
    return true;
}

constexpr int kPad = 16;  // LDS row padding: the 4 k-rows of a fragment load fall on distinct banks

template <int BM, int BN, int BK, int NBUF = 2>
constexpr int lds_elems()
{
    return NBUF * BK * (BM + kPad) + NBUF * BK * (BN + kPad);
}

// One BM x BN tile of C at (m0, n0).  FAST: the whole tile and every k slab
// are in bounds and the operands allow 16-byte vector loads (decided once per
// tile, so the k loop carries no branches between its loads).
// NBUF = 2: the slabs alternate between two LDS buffers (one barrier per
// slab); NBUF = 1: one buffer, the next slab waits in registers (two
// barriers per slab, half the LDS, so more workgroups fit on a CU).
// PF2 (with NBUF = 2): two register sets, loads run two slabs ahead.
// The accumulators start at zero and C is read once, in the epilogue
// (C := alpha * acc + beta * C): starting them at beta * C instead rounds
// every partial sum at the scale of C, which in FP32 visibly slows the
// iterative refinement of nearly cancelling Schur complements.
template <class T, int BM, int BN, int BK, bool TA, bool TB, bool FAST, int NBUF = 2, bool PF2 = false, bool BUF = true>
__device__ __forceinline__ void gemm_tile_body(const GemmProblem<T>& p, int m0, int n0, T* smem)
{
    using Op = Mfma<T>;
    using acc_t = typename Op::acc_t;
    using vec_t = typename Op::vec_t;
    constexpr int VW = Op::kVec;
    constexpr int LA = BM + kPad, LB = BN + kPad;
    constexpr int TM = BM / 32, TN = BN / 32;
    constexpr int APT = BM * BK / 256, BPT = BN * BK / 256;  // scalars per thread per slab
    constexpr int AVT = APT / VW, BVT = BPT / VW;            // vectors per thread per slab
    static_assert(APT * 256 == BM * BK && BPT * 256 == BN * BK, "tile must split over 256 threads");
    static_assert(AVT * VW == APT && BVT * VW == BPT, "slab must split into 16-byte vectors");

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    T* As = smem;
    T* Bs = smem + NBUF * BK * LA;
    const bool full = FAST || (m0 + BM <= p.m && n0 + BN <= p.n);

    acc_t acc[TM][TN];
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) acc[a][b] = acc_t{0, 0, 0, 0};

    const int nkt = (p.alpha == T(0)) ? 0 : (p.k + BK - 1) / BK;

    // Vector v of a slab: TA == false (op(A) = A, contiguous along m):
    // (i, kk) = ((v % (BM/VW)) * VW, v / (BM/VW)); TA == true (contiguous along
    // k): (i, kk) = (v / (BK/VW), (v % (BK/VW)) * VW).  Same for op(B).
    // register sets of the slabs in flight (PF2: two, so that slab kt + 2 is
    // loaded while slab kt is multiplied and slab kt + 1 waits in the other)
    constexpr int NSET = PF2 ? 2 : 1;
    T ra[NSET][AVT][VW], rb[NSET][BVT][VW];

    auto a_coord = [&](int e, int& i, int& kk) {
        const int v = tid + 256 * e;
        if (!TA) { i = (v % (BM / VW)) * VW; kk = v / (BM / VW); }
        else { kk = (v % (BK / VW)) * VW; i = v / (BK / VW); }
    };
    auto b_coord = [&](int e, int& j, int& kk) {
        const int v = tid + 256 * e;
        if (TB) { j = (v % (BN / VW)) * VW; kk = v / (BN / VW); }
        else { kk = (v % (BK / VW)) * VW; j = v / (BK / VW); }
    };
    // FAST loads (BUF): the slab's origin in each operand is uniform over the
    // workgroup (scalar registers) and each thread's place in it a 32-bit
    // byte offset fixed for the whole tile, so a load needs no per-slab
    // vector address arithmetic (the caller keeps the leading dimensions at
    // most 2^20, which bounds the offsets by 2^31).  This took the FP64
    // trailing update from 51.4 to 53.0 TF/s per launch; the batched
    // left-looking kernel's small products keep 64-bit addresses (BUF =
    // false), with which its FP64 run measured 5.31 against 6.0 ms.
    unsigned a_off[AVT], b_off[BVT];
    if (FAST && BUF) {
#pragma unroll
        for (int e = 0; e < AVT; ++e) {
            int i, kk;
            a_coord(e, i, kk);
            a_off[e] = (unsigned)(TA ? kk + i * (int)p.lda : i + kk * (int)p.lda) * (unsigned)sizeof(T);
        }
#pragma unroll
        for (int e = 0; e < BVT; ++e) {
            int j, kk;
            b_coord(e, j, kk);
            b_off[e] = (unsigned)(TB ? j + kk * (int)p.ldb : kk + j * (int)p.ldb) * (unsigned)sizeof(T);
        }
    }
    auto load = [&](int k0, int set) {
        if (FAST && BUF) {
            // the tile's coordinates are uniform but may come out of vector
            // arithmetic (the tile order's square root): made scalar here
            const int ms = __builtin_amdgcn_readfirstlane(m0), ns = __builtin_amdgcn_readfirstlane(n0);
            const int ks = __builtin_amdgcn_readfirstlane(k0);
            const T* a0 = TA ? p.A + ks + (long long)ms * p.lda : p.A + ms + (long long)ks * p.lda;
            const T* b0 = TB ? p.B + ns + (long long)ks * p.ldb : p.B + ks + (long long)ns * p.ldb;
            // buffer loads from the slab's origin: the 32-bit per-thread offset
            // goes straight into the instruction (raw buffer, gfx9 resource word 3)
            const __amdgpu_buffer_rsrc_t ar = __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(a0), 0, 0x7fffffff,
                                                                                0x00020000);
            const __amdgpu_buffer_rsrc_t br = __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(b0), 0, 0x7fffffff,
                                                                                0x00020000);
#pragma unroll
            for (int e = 0; e < AVT; ++e) {
                const vec_t v = __builtin_bit_cast(vec_t, __builtin_amdgcn_raw_buffer_load_b128(ar, a_off[e], 0, 0));
#pragma unroll
                for (int q = 0; q < VW; ++q) ra[set][e][q] = v[q];
            }
#pragma unroll
            for (int e = 0; e < BVT; ++e) {
                const vec_t v = __builtin_bit_cast(vec_t, __builtin_amdgcn_raw_buffer_load_b128(br, b_off[e], 0, 0));
#pragma unroll
                for (int q = 0; q < VW; ++q) rb[set][e][q] = v[q];
            }
            return;
        }
#pragma unroll
        for (int e = 0; e < AVT; ++e) {
            int i, kk;
            a_coord(e, i, kk);
            const long long gi = m0 + i, gk = k0 + kk;
            if (FAST) {
                const T* src = TA ? p.A + gk + gi * p.lda : p.A + gi + gk * p.lda;
                const vec_t v = *reinterpret_cast<const vec_t*>(src);
#pragma unroll
                for (int q = 0; q < VW; ++q) ra[set][e][q] = v[q];
            } else {
#pragma unroll
                for (int q = 0; q < VW; ++q) {
                    const long long ii = TA ? gi : gi + q, kq = TA ? gk + q : gk;
                    ra[set][e][q] = (ii < p.m && kq < p.k) ? (TA ? p.A[kq + ii * p.lda] : p.A[ii + kq * p.lda]) : T(0);
                }
            }
        }
#pragma unroll
        for (int e = 0; e < BVT; ++e) {
            int j, kk;
            b_coord(e, j, kk);
            const long long gj = n0 + j, gk = k0 + kk;
            if (FAST) {
                const T* src = TB ? p.B + gj + gk * p.ldb : p.B + gk + gj * p.ldb;
                const vec_t v = *reinterpret_cast<const vec_t*>(src);
#pragma unroll
                for (int q = 0; q < VW; ++q) rb[set][e][q] = v[q];
            } else {
#pragma unroll
                for (int q = 0; q < VW; ++q) {
                    const long long jj = TB ? gj + q : gj, kq = TB ? gk : gk + q;
                    rb[set][e][q] = (jj < p.n && kq < p.k) ? (TB ? p.B[jj + kq * p.ldb] : p.B[kq + jj * p.ldb]) : T(0);
                }
            }
        }
    };
    auto store = [&](int buf, int set) {
        T* as = As + buf * BK * LA;
        T* bs = Bs + buf * BK * LB;
#pragma unroll
        for (int e = 0; e < AVT; ++e) {
            int i, kk;
            a_coord(e, i, kk);
            if (!TA) {
                vec_t v;
#pragma unroll
                for (int q = 0; q < VW; ++q) v[q] = ra[set][e][q];
                *reinterpret_cast<vec_t*>(as + kk * LA + i) = v;
            } else {
#pragma unroll
                for (int q = 0; q < VW; ++q) as[(kk + q) * LA + i] = ra[set][e][q];
            }
        }
#pragma unroll
        for (int e = 0; e < BVT; ++e) {
            int j, kk;
            b_coord(e, j, kk);
            if (TB) {
                vec_t v;
#pragma unroll
                for (int q = 0; q < VW; ++q) v[q] = rb[set][e][q];
                *reinterpret_cast<vec_t*>(bs + kk * LB + j) = v;
            } else {
#pragma unroll
                for (int q = 0; q < VW; ++q) bs[(kk + q) * LB + j] = rb[set][e][q];
            }
        }
    };

    auto compute = [&](int buf) {
        const T* as = As + buf * BK * LA + wm * (BM / 2) + (lane & 15);
        const T* bs = Bs + buf * BK * LB + wn * (BN / 2) + (lane & 15);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            const int kr = kk + (lane >> 4);
            T a[TM], b[TN];
#pragma unroll
            for (int t = 0; t < TM; ++t) a[t] = as[kr * LA + t * 16];
#pragma unroll
            for (int t = 0; t < TN; ++t) b[t] = bs[kr * LB + t * 16];
            // D^T = op(B) op(A)^T: lane l % 16 of an accumulator runs along i, the
            // contiguous dimension of C
#pragma unroll
            for (int x = 0; x < TM; ++x)
#pragma unroll
                for (int y = 0; y < TN; ++y) acc[x][y] = Op::mma(b[y], a[x], acc[x][y]);
        }
    };

    if (NBUF == 2 && PF2) {
        // slab s lives in LDS buffer s % 2 and, before that, in register set s % 2
        if (nkt > 0) {
            load(0, 0);
            store(0, 0);
        }
        if (nkt > 1) load(BK, 1);
        __syncthreads();
        for (int kt = 0; kt < nkt; kt += 2) {
            if (kt + 2 < nkt) load((kt + 2) * BK, 0);
            compute(0);
            if (kt + 1 < nkt) store(1, 1);
            __syncthreads();
            if (kt + 1 >= nkt) break;
            if (kt + 3 < nkt) load((kt + 3) * BK, 1);
            compute(1);
            if (kt + 2 < nkt) store(0, 0);
            __syncthreads();
        }
    } else if (NBUF == 2) {
        if (nkt > 0) {
            load(0, 0);
            store(0, 0);
        }
        __syncthreads();
        for (int kt = 0; kt < nkt; ++kt) {
            if (kt + 1 < nkt) load((kt + 1) * BK, 0);
            compute(kt & 1);
            if (kt + 1 < nkt) store((kt & 1) ^ 1, 0);
            __syncthreads();
        }
    } else {
        if (nkt > 0) {
            load(0, 0);
            store(0, 0);
        }
        __syncthreads();
        for (int kt = 0; kt < nkt; ++kt) {
            if (kt + 1 < nkt) load((kt + 1) * BK, 0);
            compute(0);
            __syncthreads();
            if (kt + 1 < nkt) store(0, 0);
            __syncthreads();
        }
    }

    // Epilogue.  Accumulator (x, y, r) of lane l holds C(i, j) with
    // i = 16x + l % 16 and j = 16y + row(l, r) within the wavefront's block.
    // Interior tiles (in bounds and wholly on the written side of a triangle
    // mask) read a row of 16 x 16 blocks of C at a time before writing it, so
    // the reads of that row are in flight together (a write between two reads
    // would have to complete first, since the compiler cannot tell the
    // addresses apart); the other tiles go element by element.
    const int ci = wm * (BM / 2) + (lane & 15);
    bool interior = full;
    if (p.tri == kTriLowerD) interior = interior && (long long)m0 + p.diag >= n0 + BN - 1;
    if (p.tri == kTriUpperD) interior = interior && (long long)m0 + BM - 1 + p.diag <= n0;
    if (interior) {
        T* Ct = p.C + m0 + ci + (long long)(n0 + wn * (BN / 2)) * p.ldc;
        auto off = [&](int x, int y, int r) { return x * 16 + (long long)(y * 16 + Op::row(lane, r)) * p.ldc; };
        if (p.beta != T(0)) {
#pragma unroll
            for (int x = 0; x < TM; ++x) {
                T cv[TN][4];
#pragma unroll
                for (int y = 0; y < TN; ++y)
#pragma unroll
                    for (int r = 0; r < 4; ++r) cv[y][r] = Ct[off(x, y, r)];
#pragma unroll
                for (int y = 0; y < TN; ++y)
#pragma unroll
                    for (int r = 0; r < 4; ++r) Ct[off(x, y, r)] = p.alpha * acc[x][y][r] + p.beta * cv[y][r];
            }
        } else {
#pragma unroll
            for (int x = 0; x < TM; ++x)
#pragma unroll
                for (int y = 0; y < TN; ++y)
#pragma unroll
                    for (int r = 0; r < 4; ++r) Ct[off(x, y, r)] = p.alpha * acc[x][y][r];
        }
        return;
    }
    // Edge and diagonal tiles: each element masked by the bounds and the
    // triangle.  FP64: the same rows of blocks as above, all reads of a row
    // issued before its writes (element by element, each read would also
    // wait for the previous element's write, one counter covering both on
    // gfx950): 6% off the batched left-looking factorization, whose every
    // product starts with a diagonal tile.  FP32 keeps the element loop,
    // which measured 4% faster there.
    const int i0 = m0 + ci, j0 = n0 + wn * (BN / 2);
    auto live = [&](int i, int j) {
        if (!full && (i >= p.m || j >= p.n)) return false;
        if (p.tri == kTriLowerD && (long long)i + p.diag < j) return false;
        if (p.tri == kTriUpperD && (long long)i + p.diag > j) return false;
        return true;
    };
    if constexpr (sizeof(T) == 8) {
#pragma unroll
        for (int x = 0; x < TM; ++x) {
            const int i = i0 + x * 16;
            T cv[TN][4];
            if (p.beta != T(0)) {
#pragma unroll
                for (int y = 0; y < TN; ++y)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int j = j0 + y * 16 + Op::row(lane, r);
                        cv[y][r] = live(i, j) ? p.C[i + (long long)j * p.ldc] : T(0);
                    }
            }
#pragma unroll
            for (int y = 0; y < TN; ++y)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int j = j0 + y * 16 + Op::row(lane, r);
                    if (!live(i, j)) continue;
                    T v = p.alpha * acc[x][y][r];
                    if (p.beta != T(0)) v += p.beta * cv[y][r];
                    p.C[i + (long long)j * p.ldc] = v;
                }
        }
    } else {
#pragma unroll
        for (int x = 0; x < TM; ++x)
#pragma unroll
            for (int y = 0; y < TN; ++y)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = i0 + x * 16;
                    const int j = j0 + y * 16 + Op::row(lane, r);
                    if (!live(i, j)) continue;
                    T* c = p.C + i + (long long)j * p.ldc;
                    T v = p.alpha * acc[x][y][r];
                    if (p.beta != T(0)) v += p.beta * *c;
                    *c = v;
                }
    }
}

template <class T, int BM, int BN, int BK, bool TA, bool TB, int NBUF = 2, bool BUF = true>
__device__ __forceinline__ void gemm_tile(const GemmProblem<T>& p, int m0, int n0, T* smem)
{
    if (p.tri == kTriLowerD && (long long)(m0 + BM - 1) + p.diag < n0) return;
    if (p.tri == kTriUpperD && (long long)m0 + p.diag > n0 + BN - 1) return;
    const bool fast = p.vec_ab && m0 + BM <= p.m && n0 + BN <= p.n && p.k % BK == 0 &&
                      (!BUF || (p.lda <= (1 << 20) && p.ldb <= (1 << 20)));
    if (fast) gemm_tile_body<T, BM, BN, BK, TA, TB, true, NBUF, false, BUF>(p, m0, n0, smem);
    else gemm_tile_body<T, BM, BN, BK, TA, TB, false, NBUF>(p, m0, n0, smem);
}

template <class T, int BM, int BN, int BK, bool TA, bool TB, int WAVES_PER_EU>
__global__ __launch_bounds__(256, WAVES_PER_EU) void gemm_kernel(GemmProblem<T> p, TileOrder o, const int* abort)
{
    batch_offset(p, abort);
    if (abort != nullptr && *abort != 0) return;
    int ti, tj;
    if (!map_tile(o, blockIdx.x, ti, tj)) return;
    __shared__ T smem[lds_elems<BM, BN, BK>()];
    gemm_tile<T, BM, BN, BK, TA, TB>(p, ti * BM, tj * BN, smem);
}

// The same tile loop under its own symbol, launched for the triangle-masked
// (syrk-shaped) trailing updates of the factorization so that kernel traces
// report them apart from plain products.
template <class T, int BM, int BN, int BK, bool TA, bool TB, int WAVES_PER_EU, int NBUF = 2>
__global__ __launch_bounds__(256, WAVES_PER_EU) void syrk_update_kernel(GemmProblem<T> p, TileOrder o,
                                                                         const int* abort)
{
    batch_offset(p, abort);
    if (abort != nullptr && *abort != 0) return;
    int ti, tj;
    if (!map_tile(o, blockIdx.x, ti, tj)) return;
    __shared__ T smem[lds_elems<BM, BN, BK, NBUF>()];
    gemm_tile<T, BM, BN, BK, TA, TB, NBUF>(p, ti * BM, tj * BN, smem);
}

// The same tile loop once more, for the factorization's look-ahead updates of
// the next panels' columns, so that kernel traces keep them apart from the
// trailing updates above.
template <class T, int BM, int BN, int BK, bool TA, bool TB, int WAVES_PER_EU>
__global__ __launch_bounds__(256, WAVES_PER_EU) void panel_update_kernel(GemmProblem<T> p, TileOrder o,
                                                                          const int* abort)
{
    batch_offset(p, abort);
    if (abort != nullptr && *abort != 0) return;
    int ti, tj;
    if (!map_tile(o, blockIdx.x, ti, tj)) return;
    __shared__ T smem[lds_elems<BM, BN, BK>()];
    gemm_tile<T, BM, BN, BK, TA, TB>(p, ti * BM, tj * BN, smem);
}

template <class T>
struct VBatch {
    const int *m, *n, *k;
    T alpha, beta;
    const T* const* A;
    const int* lda;
    const T* const* B;
    const int* ldb;
    T* const* C;
    const int* ldc;
    int tri;  // triangle mask of every C (kTri*D, diagonal offset 0)
};

template <class T, int BM, int BN, int BK, bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_vbatched_kernel(VBatch<T> v, int offset)
{
    const int b = offset + blockIdx.z;
    GemmProblem<T> p;
    p.m = v.m[b];
    p.n = v.n[b];
    p.k = v.k[b];
    if ((int)blockIdx.x * BM >= p.m || (int)blockIdx.y * BN >= p.n) return;
    p.alpha = v.alpha;
    p.beta = v.beta;
    p.A = v.A[b];
    p.lda = v.lda[b];
    p.B = v.B[b];
    p.ldb = v.ldb[b];
    p.C = v.C[b];
    p.ldc = v.ldc[b];
    p.tri = v.tri;
    p.diag = 0;
    // 16-byte loads when this member's operands allow them
    constexpr int VW = Mfma<T>::kVec;
    p.vec_ab = ((uintptr_t)p.A % 16 == 0) && ((uintptr_t)p.B % 16 == 0) && p.lda % VW == 0 && p.ldb % VW == 0;
    __shared__ T smem[lds_elems<BM, BN, BK>()];
    gemm_tile<T, BM, BN, BK, TA, TB>(p, blockIdx.x * BM, blockIdx.y * BN, smem);
}

} // namespace gemm_detail
} // namespace kern
} // namespace ddla
