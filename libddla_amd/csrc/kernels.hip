// Hand-written CDNA4 (gfx950) kernels of the Cholesky path.
//
// Matrix products run on the matrix cores: v_mfma_f64_16x16x4_f64 for double
// and v_mfma_f32_16x16x4_f32 for float, one 64-wide wavefront per 16x16 MFMA
// tile, four wavefronts per workgroup, operands staged through double-buffered
// LDS tiles.  The diagonal-block factorization and the triangular inverses run
// in one workgroup on LDS-resident 64 x 64 blocks.
#include "kernels.h"
#include "gemm_tile.h"
#include "solve_tile.h"
#include "stage_tile.h"
#include "potf2_tile.h"
#include "panel_tile.h"
#include "batched_tile.h"
#include "split_tile.h"

#include <atomic>
#include <cfloat>
#include <cstdint>
#include <climits>
#include <stdexcept>
#include <string>

namespace ddla {
namespace kern {

namespace {

inline void check_launch(const char* what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

using namespace gemm_detail;

__global__ __launch_bounds__(256) void vbatched_validate_kernel(int na, int nb, const int* m, const int* n,
                                                                const int* k, const int* lda, const int* ldb,
                                                                const int* ldc, int batch, int* summary)
{
    __shared__ int red[4][256];
    int mm = 0, mn = 0, mk = 0, bad = INT_MAX;
    for (int b = threadIdx.x; b < batch; b += 256) {
        const int x = m[b], y = n[b], z = k[b];
        if (x < 0 || y < 0 || z < 0) { bad = min(bad, b); continue; }
        const int ra = na ? max(x, 1) : max(z, 1);
        const int rb = nb ? max(z, 1) : max(y, 1);
        if (lda[b] < ra || ldb[b] < rb || ldc[b] < max(x, 1)) { bad = min(bad, b); continue; }
        mm = max(mm, x);
        mn = max(mn, y);
        mk = max(mk, z);
    }
    red[0][threadIdx.x] = mm;
    red[1][threadIdx.x] = mn;
    red[2][threadIdx.x] = mk;
    red[3][threadIdx.x] = bad;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            for (int q = 0; q < 3; ++q) red[q][threadIdx.x] = max(red[q][threadIdx.x], red[q][threadIdx.x + s]);
            red[3][threadIdx.x] = min(red[3][threadIdx.x], red[3][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        summary[0] = red[0][0];
        summary[1] = red[1][0];
        summary[2] = red[2][0];
        summary[3] = red[3][0] == INT_MAX ? 0 : red[3][0] + 1;
    }
}

// ---------------------------------------------------------------------------
// 64 x 64 diagonal blocks: factorization and triangular inverse in LDS.
// ---------------------------------------------------------------------------
constexpr int kIB = 64;

// inv(block b) = L(b, b)^-1 for every 64 x 64 diagonal block of a factor
// (unit: the diagonal taken as ones, the stored one never used).
template <class T>
__global__ __launch_bounds__(256) void trtri_kernel(bool upper, bool unit, int n, const T* A, long long lda, T* inv,
                                                    long long bstride)
{
    using namespace potf2_detail;
    __shared__ T S[kB * kL], X[kB * kL], rdiag[kB];
    if (gridDim.y > 1) {  // problem blockIdx.y of a batch: A and inv bstride bytes apart
        A = reinterpret_cast<const T*>(reinterpret_cast<const char*>(A) + blockIdx.y * bstride);
        inv = reinterpret_cast<T*>(reinterpret_cast<char*>(inv) + blockIdx.y * bstride);
    }
    const long long j0 = (long long)blockIdx.x * kIB;
    const int w = (int)min((long long)kIB, n - j0);
    load_block<T>(upper, w, A + j0 + j0 * lda, lda, S);
    __syncthreads();
    if (threadIdx.x < kB) {
        if (unit) S[threadIdx.x * kL + threadIdx.x] = T(1);
        rdiag[threadIdx.x] = T(1) / S[threadIdx.x * kL + threadIdx.x];
    }
    __syncthreads();
    invert_lower<T>(S, rdiag, X);
    store_inverse<T>(X, inv + (long long)blockIdx.x * kIB * kIB);
}

// ---------------------------------------------------------------------------
// diag_logdet: fixed-shape two-pass reduction (deterministic summation order).
// ---------------------------------------------------------------------------
constexpr int kLogdetParts = 64;

template <class T>
__global__ __launch_bounds__(256) void logdet_partial_kernel(int n, const T* A, long long lda, double* pmag,
                                                             int* pflags)
{
    __shared__ double sm[256];
    __shared__ int sf[256];
    double m = 0;
    int f = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const double x = (double)A[(long long)i * (lda + 1)];
        if (!(x <= DBL_MAX && x >= -DBL_MAX)) { f |= 1; continue; }
        if (x == 0) { f |= 2; continue; }
        m += log(fabs(x));
        if (x < 0) f ^= 4;
    }
    sm[threadIdx.x] = m;
    sf[threadIdx.x] = f;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sm[threadIdx.x] += sm[threadIdx.x + s];
            const int a = sf[threadIdx.x], b = sf[threadIdx.x + s];
            sf[threadIdx.x] = ((a | b) & 3) | ((a ^ b) & 4);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pmag[blockIdx.x] = sm[0];
        pflags[blockIdx.x] = sf[0];
    }
}

__global__ void logdet_final_kernel(const double* pmag, const int* pflags, LogdetResult* out)
{
    if (threadIdx.x != 0) return;
    double m = 0;
    int f = 0;
    for (int b = 0; b < kLogdetParts; ++b) {
        m += pmag[b];
        f = ((f | pflags[b]) & 3) | ((f ^ pflags[b]) & 4);
    }
    out->magnitude = m;
    out->flags = f;
    out->pad = 0;
}

// ---------------------------------------------------------------------------
// Mixed-precision refinement helpers.
// ---------------------------------------------------------------------------
// One 64 x 64 tile (ib, jb) of the stored triangle: converts it to D and
// writes, for the infinity norm of the symmetric matrix, the tile's row sums
// of |a| (prow[jb * n + i]) and its column sums of the mirrored entries
// (pcol[ib * n + j]; the diagonal element is counted once, in its row).
template <class D, class S>
__global__ __launch_bounds__(256) void convert_sym_norm_kernel(int n, int lower, const S* in, long long ldi, D* out,
                                                               long long ldo, double* prow, double* pcol)
{
    const int ib = blockIdx.x, jb = blockIdx.y;
    if (lower ? jb > ib : jb < ib) return;
    __shared__ double Ts[64][65];
    const long long i0 = (long long)ib * 64, j0 = (long long)jb * 64;
    const int rows = (int)min(64LL, n - i0), cols = (int)min(64LL, n - j0);
    const bool diag = ib == jb;
    // every read of the thread's 16 elements before its writes (element by
    // element, each read would also wait for the previous element's write,
    // one counter covering both)
    S x[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const int idx = threadIdx.x + 256 * e, r = idx & 63, c = idx >> 6;
        const bool live = r < rows && c < cols && (!diag || (lower ? r >= c : r <= c));
        x[e] = live ? in[(i0 + r) + (j0 + c) * ldi] : S(0);
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const int idx = threadIdx.x + 256 * e, r = idx & 63, c = idx >> 6;
        const bool live = r < rows && c < cols && (!diag || (lower ? r >= c : r <= c));
        if (live) out[(i0 + r) + (j0 + c) * ldo] = (D)x[e];
        Ts[r][c] = live ? fabs((double)x[e]) : 0.0;
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t < 64) {
        if (t < rows) {
            double sum = 0;
            for (int c = 0; c < 64; ++c) sum += Ts[t][c];
            prow[(long long)jb * n + i0 + t] = sum;
        }
    } else if (t < 128) {
        const int c = t - 64;
        if (c < cols) {
            double sum = 0;
            for (int r = 0; r < 64; ++r)
                if (!(diag && r == c)) sum += Ts[r][c];
            pcol[(long long)ib * n + j0 + c] = sum;
        }
    }
}

// rowsum[i] = sum over the tiles of row i's block row (prow) and block column
// (pcol), in a fixed order.
__global__ void sym_rowsum_kernel(int n, int nblk, int lower, const double* prow, const double* pcol, double* rowsum)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int bi = i / 64;
        double s = 0;
        // unrolled so that several loads are in flight; the sum keeps its order
        if (lower) {
#pragma unroll 8
            for (int jb = 0; jb <= bi; ++jb) s += prow[(long long)jb * n + i];
#pragma unroll 8
            for (int ib = bi; ib < nblk; ++ib) s += pcol[(long long)ib * n + i];
        } else {
#pragma unroll 8
            for (int jb = bi; jb < nblk; ++jb) s += prow[(long long)jb * n + i];
#pragma unroll 8
            for (int ib = 0; ib <= bi; ++ib) s += pcol[(long long)ib * n + i];
        }
        rowsum[i] = s;
    }
}

// out[j] = max_i |M(i, j)|, i < m, one workgroup per column (NaN propagates).
template <class T>
__global__ __launch_bounds__(256) void colmax_kernel(int m, const T* M, long long ld, double* out)
{
    __shared__ double red[256];
    const int j = blockIdx.x;
    double v = 0;
#pragma unroll 8
    for (int i = threadIdx.x; i < m; i += 256) {
        const double x = fabs((double)M[i + (long long)j * ld]);
        v = (x > v || x != x) ? x : v;
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const double o = red[threadIdx.x + s];
            if (o > red[threadIdx.x] || o != o) red[threadIdx.x] = o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = red[0];
}

// X(i, j) += (double) W(i, j)
template <class S>
__global__ void add_converted_kernel(int m, int n, const S* W, long long ldw, double* X, long long ldx)
{
    const long long total = (long long)m * n;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t % m), j = (int)(t / m);
        X[i + j * ldx] += (double)W[i + j * ldw];
    }
}

// ---------------------------------------------------------------------------
// Setup helpers.
// ---------------------------------------------------------------------------
__device__ void philox4x64_10(uint64_t c[4], uint64_t k0, uint64_t k1)
{
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += 0x9E3779B97F4A7C15ULL;
            k1 += 0xBB67AE8584CAA73BULL;
        }
        const uint64_t a0 = 0xD2E7470EE14C6C93ULL * c[0], h0 = __umul64hi(0xD2E7470EE14C6C93ULL, c[0]);
        const uint64_t a1 = 0xCA5A826395121157ULL * c[2], h1 = __umul64hi(0xCA5A826395121157ULL, c[2]);
        const uint64_t n0 = h1 ^ c[1] ^ k0, n2 = h0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = a1;
        c[2] = n2;
        c[3] = a0;
    }
}

__global__ void fill_normal_kernel(uint64_t seed, uint64_t stream, uint64_t first, long long count, double* out)
{
    const double two_pi = 6.283185307179586476925286766559;
    const double scale = 1.0 / 9007199254740992.0;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < count;
         t += (long long)gridDim.x * blockDim.x) {
        const uint64_t e = first + (uint64_t)t;
        uint64_t c[4] = {e >> 2, stream, 0, 0};
        philox4x64_10(c, seed, 0);
        const int word = (int)(e & 3), pair = word >> 1;
        const double u1 = (double)((c[2 * pair] >> 11) + 1) * scale;
        const double u2 = (double)(c[2 * pair + 1] >> 11) * scale;
        const double r = sqrt(-2.0 * log(u1));
        out[t] = (word & 1) ? r * sin(two_pi * u2) : r * cos(two_pi * u2);
    }
}

template <class T>
__global__ void add_diagonal_kernel(int n, T* A, long long lda, T delta)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        A[i + (long long)i * lda] += delta;
}

template <class T>
__global__ __launch_bounds__(256) void transpose_kernel(int m, int n, const T* in, long long ldi, T* out,
                                                        long long ldo)
{
    __shared__ T tile[32][33];
    const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int r = ty; r < 32; r += 8) {
        const int i = i0 + tx, j = j0 + r;
        if (i < m && j < n) tile[r][tx] = in[i + (long long)j * ldi];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int j = j0 + tx, i = i0 + r;  // out(j, i) = in(i, j)
        if (i < m && j < n) out[j + (long long)i * ldo] = tile[tx][r];
    }
}

// One 64 x 64 tile of out per workgroup (64 x 4 threads), staged through LDS:
// tile[b][a] = in(r0 + a, c0 + b) for the source block (r0, c0) of the tile,
// loaded down columns of in and stored down columns of out.
//   transpose (out(i, j) = in(n-1-j, n-1-i)): r0 = n - j0 - 64, c0 = n - i0 - 64,
//     out(i0 + di, j0 + dj) = tile[63 - di][63 - dj]; in's triangle = out's;
//   rotate (out(i, j) = in(n-1-i, n-1-j)):    r0 = n - i0 - 64, c0 = n - j0 - 64,
//     out(i0 + di, j0 + dj) = tile[63 - dj][63 - di]; in's triangle = the other.
template <class T>
__global__ __launch_bounds__(256) void reverse_triangle_kernel(bool out_upper, bool transpose, int n, const T* in,
                                                               long long ldi, T* out, long long ldo, bool zero_other)
{
    __shared__ T tile[64][65];
    const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 64;
    const bool outside = out_upper ? (i0 > j0 + 63) : (i0 + 63 < j0);  // no element of out's triangle
    if (outside && !zero_other) return;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    if (!outside) {
        const bool in_upper = transpose ? out_upper : !out_upper;
        const int r0 = n - (transpose ? j0 : i0) - 64, c0 = n - (transpose ? i0 : j0) - 64;
        for (int b = ty; b < 64; b += 4) {
            const int r = r0 + tx, c = c0 + b;  // r, c < n always
            const bool stored = in_upper ? (r <= c) : (r >= c);
            if (r >= 0 && c >= 0 && stored) tile[b][tx] = in[r + (long long)c * ldi];
        }
    }
    __syncthreads();
    const int i = i0 + tx;
    if (i >= n) return;
    for (int dj = ty; dj < 64; dj += 4) {
        const int j = j0 + dj;
        if (j >= n) break;
        const bool stored = out_upper ? (i <= j) : (i >= j);
        if (stored)
            out[i + (long long)j * ldo] = transpose ? tile[63 - tx][63 - dj] : tile[63 - dj][63 - tx];
        else if (zero_other)
            out[i + (long long)j * ldo] = T(0);
    }
}

__device__ inline bool tri_selected(bool upper, bool diag, int i, int j)
{
    return upper ? (diag ? i <= j : i < j) : (diag ? i >= j : i > j);
}

// (rows / 256, min(n, 65535)) workgroups, one thread per element; columns
// grid-strided.  Elements whose size and address allow it move as 32-bit
// words, others byte by byte.
__global__ __launch_bounds__(256) void memset_tri_kernel(bool upper, bool diag, int m, int n, unsigned char* p,
                                                         long long ld_bytes, int elem, unsigned char value)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    const unsigned int word = 0x01010101u * value;
    for (int j = blockIdx.y; j < n; j += gridDim.y) {
        if (!tri_selected(upper, diag, i, j)) continue;
        unsigned char* e = p + (long long)j * ld_bytes + (long long)i * elem;
        if (elem % 4 == 0 && (uintptr_t)e % 4 == 0) {
            for (int w = 0; w < elem / 4; ++w) reinterpret_cast<unsigned int*>(e)[w] = word;
        } else {
            for (int b = 0; b < elem; ++b) e[b] = value;
        }
    }
}

__global__ __launch_bounds__(256) void memcpy_tri_kernel(bool upper, bool diag, int m, int n, unsigned char* d,
                                                         long long ldd_bytes, const unsigned char* s,
                                                         long long lds_bytes, int elem)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    for (int j = blockIdx.y; j < n; j += gridDim.y) {
        if (!tri_selected(upper, diag, i, j)) continue;
        unsigned char* e = d + (long long)j * ldd_bytes + (long long)i * elem;
        const unsigned char* f = s + (long long)j * lds_bytes + (long long)i * elem;
        if (elem % 4 == 0 && (uintptr_t)e % 4 == 0 && (uintptr_t)f % 4 == 0) {
            for (int w = 0; w < elem / 4; ++w)
                reinterpret_cast<unsigned int*>(e)[w] = reinterpret_cast<const unsigned int*>(f)[w];
        } else {
            for (int b = 0; b < elem; ++b) e[b] = f[b];
        }
    }
}

// In-place exchange of the strict lower and strict upper triangles of the
// order-n A: A(i, j) <-> A(j, i) for i > j.  The workgroup of the 32 x 32 tile
// pair (bi, bj), bi >= bj, loads the strict-lower part of tile (bi, bj) and
// the strict-upper part of tile (bj, bi) into LDS and writes each into the
// other's place: every element is read and written by one workgroup only.
template <class T>
__global__ __launch_bounds__(256) void swap_triangles_kernel(int n, T* A, long long lda)
{
    __shared__ T lo[32][33], up[32][33];  // lo[c][r] = A(i0 + r, j0 + c), up[c][r] = A(j0 + r, i0 + c)
    const int bi = blockIdx.x, bj = blockIdx.y;
    if (bi < bj) return;
    const int i0 = bi * 32, j0 = bj * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int c = ty; c < 32; c += 8) {
        const int li = i0 + tx, lj = j0 + c;  // lower element
        if (li < n && lj < n && li > lj) lo[c][tx] = A[li + (long long)lj * lda];
        const int ui = j0 + tx, uj = i0 + c;  // upper element
        if (ui < n && uj < n && ui < uj) up[c][tx] = A[ui + (long long)uj * lda];
    }
    __syncthreads();
    for (int c = ty; c < 32; c += 8) {
        const int li = i0 + tx, lj = j0 + c;  // A(li, lj) <- A(lj, li) = up[tx][c]
        if (li < n && lj < n && li > lj) A[li + (long long)lj * lda] = up[tx][c];
        const int ui = j0 + tx, uj = i0 + c;  // A(ui, uj) <- A(uj, ui) = lo[tx][c]
        if (ui < n && uj < n && ui < uj) A[ui + (long long)uj * lda] = lo[tx][c];
    }
}

template <class D, class S>
__global__ void convert_kernel(int m, int n, const S* in, long long ldi, D* out, long long ldo)
{
    const long long total = (long long)m * n;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t % m), j = (int)(t / m);
        out[i + j * ldo] = (D)in[i + j * ldi];
    }
}

int grid_for(long long work, int per_block, int cap = 65536)
{
    long long g = (work + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

bool is_n(char op) { return op == 'N' || op == 'n'; }

// Builds the tile schedule of one launch (see map_tile) and launches it.
template <class T, int BM, int BN, int BK, int WPE>
void launch_gemm_cfg(hipStream_t s, bool ta, bool tb, const GemmProblem<T>& p, const int* abort, bool lookahead,
                     int zcount)
{
    TileOrder o;
    o.ntm = (p.m + BM - 1) / BM;
    o.ntn = (p.n + BN - 1) / BN;
    o.lower_enum = (p.tri == kTriLower && p.diag == 0 && p.m == p.n && BM == BN) ? 1 : 0;
    const long long total = o.lower_enum ? (long long)o.ntm * (o.ntm + 1) / 2 : (long long)o.ntm * o.ntn;
    if (total > (long long)INT_MAX / 2) throw std::invalid_argument("gemm: problem too large for one launch");
    o.total = (int)total;
    o.group = 8;
    o.per_xcd = (o.total + 7) / 8;
    const dim3 grid((unsigned)(o.per_xcd * 8), 1, (unsigned)zcount), block(256);
    if (p.tri != kTriNoneD && !lookahead) {
        // the trailing updates: C (N T) or its mirror (T N)
        if (!ta && tb) syrk_update_kernel<T, BM, BN, BK, false, true, WPE><<<grid, block, 0, s>>>(p, o, abort);
        else if (ta && !tb) syrk_update_kernel<T, BM, BN, BK, true, false, WPE><<<grid, block, 0, s>>>(p, o, abort);
        else throw std::invalid_argument("gemm: triangle-masked products are N T or T N");
    } else if (p.tri != kTriNoneD) {
        // the look-ahead updates of the next panels' columns
        if (!ta && tb) panel_update_kernel<T, BM, BN, BK, false, true, WPE><<<grid, block, 0, s>>>(p, o, abort);
        else if (ta && !tb) panel_update_kernel<T, BM, BN, BK, true, false, WPE><<<grid, block, 0, s>>>(p, o, abort);
        else throw std::invalid_argument("gemm: triangle-masked products are N T or T N");
    } else if (!ta && !tb) {
        gemm_kernel<T, BM, BN, BK, false, false, WPE><<<grid, block, 0, s>>>(p, o, abort);
    } else if (!ta && tb) {
        gemm_kernel<T, BM, BN, BK, false, true, WPE><<<grid, block, 0, s>>>(p, o, abort);
    } else if (ta && !tb) {
        gemm_kernel<T, BM, BN, BK, true, false, WPE><<<grid, block, 0, s>>>(p, o, abort);
    } else {
        gemm_kernel<T, BM, BN, BK, true, true, WPE><<<grid, block, 0, s>>>(p, o, abort);
    }
    check_launch("gemm");
}

template <class T, int BM, int BN, int BK>
void launch_vbatched_cfg(hipStream_t s, bool ta, bool tb, const VBatch<T>& v, int batch, int max_m, int max_n)
{
    const unsigned gx = (unsigned)((max_m + BM - 1) / BM), gy = (unsigned)((max_n + BN - 1) / BN);
    if (gx == 0 || gy == 0) return;
    for (int off = 0; off < batch; off += 65535) {
        const unsigned gz = (unsigned)std::min(65535, batch - off);
        const dim3 grid(gx, gy, gz), block(256);
        if (!ta && !tb) gemm_vbatched_kernel<T, BM, BN, BK, false, false><<<grid, block, 0, s>>>(v, off);
        else if (!ta && tb) gemm_vbatched_kernel<T, BM, BN, BK, false, true><<<grid, block, 0, s>>>(v, off);
        else if (ta && !tb) gemm_vbatched_kernel<T, BM, BN, BK, true, false><<<grid, block, 0, s>>>(v, off);
        else gemm_vbatched_kernel<T, BM, BN, BK, true, true><<<grid, block, 0, s>>>(v, off);
        check_launch("gemm_vbatched");
    }
}

} // namespace

template <class T>
void gemm(hipStream_t s, char ta, char tb, int m, int n, int k, T alpha, const T* A, long long lda, const T* B,
          long long ldb, T beta, T* C, long long ldc, Tri tri, long long diag, const int* abort, bool lookahead,
          const Batch& batch)
{
    if (m <= 0 || n <= 0 || batch.count <= 0 || batch.inner <= 0) return;
    constexpr int VW = 16 / sizeof(T);
    // 16-byte loads need every problem's and sub-block's operands aligned too
    const int vec = ((uintptr_t)A % 16 == 0) && ((uintptr_t)B % 16 == 0) && lda % VW == 0 && ldb % VW == 0 &&
                    (batch.count == 1 || batch.stride % 16 == 0) && (batch.inner == 1 || batch.inner_stride % 16 == 0);
    const long long zcount = (long long)batch.count * batch.inner;
    if (zcount > 65535) throw std::invalid_argument("gemm: more than 65535 problems x sub-blocks in one launch");
    GemmProblem<T> p{m, n, k, alpha, beta, A, lda, B, ldb, C, ldc, (int)tri, diag, vec, batch.inner, batch.stride,
                     batch.inner_stride};
    const bool a_t = !is_n(ta), b_t = !is_n(tb);
    // 128 x 128 tiles unless there are fewer than 512 of them (two per CU)
    // and 64 x 64 tiles give at least twice as many
    // (counting every problem and sub-block of a batched launch)
    const long long t128 = (long long)((m + 127) / 128) * ((n + 127) / 128) * zcount;
    const long long t64 = (long long)((m + 63) / 64) * ((n + 63) / 64) * zcount;
    if (m >= 128 && n >= 128 && (t128 >= 512 || t64 < 2 * t128))
        launch_gemm_cfg<T, 128, 128, 16, 2>(s, a_t, b_t, p, abort, lookahead, (int)zcount);
    else
        launch_gemm_cfg<T, 64, 64, 16, 2>(s, a_t, b_t, p, abort, lookahead, (int)zcount);
}

template <class T>
void gemm_vbatched(hipStream_t s, char ta, char tb, const int* m, const int* n, const int* k, T alpha,
                   const T* const* A, const int* lda, const T* const* B, const int* ldb, T beta, T* const* C,
                   const int* ldc, int batch, int max_m, int max_n, Tri tri)
{
    if (batch <= 0 || max_m <= 0 || max_n <= 0) return;
    VBatch<T> v{m, n, k, alpha, beta, A, lda, B, ldb, C, ldc, (int)tri};
    launch_vbatched_cfg<T, 64, 64, 16>(s, !is_n(ta), !is_n(tb), v, batch, max_m, max_n);
}

namespace {
std::atomic<int> g_wait_limit{kDefaultWaitLimit};
} // namespace

int wait_limit() { return g_wait_limit.load(std::memory_order_relaxed); }

int set_wait_limit(int spins)
{
    return g_wait_limit.exchange(spins > 0 ? spins : kDefaultWaitLimit, std::memory_order_relaxed);
}

// rows of a split buffer: m rounded up to the 128-row column tiles, plus 256
// zero rows, so that a 256-row tile that starts inside the rounded range reads
// only written rows
int split_rows(int m) { return (m + split_detail::kCols - 1) / split_detail::kCols * split_detail::kCols + split_detail::kRows; }

namespace {
void check_split_mode(int mode)
{
    if (mode != split_detail::kSplitBf16x3 && mode != split_detail::kSplitBf16x6 && mode != split_detail::kSplitFp16x3)
        throw std::invalid_argument("split update: mode must be 1 (bf16x3), 2 (bf16x6) or 3 (fp16x3)");
}

// bytes of the fragment data of a split of `rows` rows and kb columns; the
// fp16x3 row exponents follow it
size_t split_fragment_bytes(int rows, int nks, int planes) { return (size_t)rows * nks * split_detail::kSlab * planes * 2; }
} // namespace

size_t split_buffer_bytes(int n, int mode)
{
    check_split_mode(mode);
    const int rows = split_rows(n), planes = split_detail::split_planes(mode);
    return split_fragment_bytes(rows, 512 / split_detail::kSlab, planes) + (size_t)rows * sizeof(int);
}

void split_panel(hipStream_t s, int mode, bool upper, int m, int kb, const float* X, long long ldx, void* out)
{
    check_split_mode(mode);
    if (m <= 0 || kb <= 0) return;
    if (kb > 512) throw std::invalid_argument("split_panel: at most 512 columns");
    const int rows = split_rows(m), nks = (kb + split_detail::kSlab - 1) / split_detail::kSlab;
    const int planes = split_detail::split_planes(mode);
    auto* o = static_cast<split_detail::u4_t*>(out);
    int* exps = reinterpret_cast<int*>(static_cast<char*>(out) + split_fragment_bytes(rows, nks, planes));
    const dim3 grid((unsigned)((rows + 255) / 256), (unsigned)(4 * nks)), block(256);
    if (mode == split_detail::kSplitFp16x3) {
        split_detail::split_rowexp_kernel<<<(rows + 63) / 64, block, 0, s>>>(m, kb, rows, X, ldx, upper, exps);
        check_launch("split_panel");
        split_detail::split_panel_kernel<2, true><<<grid, block, 0, s>>>(m, kb, rows, nks, X, ldx, upper, exps, o);
    } else if (mode == split_detail::kSplitBf16x6) {
        split_detail::split_panel_kernel<3, false><<<grid, block, 0, s>>>(m, kb, rows, nks, X, ldx, upper, nullptr, o);
    } else {
        split_detail::split_panel_kernel<2, false><<<grid, block, 0, s>>>(m, kb, rows, nks, X, ldx, upper, nullptr, o);
    }
    check_launch("split_panel");
}

void split_update(hipStream_t s, int mode, bool upper, int m, int n2, int kb, const void* P, int m_split,
                  int a_row, int b_row, float* C, long long ldc, bool tri, long long diag, const int* abort)
{
    check_split_mode(mode);
    if (upper) throw std::invalid_argument("split_update: the split products update the lower view only");
    if (m <= 0 || n2 <= 0 || kb <= 0) return;
    using split_detail::kCols;
    using split_detail::kRows;
    const int planes = split_detail::split_planes(mode), rows = split_rows(m_split);
    const int nI = (m + kRows - 1) / kRows, nt = (n2 + kCols - 1) / kCols;
    if (kb > 512 || a_row % 16 != 0 || b_row % 16 != 0 || a_row < 0 || b_row < 0 ||
        a_row + (long long)nI * kRows > rows || b_row + (long long)nt * kCols > rows)
        throw std::invalid_argument("split_update: rows outside the split buffer");
    if ((long long)kCols * ldc >= INT_MAX) throw std::invalid_argument("split_update: leading dimension too large");
    const int nks = (kb + split_detail::kSlab - 1) / split_detail::kSlab;
    const long long rb_units = (long long)nks * planes * split_detail::kFragU4;  // 16-byte units per 16-row block
    const auto* base = static_cast<const split_detail::u4_t*>(P);
    const int* exps = reinterpret_cast<const int*>(static_cast<const char*>(P) + split_fragment_bytes(rows, nks, planes));
    split_detail::SplitProblem p{m, n2, nks, base + (a_row / 16) * rb_units, base + (b_row / 16) * rb_units, C,
                                 ldc, tri ? 1 : 0, diag, exps + a_row, exps + b_row};
    split_detail::SplitOrder o;
    o.nI = nI;
    o.nt = nt;
    o.lower = (tri && diag == 0 && m == n2) ? 1 : 0;
    long long total = 0;
    if (o.lower) {
        for (int g = 0; g < (nt + 7) / 8; ++g) total += split_detail::split_group_tiles(g, nI, nt);
    } else {
        total = (long long)nI * nt;
    }
    if (total > (long long)INT_MAX / 2) throw std::invalid_argument("split_update: problem too large for one launch");
    o.total = (int)total;
    o.per_xcd = (o.total + 7) / 8;
    const dim3 grid((unsigned)(o.per_xcd * 8)), block(256);
    if (mode == split_detail::kSplitBf16x3) split_detail::split_syrk_kernel<split_detail::kSplitBf16x3><<<grid, block, 0, s>>>(p, o, abort);
    else if (mode == split_detail::kSplitBf16x6) split_detail::split_syrk_kernel<split_detail::kSplitBf16x6><<<grid, block, 0, s>>>(p, o, abort);
    else split_detail::split_syrk_kernel<split_detail::kSplitFp16x3><<<grid, block, 0, s>>>(p, o, abort);
    check_launch("split_update");
}

void vbatched_validate(hipStream_t s, int na, int nb, const int* m, const int* n, const int* k, const int* lda,
                       const int* ldb, const int* ldc, int batch, int* summary)
{
    vbatched_validate_kernel<<<1, 256, 0, s>>>(na, nb, m, n, k, lda, ldb, ldc, batch, summary);
    check_launch("vbatched_validate");
}

// panel work: [ kMaxCols inverses | kSyncInts counters per 64-column block | error word ]
template <class T>
size_t panel_work_bytes(int n)
{
    using namespace panel_detail;
    const size_t nb = (size_t)(n + kB - 1) / kB;
    return ((size_t)kMaxCols * kB * kB * sizeof(T) + 255) / 256 * 256 + ((nb * kSyncInts + 1) * sizeof(int) + 255) / 256 * 256;
}

namespace {
template <class T>
int* panel_sync(void* work)
{
    using namespace panel_detail;
    return reinterpret_cast<int*>(static_cast<char*>(work) + ((size_t)kMaxCols * kB * kB * sizeof(T) + 255) / 256 * 256);
}
} // namespace

template <class T>
int* panel_error_word(int n, void* work)
{
    using namespace panel_detail;
    return panel_sync<T>(work) + (size_t)((n + kB - 1) / kB) * kSyncInts;
}

namespace {
// p[0, count) := 0 for each problem y (problem y's words stride bytes after problem 0's)
__global__ void zero_ints_batched_kernel(int* p, int count, long long stride)
{
    int* q = reinterpret_cast<int*>(reinterpret_cast<char*>(p) + blockIdx.y * stride);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) q[i] = 0;
}
} // namespace

template <class T>
void panel_reset(hipStream_t s, int n, void* work, const Batch& batch)
{
    using namespace panel_detail;
    const int count = ((n + kB - 1) / kB) * kSyncInts + 1;
    if (batch.count == 1) {
        solve_detail::zero_ints_kernel<<<grid_for(count, 256), 256, 0, s>>>(panel_sync<T>(work), count);
    } else if (batch.count > 1) {
        const dim3 grid((unsigned)grid_for(count, 256), (unsigned)batch.count);
        zero_ints_batched_kernel<<<grid, 256, 0, s>>>(panel_sync<T>(work), count, batch.stride);
    }
    check_launch("panel_reset");
}

template <class T>
void panel_factor(hipStream_t s, bool upper, int n, T* A, long long lda, int k, int kb, int* info, void* work,
                  const Batch& batch)
{
    using namespace panel_detail;
    if (kb <= 0 || k < 0 || k + kb > n) return;
    if (kb > kMaxCols * kB) throw std::invalid_argument("panel_factor: panel wider than 768 columns");
    if (k % kB != 0) throw std::invalid_argument("panel_factor: panel must start on a 64-column boundary");
    PanelArgs<T> a{A, lda, upper ? 1 : 0, n, k, kb, info, static_cast<T*>(work),
                   panel_sync<T>(work) + (size_t)(k / kB) * kSyncInts, panel_error_word<T>(n, work), wait_limit(),
                   batch.stride};
    const int rows = (n - k + kB - 1) / kB;
    if (batch.count <= 0) return;
    panel_kernel<T><<<dim3((unsigned)rows, (unsigned)batch.count), 256, 0, s>>>(a);
    check_launch("panel_factor");
}

template <class T>
void trtri_diag_blocks(hipStream_t s, bool upper, int n, const T* A, long long lda, T* inv, bool unit,
                       const Batch& batch)
{
    if (n <= 0 || batch.count <= 0) return;
    const dim3 grid((unsigned)((n + kIB - 1) / kIB), (unsigned)batch.count);
    trtri_kernel<T><<<grid, 256, 0, s>>>(upper, unit, n, A, lda, inv, batch.stride);
    check_launch("trtri_diag_blocks");
}

namespace {
size_t pad256(size_t b) { return (b + 255) / 256 * 256; }

// [count | D_j of every member (batch x max_blocks x 64 x 64) | log sums
// (batch x 16 doubles) | m | k | kq | ld | ld64 | pm | pn | cm | cn | P | Q | C
// | D] of the blocked path
size_t blocked_bytes(size_t elem, int batch, int max_blocks)
{
    return 256 + pad256((size_t)batch * max_blocks * 64 * 64 * elem) +
           pad256((size_t)batch * batched_detail::kMaxBlocks * sizeof(double)) + 9 * pad256((size_t)batch * sizeof(int)) +
           4 * pad256((size_t)batch * sizeof(void*));
}

size_t fused_bytes(size_t elem, int batch, int grid, int max_blocks)
{
    return pad256((size_t)grid * max_blocks * 64 * 64 * elem) + pad256((size_t)batch * sizeof(int)) + 256;
}
} // namespace

size_t batched_work_bytes(size_t elem, int batch, int grid, int max_blocks, bool blocked)
{
    return fused_bytes(elem, batch, grid, max_blocks) + (blocked ? blocked_bytes(elem, batch, max_blocks) : 0);
}

namespace {
__global__ void fill_iota_kernel(int* p, int count)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) p[i] = i;
}

template <class T>
__global__ __launch_bounds__(256) void batched_copy_kernel(int batch, const int* n, int cols, T* const* src,
                                                           const int* lds, T* const* dst, const int* ldd)
{
    const int i = blockIdx.y;
    if (i >= batch) return;
    const int m = n[i], c = cols < 0 ? m : cols;
    const long long total = (long long)m * c;
    const T* a = src[i];
    T* b = dst[i];
    const long long la = lds[i], lb = ldd[i];
    for (long long t = blockIdx.x * 256LL + threadIdx.x; t < total; t += (long long)gridDim.x * 256) {
        const long long r = t % m, col = t / m;
        b[r + col * lb] = a[r + col * la];
    }
}
} // namespace

template <class T>
void batched_pposv(hipStream_t s, bool upper, int batch, int nrhs, const int* n, T* const* A, const int* lda,
                   T* const* B, const int* ldb, int* info, double* logdet, int grid, int max_blocks, bool ordered,
                   int occupancy, int split_blocks, void* work, bool left_looking, int step_blocks)
{
    using namespace batched_detail;
    if (batch <= 0) return;
    if (max_blocks < 1 || max_blocks > kMaxBlocks) throw std::invalid_argument("batched_pposv: order above 1024");
    if (grid < 1) throw std::invalid_argument("batched_pposv: empty grid");
    char* w = static_cast<char*>(work);
    T* xinv = reinterpret_cast<T*>(w);
    const size_t inv = ((size_t)grid * max_blocks * 64 * 64 * sizeof(T) + 255) / 256 * 256;
    int* order = reinterpret_cast<int*>(w + inv);
    int* queue = reinterpret_cast<int*>(w + inv + ((size_t)batch * sizeof(int) + 255) / 256 * 256);
    if (ordered) {
        order_by_blocks_kernel<<<1, kOrderThreads, 0, s>>>(batch, max_blocks, n, order);
    } else {
        fill_iota_kernel<<<grid_for(batch, 256), 256, 0, s>>>(order, batch);
    }
    check_launch("batched_pposv");
    // blocked path for the members of more than split_blocks blocks (needs the sorted order)
    if (left_looking) {
        solve_detail::zero_ints_kernel<<<1, 64, 0, s>>>(queue, 1);
        check_launch("batched_pposv");
        BatchArgs<T> a{batch, nrhs, upper ? 1 : 0, max_blocks, n, A, lda, B, ldb, info, logdet, order, queue, xinv,
                       nullptr};
        if (occupancy == 2) batched_ll_kernel<T, 2><<<grid, kThreads, 0, s>>>(a);
        else batched_ll_kernel<T, 1><<<grid, kThreads, 0, s>>>(a);
        check_launch("batched_pposv");
        return;
    }
    const int* first = nullptr;
    if (ordered && split_blocks >= 1 && split_blocks < max_blocks) {
        char* bw = w + fused_bytes(sizeof(T), batch, grid, max_blocks);
        int* count = reinterpret_cast<int*>(bw);
        bw += 256;
        T* dinv = reinterpret_cast<T*>(bw);
        bw += pad256((size_t)batch * max_blocks * 64 * 64 * sizeof(T));
        double* logparts = reinterpret_cast<double*>(bw);
        bw += pad256((size_t)batch * kMaxBlocks * sizeof(double));
        StepArrays<T> sa;
        int** ints[] = {&sa.m, &sa.k, &sa.kq, &sa.ld, &sa.ld64, &sa.pm, &sa.pn, &sa.cm, &sa.cn};
        for (int** q : ints) {
            *q = reinterpret_cast<int*>(bw);
            bw += pad256((size_t)batch * sizeof(int));
        }
        T*** ptrs[] = {&sa.P, &sa.Q, &sa.C, &sa.D};
        for (T*** q : ptrs) {
            *q = reinterpret_cast<T**>(bw);
            bw += pad256((size_t)batch * sizeof(void*));
        }
        count_large_kernel<<<1, 256, 0, s>>>(batch, max_blocks, n, split_blocks, count);
        check_launch("batched_pposv");
        BlockedArgs<T> ba{batch, upper ? 1 : 0, max_blocks, nrhs, n, A, lda, ldb, info, order, count, dinv, logparts};
        const int S = step_blocks == 2 ? 2 : 1;
        for (int j = 0; j < max_blocks; ++j) {
            blocked_panel_kernel<T><<<batch, kThreads, 0, s>>>(ba, j);
            check_launch("batched_pposv");
            if (j + 1 < max_blocks) {
                const int J = j - j % S;  // the step's first block column
                blocked_prep_kernel<T><<<(batch + 255) / 256, 256, 0, s>>>(ba, j, J, sa);
                check_launch("batched_pposv");
                const int mx = (max_blocks - 1 - j) * 64;
                // panel: L(r, j) = A(r, j) D_j^T in place (lower), U = D_j U in place (upper)
                if (upper)
                    gemm_vbatched<T>(s, 'N', 'N', sa.pm, sa.pn, sa.k, T(1), sa.D, sa.ld64, sa.P, sa.ld, T(0), sa.P,
                                     sa.ld, batch, 64, mx);
                else
                    gemm_vbatched<T>(s, 'N', 'T', sa.pm, sa.pn, sa.k, T(1), sa.P, sa.ld, sa.D, sa.ld64, T(0), sa.P,
                                     sa.ld, batch, mx, 64);
                if ((j + 1) % S != 0) {
                    // the step's next block column alone, by this block column
                    gemm_vbatched<T>(s, upper ? 'T' : 'N', upper ? 'N' : 'T', sa.cm, sa.cn, sa.k, T(-1), sa.P, sa.ld,
                                     sa.P, sa.ld, T(1), sa.C, sa.ld, batch, upper ? 64 : mx, upper ? mx : 64,
                                     upper ? kTriUpper : kTriLower);
                } else {
                    // trailing update by the whole step's panel (k = 64 (j + 1 - J))
                    gemm_vbatched<T>(s, upper ? 'T' : 'N', upper ? 'N' : 'T', sa.m, sa.m, sa.kq, T(-1), sa.Q, sa.ld,
                                     sa.Q, sa.ld, T(1), sa.C, sa.ld, batch, mx, mx, upper ? kTriUpper : kTriLower);
                }
            }
        }
        blocked_solve_kernel<T><<<batch, kThreads, 0, s>>>(ba, B, logdet);
        check_launch("batched_pposv");
        first = count;
    }
    solve_detail::zero_ints_kernel<<<1, 64, 0, s>>>(queue, 1);
    check_launch("batched_pposv");
    BatchArgs<T> a{batch, nrhs, upper ? 1 : 0, max_blocks, n, A, lda, B, ldb, info, logdet, order, queue, xinv, first};
    if (occupancy == 2) batched_pposv_kernel<T, 2><<<grid, kThreads, 0, s>>>(a);
    else batched_pposv_kernel<T, 1><<<grid, kThreads, 0, s>>>(a);
    check_launch("batched_pposv");
}

template <class T>
void batched_copy(hipStream_t s, int batch, const int* n, int cols, T* const* src, const int* lds, T* const* dst,
                  const int* ldd, int max_n)
{
    if (batch <= 0 || max_n <= 0) return;
    const long long per = (long long)max_n * (cols < 0 ? max_n : cols);
    const unsigned gx = (unsigned)std::max(1LL, std::min(64LL, (per + 255) / 256));
    for (int off = 0; off < batch; off += 65535) {
        const int nb = std::min(65535, batch - off);
        batched_copy_kernel<T><<<dim3(gx, nb), 256, 0, s>>>(nb, n + off, cols, src + off, lds + off, dst + off,
                                                            ldd + off);
        check_launch("batched_copy");
    }
}

size_t solve_work_bytes(int nrhs)
{
    const size_t slabs = (size_t)(nrhs + solve_detail::kSB - 1) / solve_detail::kSB;
    return (2 * slabs * sizeof(int) + 255) / 256 * 256;
}

template <class T>
void potrs_sweeps(hipStream_t s, bool upper, int n, int nrhs, const T* L, long long ldl, const T* dinv, T* B,
                  long long ldb, void* work, int* error)
{
    if (n <= 0 || nrhs <= 0) return;
    if (trsm_splits(n, nrhs)) {
        trsm<T>(s, upper, false, n, nrhs, L, ldl, dinv, B, ldb, work, error);
        trsm<T>(s, upper, true, n, nrhs, L, ldl, dinv, B, ldb, work, error);
        return;
    }
    using namespace solve_detail;
    const int nblk = (n + kSB - 1) / kSB, slabs = (nrhs + kSB - 1) / kSB;
    if (slabs > 65535) throw std::invalid_argument("potrs: too many right-hand sides for one launch");
    SolveArgs<T> a{n, nrhs, nblk, L, ldl, upper ? 1 : 0, dinv, B, ldb, static_cast<int*>(work), error, wait_limit()};
    const dim3 grid((unsigned)nblk, (unsigned)slabs), block(kThreads);
    zero_ints_kernel<<<1, 256, 0, s>>>(a.sync, 2 * slabs);
    check_launch("potrs");
    solve_sweep_kernel<T, false><<<grid, block, 0, s>>>(a);
    check_launch("potrs");
    zero_ints_kernel<<<1, 256, 0, s>>>(a.sync, 2 * slabs);
    check_launch("potrs");
    solve_sweep_kernel<T, true><<<grid, block, 0, s>>>(a);
    check_launch("potrs");
}

bool stage_solve(int n, int nrhs) { return n >= kStageMinOrder && nrhs > 0 && !trsm_splits(n, nrhs); }

namespace {
// Layout of the staged solve's work: [E | M forward | M backward | accumulators | counters].
struct StageLayout {
    size_t e, mf, mb, acc, sync, total;
};
StageLayout stage_layout(int n, int nrhs, size_t elem)
{
    using namespace stage_detail;
    const size_t nst = (size_t)(n + kS - 1) / kS, slabs = (size_t)(nrhs + kSB - 1) / kSB;
    const size_t nblk = (size_t)(n + kSB - 1) / kSB;
    const size_t block = nst * kS * kS * elem;
    StageLayout l{};
    size_t off = 0;
    auto put = [&](size_t bytes) {
        const size_t at = off;
        off += (bytes + 255) / 256 * 256;
        return at;
    };
    l.e = put(block);
    l.mf = put(block);
    l.mb = put(block);
    l.acc = put(slabs * (size_t)n * kSB * elem);
    l.sync = put(slabs * (size_t)sync_ints((int)nst, (int)nblk) * sizeof(int));
    l.total = off;
    return l;
}

template <class T>
stage_detail::StageArgs<T> stage_args(bool upper, int n, int nrhs, const T* L, long long ldl, const T* dinv, T* B,
                                      long long ldb, void* work, int* error)
{
    using namespace stage_detail;
    const StageLayout l = stage_layout(n, nrhs, sizeof(T));
    char* w = static_cast<char*>(work);
    StageArgs<T> a{};
    a.n = n;
    a.nrhs = nrhs;
    a.nst = (n + kS - 1) / kS;
    a.nblk = (n + kSB - 1) / kSB;
    a.L = L;
    a.ldl = ldl;
    a.upper = upper ? 1 : 0;
    a.dinv = dinv;
    a.E = reinterpret_cast<T*>(w + l.e);
    a.M = reinterpret_cast<T*>(w + l.mf);
    a.B = B;
    a.ldb = ldb;
    a.acc = reinterpret_cast<T*>(w + l.acc);
    a.sync = reinterpret_cast<int*>(w + l.sync);
    a.error = error;
    a.spin_limit = wait_limit();
    return a;
}
} // namespace

size_t stage_work_bytes(int n, int nrhs, size_t elem) { return stage_solve(n, nrhs) ? stage_layout(n, nrhs, elem).total : 0; }

template <class T>
void stage_prepare(hipStream_t s, bool upper, int n, const T* L, long long ldl, const T* dinv, void* work)
{
    using namespace stage_detail;
    StageArgs<T> a = stage_args<T>(upper, n, 1, L, ldl, dinv, nullptr, 0, work, nullptr);
    const StageLayout l = stage_layout(n, 1, sizeof(T));
    stage_inverse_kernel<T><<<dim3((unsigned)a.nst, kBlocks), kThreads, 0, s>>>(a);
    check_launch("stage_prepare");
    if (a.nst < 2) return;
    const dim3 grid((unsigned)(a.nst - 1), kBlocks, kBlocks);
    stage_product_kernel<T, false><<<grid, kThreads, 0, s>>>(a);
    check_launch("stage_prepare");
    a.M = reinterpret_cast<T*>(static_cast<char*>(work) + l.mb);
    stage_product_kernel<T, true><<<grid, kThreads, 0, s>>>(a);
    check_launch("stage_prepare");
}

template <class T>
void stage_sweeps(hipStream_t s, bool upper, int n, int nrhs, const T* L, long long ldl, T* B, long long ldb,
                  void* work, int* error)
{
    using namespace stage_detail;
    if (n <= 0 || nrhs <= 0) return;
    if (!stage_solve(n, nrhs)) throw std::logic_error("stage_sweeps: order or right-hand sides out of range");
    StageArgs<T> a = stage_args<T>(upper, n, nrhs, L, ldl, nullptr, B, ldb, work, error);
    const StageLayout l = stage_layout(n, nrhs, sizeof(T));
    const int slabs = (nrhs + kSB - 1) / kSB;
    const int counters = slabs * sync_ints(a.nst, a.nblk);
    const dim3 grid((unsigned)a.nblk, (unsigned)slabs);
    solve_detail::zero_ints_kernel<<<(counters + 255) / 256, 256, 0, s>>>(a.sync, counters);
    check_launch("stage_sweeps");
    stage_sweep_kernel<T, false><<<grid, kThreads, 0, s>>>(a);
    check_launch("stage_sweeps");
    a.M = reinterpret_cast<T*>(static_cast<char*>(work) + l.mb);
    solve_detail::zero_ints_kernel<<<(counters + 255) / 256, 256, 0, s>>>(a.sync, counters);
    check_launch("stage_sweeps");
    stage_sweep_kernel<T, true><<<grid, kThreads, 0, s>>>(a);
    check_launch("stage_sweeps");
}

template <class T>
void trsm_sweep(hipStream_t s, bool upper, bool backward, int n, int nrhs, const T* L, long long ldl, const T* dinv,
                T* B, long long ldb, void* work, int* error)
{
    if (n <= 0 || nrhs <= 0) return;
    using namespace solve_detail;
    const int nblk = (n + kSB - 1) / kSB, slabs = (nrhs + kSB - 1) / kSB;
    if (slabs > 65535) throw std::invalid_argument("trsm_sweep: too many right-hand sides for one launch");
    SolveArgs<T> a{n, nrhs, nblk, L, ldl, upper ? 1 : 0, dinv, B, ldb, static_cast<int*>(work), error, wait_limit()};
    const dim3 grid((unsigned)nblk, (unsigned)slabs), block(kThreads);
    zero_ints_kernel<<<1, 256, 0, s>>>(a.sync, 2 * slabs);
    check_launch("trsm_sweep");
    if (backward) solve_sweep_kernel<T, true><<<grid, block, 0, s>>>(a);
    else solve_sweep_kernel<T, false><<<grid, block, 0, s>>>(a);
    check_launch("trsm_sweep");
}

bool trsm_splits(int n, int nrhs) { return nrhs >= kTrsmMinColumns && n > kTrsmLeaf; }

template <class T>
void trsm(hipStream_t s, bool upper, bool backward, int n, int nrhs, const T* L, long long ldl, const T* dinv, T* B,
          long long ldb, void* work, int* error)
{
    if (n <= 0 || nrhs <= 0) return;
    if (!trsm_splits(n, nrhs)) {
        trsm_sweep<T>(s, upper, backward, n, nrhs, L, ldl, dinv, B, ldb, work, error);
        return;
    }
    // L = [L11 0; L21 L22] with n1 a multiple of the 64-row diagonal blocks, so
    // that dinv of L22 starts at block n1 / 64.  L21 (L view, n2 x n1) is
    // stored at L + n1 as an n2 x n1 matrix (lower) or at L + n1 * ldl as its
    // n1 x n2 transpose (upper).
    const int n1 = (n / 2 + 63) / 64 * 64, n2 = n - n1;
    const T* L21 = upper ? L + (long long)n1 * ldl : L + n1;
    const T* L22 = L + n1 + (long long)n1 * ldl;
    const T* dinv22 = dinv + (long long)(n1 / 64) * 4096;
    T* B2 = B + n1;
    if (!backward) {
        // L11 Y1 = B1;  B2 -= L21 Y1;  L22 Y2 = B2
        trsm<T>(s, upper, false, n1, nrhs, L, ldl, dinv, B, ldb, work, error);
        gemm<T>(s, upper ? 'T' : 'N', 'N', n2, nrhs, n1, T(-1), L21, ldl, B, ldb, T(1), B2, ldb);
        trsm<T>(s, upper, false, n2, nrhs, L22, ldl, dinv22, B2, ldb, work, error);
    } else {
        // L22^T X2 = B2;  B1 -= L21^T X2;  L11^T X1 = B1
        trsm<T>(s, upper, true, n2, nrhs, L22, ldl, dinv22, B2, ldb, work, error);
        gemm<T>(s, upper ? 'N' : 'T', 'N', n1, nrhs, n2, T(-1), L21, ldl, B2, ldb, T(1), B, ldb);
        trsm<T>(s, upper, true, n1, nrhs, L, ldl, dinv, B, ldb, work, error);
    }
}

namespace {
// M := alpha M (m x n); alpha == 0 writes zeros without reading M.
template <class T>
__global__ void scale_kernel(int m, int n, T alpha, T* M, long long ld)
{
    const long long total = (long long)m * n;
    for (long long t = blockIdx.x * 256LL + threadIdx.x; t < total; t += (long long)gridDim.x * 256) {
        const long long i = t % m, j = t / m;
        M[i + j * ld] = alpha == T(0) ? T(0) : alpha * M[i + j * ld];
    }
}

// out = the order-n identity.  One thread per element of out.
template <class T>
__global__ void identity_kernel(int n, T* out, long long ldo)
{
    const long long total = (long long)n * n;
    for (long long t = blockIdx.x * 256LL + threadIdx.x; t < total; t += (long long)gridDim.x * 256) {
        const int i = (int)(t % n), j = (int)(t / n);
        out[i + j * ldo] = i == j ? T(1) : T(0);
    }
}

// out = tril(in), strict upper triangle zero.  One thread per element of out.
template <class T>
__global__ void tril_copy_kernel(int n, const T* in, long long ldi, T* out, long long ldo)
{
    const long long total = (long long)n * n;
    for (long long t = blockIdx.x * 256LL + threadIdx.x; t < total; t += (long long)gridDim.x * 256) {
        const int i = (int)(t % n), j = (int)(t / n);
        out[i + j * ldo] = i >= j ? in[i + j * ldi] : T(0);
    }
}

// out(i, j) = scale * in(max(i, j), min(i, j)).  Workgroup (bi, bj) writes
// the 32 x 32 tile (bi, bj) of out from the lower-triangle tile (max, min) of
// in, staged through LDS so that the mirrored read and the write are both
// coalesced.
template <class T>
__global__ __launch_bounds__(256) void mirror_lower_kernel(int n, const T* in, long long ldi, T scale, T* out,
                                                           long long ldo)
{
    __shared__ T tile[32][33];
    const int bi = blockIdx.x, bj = blockIdx.y;
    const long long si = (long long)max(bi, bj) * 32, sj = (long long)min(bi, bj) * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int r = ty; r < 32; r += 8)
        if (si + tx < n && sj + r < n) tile[r][tx] = in[(si + tx) + (sj + r) * ldi];  // tile[col][row]
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const long long i = (long long)bi * 32 + tx, j = (long long)bj * 32 + r;
        if (i < n && j < n) {
            const bool mirrored = bi < bj || (bi == bj && tx < r);
            out[i + j * ldo] = scale * (mirrored ? tile[tx][r] : tile[r][tx]);
        }
    }
}

// Y(i, j) = sum_{k >= i} L(k, i) W(k, j), i < n, j < c <= kTmvColumns: the
// transposed product with the lower triangle of L (lower storage; the strict
// upper triangle is never read).  One wavefront per row i reads column i of L
// from the diagonal down, coalesced; fixed lane partition and butterfly, so
// the summation order is deterministic.
template <class T>
__global__ __launch_bounds__(256) void lower_tmv_kernel(int n, int c, const T* L, long long ldl, const T* W,
                                                        long long ldw, T* Y, long long ldy)
{
    const int lane = threadIdx.x & 63;
    const long long i = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    T acc[kTmvColumns];
#pragma unroll
    for (int j = 0; j < kTmvColumns; ++j) acc[j] = T(0);
    const T* col = L + i * ldl;
    for (long long k = i + lane; k < n; k += 64) {
        const T l = col[k];
#pragma unroll
        for (int j = 0; j < kTmvColumns; ++j)
            if (j < c) acc[j] = fma(l, W[k + j * ldw], acc[j]);
    }
#pragma unroll
    for (int j = 0; j < kTmvColumns; ++j) {
        if (j >= c) break;
        T v = acc[j];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) Y[i + j * ldy] = v;
    }
}

template <class T>
__global__ void diag_adjoint_kernel(int n, const T* L, long long ldl, double g, T* Lbar, long long ldlb)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        Lbar[i + i * ldlb] += (T)(2.0 * g / (double)L[i + i * ldl]);
}
} // namespace

template <class T>
void tril_copy(hipStream_t s, int n, const T* in, long long ldi, T* out, long long ldo)
{
    if (n <= 0) return;
    tril_copy_kernel<T><<<grid_for((long long)n * n, 256, 1 << 16), 256, 0, s>>>(n, in, ldi, out, ldo);
    check_launch("tril_copy");
}

template <class T>
void scale_matrix(hipStream_t s, int m, int n, T alpha, T* M, long long ld)
{
    if (m <= 0 || n <= 0) return;
    scale_kernel<T><<<grid_for((long long)m * n, 256, 1 << 16), 256, 0, s>>>(m, n, alpha, M, ld);
    check_launch("scale_matrix");
}

template <class T>
void set_identity(hipStream_t s, int n, T* out, long long ldo)
{
    if (n <= 0) return;
    identity_kernel<T><<<grid_for((long long)n * n, 256, 1 << 16), 256, 0, s>>>(n, out, ldo);
    check_launch("set_identity");
}

template <class T>
void mirror_lower(hipStream_t s, int n, const T* in, long long ldi, T scale, T* out, long long ldo)
{
    if (n <= 0) return;
    if (in == out) throw std::invalid_argument("mirror_lower: output must not alias the input");
    const unsigned tiles = (unsigned)((n + 31) / 32);
    if (tiles > 65535) throw std::invalid_argument("mirror_lower: order too large");
    mirror_lower_kernel<T><<<dim3(tiles, tiles), 256, 0, s>>>(n, in, ldi, scale, out, ldo);
    check_launch("mirror_lower");
}

template <class T>
void lower_tmv(hipStream_t s, int n, int c, const T* L, long long ldl, const T* W, long long ldw, T* Y,
               long long ldy)
{
    if (n <= 0 || c <= 0) return;
    if (c > kTmvColumns) throw std::invalid_argument("lower_tmv: too many columns");
    lower_tmv_kernel<T><<<(unsigned)((n + 3) / 4), 256, 0, s>>>(n, c, L, ldl, W, ldw, Y, ldy);
    check_launch("lower_tmv");
}

template <class T>
void diag_adjoint(hipStream_t s, int n, const T* L, long long ldl, double g, T* Lbar, long long ldlb)
{
    if (n <= 0) return;
    diag_adjoint_kernel<T><<<grid_for(n, 256), 256, 0, s>>>(n, L, ldl, g, Lbar, ldlb);
    check_launch("diag_adjoint");
}

size_t sym_norm_work_bytes(int n)
{
    const size_t nblk = (size_t)(n + 63) / 64;
    return ((2 * nblk + 1) * (size_t)n * sizeof(double) + 255) / 256 * 256;
}

template <class D, class S>
void convert_sym_with_norm(hipStream_t s, int n, bool lower, const S* in, long long ldi, D* out, long long ldo,
                           void* work, double* rowmax)
{
    if (n <= 0) return;
    const int nblk = (n + 63) / 64;
    if (nblk > 65535) throw std::invalid_argument("convert_sym_with_norm: order too large");
    double* prow = static_cast<double*>(work);
    double* pcol = prow + (size_t)nblk * n;
    double* rowsum = pcol + (size_t)nblk * n;
    convert_sym_norm_kernel<D, S><<<dim3(nblk, nblk), 256, 0, s>>>(n, lower ? 1 : 0, in, ldi, out, ldo, prow, pcol);
    check_launch("convert_sym_with_norm");
    sym_rowsum_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, nblk, lower ? 1 : 0, prow, pcol, rowsum);
    check_launch("convert_sym_with_norm");
    colmax_kernel<double><<<1, 256, 0, s>>>(n, rowsum, n, rowmax);
    check_launch("convert_sym_with_norm");
}

template <class T>
void column_max(hipStream_t s, int m, int n, const T* M, long long ld, double* out)
{
    if (m <= 0 || n <= 0) return;
    colmax_kernel<T><<<n, 256, 0, s>>>(m, M, ld, out);
    check_launch("column_max");
}

template <class S>
void add_converted(hipStream_t s, int m, int n, const S* W, long long ldw, double* X, long long ldx)
{
    if (m <= 0 || n <= 0) return;
    add_converted_kernel<S><<<grid_for((long long)m * n, 256, 1 << 20), 256, 0, s>>>(m, n, W, ldw, X, ldx);
    check_launch("add_converted");
}

template <class T>
void symm_residual(hipStream_t s, bool lower, int n, int nrhs, const T* S, long long lds, const T* X, long long ldx,
                   const T* B, long long ldb, T* R, long long ldr)
{
    if (n <= 0 || nrhs <= 0) return;
    using namespace solve_detail;
    const int nblk = (n + kSB - 1) / kSB, slabs = (nrhs + kSB - 1) / kSB;
    if (slabs > 65535) throw std::invalid_argument("symm_residual: too many right-hand sides for one launch");
    SymmArgs<T> a{n, nrhs, nblk, S, lds, lower ? 1 : 0, X, ldx, B, ldb, R, ldr};
    symm_residual_kernel<T><<<dim3((unsigned)nblk, (unsigned)slabs), kThreads, 0, s>>>(a);
    check_launch("symm_residual");
}

size_t logdet_work_bytes() { return kLogdetParts * (sizeof(double) + sizeof(int)); }

template <class T>
void diag_logdet(hipStream_t s, int n, const T* A, long long lda, void* work, LogdetResult* out)
{
    double* pmag = static_cast<double*>(work);
    int* pflags = reinterpret_cast<int*>(pmag + kLogdetParts);
    logdet_partial_kernel<T><<<kLogdetParts, 256, 0, s>>>(n, A, lda, pmag, pflags);
    check_launch("diag_logdet");
    logdet_final_kernel<<<1, 64, 0, s>>>(pmag, pflags, out);
    check_launch("diag_logdet");
}

void fill_normal(hipStream_t s, uint64_t seed, uint64_t stream, uint64_t first, long long count, double* out)
{
    if (count <= 0) return;
    fill_normal_kernel<<<grid_for(count, 256, 1 << 20), 256, 0, s>>>(seed, stream, first, count, out);
    check_launch("fill_normal");
}

template <class T>
void add_diagonal(hipStream_t s, int n, T* A, long long lda, T delta)
{
    if (n <= 0) return;
    add_diagonal_kernel<T><<<grid_for(n, 256), 256, 0, s>>>(n, A, lda, delta);
    check_launch("add_diagonal");
}

template <class T>
void reverse_triangle(hipStream_t s, bool out_upper, bool transpose, int n, const T* in, long long ldi, T* out,
                      long long ldo, bool zero_other)
{
    if (n <= 0) return;
    if (ldi < n || ldo < n) throw std::invalid_argument("reverse_triangle: leading dimension below the order");
    const unsigned tiles = (unsigned)((n + 63) / 64);
    if (tiles > 65535) throw std::invalid_argument("reverse_triangle: order too large");
    reverse_triangle_kernel<T><<<dim3(tiles, tiles), 256, 0, s>>>(out_upper, transpose, n, in, ldi, out, ldo,
                                                                  zero_other);
    check_launch("reverse_triangle");
}

void memset_tri_bytes(hipStream_t s, bool upper, bool diag, int m, int n, void* p, long long ld, std::size_t elem,
                      unsigned char value)
{
    if (m <= 0 || n <= 0) return;
    const dim3 grid((unsigned)((m + 255) / 256), (unsigned)std::min(n, 65535));
    memset_tri_kernel<<<grid, 256, 0, s>>>(upper, diag, m, n, static_cast<unsigned char*>(p),
                                          ld * (long long)elem, (int)elem, value);
    check_launch("memset_tri");
}

void memcpy_tri_bytes(hipStream_t s, bool upper, bool diag, int m, int n, void* dst, long long ldd, const void* src,
                      long long lds, std::size_t elem)
{
    if (m <= 0 || n <= 0) return;
    const dim3 grid((unsigned)((m + 255) / 256), (unsigned)std::min(n, 65535));
    memcpy_tri_kernel<<<grid, 256, 0, s>>>(upper, diag, m, n, static_cast<unsigned char*>(dst),
                                          ldd * (long long)elem, static_cast<const unsigned char*>(src),
                                          lds * (long long)elem, (int)elem);
    check_launch("memcpy_tri");
}

template <class T>
void swap_triangles(hipStream_t s, int n, T* A, long long lda)
{
    if (n <= 1) return;
    if (lda < n) throw std::invalid_argument("swap_triangles: leading dimension below the order");
    const unsigned tiles = (unsigned)((n + 31) / 32);
    if (tiles > 65535) throw std::invalid_argument("swap_triangles: order too large");
    swap_triangles_kernel<T><<<dim3(tiles, tiles), 256, 0, s>>>(n, A, lda);
    check_launch("swap_triangles");
}

template <class T>
void transpose(hipStream_t s, int m, int n, const T* in, long long ldi, T* out, long long ldo)
{
    if (m <= 0 || n <= 0) return;
    const dim3 grid((unsigned)((m + 31) / 32), (unsigned)((n + 31) / 32));
    if (grid.y > 65535) throw std::invalid_argument("transpose: too many columns");
    transpose_kernel<T><<<grid, 256, 0, s>>>(m, n, in, ldi, out, ldo);
    check_launch("transpose");
}

template <class D, class S>
void convert(hipStream_t s, int m, int n, const S* in, long long ldi, D* out, long long ldo)
{
    if (m <= 0 || n <= 0) return;
    convert_kernel<D, S><<<grid_for((long long)m * n, 256, 1 << 20), 256, 0, s>>>(m, n, in, ldi, out, ldo);
    check_launch("convert");
}

#define DDLA_KERN_INSTANTIATE(T)                                                                              \
    template void gemm<T>(hipStream_t, char, char, int, int, int, T, const T*, long long, const T*, long long, \
                          T, T*, long long, Tri, long long, const int*, bool, const Batch&);                  \
    template void gemm_vbatched<T>(hipStream_t, char, char, const int*, const int*, const int*, T,            \
                                   const T* const*, const int*, const T* const*, const int*, T, T* const*,    \
                                   const int*, int, int, int, Tri);                                           \
    template size_t panel_work_bytes<T>(int);                                                                 \
    template void panel_reset<T>(hipStream_t, int, void*, const Batch&);                                      \
    template int* panel_error_word<T>(int, void*);                                                            \
    template void panel_factor<T>(hipStream_t, bool, int, T*, long long, int, int, int*, void*, const Batch&); \
    template void batched_pposv<T>(hipStream_t, bool, int, int, const int*, T* const*, const int*, T* const*,  \
                                   const int*, int*, double*, int, int, bool, int, int, void*, bool, int);     \
    template void batched_copy<T>(hipStream_t, int, const int*, int, T* const*, const int*, T* const*,         \
                                  const int*, int);                                                            \
    template void trtri_diag_blocks<T>(hipStream_t, bool, int, const T*, long long, T*, bool, const Batch&);  \
    template void scale_matrix<T>(hipStream_t, int, int, T, T*, long long);                                   \
    template void diag_logdet<T>(hipStream_t, int, const T*, long long, void*, LogdetResult*);                \
    template void potrs_sweeps<T>(hipStream_t, bool, int, int, const T*, long long, const T*, T*, long long,   \
                                  void*, int*);                                                               \
    template void add_diagonal<T>(hipStream_t, int, T*, long long, T);                                        \
    template void stage_prepare<T>(hipStream_t, bool, int, const T*, long long, const T*, void*);                \
    template void stage_sweeps<T>(hipStream_t, bool, int, int, const T*, long long, T*, long long, void*, int*);  \
    template void trsm_sweep<T>(hipStream_t, bool, bool, int, int, const T*, long long, const T*, T*, long long, \
                                void*, int*);                                                                 \
    template void trsm<T>(hipStream_t, bool, bool, int, int, const T*, long long, const T*, T*, long long,       \
                          void*, int*);                                                                       \
    template void tril_copy<T>(hipStream_t, int, const T*, long long, T*, long long);                         \
    template void mirror_lower<T>(hipStream_t, int, const T*, long long, T, T*, long long);                   \
    template void set_identity<T>(hipStream_t, int, T*, long long);                                           \
    template void lower_tmv<T>(hipStream_t, int, int, const T*, long long, const T*, long long, T*, long long); \
    template void diag_adjoint<T>(hipStream_t, int, const T*, long long, double, T*, long long);              \
    template void transpose<T>(hipStream_t, int, int, const T*, long long, T*, long long);                     \
    template void reverse_triangle<T>(hipStream_t, bool, bool, int, const T*, long long, T*, long long, bool); \
    template void swap_triangles<T>(hipStream_t, int, T*, long long);
DDLA_KERN_INSTANTIATE(double)
DDLA_KERN_INSTANTIATE(float)
#undef DDLA_KERN_INSTANTIATE
template void convert<float, double>(hipStream_t, int, int, const double*, long long, float*, long long);
template void convert_sym_with_norm<float, double>(hipStream_t, int, bool, const double*, long long, float*, long long,
                                                   void*, double*);
template void column_max<double>(hipStream_t, int, int, const double*, long long, double*);
template void column_max<float>(hipStream_t, int, int, const float*, long long, double*);
template void add_converted<float>(hipStream_t, int, int, const float*, long long, double*, long long);
template void symm_residual<double>(hipStream_t, bool, int, int, const double*, long long, const double*, long long,
                                    const double*, long long, double*, long long);
template void convert<double, float>(hipStream_t, int, int, const float*, long long, double*, long long);
template void convert<double, double>(hipStream_t, int, int, const double*, long long, double*, long long);
template void convert<float, float>(hipStream_t, int, int, const float*, long long, float*, long long);

} // namespace kern
} // namespace ddla
