// Hand-written CDNA4 (gfx950) kernels of the Cholesky path.
//
// Matrix products run on the matrix cores: v_mfma_f64_16x16x4_f64 for double
// and v_mfma_f32_16x16x4_f32 for float, one 64-wide wavefront per 16x16 MFMA
// tile, four wavefronts per workgroup, operands staged through double-buffered
// LDS tiles.  The diagonal-block factorization and the triangular inverses run
// in one workgroup on LDS-resident 64 x 64 blocks.
#include "kernels.h"

#include <cfloat>
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

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

// Operand layout of both 16x16x4 instructions: lane l supplies A[l%16][l/16]
// and B[l/16][l%16].  The accumulator layout differs between the two (checked
// on gfx950): see row().
template <class T> struct Mfma;
template <> struct Mfma<double> {
    using acc_t = d4_t;
    __device__ static acc_t mma(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // lane l, accumulator r holds D[l/16 + 4r][l%16]
    __device__ static int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct Mfma<float> {
    using acc_t = f4_t;
    __device__ static acc_t mma(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // lane l, accumulator r holds D[4(l/16) + r][l%16]
    __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

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
};

constexpr int kPad = 16;  // LDS row padding (elements): keeps the 4 k-rows of a fragment load on distinct banks

template <int BM, int BN, int BK>
constexpr int gemm_lds_elems()
{
    return 2 * BK * (BM + kPad) + 2 * BK * (BN + kPad);
}

// One BM x BN tile of C at (m0, n0).  256 threads = 4 wavefronts in a 2 x 2
// arrangement; each wavefront owns a (BM/2) x (BN/2) sub-tile made of
// (BM/32) x (BN/32) MFMA tiles.  The k loop streams BK-deep slabs of op(A) and
// op(B) through registers into two alternating LDS buffers (one barrier per
// slab): the global loads of slab t+1 are in flight while slab t is multiplied.
template <class T, int BM, int BN, int BK, bool TA, bool TB>
__device__ __forceinline__ void gemm_tile(const GemmProblem<T>& p, int m0, int n0, T* smem)
{
    constexpr int LA = BM + kPad, LB = BN + kPad;
    constexpr int TM = BM / 32, TN = BN / 32;
    constexpr int APT = BM * BK / 256, BPT = BN * BK / 256;
    static_assert(APT * 256 == BM * BK && BPT * 256 == BN * BK, "tile must split over 256 threads");
    using Op = Mfma<T>;
    using acc_t = typename Op::acc_t;

    if (p.tri == kTriLower && (long long)(m0 + BM - 1) + p.diag < n0) return;
    if (p.tri == kTriUpper && (long long)m0 + p.diag > n0 + BN - 1) return;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    T* As = smem;
    T* Bs = smem + 2 * BK * LA;

    acc_t acc[TM][TN];
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) acc[a][b] = acc_t{0, 0, 0, 0};

    const int nkt = (p.alpha == T(0)) ? 0 : (p.k + BK - 1) / BK;
    T ra[APT], rb[BPT];

    auto a_index = [&](int e, int& i, int& kk) {
        const int lin = tid + 256 * e;
        if (!TA) { i = lin % BM; kk = lin / BM; } else { kk = lin % BK; i = lin / BK; }
    };
    auto b_index = [&](int e, int& j, int& kk) {
        const int lin = tid + 256 * e;
        if (!TB) { kk = lin % BK; j = lin / BK; } else { j = lin % BN; kk = lin / BN; }
    };
    auto load = [&](int k0) {
#pragma unroll
        for (int e = 0; e < APT; ++e) {
            int i, kk;
            a_index(e, i, kk);
            const int gi = m0 + i, gk = k0 + kk;
            ra[e] = (gi < p.m && gk < p.k)
                        ? (TA ? p.A[gk + (long long)gi * p.lda] : p.A[gi + (long long)gk * p.lda])
                        : T(0);
        }
#pragma unroll
        for (int e = 0; e < BPT; ++e) {
            int j, kk;
            b_index(e, j, kk);
            const int gj = n0 + j, gk = k0 + kk;
            rb[e] = (gj < p.n && gk < p.k)
                        ? (TB ? p.B[gj + (long long)gk * p.ldb] : p.B[gk + (long long)gj * p.ldb])
                        : T(0);
        }
    };
    auto store = [&](int buf) {
        T* as = As + buf * BK * LA;
        T* bs = Bs + buf * BK * LB;
#pragma unroll
        for (int e = 0; e < APT; ++e) {
            int i, kk;
            a_index(e, i, kk);
            as[kk * LA + i] = ra[e];
        }
#pragma unroll
        for (int e = 0; e < BPT; ++e) {
            int j, kk;
            b_index(e, j, kk);
            bs[kk * LB + j] = rb[e];
        }
    };

    if (nkt > 0) {
        load(0);
        store(0);
    }
    __syncthreads();
    for (int kt = 0; kt < nkt; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nkt) load((kt + 1) * BK);
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
#pragma unroll
            for (int x = 0; x < TM; ++x)
#pragma unroll
                for (int y = 0; y < TN; ++y) acc[x][y] = Op::mma(a[x], b[y], acc[x][y]);
        }
        if (kt + 1 < nkt) store(buf ^ 1);
        __syncthreads();
    }

#pragma unroll
    for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TN; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = m0 + wm * (BM / 2) + x * 16 + Op::row(lane, r);
                const int j = n0 + wn * (BN / 2) + y * 16 + (lane & 15);
                if (i >= p.m || j >= p.n) continue;
                if (p.tri == kTriLower && (long long)i + p.diag < j) continue;
                if (p.tri == kTriUpper && (long long)i + p.diag > j) continue;
                T* c = p.C + i + (long long)j * p.ldc;
                T v = p.alpha * acc[x][y][r];
                if (p.beta != T(0)) v += p.beta * *c;
                *c = v;
            }
}

template <class T, int BM, int BN, int BK, bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_kernel(GemmProblem<T> p, const int* abort)
{
    if (abort != nullptr && *abort != 0) return;
    __shared__ T smem[gemm_lds_elems<BM, BN, BK>()];
    gemm_tile<T, BM, BN, BK, TA, TB>(p, blockIdx.x * BM, blockIdx.y * BN, smem);
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
    p.tri = kTriNone;
    p.diag = 0;
    __shared__ T smem[gemm_lds_elems<BM, BN, BK>()];
    gemm_tile<T, BM, BN, BK, TA, TB>(p, blockIdx.x * BM, blockIdx.y * BN, smem);
}

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
constexpr int kLd = kIB + 1;

// Column c of X = L^-1 by forward substitution, one thread per column.
template <class T>
__device__ void lower_inverse_lds(const T* Ls, T* Xs)
{
    const int c = threadIdx.x;
    if (c < kIB) {
        for (int i = 0; i < c; ++i) Xs[c * kLd + i] = T(0);
        Xs[c * kLd + c] = T(1) / Ls[c + c * kLd];
        for (int i = c + 1; i < kIB; ++i) {
            T s = T(0);
            for (int q = c; q < i; ++q) s += Ls[i + q * kLd] * Xs[c * kLd + q];
            Xs[c * kLd + i] = -s / Ls[i + i * kLd];
        }
    }
}

template <class T>
__device__ void load_lower_block(bool upper, int w, const T* A, long long lda, T* Ls)
{
    for (int idx = threadIdx.x; idx < kIB * kIB; idx += blockDim.x) {
        const int i = idx & (kIB - 1), j = idx >> 6;
        T v = T(0);
        if (i < w && j < w) {
            if (i >= j) v = upper ? A[j + (long long)i * lda] : A[i + (long long)j * lda];
        } else if (i == j) {
            v = T(1);
        }
        Ls[i + j * kLd] = v;
    }
}

template <class T>
__device__ void store_inverse(const T* Xs, T* inv)
{
    for (int idx = threadIdx.x; idx < kIB * kIB; idx += blockDim.x) {
        const int i = idx & (kIB - 1), j = idx >> 6;
        inv[idx] = Xs[j * kLd + i];  // inv[i + 64 j] = X(i, j)
    }
}

template <class T>
__global__ __launch_bounds__(256) void potf2_inv_kernel(bool upper, int w, T* A, long long lda, T* inv, int* info,
                                                        int info_base)
{
    if (*info != 0) return;
    __shared__ T Ls[kIB * kLd];
    __shared__ T Xs[kIB * kLd];
    __shared__ int fail;
    const int tid = threadIdx.x;
    load_lower_block(upper, w, A, lda, Ls);
    if (tid == 0) fail = 0;
    __syncthreads();
    for (int j = 0; j < w; ++j) {
        if (tid == 0) {
            const T d = Ls[j + j * kLd];
            if (!(d > T(0))) fail = j + 1;
            else Ls[j + j * kLd] = sqrt(d);
        }
        __syncthreads();
        if (fail != 0) break;
        const T rinv = T(1) / Ls[j + j * kLd];
        for (int i = j + 1 + tid; i < w; i += 256) Ls[i + j * kLd] *= rinv;
        __syncthreads();
        const int r = w - j - 1;
        for (int idx = tid; idx < r * r; idx += 256) {
            const int c = j + 1 + idx / r, i = j + 1 + idx % r;
            if (i >= c) Ls[i + c * kLd] -= Ls[i + j * kLd] * Ls[c + j * kLd];
        }
        __syncthreads();
    }
    for (int idx = tid; idx < kIB * kIB; idx += 256) {
        const int i = idx & (kIB - 1), j = idx >> 6;
        if (i < w && j < w && i >= j) {
            if (upper) A[j + (long long)i * lda] = Ls[i + j * kLd];
            else A[i + (long long)j * lda] = Ls[i + j * kLd];
        }
    }
    if (fail != 0) {
        if (tid == 0) *info = info_base + fail;
        return;
    }
    lower_inverse_lds(Ls, Xs);
    __syncthreads();
    store_inverse(Xs, inv);
}

template <class T>
__global__ __launch_bounds__(256) void trtri_kernel(bool upper, int n, const T* A, long long lda, T* inv)
{
    __shared__ T Ls[kIB * kLd];
    __shared__ T Xs[kIB * kLd];
    const long long j0 = (long long)blockIdx.x * kIB;
    const int w = (int)min((long long)kIB, n - j0);
    load_lower_block(upper, w, A + j0 + j0 * lda, lda, Ls);
    __syncthreads();
    lower_inverse_lds(Ls, Xs);
    __syncthreads();
    store_inverse(Xs, inv + (long long)blockIdx.x * kIB * kIB);
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

template <class T, int BM, int BN, int BK>
void launch_gemm_cfg(hipStream_t s, bool ta, bool tb, const GemmProblem<T>& p, const int* abort)
{
    const long long gx = (p.m + BM - 1) / BM, gy = (p.n + BN - 1) / BN;
    if (gy > 65535 || gx > INT_MAX) throw std::invalid_argument("gemm: problem too large for one launch");
    const dim3 grid((unsigned)gx, (unsigned)gy), block(256);
    if (!ta && !tb) gemm_kernel<T, BM, BN, BK, false, false><<<grid, block, 0, s>>>(p, abort);
    else if (!ta && tb) gemm_kernel<T, BM, BN, BK, false, true><<<grid, block, 0, s>>>(p, abort);
    else if (ta && !tb) gemm_kernel<T, BM, BN, BK, true, false><<<grid, block, 0, s>>>(p, abort);
    else gemm_kernel<T, BM, BN, BK, true, true><<<grid, block, 0, s>>>(p, abort);
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
          long long ldb, T beta, T* C, long long ldc, Tri tri, long long diag, const int* abort)
{
    if (m <= 0 || n <= 0) return;
    GemmProblem<T> p{m, n, k, alpha, beta, A, lda, B, ldb, C, ldc, (int)tri, diag};
    const bool a_t = !is_n(ta), b_t = !is_n(tb);
    if (m >= 128 && n >= 128) launch_gemm_cfg<T, 128, 128, 16>(s, a_t, b_t, p, abort);
    else launch_gemm_cfg<T, 64, 64, 16>(s, a_t, b_t, p, abort);
}

template <class T>
void gemm_vbatched(hipStream_t s, char ta, char tb, const int* m, const int* n, const int* k, T alpha,
                   const T* const* A, const int* lda, const T* const* B, const int* ldb, T beta, T* const* C,
                   const int* ldc, int batch, int max_m, int max_n)
{
    if (batch <= 0 || max_m <= 0 || max_n <= 0) return;
    VBatch<T> v{m, n, k, alpha, beta, A, lda, B, ldb, C, ldc};
    launch_vbatched_cfg<T, 64, 64, 16>(s, !is_n(ta), !is_n(tb), v, batch, max_m, max_n);
}

void vbatched_validate(hipStream_t s, int na, int nb, const int* m, const int* n, const int* k, const int* lda,
                       const int* ldb, const int* ldc, int batch, int* summary)
{
    vbatched_validate_kernel<<<1, 256, 0, s>>>(na, nb, m, n, k, lda, ldb, ldc, batch, summary);
    check_launch("vbatched_validate");
}

template <class T>
void potf2_inv(hipStream_t s, bool upper, int w, T* A, long long lda, T* inv, int* info, int info_base)
{
    if (w <= 0 || w > kIB) throw std::invalid_argument("potf2_inv: block order out of range");
    potf2_inv_kernel<T><<<1, 256, 0, s>>>(upper, w, A, lda, inv, info, info_base);
    check_launch("potf2_inv");
}

template <class T>
void trtri_diag_blocks(hipStream_t s, bool upper, int n, const T* A, long long lda, T* inv)
{
    if (n <= 0) return;
    trtri_kernel<T><<<(n + kIB - 1) / kIB, 256, 0, s>>>(upper, n, A, lda, inv);
    check_launch("trtri_diag_blocks");
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
                          T, T*, long long, Tri, long long, const int*);                                      \
    template void gemm_vbatched<T>(hipStream_t, char, char, const int*, const int*, const int*, T,            \
                                   const T* const*, const int*, const T* const*, const int*, T, T* const*,    \
                                   const int*, int, int, int);                                                \
    template void potf2_inv<T>(hipStream_t, bool, int, T*, long long, T*, int*, int);                         \
    template void trtri_diag_blocks<T>(hipStream_t, bool, int, const T*, long long, T*);                      \
    template void diag_logdet<T>(hipStream_t, int, const T*, long long, void*, LogdetResult*);                \
    template void add_diagonal<T>(hipStream_t, int, T*, long long, T);                                        \
    template void transpose<T>(hipStream_t, int, int, const T*, long long, T*, long long);
DDLA_KERN_INSTANTIATE(double)
DDLA_KERN_INSTANTIATE(float)
#undef DDLA_KERN_INSTANTIATE
template void convert<float, double>(hipStream_t, int, int, const double*, long long, float*, long long);
template void convert<double, float>(hipStream_t, int, int, const float*, long long, double*, long long);
template void convert<double, double>(hipStream_t, int, int, const double*, long long, double*, long long);

} // namespace kern
} // namespace ddla
