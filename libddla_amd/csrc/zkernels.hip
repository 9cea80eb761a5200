// Complex kernels of the Hermitian Cholesky path (declarations and contracts
// in zkernels.h).  There is no counterpart kernel in the reference, which calls
// vendor routines (hipSOLVER zpotrf / hipBLAS ztrsm, OpenBLAS on the CPU) for
// the complex types; the semantics followed are those of LAPACK ?potrf /
// ?potrs and of the reference's gemmVbatched and diag_logdet
// (src/vbatched/gemm_template_kernel_vbatched.hpp:46-105,
// src/diag_logdet_impl.h:50-113).
//
// Complex arithmetic runs on the vector ALUs: on MI355X the FP64 and FP32
// vector FMA rates equal the matrix-core rates of those types (78.6 and
// 157.3 TF/s), and a complex product is four real FMAs per lane with no
// operand shuffling.  GEMM tiles are 64 x 64 per 256-thread workgroup, a 4 x 4
// block of complex accumulators per thread, op(A) and op(B) staged through LDS
// 16 columns deep with the next slab's loads in flight.
#include "zkernels.h"

#include <cfloat>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace ddla {
namespace kern {

namespace {

inline void zcheck_launch(const char* what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <class R>
struct Cx {
    R re, im;
};

template <class R>
__device__ __forceinline__ Cx<R> cmul(Cx<R> a, Cx<R> b)
{
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class R>
__device__ __forceinline__ void cfma(Cx<R>& c, Cx<R> a, Cx<R> b)  // c += a b
{
    c.re = fma(a.re, b.re, c.re);
    c.re = fma(-a.im, b.im, c.re);
    c.im = fma(a.re, b.im, c.im);
    c.im = fma(a.im, b.re, c.im);
}
template <class R>
__device__ __forceinline__ Cx<R> conj(Cx<R> a)
{
    return {a.re, -a.im};
}

// ---------------------------------------------------------------------------
// GEMM tile
// ---------------------------------------------------------------------------
constexpr int kT = 64;   // tile rows / columns
constexpr int kK = 16;   // slab depth
constexpr int kTh = 256;
constexpr int kPerT = kT * kK / kTh;  // operand elements per thread per slab (4)

template <class R>
struct ZProblem {
    int m, n, k;
    Cx<R> alpha, beta;
    const void* A;
    long long lda;
    int opa;
    const void* B;
    long long ldb;
    int opb;
    Cx<R>* C;
    long long ldc;
    int tri;
    long long diag;
};

// Element of a real or complex operand (index idx of its array).
template <class R, bool CPX>
__device__ __forceinline__ Cx<R> operand(const void* P, long long idx, bool conjugate)
{
    if constexpr (CPX) {
        const Cx<R> v = static_cast<const Cx<R>*>(P)[idx];
        return conjugate ? Cx<R>{v.re, -v.im} : v;
    } else {
        return Cx<R>{static_cast<const R*>(P)[idx], R(0)};
    }
}

// Tile (ti, tj) of the problem.
template <class R, bool ACPX, bool BCPX>
__device__ void ztile(const ZProblem<R>& p, int ti, int tj, Cx<R> (*As)[kT + 1], Cx<R> (*Bs)[kT + 1])
{
    const int m0 = ti * kT, n0 = tj * kT;
    if (m0 >= p.m || n0 >= p.n) return;
    if (p.tri == 1 && (long long)(m0 + kT - 1) + p.diag < n0) return;  // wholly above the written lower part
    if (p.tri == 2 && (long long)m0 + p.diag > n0 + kT - 1) return;    // wholly below the written upper part
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    Cx<R> acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = Cx<R>{R(0), R(0)};
    const bool product = p.k > 0 && (p.alpha.re != R(0) || p.alpha.im != R(0));
    if (product) {
        const bool a_n = p.opa == 0, b_n = p.opb == 0;
        const bool a_conj = p.opa == 2, b_conj = p.opb == 2;
        Cx<R> va[kPerT], vb[kPerT];
        // op(A)(i, kk) and op(B)(kk, j) of slab k0, clamped and masked
        auto load = [&](int k0) {
#pragma unroll
            for (int e = 0; e < kPerT; ++e) {
                const int idx = tid + kTh * e;
                const int i = a_n ? (idx & (kT - 1)) : (idx >> 4);
                const int kk = a_n ? (idx >> 6) : (idx & (kK - 1));
                const int gi = m0 + i, gk = k0 + kk;
                const bool in = gi < p.m && gk < p.k;
                const long long ci = min(gi, p.m - 1), ck = min(gk, p.k - 1);
                const Cx<R> v = operand<R, ACPX>(p.A, a_n ? ci + ck * p.lda : ck + ci * p.lda, a_conj);
                va[e] = in ? v : Cx<R>{R(0), R(0)};
            }
#pragma unroll
            for (int e = 0; e < kPerT; ++e) {
                const int idx = tid + kTh * e;
                const int kk = b_n ? (idx & (kK - 1)) : (idx >> 6);
                const int j = b_n ? (idx >> 4) : (idx & (kT - 1));
                const int gj = n0 + j, gk = k0 + kk;
                const bool in = gj < p.n && gk < p.k;
                const long long cj = min(gj, p.n - 1), ck = min(gk, p.k - 1);
                const Cx<R> v = operand<R, BCPX>(p.B, b_n ? ck + cj * p.ldb : cj + ck * p.ldb, b_conj);
                vb[e] = in ? v : Cx<R>{R(0), R(0)};
            }
        };
        auto store = [&]() {
#pragma unroll
            for (int e = 0; e < kPerT; ++e) {
                const int idx = tid + kTh * e;
                const int i = a_n ? (idx & (kT - 1)) : (idx >> 4);
                const int kk = a_n ? (idx >> 6) : (idx & (kK - 1));
                As[kk][i] = va[e];
            }
#pragma unroll
            for (int e = 0; e < kPerT; ++e) {
                const int idx = tid + kTh * e;
                const int kk = b_n ? (idx & (kK - 1)) : (idx >> 6);
                const int j = b_n ? (idx >> 4) : (idx & (kT - 1));
                Bs[kk][j] = vb[e];
            }
        };
        load(0);
        for (int k0 = 0; k0 < p.k; k0 += kK) {
            store();
            __syncthreads();
            if (k0 + kK < p.k) load(k0 + kK);
#pragma unroll 4
            for (int kk = 0; kk < kK; ++kk) {
                Cx<R> a[4], b[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) a[r] = As[kk][tx * 4 + r];
#pragma unroll
                for (int c = 0; c < 4; ++c) b[c] = Bs[kk][ty * 4 + c];
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int c = 0; c < 4; ++c) cfma(acc[r][c], a[r], b[c]);
            }
            __syncthreads();
        }
    }
    const bool beta_zero = p.beta.re == R(0) && p.beta.im == R(0);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int j = n0 + ty * 4 + c;
        if (j >= p.n) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = m0 + tx * 4 + r;
            if (i >= p.m) continue;
            if (p.tri == 1 && (long long)i + p.diag < j) continue;
            if (p.tri == 2 && (long long)i + p.diag > j) continue;
            Cx<R>* cp = p.C + i + (long long)j * p.ldc;
            Cx<R> v = cmul(p.alpha, acc[r][c]);
            if (!beta_zero) cfma(v, p.beta, *cp);
            *cp = v;
        }
    }
}

template <class R, bool ACPX, bool BCPX>
__global__ __launch_bounds__(kTh) void zgemm_kernel(ZProblem<R> p, const int* abort)
{
    __shared__ Cx<R> As[kK][kT + 1];
    __shared__ Cx<R> Bs[kK][kT + 1];
    if (abort != nullptr && *abort != 0) return;
    ztile<R, ACPX, BCPX>(p, blockIdx.x, blockIdx.y, As, Bs);
}

template <class R>
struct ZBatch {
    const int *m, *n, *k, *lda, *ldb, *ldc;
    const void* const* A;
    const void* const* B;
    Cx<R>* const* C;
    Cx<R> alpha, beta;
    int opa, opb, batch, z0;
};

template <class R, bool ACPX, bool BCPX>
__global__ __launch_bounds__(kTh) void zgemm_vbatched_kernel(ZBatch<R> v)
{
    __shared__ Cx<R> As[kK][kT + 1];
    __shared__ Cx<R> Bs[kK][kT + 1];
    const int b = v.z0 + blockIdx.z;
    if (b >= v.batch) return;
    ZProblem<R> p;
    p.m = v.m[b];
    p.n = v.n[b];
    p.k = v.k[b];
    if (p.m <= 0 || p.n <= 0) return;
    p.alpha = v.alpha;
    p.beta = v.beta;
    p.A = v.A[b];
    p.lda = v.lda[b];
    p.opa = v.opa;
    p.B = v.B[b];
    p.ldb = v.ldb[b];
    p.opb = v.opb;
    p.C = v.C[b];
    p.ldc = v.ldc[b];
    p.tri = 0;
    p.diag = 0;
    ztile<R, ACPX, BCPX>(p, blockIdx.x, blockIdx.y, As, Bs);
}

// ---------------------------------------------------------------------------
// 64 x 64 Hermitian diagonal block: factor (lower, right-looking, in LDS) and
// inverse of the factor.
// ---------------------------------------------------------------------------
constexpr int kD = 64;
constexpr int kDL = kD + 1;

// X := S^-1 for the order-w lower-triangular S (real positive diagonal) in
// LDS; thread c < 64 forms column c by forward substitution.  Rows and
// columns from w on: X = identity there.
template <class R>
__device__ void invert_lower(int w, Cx<R> (*S)[kDL], Cx<R> (*X)[kDL])
{
    const int c = threadIdx.x;
    if (c < kD) {
        for (int i = 0; i < kD; ++i) {
            if (i < c || c >= w || i >= w) {
                X[i][c] = Cx<R>{(i == c) ? R(1) : R(0), R(0)};
                continue;
            }
            Cx<R> acc{i == c ? R(1) : R(0), R(0)};
            for (int k = c; k < i; ++k) {
                const Cx<R> l = S[i][k], x = X[k][c];
                acc.re -= l.re * x.re - l.im * x.im;
                acc.im -= l.re * x.im + l.im * x.re;
            }
            const R d = S[i][i].re;
            X[i][c] = Cx<R>{acc.re / d, acc.im / d};
        }
    }
}

template <class R>
__device__ void store_inverse(Cx<R> (*X)[kDL], Cx<R>* inv)
{
    for (int idx = threadIdx.x; idx < kD * kD; idx += blockDim.x) {
        const int r = idx & (kD - 1), c = idx >> 6;
        inv[idx] = r >= c ? X[r][c] : Cx<R>{R(0), R(0)};
    }
}

template <class R>
__global__ __launch_bounds__(256) void zpotf2_kernel(int n, Cx<R>* A, long long lda, int k0, Cx<R>* inv, int* info)
{
    __shared__ Cx<R> S[kD][kDL];
    __shared__ Cx<R> X[kD][kDL];
    __shared__ int s_fail;
    if (*info != 0) return;
    const int w = min(kD, n - k0), tid = threadIdx.x;
    Cx<R>* D = A + k0 + (long long)k0 * lda;
    for (int idx = tid; idx < kD * kD; idx += 256) {
        const int r = idx & (kD - 1), c = idx >> 6;
        S[r][c] = (r < w && c < w && r >= c) ? D[r + (long long)c * lda] : Cx<R>{R(0), R(0)};
    }
    if (tid == 0) s_fail = 0;
    __syncthreads();
    for (int j = 0; j < w; ++j) {
        const R d = S[j][j].re;
        if (!(d > R(0))) {  // not positive, or NaN (LAPACK zpotrf's test on the real diagonal)
            if (tid == 0) s_fail = j + 1;
            break;  // uniform: every thread reads the same S[j][j]
        }
        const R l = sqrt(d);
        __syncthreads();
        if (tid == 0) S[j][j] = Cx<R>{l, R(0)};
        for (int i = j + 1 + tid; i < w; i += 256) S[i][j] = Cx<R>{S[i][j].re / l, S[i][j].im / l};
        __syncthreads();
        // S(i, c) -= S(i, j) conj(S(c, j)), j < c <= i < w
        const int t = w - j - 1;
        for (int idx = tid; idx < t * t; idx += 256) {
            const int ii = idx % t, cc = idx / t;
            if (ii < cc) continue;
            const int i = j + 1 + ii, c = j + 1 + cc;
            const Cx<R> a = S[i][j], b = S[c][j];
            S[i][c].re -= a.re * b.re + a.im * b.im;
            S[i][c].im -= a.im * b.re - a.re * b.im;
        }
        __syncthreads();
    }
    __syncthreads();
    // the factored part (and, on failure, the partly updated rest) goes back
    for (int idx = tid; idx < kD * kD; idx += 256) {
        const int r = idx & (kD - 1), c = idx >> 6;
        if (r < w && c < w && r >= c) D[r + (long long)c * lda] = S[r][c];
    }
    if (s_fail != 0) {
        if (tid == 0) *info = k0 + s_fail;
        return;
    }
    invert_lower<R>(w, S, X);
    __syncthreads();
    store_inverse<R>(X, inv);
}

template <class R>
__global__ __launch_bounds__(256) void ztrtri_kernel(bool upper, int n, const Cx<R>* A, long long lda, Cx<R>* inv)
{
    __shared__ Cx<R> S[kD][kDL];
    __shared__ Cx<R> X[kD][kDL];
    const long long k0 = (long long)blockIdx.x * kD;
    const int w = (int)min((long long)kD, n - k0);
    for (int idx = threadIdx.x; idx < kD * kD; idx += 256) {
        const int r = idx & (kD - 1), c = idx >> 6;
        Cx<R> v{R(0), R(0)};
        if (r < w && c < w && r >= c) {
            // L view: L(i, j) = A[i + j lda], or conj(A[j + i lda]) for an upper factor
            const long long i = k0 + r, j = k0 + c;
            v = upper ? conj(A[j + i * lda]) : A[i + j * lda];
        }
        S[r][c] = v;
    }
    __syncthreads();
    invert_lower<R>(w, S, X);
    __syncthreads();
    store_inverse<R>(X, inv + (long long)blockIdx.x * kD * kD);
}

// ---------------------------------------------------------------------------
// Transposes and the Hermitian triangle swap, 32 x 32 tiles through LDS.
// ---------------------------------------------------------------------------
template <class R>
__global__ __launch_bounds__(256) void zconj_transpose_kernel(int m, int n, const Cx<R>* in, long long ldi, Cx<R>* out,
                                                              long long ldo)
{
    __shared__ Cx<R> t[32][33];
    const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32, tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {
        const int i = i0 + tx, j = j0 + r;
        if (i < m && j < n) t[r][tx] = in[i + (long long)j * ldi];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int j = j0 + tx, i = i0 + r;  // out(j, i) = conj(in(i, j))
        if (i < m && j < n) out[j + (long long)i * ldo] = conj(t[tx][r]);
    }
}

// Tile pair (bi, bj), bi >= bj, of the 32 x 32 tiling: lower tile (bi, bj)
// and upper tile (bj, bi) exchange with conjugation (diagonal tiles: their
// strict triangles).
template <class R>
__global__ __launch_bounds__(256) void zswap_triangles_kernel(int n, Cx<R>* A, long long lda)
{
    __shared__ Cx<R> lo[32][33], up[32][33];
    // blockIdx.x enumerates the lower triangle of tiles
    const int t = blockIdx.x;
    int bi = (int)((sqrt(8.0 * t + 1.0) - 1.0) / 2.0);
    while ((bi + 1) * (bi + 2) / 2 <= t) ++bi;
    while (bi * (bi + 1) / 2 > t) --bi;
    const int bj = t - bi * (bi + 1) / 2;
    const int i0 = bi * 32, j0 = bj * 32, tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {
        const int i = i0 + tx, j = j0 + r;  // lower tile element (i, j)
        if (i < n && j < n) lo[tx][r] = A[i + (long long)j * lda];
        const int ui = j0 + tx, uj = i0 + r;  // upper tile element (ui, uj)
        if (ui < n && uj < n) up[tx][r] = A[ui + (long long)uj * lda];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        // lower (i, j) <- conj(upper (j, i)); upper (ui, uj) <- conj(lower (uj, ui))
        const int i = i0 + tx, j = j0 + r;
        if (i < n && j < n && i > j) A[i + (long long)j * lda] = conj(up[r][tx]);
        const int ui = j0 + tx, uj = i0 + r;
        if (ui < n && uj < n && uj > ui) A[ui + (long long)uj * lda] = conj(lo[r][tx]);
    }
}

// ---------------------------------------------------------------------------
// log-determinant of the diagonal: deterministic two-pass reduction
// ---------------------------------------------------------------------------
constexpr int kZParts = 64;

template <class R>
__global__ __launch_bounds__(256) void zlogdet_partial_kernel(int n, const Cx<R>* A, long long lda, double* pmag,
                                                              double* pphase, int* pflags)
{
    __shared__ double sm[256], sp[256];
    __shared__ int sf[256];
    double mag = 0, ph = 0;
    int f = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const Cx<R> z = A[(long long)i * (lda + 1)];
        const double x = (double)z.re, y = (double)z.im;
        if (!(x <= DBL_MAX && x >= -DBL_MAX && y <= DBL_MAX && y >= -DBL_MAX)) {
            f |= 1;
            continue;
        }
        const double ax = fabs(x), ay = fabs(y);
        const double hi = ax > ay ? ax : ay, lo = ax > ay ? ay : ax;
        if (hi == 0) {
            f |= 2;
            continue;
        }
        const double ratio = lo / hi;
        mag += log(hi) + 0.5 * log1p(ratio * ratio);  // log |z| without overflowing |z|
        if (y == 0) {
            if (x < 0) f ^= 4;
        } else {
            ph += atan2(y, x);
        }
    }
    sm[threadIdx.x] = mag;
    sp[threadIdx.x] = ph;
    sf[threadIdx.x] = f;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sm[threadIdx.x] += sm[threadIdx.x + s];
            sp[threadIdx.x] += sp[threadIdx.x + s];
            const int a = sf[threadIdx.x], b = sf[threadIdx.x + s];
            sf[threadIdx.x] = ((a | b) & 3) | ((a ^ b) & 4);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pmag[blockIdx.x] = sm[0];
        pphase[blockIdx.x] = sp[0];
        pflags[blockIdx.x] = sf[0];
    }
}

__global__ void zlogdet_final_kernel(const double* pmag, const double* pphase, const int* pflags, double* out)
{
    if (threadIdx.x != 0) return;
    double m = 0, ph = 0;
    int f = 0;
    for (int b = 0; b < kZParts; ++b) {
        m += pmag[b];
        ph += pphase[b];
        f = ((f | pflags[b]) & 3) | ((f ^ pflags[b]) & 4);
    }
    out[0] = m;
    out[1] = ph;
    *reinterpret_cast<int*>(out + 2) = f;
}

template <class R>
void launch_zgemm(hipStream_t s, bool ac, bool bc, const ZProblem<R>& p, const int* abort)
{
    const dim3 grid((unsigned)((p.m + kT - 1) / kT), (unsigned)((p.n + kT - 1) / kT));
    if (grid.y > 65535) throw std::invalid_argument("zgemm: too many column tiles");
    if (ac && bc) zgemm_kernel<R, true, true><<<grid, kTh, 0, s>>>(p, abort);
    else if (ac) zgemm_kernel<R, true, false><<<grid, kTh, 0, s>>>(p, abort);
    else if (bc) zgemm_kernel<R, false, true><<<grid, kTh, 0, s>>>(p, abort);
    else throw std::invalid_argument("zgemm: at least one complex operand");
    zcheck_launch("zgemm");
}

} // namespace

int op_code(char op)
{
    switch (op) {
    case 'N': case 'n': return 0;
    case 'T': case 't': return 1;
    case 'C': case 'c': return 2;
    default: throw std::invalid_argument("invalid transpose operation");
    }
}

template <class R>
void zgemm(hipStream_t s, bool a_complex, bool b_complex, char ta, char tb, int m, int n, int k,
           std::complex<R> alpha, const void* A, long long lda, const void* B, long long ldb, std::complex<R> beta,
           std::complex<R>* C, long long ldc, int tri, long long diag, const int* abort)
{
    if (m <= 0 || n <= 0) return;
    ZProblem<R> p;
    p.m = m;
    p.n = n;
    p.k = k < 0 ? 0 : k;
    p.alpha = {alpha.real(), alpha.imag()};
    p.beta = {beta.real(), beta.imag()};
    p.A = A;
    p.lda = lda;
    p.opa = op_code(ta);
    p.B = B;
    p.ldb = ldb;
    p.opb = op_code(tb);
    p.C = reinterpret_cast<Cx<R>*>(C);
    p.ldc = ldc;
    p.tri = tri;
    p.diag = diag;
    launch_zgemm<R>(s, a_complex, b_complex, p, abort);
}

template <class R>
void zgemm_vbatched(hipStream_t s, bool a_complex, bool b_complex, char ta, char tb, const int* m, const int* n,
                    const int* k, std::complex<R> alpha, const void* const* A, const int* lda, const void* const* B,
                    const int* ldb, std::complex<R> beta, std::complex<R>* const* C, const int* ldc, int batch,
                    int max_m, int max_n)
{
    if (batch <= 0 || max_m <= 0 || max_n <= 0) return;
    ZBatch<R> v;
    v.m = m;
    v.n = n;
    v.k = k;
    v.lda = lda;
    v.ldb = ldb;
    v.ldc = ldc;
    v.A = A;
    v.B = B;
    v.C = reinterpret_cast<Cx<R>* const*>(C);
    v.alpha = {alpha.real(), alpha.imag()};
    v.beta = {beta.real(), beta.imag()};
    v.opa = op_code(ta);
    v.opb = op_code(tb);
    v.batch = batch;
    const unsigned gx = (unsigned)((max_m + kT - 1) / kT), gy = (unsigned)((max_n + kT - 1) / kT);
    if (gy > 65535) throw std::invalid_argument("zgemm_vbatched: too many column tiles");
    for (int z0 = 0; z0 < batch; z0 += 65535) {
        v.z0 = z0;
        const dim3 grid(gx, gy, (unsigned)std::min(65535, batch - z0));
        if (a_complex && b_complex) zgemm_vbatched_kernel<R, true, true><<<grid, kTh, 0, s>>>(v);
        else if (a_complex) zgemm_vbatched_kernel<R, true, false><<<grid, kTh, 0, s>>>(v);
        else if (b_complex) zgemm_vbatched_kernel<R, false, true><<<grid, kTh, 0, s>>>(v);
        else throw std::invalid_argument("zgemm_vbatched: at least one complex operand");
        zcheck_launch("zgemm_vbatched");
    }
}

template <class R>
void zpotf2(hipStream_t s, int n, std::complex<R>* A, long long lda, int k0, std::complex<R>* inv, int* info)
{
    zpotf2_kernel<R><<<1, 256, 0, s>>>(n, reinterpret_cast<Cx<R>*>(A), lda, k0, reinterpret_cast<Cx<R>*>(inv),
                                       info);
    zcheck_launch("zpotf2");
}

template <class R>
void ztrtri_diag_blocks(hipStream_t s, bool upper, int n, const std::complex<R>* A, long long lda,
                        std::complex<R>* inv)
{
    if (n <= 0) return;
    ztrtri_kernel<R><<<(n + kD - 1) / kD, 256, 0, s>>>(upper, n, reinterpret_cast<const Cx<R>*>(A), lda,
                                                       reinterpret_cast<Cx<R>*>(inv));
    zcheck_launch("ztrtri_diag_blocks");
}

template <class R>
void zconj_transpose(hipStream_t s, int m, int n, const std::complex<R>* in, long long ldi, std::complex<R>* out,
                     long long ldo)
{
    if (m <= 0 || n <= 0) return;
    const dim3 grid((unsigned)((m + 31) / 32), (unsigned)((n + 31) / 32));
    if (grid.y > 65535) throw std::invalid_argument("zconj_transpose: too many columns");
    zconj_transpose_kernel<R><<<grid, 256, 0, s>>>(m, n, reinterpret_cast<const Cx<R>*>(in), ldi,
                                                   reinterpret_cast<Cx<R>*>(out), ldo);
    zcheck_launch("zconj_transpose");
}

template <class R>
void zswap_triangles(hipStream_t s, int n, std::complex<R>* A, long long lda)
{
    if (n <= 1) return;
    const long long nt = (n + 31) / 32, tiles = nt * (nt + 1) / 2;
    if (tiles > 0x7fffffffLL) throw std::invalid_argument("zswap_triangles: order too large");
    zswap_triangles_kernel<R><<<(unsigned)tiles, 256, 0, s>>>(n, reinterpret_cast<Cx<R>*>(A), lda);
    zcheck_launch("zswap_triangles");
}

std::size_t zlogdet_work_bytes() { return (std::size_t)kZParts * (2 * sizeof(double) + sizeof(int)); }

template <class R>
void zdiag_logdet(hipStream_t s, int n, const std::complex<R>* A, long long lda, void* work, double* out)
{
    double* pmag = static_cast<double*>(work);
    double* pphase = pmag + kZParts;
    int* pflags = reinterpret_cast<int*>(pphase + kZParts);
    zlogdet_partial_kernel<R><<<kZParts, 256, 0, s>>>(n, reinterpret_cast<const Cx<R>*>(A), lda, pmag, pphase,
                                                       pflags);
    zcheck_launch("zdiag_logdet");
    zlogdet_final_kernel<<<1, 64, 0, s>>>(pmag, pphase, pflags, out);
    zcheck_launch("zdiag_logdet");
}

#define DDLA_ZINSTANTIATE(R)                                                                                      \
    template void zgemm<R>(hipStream_t, bool, bool, char, char, int, int, int, std::complex<R>, const void*,     \
                           long long, const void*, long long, std::complex<R>, std::complex<R>*, long long, int,  \
                           long long, const int*);                                                               \
    template void zgemm_vbatched<R>(hipStream_t, bool, bool, char, char, const int*, const int*, const int*,     \
                                    std::complex<R>, const void* const*, const int*, const void* const*,         \
                                    const int*, std::complex<R>, std::complex<R>* const*, const int*, int, int,  \
                                    int);                                                                        \
    template void zpotf2<R>(hipStream_t, int, std::complex<R>*, long long, int, std::complex<R>*, int*);         \
    template void ztrtri_diag_blocks<R>(hipStream_t, bool, int, const std::complex<R>*, long long,              \
                                        std::complex<R>*);                                                       \
    template void zconj_transpose<R>(hipStream_t, int, int, const std::complex<R>*, long long, std::complex<R>*, \
                                     long long);                                                                 \
    template void zswap_triangles<R>(hipStream_t, int, std::complex<R>*, long long);                             \
    template void zdiag_logdet<R>(hipStream_t, int, const std::complex<R>*, long long, void*, double*);
DDLA_ZINSTANTIATE(float)
DDLA_ZINSTANTIATE(double)
#undef DDLA_ZINSTANTIATE

} // namespace kern
} // namespace ddla
