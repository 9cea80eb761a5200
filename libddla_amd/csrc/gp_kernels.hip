// Element-wise and reduction kernels of the batched Gaussian-process
// likelihood (gp.cpp).  Grid y (or one workgroup per problem) runs over the
// independent problems; problem b's working region sits b * stride bytes
// after problem 0's, and the caller's arrays have strides of their own.
#include <hip/hip_runtime.h>

#include "kernels.h"

#include <algorithm>
#include <stdexcept>
#include <string>

namespace ddla {
namespace kern {

namespace {

void check_launch(const char* what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <class P>
__device__ inline P* moved(P* p, long long bytes)
{
    return reinterpret_cast<P*>(reinterpret_cast<char*>(const_cast<void*>(static_cast<const void*>(p))) + bytes);
}

// L_b = the lower triangle of K_b (n x n, both column-major with leading
// dimension n; L_b's strict upper triangle is left as it is), y copied,
// info_b = 0.
__global__ __launch_bounds__(256) void gp_load_kernel(int n, const double* K, long long k_stride, const double* y,
                                                      long long y_stride, double* L, double* ybuf, int* info,
                                                      long long stride)
{
    const long long b = blockIdx.y;
    const double* Kb = K + b * k_stride;
    double* Lb = moved(L, b * stride);
    // the lower triangle, column by column (the factorization reads nothing else)
    for (long long j = blockIdx.x; j < n; j += gridDim.x)
        for (long long i = j + threadIdx.x; i < n; i += 256) Lb[i + j * n] = Kb[i + j * n];
    if (blockIdx.x == 0) {
        double* yb = moved(ybuf, b * stride);
        for (int i = threadIdx.x; i < n; i += 256) yb[i] = y[b * y_stride + i];
        if (threadIdx.x == 0) *moved(info, b * stride) = 0;
    }
}

// X_b := the block-diagonal matrix of the 64 x 64 inverses D_j (dinv, zeros
// above their diagonals), zero elsewhere.
__global__ __launch_bounds__(256) void gp_place_diag_kernel(int n, const double* dinv, double* X, long long stride)
{
    const long long b = blockIdx.y;
    const double* D = moved(dinv, b * stride);
    double* Xb = moved(X, b * stride);
    const long long total = (long long)n * n;
    for (long long t = blockIdx.x * 256LL + threadIdx.x; t < total; t += (long long)gridDim.x * 256) {
        const int i = (int)(t % n), j = (int)(t / n);
        const int bi = i >> 6, bj = j >> 6;
        Xb[t] = bi == bj ? D[(long long)bj * 4096 + (i & 63) + 64 * (j & 63)] : 0.0;
    }
}

// One workgroup per problem: loss_b = z^T z / 2 + sum_i log L_ii (fixed
// summation order), also kept at loss_slot; info copied out.
__global__ __launch_bounds__(256) void gp_loss_kernel(int n, const double* L, const double* z, const int* info,
                                                      double* loss_slot, long long stride, double* loss_out,
                                                      int* info_out)
{
    __shared__ double sq[256], sl[256];
    const long long b = blockIdx.x;
    const double* Lb = moved(L, b * stride);
    const double* zb = moved(z, b * stride);
    double q = 0.0, l = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        q = fma(zb[i], zb[i], q);
        l += log(Lb[(long long)i * (n + 1)]);
    }
    sq[threadIdx.x] = q;
    sl[threadIdx.x] = l;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            sq[threadIdx.x] += sq[threadIdx.x + o];
            sl[threadIdx.x] += sl[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double v = 0.5 * sq[0] + sl[0];
        *moved(loss_slot, b * stride) = v;
        loss_out[b] = v;
        info_out[b] = *moved(info, b * stride);
    }
}

// Kbar_b(i, j) = g_b (M(max(i, j), min(i, j)) - a_i a_j) / 2 and ybar_b = g_b a
// (M = tril(K^-1), a = K^-1 y).  Workgroup (bi, bj, b) writes the 32 x 32
// tile (bi, bj) of Kbar_b from the lower tile (max, min) of M_b staged in LDS,
// so both the reads and the writes run along columns.
__global__ __launch_bounds__(256) void gp_grad_kernel(int n, const double* M, const double* a, long long stride,
                                                      const double* g, double* Kbar, long long kbar_stride,
                                                      double* ybar, long long ybar_stride)
{
    __shared__ double tile[32][33];
    const long long b = blockIdx.z;
    const double* Mb = moved(M, b * stride);
    const double* ab = moved(a, b * stride);
    const double gb = g[b];
    double* Kb = Kbar + b * kbar_stride;
    const int bi = blockIdx.x, bj = blockIdx.y;
    const long long si = (long long)max(bi, bj) * 32, sj = (long long)min(bi, bj) * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int r = ty; r < 32; r += 8)
        if (si + tx < n && sj + r < n) tile[r][tx] = Mb[(si + tx) + (sj + r) * (long long)n];  // tile[col][row]
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const long long i = (long long)bi * 32 + tx, j = (long long)bj * 32 + r;
        if (i < n && j < n) {
            const bool mirrored = bi < bj || (bi == bj && tx < r);
            const double m = mirrored ? tile[tx][r] : tile[r][tx];
            Kb[i + j * n] = 0.5 * gb * (m - ab[i] * ab[j]);
        }
    }
    if (bi == 0 && bj == 0 && ybar != nullptr)
        for (int i = threadIdx.x; i < n; i += 256) ybar[b * ybar_stride + i] = gb * ab[i];
}

unsigned grid_x(long long work) { return (unsigned)std::max(1LL, std::min(4096LL, (work + 255) / 256)); }

} // namespace

void gp_load(hipStream_t s, int batch, int n, const double* K, long long k_stride, const double* y, long long y_stride,
             double* L, double* ybuf, int* info, long long stride)
{
    if (batch <= 0 || n <= 0) return;
    gp_load_kernel<<<dim3((unsigned)std::min(n, 1024), (unsigned)batch), 256, 0, s>>>(n, K, k_stride, y, y_stride,
                                                                                      L, ybuf, info, stride);
    check_launch("gp_load");
}

void gp_place_diag(hipStream_t s, int batch, int n, const double* dinv, double* X, long long stride)
{
    if (batch <= 0 || n <= 0) return;
    gp_place_diag_kernel<<<dim3(grid_x((long long)n * n), (unsigned)batch), 256, 0, s>>>(n, dinv, X, stride);
    check_launch("gp_place_diag");
}

void gp_loss(hipStream_t s, int batch, int n, const double* L, const double* z, const int* info, double* loss_slot,
             long long stride, double* loss_out, int* info_out)
{
    if (batch <= 0) return;
    gp_loss_kernel<<<(unsigned)batch, 256, 0, s>>>(n, L, z, info, loss_slot, stride, loss_out, info_out);
    check_launch("gp_loss");
}

void gp_grad(hipStream_t s, int batch, int n, const double* M, const double* a, long long stride, const double* g,
             double* Kbar, long long kbar_stride, double* ybar, long long ybar_stride)
{
    if (batch <= 0 || n <= 0) return;
    const unsigned tiles = (unsigned)((n + 31) / 32);
    gp_grad_kernel<<<dim3(tiles, tiles, (unsigned)batch), 256, 0, s>>>(n, M, a, stride, g, Kbar, kbar_stride, ybar,
                                                                      ybar_stride);
    check_launch("gp_grad");
}

} // namespace kern
} // namespace ddla
