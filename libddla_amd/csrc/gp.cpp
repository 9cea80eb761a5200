// Batched Gaussian-process negative log-likelihood: forward and backward for
// many independent problems of one order, with every launch serving all of
// them.  No counterpart in the reference; per problem it computes what the
// composed ppotrf / ppotrs / diag_logdet / potri path does:
//   forward:  K = L L^T (the blocked factorization with its look-ahead, the
//             panel launch and each trailing update covering every problem),
//             X = L^-1 by a bottom-up merge of inverted diagonal blocks
//             (X21 = -X22 (L21 X11), two products per merge, every merge of a
//             level in one launch), z = X y,
//             loss = z^T z / 2 + sum log L_ii  (= y^T K^-1 y / 2 + log det K / 2);
//   backward: K^-1 = X^T X (lower triangle), a = X^T z = K^-1 y,
//             Kbar = g (K^-1 - a a^T) / 2, ybar = g a
//             (Rasmussen & Williams 2006, eq. 5.9).
// Problem b's working region sits b * S bytes after problem 0's, S =
// gpBatchedWorkspaceSize(n), so each kernel moves all of its pointers by the
// same offset (kern::Batch).
#include "internal.h"
#include "kernels.h"

#include <algorithm>
#include <string>
#include <vector>

namespace ddla {

template <class T>
void potrf_batched_dispatch(DdlaStream* h, int n, T* A, long long lda, int* info, void* work, const kern::Batch& batch);
void lower_gram_batched(hipStream_t s, int n, const double* X, double* M, const kern::Batch& batch);
double lower_gram_flops(int n);

namespace {

constexpr int kIB = 64;

// One problem's region: [L | X | M (n x n each) | inverted diagonal blocks |
// y | z | a | loss | info | panel work], 256-byte aligned pieces.
struct GpLayout {
    std::size_t L, X, M, dinv, y, z, a, loss, info, panel, bytes;
};

GpLayout gp_layout(int n)
{
    GpLayout l{};
    const std::size_t nn = (std::size_t)n * n, blocks = (std::size_t)(n + kIB - 1) / kIB;
    std::size_t o = 0;
    auto put = [&](std::size_t& at, std::size_t b) {
        at = o;
        o += (b + 255) / 256 * 256;
    };
    put(l.L, nn * sizeof(double));
    put(l.X, nn * sizeof(double));
    put(l.M, nn * sizeof(double));
    put(l.dinv, blocks * kIB * kIB * sizeof(double));
    put(l.y, n * sizeof(double));
    put(l.z, n * sizeof(double));
    put(l.a, n * sizeof(double));
    put(l.loss, 2 * sizeof(double));
    put(l.info, sizeof(int));
    put(l.panel, kern::panel_work_bytes<double>(n));
    l.bytes = o;
    return l;
}

void check_args(const DdlaHandle_t& h, const char* who, int batch, int n, const void* work)
{
    require_gpu(h, who);
    if (batch < 0 || batch > 65535) throw std::invalid_argument(std::string(who) + ": batch must be in [0, 65535]");
    if (n < 1) throw std::invalid_argument(std::string(who) + ": the order must be positive");
    // the merges of L^-1 launch every problem's pairs of a level at once (grid z <= 65535)
    if ((long long)batch * std::max(1, n / (2 * kIB)) > 65535)
        throw std::invalid_argument(std::string(who) + ": batch * (n / 128) must not exceed 65535");
    if (batch > 0 && (work == nullptr || (uintptr_t)work % 256 != 0))
        throw std::invalid_argument(std::string(who) + ": work must be a 256-byte aligned device buffer");
}

// Chunks of a merge of order-h blocks: X11 and X22 are lower triangular,
// so T = L21 X11 is formed in `chunks` column slices (slice c needs rows c w..
// of X11) and X21 = -X22 T in as many row slices (slice r needs columns ..
// (r + 1) w of X22): 2 h^3 (1 + 1 / chunks) flops instead of 4 h^3.
int merge_chunks(long long h) { return h >= 1024 ? 4 : h >= 256 ? 2 : 1; }

// X = L^-1 for every problem, X holding the inverted 64 x 64 diagonal blocks
// on entry: at level h (64, 128, ...) each aligned pair of order-h blocks
// becomes one of order 2h by X21 = -X22 (L21 X11), T = L21 X11 staged in M;
// the pairs of a level are equally spaced (2 h (n + 1) elements apart), so
// one launch per slice takes them all, plus the products of a last, shorter
// pair.
void lower_inverse_batched(hipStream_t s, int n, const double* L, double* X, double* M, const kern::Batch& b)
{
    const long long ld = n;
    for (long long h = kIB; h < n; h *= 2) {
        const int hh = (int)h;
        const int full = (int)(n / (2 * h));
        if (full > 0) {
            const kern::Batch bi{b.count, b.stride, full, 2 * h * (ld + 1) * (long long)sizeof(double)};
            const int S = merge_chunks(h);
            const long long w = h / S;
            for (int c = 0; c < S; ++c) {  // T(:, c w..) = L21(:, c w..) X11(c w.., c w..)
                const long long k0 = c * w;
                kern::gemm<double>(s, 'N', 'N', hh, (int)w, (int)(h - k0), 1.0, L + h + k0 * ld, ld,
                                   X + k0 + k0 * ld, ld, 0.0, M + h + k0 * ld, ld, kern::kTriNone, 0, nullptr, false,
                                   bi);
            }
            for (int r = 0; r < S; ++r) {  // X21(r w.., :) = -X22(r w.., ..(r + 1) w) T(..(r + 1) w, :)
                const long long r0 = r * w;
                kern::gemm<double>(s, 'N', 'N', (int)w, hh, (int)(r0 + w), -1.0, X + h + r0 + h * ld, ld, M + h, ld,
                                   0.0, X + h + r0, ld, kern::kTriNone, 0, nullptr, false, bi);
            }
        }
        const long long o = (long long)full * 2 * h;
        if (o + h < n) {
            const int n2 = (int)(n - o - h);
            kern::gemm<double>(s, 'N', 'N', n2, hh, hh, 1.0, L + (o + h) + o * ld, ld, X + o + o * ld, ld, 0.0,
                               M + (o + h) + o * ld, ld, kern::kTriNone, 0, nullptr, false, b);
            kern::gemm<double>(s, 'N', 'N', n2, hh, n2, -1.0, X + (o + h) + (o + h) * ld, ld, M + (o + h) + o * ld,
                               ld, 0.0, X + (o + h) + o * ld, ld, kern::kTriNone, 0, nullptr, false, b);
        }
    }
}

// flops of lower_inverse_batched for one problem
double lower_inverse_flops(int n)
{
    double f = 0;
    for (long long h = kIB; h < n; h *= 2) {
        const long long full = n / (2 * h), o = full * 2 * h;
        f += full * 2.0 * h * h * h * (1.0 + 1.0 / merge_chunks(h));
        if (o + h < n) {
            const double n2 = (double)(n - o - h);
            f += 2.0 * n2 * h * h + 2.0 * n2 * h * n2;
        }
    }
    return f;
}

} // namespace

std::size_t gpBatchedWorkspaceSize(int n)
{
    if (n < 1) throw std::invalid_argument("gpBatchedWorkspaceSize: the order must be positive");
    return gp_layout(n).bytes;
}

void gpBatchedForward(const DdlaHandle_t& h, int batch, int n, const double* K, long long k_stride, const double* y,
                      long long y_stride, void* work, double* loss, int* info)
{
    check_args(h, "gpBatchedForward", batch, n, work);
    if (batch == 0) return;
    if (!K || !y || !loss || !info) throw std::invalid_argument("gpBatchedForward: null device array");
    if ((batch > 1 && k_stride < (long long)n * n) || (batch > 1 && y_stride < n))
        throw std::invalid_argument("gpBatchedForward: problems overlap (stride below the problem size)");

    const GpLayout l = gp_layout(n);
    const kern::Batch b{batch, (long long)l.bytes};
    char* base = static_cast<char*>(work);
    double* L = reinterpret_cast<double*>(base + l.L);
    double* X = reinterpret_cast<double*>(base + l.X);
    double* M = reinterpret_cast<double*>(base + l.M);
    double* dinv = reinterpret_cast<double*>(base + l.dinv);
    double* yb = reinterpret_cast<double*>(base + l.y);
    double* z = reinterpret_cast<double*>(base + l.z);
    double* slot = reinterpret_cast<double*>(base + l.loss);
    int* d_info = reinterpret_cast<int*>(base + l.info);
    void* panel = base + l.panel;
    hipStream_t s = h->stream;
    kern::gp_load(s, batch, n, K, k_stride, y, y_stride, L, yb, d_info, (long long)l.bytes);
    potrf_batched_dispatch<double>(h, n, L, n, d_info, panel, b);
    {
        LaunchTimer timer(h, kClassOther, lower_inverse_flops(n) * batch);
        kern::trtri_diag_blocks<double>(s, false, n, L, n, dinv, false, b);
        kern::gp_place_diag(s, batch, n, dinv, X, (long long)l.bytes);
        lower_inverse_batched(s, n, L, X, M, b);
    }
    kern::gemm<double>(s, 'N', 'N', n, 1, n, 1.0, X, n, yb, n, 0.0, z, n, kern::kTriNone, 0, nullptr, false, b);
    kern::gp_loss(s, batch, n, L, z, d_info, slot, (long long)l.bytes, loss, info);
    // each problem's panel wait error word, gathered at the region stride
    std::vector<int> err(batch, 0);
    DDLA_HIP(hipMemcpy2DAsync(err.data(), sizeof(int), kern::panel_error_word<double>(n, panel), l.bytes, sizeof(int),
                              batch, hipMemcpyDeviceToHost, s));
    DDLA_HIP(hipStreamSynchronize(s));
    for (int e : err)
        if (e != 0) throw std::runtime_error("gpBatchedForward: a panel factorization wait exceeded its limit");
}

void gpBatchedBackward(const DdlaHandle_t& h, int batch, int n, void* work, const double* g, double* Kbar,
                       long long kbar_stride, double* ybar, long long ybar_stride)
{
    check_args(h, "gpBatchedBackward", batch, n, work);
    if (batch == 0) return;
    if (!g || !Kbar) throw std::invalid_argument("gpBatchedBackward: null device array");
    if ((batch > 1 && kbar_stride < (long long)n * n) || (ybar && batch > 1 && ybar_stride < n))
        throw std::invalid_argument("gpBatchedBackward: problems overlap (stride below the problem size)");
    const GpLayout l = gp_layout(n);
    const kern::Batch b{batch, (long long)l.bytes};
    char* base = static_cast<char*>(work);
    double* X = reinterpret_cast<double*>(base + l.X);
    double* M = reinterpret_cast<double*>(base + l.M);
    double* z = reinterpret_cast<double*>(base + l.z);
    double* a = reinterpret_cast<double*>(base + l.a);
    hipStream_t s = h->stream;
    {
        LaunchTimer timer(h, kClassOther, lower_gram_flops(n) * batch);
        lower_gram_batched(s, n, X, M, b);  // M = tril(X^T X) = tril(K^-1)
    }
    kern::gemm<double>(s, 'T', 'N', n, 1, n, 1.0, X, n, z, n, 0.0, a, n, kern::kTriNone, 0, nullptr, false, b);
    kern::gp_grad(s, batch, n, M, a, (long long)l.bytes, g, Kbar, kbar_stride, ybar, ybar_stride);
}

} // namespace ddla
