// Cholesky path on one GPU: ppotrf / ppotrs / pposv (check / workspace size /
// run and the public drivers), local potrf, diag_logdet and gemmVbatched.
//
// Reference behaviour followed (1 x 1 grid):
//   ppotrf   src/Cholesky/ppotrf.cpp:21-54, :326-349
//   ppotrs   src/Cholesky/ppotrs.cpp:14-85 (trans mapping :44-46)
//   pposv    src/Cholesky/pposv.cpp:15-63 (solve only if info == 0)
//   sizes    src/Cholesky/cholesky_workspace.h:133-154, :207-249
//   potrf    src/Cholesky/potrf.cpp:123-193
//   logdet   src/diag_logdet.cpp:42-55, src/diag_logdet_impl.h:22-113
//   vbatched src/BLAS/gemm_vbatched.cpp:74-160, :350-412
// The arithmetic is this project's: the reference calls a vendor ?potrf /
// ?trsm here; this file schedules the hand-written kernels of kernels.hip in
// the blocked right-looking order of src/Cholesky/ppotrf.cpp:108-322.
#include "internal.h"
#include "kernels.h"
#include "zkernels.h"

#include <atomic>
#include <cmath>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <type_traits>

namespace ddla {

namespace {

using kern::kTriLower;
using kern::kTriNone;

constexpr int kIB = 64;  // diagonal block of potf2 and of the solves

// Outer panel width of the factorization: the depth of the trailing-update
// GEMM.  Not the descriptor's NB -- at 1 x 1 the block size of the
// distribution does not change the arithmetic (the reference hands the whole
// matrix to one ?potrf call).  Deeper updates amortize each tile's C
// read-modify-write better, wider panels lengthen the panel chain: FP64 at
// n = 32768 ran 209.3 / 205.0 / 206.4 ms with 512 / 768 / 1024, while at
// n = 4096 (panel-bound) 512 is best.  The split-product path keeps 512 (its
// panel buffers hold 512 columns).
// A process-wide override (set_panel_width; 0 = the rule above) lets tests
// run the 768-column panels at small orders.
std::atomic<int> g_panel_width{0};
template <class T>
int outer_block(int n, bool split)
{
    const int w = g_panel_width.load(std::memory_order_relaxed);
    if (w > 0) return split ? std::min(w, 512) : w;
    return std::is_same<T, double>::value && n >= 16384 ? 768 : 512;
}

// The blocked factorization always runs on the lower triangle (uplo 'U' is
// brought there by potrf_blocked): element (i, j), i >= j, of L is
// A[i + j*lda].  The helpers below translate operations on L into kernel
// calls.
template <class T>
struct LView {
    T* A;
    long long lda;
    T* at(long long r, long long c) const { return A + r + c * lda; }  // stored (r, c)
};

// C_L(ci.., cj..) (M x N) := alpha * A_L(ai.., ak..) * B_L(bj.., bk..)^T + beta * C_L,
// writing only the lower triangle of the global L view when mask is set.
template <class T>
void update_lower(DdlaStream* h, hipStream_t st, int cls, const LView<T>& v, int M, int N, int K, T alpha,
                  long long ai, long long ak, long long bj, long long bk, T beta, long long ci, long long cj,
                  bool mask, const int* abort, double flops, const kern::Batch& batch = kern::Batch())
{
    if (M <= 0 || N <= 0) return;
    LaunchTimer timer(h, cls, flops * batch.count, 0, st);
    const bool lookahead = cls == kClassLookahead;
    kern::gemm<T>(st, 'N', 'T', M, N, K, alpha, v.at(ai, ak), v.lda, v.at(bj, bk), v.lda, beta, v.at(ci, cj), v.lda,
                  mask ? kTriLower : kTriNone, ci - cj, abort, lookahead, batch);
}

// Factors the panel of columns [k, k + kb) of the order-n L view (its
// diagonal block and everything below it), assuming every earlier trailing
// update has been applied to these columns: one persistent launch.
template <class T>
void factor_panel(DdlaStream* h, hipStream_t st, const LView<T>& v, int n, int k, int kb, int* info, void* work,
                  const kern::Batch& batch = kern::Batch())
{
    const double below = (double)n - k - kb;
    LaunchTimer timer(h, kClassPanel, ((double)kb * kb * kb / 3.0 + below * kb * kb) * batch.count, 0, st);
    kern::panel_factor<T>(st, false, n, v.A, v.lda, k, kb, info, work, batch);
}

// Split-product trailing updates (FP32 factorization only): after panel p is
// factored, its rows below the diagonal block are split into bfloat16 or
// half-precision planes in buf[p % 3] (split_buffer_bytes(n, mode) bytes
// each), and every
// trailing update of step p reads that buffer.  Three buffers: the split of
// panel p + 3 waits (through the look-ahead chain) for step p's update of the
// rest, the last reader of buf[p % 3].
struct SplitPlan {
    int mode;  // 1 bf16x3, 2 bf16x6, 3 fp16x3
    void* buf[3];
};

// C_L(r + a_row.., r + b_row..) -= P P^T over the split panel of step k (rows
// r.. of the L view split into buf).
template <class T>
void split_update_lower(DdlaStream* h, hipStream_t st, int cls, const LView<T>& v, const SplitPlan& sp,
                        const void* buf, int m_split, int M, int N, int kb, long long r, int a_row, int b_row,
                        const int* abort, double flops)
{
    if (M <= 0 || N <= 0) return;
    LaunchTimer timer(h, cls, flops, 0, st);
    if constexpr (std::is_same<T, float>::value) {
        kern::split_update(st, sp.mode, false, M, N, kb, buf, m_split, a_row, b_row, v.at(r + a_row, r + b_row),
                           v.lda, true, (long long)a_row - b_row, abort);
    } else {
        throw std::logic_error("split-product updates apply to the FP32 factorization");
    }
}

} // namespace

namespace sched {

// The stream schedule of potrf_blocked for order n and panel width NB: panel
// p holds columns [p NB, (p + 1) NB) (clipped to n).  After panel 0 is
// factored, step p (panel p factored) applies panel p to the columns of
// panel p + 1 on the look stream, which the side stream then factors; to
// panel p + 2's columns on the look stream, after step p - 1's update of the
// rest (which also wrote them); and to every column from panel p + 3 on the
// main stream.  Three events order the streams; the schedule ends with all
// work ordered on the main stream.
std::vector<Step> potrf_schedule(int n, int NB)
{
    std::vector<Step> s;
    if (n <= 0 || NB <= 0) return s;
    const int np = (n + NB - 1) / NB;
    auto col = [&](int q) { return std::min<long long>(n, (long long)q * NB); };
    auto launch = [&](int op, int stream, int a, int b = 0, int c = 0) { s.push_back(Step{op, stream, a, b, c}); };
    auto update = [&](int stream, int p, long long lo, long long hi) {
        launch(kUpdate, stream, p, (int)lo, (int)hi);
    };
    launch(kReset, kMain, 0);
    launch(kFactor, kMain, 0);
    launch(kRecord, kMain, kEvPanel);
    for (int p = 0; p + 1 < np; ++p) {
        launch(kWait, kLook, kEvPanel);
        launch(kWait, kMain, kEvPanel);
        update(kLook, p, col(p + 1), col(p + 2));
        launch(kRecord, kLook, kEvLook);
        launch(kWait, kSide, kEvLook);
        launch(kFactor, kSide, p + 1);
        launch(kRecord, kSide, kEvPanel);
        if (p + 2 < np) {
            if (p > 0) launch(kWait, kLook, kEvRest);
            update(kLook, p, col(p + 2), col(p + 3));
        }
        update(kMain, p, col(p + 3), n);
        launch(kRecord, kMain, kEvRest);
    }
    launch(kRecord, kLook, kEvLook);
    launch(kWait, kMain, kEvLook);
    launch(kWait, kMain, kEvPanel);
    return s;
}

} // namespace sched

int potrf_schedule_steps(int n, int panel_width, int* steps, int capacity)
{
    const int NB = panel_width > 0 ? panel_width : outer_block<double>(n, false);
    const std::vector<sched::Step> s = sched::potrf_schedule(n, NB);
    const int count = (int)s.size();
    for (int i = 0; i < count && i < capacity; ++i) {
        const sched::Step& t = s[i];
        const int v[5] = {t.op, t.stream, t.a, t.b, t.c};
        std::copy(v, v + 5, steps + 5 * i);
    }
    return count;
}

namespace {

// Blocked right-looking Cholesky of the order-n L view with a look-ahead of
// two panels; *info (device) must be zero on entry and receives the 1-based
// index of the first failed pivot; work holds kern::panel_work_bytes<T>(n)
// bytes.  Step p (panel p factored) applies panel p to
//   * the columns of panel p + 1, then panel p + 2, on the `look` stream;
//     panel p + 1 is factored on the `side` stream as soon as its columns are
//     done, and panel p + 2's update waits for the previous step's update of
//     the rest (which also wrote those columns);
//   * every column beyond panel p + 2 on the handle's stream, right after the
//     previous step's.
// So the handle's stream runs the large updates back to back while the next
// panels are updated and factored beside them; concurrent launches write
// disjoint columns and read only finished panels.  Returns with all work
// ordered on the handle's stream.  split (FP32 only, optional): the trailing
// updates run as split products.  batch (uplo L, no split): batch.count
// independent problems of order n at a common stride (A, info and work of
// problem b sit b * batch.stride bytes after problem 0's), factored in step
// with each launch serving all of them.
template <class T>
void potrf_blocked(DdlaStream* h, bool upper, int n, T* A, long long lda, int* info, void* work,
                   const SplitPlan* split = nullptr, const kern::Batch& batch = kern::Batch())
{
    if (batch.count != 1 && (upper || split != nullptr || batch.inner != 1))
        throw std::logic_error("potrf_blocked: a batch of problems is factored in the lower view, without splits");
    if (upper) {
        // The factorization's kernels are tuned for the lower view; run in
        // the upper view (transposed operand access throughout) the same
        // factorization took about twice as long.  So the two strict
        // triangles are exchanged in place, the lower one (now A's upper
        // triangle transposed) is factored, and the exchange is undone: the
        // factor lands in the upper triangle and the caller's strict lower
        // triangle is restored bit for bit (two HBM passes over the matrix).
        const double bytes = 2.0 * n * (n - 1) * sizeof(T);
        {
            LaunchTimer timer(h, kClassOther, 0, bytes);
            kern::swap_triangles<T>(h->stream, n, A, lda);
        }
        potrf_blocked<T>(h, false, n, A, lda, info, work, split);
        LaunchTimer timer(h, kClassOther, 0, bytes);
        kern::swap_triangles<T>(h->stream, n, A, lda);
        return;
    }
    const LView<T> v{A, lda};
    const int NB = outer_block<T>(n, split != nullptr);
    const int np = (n + NB - 1) / NB;
    hipStream_t main = h->stream, side = h->side, look = h->look;
    auto col = [&](int q) { return std::min(n, q * NB); };  // first column of panel q (n past the end)
    // factors panel q and splits its rows below the diagonal block into buf[q % 3]
    auto factor = [&](hipStream_t st, int q) {
        const int k = col(q), kb = col(q + 1) - k;
        factor_panel<T>(h, st, v, n, k, kb, info, work, batch);
        const long long r = (long long)k + kb;
        if (!split || r >= n) return;
        if constexpr (std::is_same<T, float>::value)
            kern::split_panel(st, split->mode, false, n - (int)r, kb, v.at(r, k), lda, split->buf[q % 3]);
    };
    // C_L(i, j) -= sum over panel p's columns of L(i, .) L(j, .) for columns
    // c_lo <= j < c_hi and rows i >= j (lower trapezoid), on stream st
    auto update = [&](hipStream_t st, int p, int c_lo, int c_hi) {
        const int k = col(p), kb = col(p + 1) - k;
        const long long r = (long long)k + kb;
        const int M = n - c_lo, N = c_hi - c_lo;
        if (M <= 0 || N <= 0) return;
        const double flops = (double)N * (N + 1) * kb + 2.0 * (M - N) * N * kb;
        const int cls = st == main ? kClassTrailingUpdate : kClassLookahead;
        if (split)
            split_update_lower<T>(h, st, cls, v, *split, split->buf[p % 3], n - (int)r, M, N, kb, r, c_lo - (int)r,
                                  c_lo - (int)r, info, flops);
        else
            update_lower<T>(h, st, cls, v, M, N, kb, T(-1), c_lo, k, c_lo, k, T(1), c_lo, c_lo, true, info, flops,
                            batch);
    };
    (void)np;
    const hipStream_t streams[3] = {main, side, look};
    const hipEvent_t events[3] = {h->ev_panel, h->ev_look, h->ev_rest};
    for (const sched::Step& st : sched::potrf_schedule(n, NB)) {
        switch (st.op) {
        case sched::kReset: kern::panel_reset<T>(streams[st.stream], n, work, batch); break;
        case sched::kFactor: factor(streams[st.stream], st.a); break;
        case sched::kUpdate: update(streams[st.stream], st.a, st.b, st.c); break;
        case sched::kRecord: DDLA_HIP(hipEventRecord(events[st.a], streams[st.stream])); break;
        case sched::kWait: DDLA_HIP(hipStreamWaitEvent(streams[st.stream], events[st.a], 0)); break;
        }
    }
}

// Left solve A X = B with A = L L^T given by its L view; B is n x nrhs: the
// forward and the backward sweep, one launch each.  stage_work (optional):
// the staged solve's work (kern::stage_work_bytes), used when
// kern::stage_solve(n, nrhs).
template <class T>
void solve_left(DdlaStream* h, bool upper, int n, int nrhs, const T* F, long long ldf, T* B, long long ldb,
                const T* dinv, void* sync, int* error, void* stage_work)
{
    if (stage_work != nullptr && kern::stage_solve(n, nrhs)) {
        {
            LaunchTimer timer(h, kClassTrtri, 2.0 * n * 512.0 * 512.0 + (double)n * 512.0 * 512.0 / 3.0);
            kern::stage_prepare<T>(h->stream, upper, n, F, ldf, dinv, stage_work);
        }
        LaunchTimer timer(h, kClassSolve, 2.0 * n * n * nrhs, (double)n * n * sizeof(T));
        kern::stage_sweeps<T>(h->stream, upper, n, nrhs, F, ldf, B, ldb, stage_work, error);
        return;
    }
    LaunchTimer timer(h, kClassSolve, 2.0 * n * n * nrhs, (double)n * n * sizeof(T));
    kern::potrs_sweeps<T>(h->stream, upper, n, nrhs, F, ldf, dinv, B, ldb, sync, error);
}

std::size_t diag_inverse_elems(int n) { return (std::size_t)((n + kIB - 1) / kIB) * kIB * kIB; }

void bad_workspace(const char* who)
{
    throw std::invalid_argument(std::string(who) + ": backend workspace smaller than the workspace query");
}

char upper_char(char c) { return (c >= 'a' && c <= 'z') ? char(c - 'a' + 'A') : c; }

} // namespace

int set_panel_width(int cols)
{
    const int w = cols <= 0 ? 0 : std::min(768, std::max(64, cols / 64 * 64));
    return g_panel_width.exchange(w, std::memory_order_relaxed);
}

namespace cholesky_impl {

void ppotrf_check(const DdlaHandle_t& h, char uplo, int n, const void* A, int ia, int ja, const int* d)
{
    check_desc(d, h);
    require_gpu(h, "ppotrf");
    if ((uplo != 'L' && uplo != 'U') || n < 0 || n > d[DDLA_M_] || n > d[DDLA_N_] || d[DDLA_MB_] != d[DDLA_NB_] ||
        ia != 1 || ja != 1)
        throw std::invalid_argument("ppotrf: invalid triangle, dimensions, blocks or offsets");
    if (n && !A) throw std::invalid_argument("ppotrf: null local matrix");
}

template <class T>
void ppotrf_workspace_size(const DdlaHandle_t&, char, int n, T*, const int*, std::size_t* backend_bytes,
                           std::size_t* host_bytes)
{
    *backend_bytes = *host_bytes = 0;
    if (n <= 0) return;
    if constexpr (is_complex<T>::value) {
        *backend_bytes = complex_impl::ppotrf_bytes<typename real_type<T>::type>(n);
    } else {
        // [ info word | panel work: diagonal-block inverses, panel counters, error word ]
        *backend_bytes = detail::padded_bytes<int>(1) + detail::padded_bytes<char>(kern::panel_work_bytes<T>(n));
    }
}

template <class T>
void ppotrf_run(const DdlaHandle_t& h, char uplo, int n, T* A, int, int, const int* d, int& info, void* backend_work,
                std::size_t backend_bytes, void*, std::size_t)
{
    if (n == 0) {
        info = 0;
        return;
    }
    if constexpr (is_complex<T>::value) {
        complex_impl::ppotrf_run<typename real_type<T>::type>(h, uplo, n, A, d, info, backend_work, backend_bytes);
        return;
    } else {
        std::size_t need = 0, host = 0;
        ppotrf_workspace_size<T>(h, uplo, n, A, d, &need, &host);
        if (backend_bytes < need || backend_work == nullptr) bad_workspace("ppotrf");
        char* w = static_cast<char*>(backend_work);
        int* d_info = detail::take<int>(w, 1);
        char* panel = detail::take<char>(w, kern::panel_work_bytes<T>(n));
        fill_int(h->stream, d_info, 0);
        potrf_blocked<T>(h, uplo == 'U', n, A, d[DDLA_LLD_], d_info, panel);
        int error = 0;
        DDLA_HIP(hipMemcpyAsync(&info, d_info, sizeof(int), hipMemcpyDeviceToHost, h->stream));
        DDLA_HIP(hipMemcpyAsync(&error, kern::panel_error_word<T>(n, panel), sizeof(int), hipMemcpyDeviceToHost,
                                h->stream));
        DDLA_HIP(hipStreamSynchronize(h->stream));
        if (error != 0) throw std::runtime_error("ppotrf: a panel factorization wait exceeded its limit");
    }
}

void ppotrs_check(const DdlaHandle_t& h, char side, char uplo, char trans, int n, int nrhs, const void* A,
                  const int* da, const void* B, const int* db)
{
    check_desc(da, h);
    check_desc(db, h);
    require_gpu(h, "ppotrs");
    if ((side != 'L' && side != 'R') || (uplo != 'L' && uplo != 'U') || (trans != 'N' && trans != 'C') || n < 0 ||
        nrhs < 0 || n > da[DDLA_M_] || n > da[DDLA_N_] || da[DDLA_MB_] != da[DDLA_NB_])
        throw std::invalid_argument("ppotrs: invalid flags, dimensions or factor blocks");
    if (db[DDLA_MB_] != da[DDLA_MB_] || (side == 'L' && db[DDLA_RSRC_] != da[DDLA_RSRC_]) ||
        (side == 'R' && (db[DDLA_NB_] != da[DDLA_NB_] || db[DDLA_CSRC_] != da[DDLA_CSRC_])))
        throw std::invalid_argument("ppotrs: B must be distributed like A along the solved dimension");
    const int rows = side == 'L' ? n : nrhs, cols = side == 'L' ? nrhs : n;
    if (rows > db[DDLA_M_] || cols > db[DDLA_N_]) throw std::invalid_argument("ppotrs: invalid right-hand side");
    if (n && !A) throw std::invalid_argument("ppotrs: null local factor");
    if (rows && cols && !B) throw std::invalid_argument("ppotrs: null local right-hand side");
}

// Backend bytes of a ppotrs run; staged: with the staged solve's work when
// the order and right-hand sides take that route.
template <class T>
std::size_t ppotrs_bytes(char side, int n, int nrhs, bool staged)
{
    if (n == 0 || nrhs == 0) return 0;
    // [ inverses of the 64 x 64 diagonal blocks | sweep counters | error word | staged-solve work |
    //   transposed B (side 'R') ]
    std::size_t b = detail::padded_bytes<T>(diag_inverse_elems(n)) + kern::solve_work_bytes(nrhs) +
                    detail::padded_bytes<int>(1);
    if (staged) b += detail::padded_bytes<char>(kern::stage_work_bytes(n, nrhs, sizeof(T)));
    if (side == 'R') b += detail::padded_bytes<T>((std::size_t)n * nrhs);
    return b;
}

template <class T>
void ppotrs_workspace_size(const DdlaHandle_t&, char side, int n, int nrhs, const int*, const int*,
                           std::size_t* backend_bytes, std::size_t* host_bytes)
{
    *host_bytes = 0;
    if constexpr (is_complex<T>::value)
        *backend_bytes = complex_impl::ppotrs_bytes<typename real_type<T>::type>(side, n, nrhs);
    else
        *backend_bytes = ppotrs_bytes<T>(side, n, nrhs, true);
}

template <class T>
void ppotrs_exec(const DdlaHandle_t& h, char side, char uplo, int n, int nrhs, const T* A, const int* da, T* B,
                 const int* db, void* backend_work, std::size_t backend_bytes, bool staged)
{
    if (n == 0 || nrhs == 0) return;
    if (backend_bytes < ppotrs_bytes<T>(side, n, nrhs, staged) || backend_work == nullptr) bad_workspace("ppotrs");
    const bool upper = uplo == 'U';
    const long long lda = da[DDLA_LLD_], ldb = db[DDLA_LLD_];
    char* w = static_cast<char*>(backend_work);
    T* dinv = detail::take<T>(w, diag_inverse_elems(n));
    void* sync = w;
    w += kern::solve_work_bytes(nrhs);
    int* error = detail::take<int>(w, 1);
    void* stage_work = nullptr;
    if (staged && kern::stage_solve(n, nrhs))
        stage_work = detail::take<char>(w, kern::stage_work_bytes(n, nrhs, sizeof(T)));
    fill_int(h->stream, error, 0);
    {
        LaunchTimer timer(h, kClassTrtri, (double)n * kIB * kIB / 3.0);
        kern::trtri_diag_blocks<T>(h->stream, upper, n, A, lda, dinv);
    }
    // A = L L^T is symmetric: trans 'N' and 'C' solve the same system.
    // side 'R' (X A = B, B nrhs x n) is A X^T = B^T.
    if (side == 'L') {
        solve_left<T>(h, upper, n, nrhs, A, lda, B, ldb, dinv, sync, error, stage_work);
    } else {
        T* bt = detail::take<T>(w, (std::size_t)n * nrhs);
        kern::transpose<T>(h->stream, nrhs, n, B, ldb, bt, n);
        solve_left<T>(h, upper, n, nrhs, A, lda, bt, n, dinv, sync, error, stage_work);
        kern::transpose<T>(h->stream, n, nrhs, bt, n, B, ldb);
    }
    int host_error = 0;
    DDLA_HIP(hipMemcpyAsync(&host_error, error, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    DDLA_HIP(hipStreamSynchronize(h->stream));
    if (host_error != 0) throw std::runtime_error("ppotrs: a triangular-solve sweep exceeded its wait limit");
}

template <class T>
void ppotrs_run(const DdlaHandle_t& h, char side, char uplo, char, int n, int nrhs, const T* A, const int* da, T* B,
                const int* db, void* backend_work, std::size_t backend_bytes, void*, std::size_t)
{
    if constexpr (is_complex<T>::value)
        complex_impl::ppotrs_run<typename real_type<T>::type>(h, side, uplo, n, nrhs, A, da, B, db, backend_work,
                                                              backend_bytes);
    else
        ppotrs_exec<T>(h, side, uplo, n, nrhs, A, da, B, db, backend_work, backend_bytes, true);
}

void pposv_check(const DdlaHandle_t& h, char side, char uplo, char trans, int n, int nrhs, const void* A, int ia,
                 int ja, const int* da, const void* B, int ib, int jb, const int* db)
{
    ppotrf_check(h, uplo, n, A, ia, ja, da);
    ppotrs_check(h, side, uplo, trans, n, nrhs, A, da, B, db);
    if (ib != 1 || jb != 1) throw std::invalid_argument("pposv: offsets must be one");
}

template <class T>
void pposv_workspace_size(const DdlaHandle_t& h, char side, char uplo, int n, int nrhs, T* A, const int* da,
                          const int* db, std::size_t* backend_bytes, std::size_t* host_bytes)
{
    std::size_t sb = 0, sh = 0;
    ppotrf_workspace_size<T>(h, uplo, n, A, da, backend_bytes, host_bytes);
    ppotrs_workspace_size<T>(h, side, n, nrhs, da, db, &sb, &sh);
    *backend_bytes = std::max(*backend_bytes, sb);
    *host_bytes = std::max(*host_bytes, sh);
}

template <class T>
void pposv_run(const DdlaHandle_t& h, char side, char uplo, char trans, int n, int nrhs, T* A, int ia, int ja,
               const int* da, T* B, int, int, const int* db, int& info, void* backend_work,
               std::size_t backend_bytes, void* host_work, std::size_t host_bytes)
{
    if (n == 0) {
        info = 0;
        return;
    }
    ppotrf_run<T>(h, uplo, n, A, ia, ja, da, info, backend_work, backend_bytes, host_work, host_bytes);
    if (info == 0)
        ppotrs_run<T>(h, side, uplo, trans, n, nrhs, A, da, B, db, backend_work, backend_bytes, host_work,
                      host_bytes);
}

} // namespace cholesky_impl

template <class T>
std::size_t pposv_unstaged_bytes(char side, int n, int nrhs)
{
    std::size_t fb = 0, fh = 0;
    cholesky_impl::ppotrf_workspace_size<T>(nullptr, 'L', n, static_cast<T*>(nullptr), nullptr, &fb, &fh);
    return std::max(fb, cholesky_impl::ppotrs_bytes<T>(side, n, nrhs, false));
}

template <class T>
void pposv_run_unstaged(const DdlaHandle_t& h, char side, char uplo, int n, int nrhs, T* A, const int* da, T* B,
                        const int* db, int& info, void* backend_work, std::size_t backend_bytes)
{
    if (n == 0) {
        info = 0;
        return;
    }
    cholesky_impl::ppotrf_run<T>(h, uplo, n, A, 1, 1, da, info, backend_work, backend_bytes, nullptr, 0);
    if (info == 0)
        cholesky_impl::ppotrs_exec<T>(h, side, uplo, n, nrhs, A, da, B, db, backend_work, backend_bytes, false);
}

template std::size_t pposv_unstaged_bytes<double>(char, int, int);
template void pposv_run_unstaged<double>(const DdlaHandle_t&, char, char, int, int, double*, const int*, double*,
                                         const int*, int&, void*, std::size_t);


// ---------------------------------------------------------------------------
// Public drivers: check, size, scratch, run, synchronize.
// ---------------------------------------------------------------------------
template <class T>
void ppotrf(const DdlaHandle_t& h, char uplo, int n, T* A, int ia, int ja, const int* descA, int& info)
{
    cholesky_impl::ppotrf_check(h, uplo, n, A, ia, ja, descA);
    if (n == 0) {
        info = 0;
        return;
    }
    std::size_t bb = 0, hb = 0;
    cholesky_impl::ppotrf_workspace_size<T>(h, uplo, n, A, descA, &bb, &hb);
    void* work = h->ensure_scratch(bb);
    cholesky_impl::ppotrf_run<T>(h, uplo, n, A, ia, ja, descA, info, work, bb, nullptr, 0);
}

template <class T>
void ppotrs(const DdlaHandle_t& h, char side, char uplo, char trans, int n, int nrhs, const T* A, const int* descA,
            T* B, const int* descB)
{
    cholesky_impl::ppotrs_check(h, side, uplo, trans, n, nrhs, A, descA, B, descB);
    if (n == 0 || nrhs == 0) return;
    std::size_t bb = 0, hb = 0;
    cholesky_impl::ppotrs_workspace_size<T>(h, side, n, nrhs, descA, descB, &bb, &hb);
    void* work = h->ensure_scratch(bb);
    cholesky_impl::ppotrs_run<T>(h, side, uplo, trans, n, nrhs, A, descA, B, descB, work, bb, nullptr, 0);
    DDLA_CHECK(ddlaSynchronize(h));
}

template <class T>
void pposv(const DdlaHandle_t& h, char side, char uplo, char trans, int n, int nrhs, T* A, int ia, int ja,
           const int* descA, T* B, int ib, int jb, const int* descB, int& info)
{
    cholesky_impl::pposv_check(h, side, uplo, trans, n, nrhs, A, ia, ja, descA, B, ib, jb, descB);
    if (n == 0) {
        info = 0;
        return;
    }
    std::size_t bb = 0, hb = 0;
    cholesky_impl::pposv_workspace_size<T>(h, side, uplo, n, nrhs, A, descA, descB, &bb, &hb);
    void* work = h->ensure_scratch(bb);
    cholesky_impl::pposv_run<T>(h, side, uplo, trans, n, nrhs, A, ia, ja, descA, B, ib, jb, descB, info, work, bb,
                                nullptr, 0);
    DDLA_CHECK(ddlaSynchronize(h));
}

template <class T>
void potrf(const DdlaHandle_t& h, char uplo, int n, T* A, int lda, int* info)
{
    require_gpu(h, "potrf");
    const char u = upper_char(uplo);
    if (u != 'L' && u != 'U') throw std::invalid_argument("potrf: uplo must be 'U' or 'L'");
    if (n < 0) throw std::runtime_error("potrf: negative matrix size");
    if (info == nullptr) throw std::runtime_error("potrf: null info");
    if (n == 0) {
        fill_int(h->stream, info, 0);
        return;
    }
    if (A == nullptr) throw std::runtime_error("potrf: null A");
    if (lda < std::max(1, n)) throw std::runtime_error("potrf: lda too small");
    if constexpr (is_complex<T>::value) {
        complex_impl::potrf<typename real_type<T>::type>(h, u == 'U', n, A, lda, info);
    } else {
        void* work = h->ensure_scratch(kern::panel_work_bytes<T>(n));
        fill_int(h->stream, info, 0);
        potrf_blocked<T>(h, u == 'U', n, A, lda, info, work);
    }
}

template <class T>
void diag_logdet(const DdlaHandle_t& h, int n, const T* a, int lda, std::complex<typename real_type<T>::type>& result)
{
    using R = typename real_type<T>::type;
    require_gpu(h, "diag_logdet");
    if (n < 0 || lda < std::max(1, n) || (n > 0 && !a))
        throw std::invalid_argument("diag_logdet: invalid order, leading dimension or pointer");
    if (n > 0) {
        const std::uint64_t last = std::uint64_t(n - 1) * std::uint64_t(lda) + n - 1;
        if (last > std::uint64_t(std::numeric_limits<std::ptrdiff_t>::max()) / sizeof(T))
            throw std::overflow_error("diag_logdet: matrix address span is not representable");
    }
    if (n == 0) {
        result = {R(0), R(0)};
        return;
    }
    if constexpr (is_complex<T>::value) {
        complex_impl::diag_logdet<R>(h, n, a, lda, result);
        return;
    } else {
        const std::size_t part = detail::padded_bytes<char>(kern::logdet_work_bytes());
        char* w = static_cast<char*>(h->ensure_scratch(part + detail::padded_bytes<kern::LogdetResult>(1)));
        auto* d_out = reinterpret_cast<kern::LogdetResult*>(w + part);
        kern::diag_logdet<T>(h->stream, n, a, lda, w, d_out);
        kern::LogdetResult out{};
        DDLA_HIP(hipMemcpyAsync(&out, d_out, sizeof(out), hipMemcpyDeviceToHost, h->stream));
        DDLA_HIP(hipStreamSynchronize(h->stream));
        if (out.flags & 1) throw std::domain_error("diag_logdet: nonfinite diagonal");
        if (out.flags & 2) {
            result = {-std::numeric_limits<R>::infinity(), R(0)};
            return;
        }
        const double pi = std::acos(-1.0);
        double phase = std::remainder(((out.flags >> 2) & 1) * pi, 2 * pi);
        if (phase <= -pi) phase = pi;
        result = {static_cast<R>(out.magnitude), static_cast<R>(phase)};
    }
}

namespace {

bool valid_op(char op) { return op == 'N' || op == 'n' || op == 'T' || op == 't' || op == 'C' || op == 'c'; }
bool is_n_op(char op) { return op == 'N' || op == 'n'; }

void check_vbatched_handle(const DdlaHandle_t& h, const char* who)
{
    if (h == nullptr || !h->initialized || h->backend != DdlaBackend::GPU)
        throw std::invalid_argument(std::string(who) + ": invalid DDLA handle");
}

// Validates one stage's device dimension arrays (one reduction kernel and a
// host read of its 4-int summary) and returns the maxima of m, n, k.
void vbatched_maxima(const DdlaHandle_t& h, const char* who, char ta, char tb, const int* m, const int* n,
                     const int* k, const int* lda, const int* ldb, const int* ldc, int batch, int* summary, int out[3])
{
    if (!m || !n || !k || !lda || !ldb || !ldc) throw std::invalid_argument(std::string(who) + ": null dimension array");
    kern::vbatched_validate(h->stream, is_n_op(ta), is_n_op(tb), m, n, k, lda, ldb, ldc, batch, summary);
    int host[4] = {0, 0, 0, 0};
    DDLA_HIP(hipMemcpyAsync(host, summary, sizeof(host), hipMemcpyDeviceToHost, h->stream));
    DDLA_HIP(hipStreamSynchronize(h->stream));
    if (host[3] != 0)
        throw std::invalid_argument(std::string(who) + ": invalid dimensions at batch " + std::to_string(host[3] - 1));
    out[0] = host[0];
    out[1] = host[1];
    out[2] = host[2];
}

} // namespace


// pdiag_logdet at a 1 x 1 grid (src/PBLAS/pdiag_logdet.cpp:11-48): the
// descriptor checks, then diag_logdet on the local matrix with its LLD.
template <class T>
void pdiag_logdet(const DdlaHandle_t& h, int n, const T* a, const int* desc,
                  std::complex<typename real_type<T>::type>& result)
{
    require_gpu(h, "pdiag_logdet");
    check_desc(desc, h);
    if (n < 0 || n > desc[DDLA_M_] || n > desc[DDLA_N_])
        throw std::invalid_argument("pdiag_logdet: order exceeds descriptor");
    if (n > 0 && !a) throw std::invalid_argument("pdiag_logdet: null local matrix");
    diag_logdet<T>(h, n, a, desc[DDLA_LLD_], result);
}

template <class T>
void gemmVbatched(const DdlaHandle_t& h, char transA, char transB, int* d_m, int* d_n, int* d_k, T alpha,
                  const T* const* d_A, int* d_lda, const T* const* d_B, int* d_ldb, T beta, T** d_C, int* d_ldc,
                  int batch)
{
    check_vbatched_handle(h, "gemmVbatched");
    if (batch < 0) throw std::invalid_argument("gemmVbatched: negative batch count");
    if (batch == 0) return;
    if (d_A == nullptr || d_B == nullptr || d_C == nullptr)
        throw std::invalid_argument("gemmVbatched: null matrix pointer array");
    if (!valid_op(transA) || !valid_op(transB)) throw std::invalid_argument("gemmVbatched: invalid transpose operation");
    int* summary = static_cast<int*>(h->ensure_scratch(detail::padded_bytes<int>(4)));
    int mx[3];
    vbatched_maxima(h, "gemmVbatched", transA, transB, d_m, d_n, d_k, d_lda, d_ldb, d_ldc, batch, summary, mx);
    if constexpr (is_complex<T>::value)
        kern::zgemm_vbatched<typename real_type<T>::type>(
            h->stream, true, true, transA, transB, d_m, d_n, d_k, alpha, reinterpret_cast<const void* const*>(d_A),
            d_lda, reinterpret_cast<const void* const*>(d_B), d_ldb, beta, d_C, d_ldc, batch, mx[0], mx[1]);
    else
        kern::gemm_vbatched<T>(h->stream, transA, transB, d_m, d_n, d_k, alpha, d_A, d_lda, d_B, d_ldb, beta, d_C,
                               d_ldc, batch, mx[0], mx[1]);
}

// src/BLAS/gemm_vbatched.cpp:420-523 and
// src/vbatched/gemm_template_kernel_vbatched.hpp:439-497: per host segment, in
// order, stage 0 C0 = alpha0 op(A0) op(B0) (beta0 must be zero) and then
// stage 1 C1 = alpha1 op(C0) op(AB1) + beta1 C1 (C0_left) or
// alpha1 op(AB1) op(C0) + beta1 C1, with stage-1 leading dimensions lda1 for
// the left operand and ldb1 for the right one.  Segments may reuse C0
// storage: stream order runs each segment's two stages before the next.
template <class T>
void gemmVbatched2s(const DdlaHandle_t& h, char transA_0, char transB_0, int* d_m_0, int* d_n_0, int* d_k_0,
                    T alpha_0, const T* const* d_A_0, int* d_lda_0, const T* const* d_B_0, int* d_ldb_0, T beta_0,
                    T** d_C_0, int* d_ldc_0, char transA_1, char transB_1, int* d_m_1, int* d_n_1, int* d_k_1,
                    T alpha_1, const T* const* d_AB_1, int* d_lda_1, int* d_ldb_1, T beta_1, T** d_C_1,
                    int* d_ldc_1, bool C0_left, int batch, const int* segment_sizes, int segment_count)
{
    check_vbatched_handle(h, "gemmVbatched2s");
    if (batch < 0) throw std::invalid_argument("gemmVbatched2s: negative batch count");
    if (batch == 0) return;
    if (beta_0 != T(0)) throw std::invalid_argument("gemmVbatched2s: stage-zero beta must be zero");
    if (segment_sizes == nullptr || segment_count <= 0)
        throw std::invalid_argument("gemmVbatched2s: missing segment sizes");
    long long total = 0;
    for (int s = 0; s < segment_count; ++s) {
        if (segment_sizes[s] <= 0) throw std::invalid_argument("gemmVbatched2s: segment sizes must be positive");
        total += segment_sizes[s];
    }
    if (total != batch) throw std::invalid_argument("gemmVbatched2s: segment total does not match batch count");
    if (!d_A_0 || !d_B_0 || !d_C_0 || !d_AB_1 || !d_C_1)
        throw std::invalid_argument("gemmVbatched2s: null matrix pointer array");
    if (!valid_op(transA_0) || !valid_op(transB_0) || !valid_op(transA_1) || !valid_op(transB_1))
        throw std::invalid_argument("gemmVbatched2s: invalid transpose operation");
    int* summary = static_cast<int*>(h->ensure_scratch(detail::padded_bytes<int>(4)));
    int mx0[3], mx1[3];
    vbatched_maxima(h, "gemmVbatched2s", transA_0, transB_0, d_m_0, d_n_0, d_k_0, d_lda_0, d_ldb_0, d_ldc_0, batch,
                    summary, mx0);
    vbatched_maxima(h, "gemmVbatched2s", transA_1, transB_1, d_m_1, d_n_1, d_k_1, d_lda_1, d_ldb_1, d_ldc_1, batch,
                    summary, mx1);
    const T* const* C0 = const_cast<const T* const*>(d_C_0);
    for (int s = 0, c = 0; s < segment_count; c += segment_sizes[s], ++s) {
        const int nb = segment_sizes[s];
        const T* const* left = C0_left ? C0 + c : d_AB_1 + c;
        const T* const* right = C0_left ? d_AB_1 + c : C0 + c;
        if constexpr (is_complex<T>::value) {
            using R = typename real_type<T>::type;
            auto vp = [](const T* const* p) { return reinterpret_cast<const void* const*>(p); };
            kern::zgemm_vbatched<R>(h->stream, true, true, transA_0, transB_0, d_m_0 + c, d_n_0 + c, d_k_0 + c,
                                    alpha_0, vp(d_A_0 + c), d_lda_0 + c, vp(d_B_0 + c), d_ldb_0 + c, beta_0,
                                    d_C_0 + c, d_ldc_0 + c, nb, mx0[0], mx0[1]);
            kern::zgemm_vbatched<R>(h->stream, true, true, transA_1, transB_1, d_m_1 + c, d_n_1 + c, d_k_1 + c,
                                    alpha_1, vp(left), d_lda_1 + c, vp(right), d_ldb_1 + c, beta_1, d_C_1 + c,
                                    d_ldc_1 + c, nb, mx1[0], mx1[1]);
        } else {
            kern::gemm_vbatched<T>(h->stream, transA_0, transB_0, d_m_0 + c, d_n_0 + c, d_k_0 + c, alpha_0,
                                   d_A_0 + c, d_lda_0 + c, d_B_0 + c, d_ldb_0 + c, beta_0, d_C_0 + c, d_ldc_0 + c,
                                   nb, mx0[0], mx0[1]);
            kern::gemm_vbatched<T>(h->stream, transA_1, transB_1, d_m_1 + c, d_n_1 + c, d_k_1 + c, alpha_1, left,
                                   d_lda_1 + c, right, d_ldb_1 + c, beta_1, d_C_1 + c, d_ldc_1 + c, nb, mx1[0],
                                   mx1[1]);
        }
    }
}

// ---------------------------------------------------------------------------
// Mixed real/complex gemmVbatched and gemmVbatched2s (the reference's
// include/ddla/ddla.h:1878-1960): one operand real, output and scalars
// complex; 'C' on a real operand is 'T'.  Same validation and zero-product
// rules as the one-type overloads.
// ---------------------------------------------------------------------------
namespace {

template <class R>
void mixed_vbatched(const DdlaHandle_t& h, const char* who, bool a_complex, bool b_complex, char transA, char transB,
                    int* d_m, int* d_n, int* d_k, std::complex<R> alpha, const void* const* d_A, int* d_lda,
                    const void* const* d_B, int* d_ldb, std::complex<R> beta, std::complex<R>** d_C, int* d_ldc,
                    int batch)
{
    check_vbatched_handle(h, who);
    if (batch < 0) throw std::invalid_argument(std::string(who) + ": negative batch count");
    if (batch == 0) return;
    if (d_A == nullptr || d_B == nullptr || d_C == nullptr)
        throw std::invalid_argument(std::string(who) + ": null matrix pointer array");
    if (!valid_op(transA) || !valid_op(transB))
        throw std::invalid_argument(std::string(who) + ": invalid transpose operation");
    int* summary = static_cast<int*>(h->ensure_scratch(detail::padded_bytes<int>(4)));
    int mx[3];
    vbatched_maxima(h, who, transA, transB, d_m, d_n, d_k, d_lda, d_ldb, d_ldc, batch, summary, mx);
    kern::zgemm_vbatched<R>(h->stream, a_complex, b_complex, transA, transB, d_m, d_n, d_k, alpha, d_A, d_lda, d_B,
                            d_ldb, beta, d_C, d_ldc, batch, mx[0], mx[1]);
}

// Stage 0 with operand kinds (a0, b0), stage 1 with the complex intermediate
// C0 and an external operand AB1 of kind ab1 (complex or real).
template <class R>
void mixed_vbatched2s(const DdlaHandle_t& h, bool a0_complex, bool b0_complex, bool ab1_complex, char transA_0,
                      char transB_0, int* d_m_0, int* d_n_0, int* d_k_0, std::complex<R> alpha_0,
                      const void* const* d_A_0, int* d_lda_0, const void* const* d_B_0, int* d_ldb_0,
                      std::complex<R> beta_0, std::complex<R>** d_C_0, int* d_ldc_0, char transA_1, char transB_1,
                      int* d_m_1, int* d_n_1, int* d_k_1, std::complex<R> alpha_1, const void* const* d_AB_1,
                      int* d_lda_1, int* d_ldb_1, std::complex<R> beta_1, std::complex<R>** d_C_1, int* d_ldc_1,
                      bool C0_left, int batch, const int* segment_sizes, int segment_count)
{
    const char* who = "gemmVbatched2s";
    check_vbatched_handle(h, who);
    if (batch < 0) throw std::invalid_argument("gemmVbatched2s: negative batch count");
    if (batch == 0) return;
    if (beta_0 != std::complex<R>(0)) throw std::invalid_argument("gemmVbatched2s: stage-zero beta must be zero");
    if (segment_sizes == nullptr || segment_count <= 0)
        throw std::invalid_argument("gemmVbatched2s: missing segment sizes");
    long long total = 0;
    for (int s = 0; s < segment_count; ++s) {
        if (segment_sizes[s] <= 0) throw std::invalid_argument("gemmVbatched2s: segment sizes must be positive");
        total += segment_sizes[s];
    }
    if (total != batch) throw std::invalid_argument("gemmVbatched2s: segment total does not match batch count");
    if (!d_A_0 || !d_B_0 || !d_C_0 || !d_AB_1 || !d_C_1)
        throw std::invalid_argument("gemmVbatched2s: null matrix pointer array");
    if (!valid_op(transA_0) || !valid_op(transB_0) || !valid_op(transA_1) || !valid_op(transB_1))
        throw std::invalid_argument("gemmVbatched2s: invalid transpose operation");
    int* summary = static_cast<int*>(h->ensure_scratch(detail::padded_bytes<int>(4)));
    int mx0[3], mx1[3];
    vbatched_maxima(h, who, transA_0, transB_0, d_m_0, d_n_0, d_k_0, d_lda_0, d_ldb_0, d_ldc_0, batch, summary, mx0);
    vbatched_maxima(h, who, transA_1, transB_1, d_m_1, d_n_1, d_k_1, d_lda_1, d_ldb_1, d_ldc_1, batch, summary, mx1);
    const void* const* C0 = reinterpret_cast<const void* const*>(d_C_0);
    for (int s = 0, c = 0; s < segment_count; c += segment_sizes[s], ++s) {
        const int nb = segment_sizes[s];
        kern::zgemm_vbatched<R>(h->stream, a0_complex, b0_complex, transA_0, transB_0, d_m_0 + c, d_n_0 + c,
                                d_k_0 + c, alpha_0, d_A_0 + c, d_lda_0 + c, d_B_0 + c, d_ldb_0 + c, beta_0, d_C_0 + c,
                                d_ldc_0 + c, nb, mx0[0], mx0[1]);
        const void* const* left = C0_left ? C0 + c : d_AB_1 + c;
        const void* const* right = C0_left ? d_AB_1 + c : C0 + c;
        const bool left_complex = C0_left ? true : ab1_complex, right_complex = C0_left ? ab1_complex : true;
        kern::zgemm_vbatched<R>(h->stream, left_complex, right_complex, transA_1, transB_1, d_m_1 + c, d_n_1 + c,
                                d_k_1 + c, alpha_1, left, d_lda_1 + c, right, d_ldb_1 + c, beta_1, d_C_1 + c,
                                d_ldc_1 + c, nb, mx1[0], mx1[1]);
    }
}

template <class P>
const void* const* vpa(const P* const* p)
{
    return reinterpret_cast<const void* const*>(p);
}

} // namespace

template <typename R>
void gemmVbatched(const DdlaHandle_t& h, char transA, char transB, int* m, int* n, int* k, std::complex<R> alpha,
                  const R* const* A, int* lda, const std::complex<R>* const* B, int* ldb, std::complex<R> beta,
                  std::complex<R>** C, int* ldc, int batch_count)
{
    mixed_vbatched<R>(h, "gemmVbatched", false, true, transA, transB, m, n, k, alpha, vpa(A), lda, vpa(B), ldb, beta,
                      C, ldc, batch_count);
}

template <typename R>
void gemmVbatched(const DdlaHandle_t& h, char transA, char transB, int* m, int* n, int* k, std::complex<R> alpha,
                  const std::complex<R>* const* A, int* lda, const R* const* B, int* ldb, std::complex<R> beta,
                  std::complex<R>** C, int* ldc, int batch_count)
{
    mixed_vbatched<R>(h, "gemmVbatched", true, false, transA, transB, m, n, k, alpha, vpa(A), lda, vpa(B), ldb, beta,
                      C, ldc, batch_count);
}

#define DDLA_MIXED2S(A0T, B0T, AB1T, A0C, B0C, AB1C)                                                             \
    template <typename R>                                                                                        \
    void gemmVbatched2s(const DdlaHandle_t& h, char transA_0, char transB_0, int* m_0, int* n_0, int* k_0,      \
                        std::complex<R> alpha_0, const A0T* const* A_0, int* lda_0, const B0T* const* B_0,       \
                        int* ldb_0, std::complex<R> beta_0, std::complex<R>** C_0, int* ldc_0, char transA_1,   \
                        char transB_1, int* m_1, int* n_1, int* k_1, std::complex<R> alpha_1,                   \
                        const AB1T* const* AB_1, int* lda_1, int* ldb_1, std::complex<R> beta_1,                \
                        std::complex<R>** C_1, int* ldc_1, bool C0_left, int batch_count,                       \
                        const int* segment_sizes, int segment_count)                                             \
    {                                                                                                            \
        mixed_vbatched2s<R>(h, A0C, B0C, AB1C, transA_0, transB_0, m_0, n_0, k_0, alpha_0, vpa(A_0), lda_0,     \
                            vpa(B_0), ldb_0, beta_0, C_0, ldc_0, transA_1, transB_1, m_1, n_1, k_1, alpha_1,    \
                            vpa(AB_1), lda_1, ldb_1, beta_1, C_1, ldc_1, C0_left, batch_count, segment_sizes,   \
                            segment_count);                                                                      \
    }
DDLA_MIXED2S(R, std::complex<R>, R, false, true, false)
DDLA_MIXED2S(std::complex<R>, R, R, true, false, false)
DDLA_MIXED2S(R, std::complex<R>, std::complex<R>, false, true, true)
DDLA_MIXED2S(std::complex<R>, R, std::complex<R>, true, false, true)
#undef DDLA_MIXED2S

template <class T>
void potrf_blocked_dispatch(DdlaStream* h, bool upper, int n, T* A, long long lda, int* info, void* work)
{
    potrf_blocked<T>(h, upper, n, A, lda, info, work);
}
template void potrf_blocked_dispatch<float>(DdlaStream*, bool, int, float*, long long, int*, void*);
template void potrf_blocked_dispatch<double>(DdlaStream*, bool, int, double*, long long, int*, void*);

template <class T>
void potrf_batched_dispatch(DdlaStream* h, int n, T* A, long long lda, int* info, void* work, const kern::Batch& batch)
{
    potrf_blocked<T>(h, false, n, A, lda, info, work, nullptr, batch);
}
template void potrf_batched_dispatch<double>(DdlaStream*, int, double*, long long, int*, void*, const kern::Batch&);

void potrf_split_dispatch(DdlaStream* h, bool upper, int n, float* A, long long lda, int* info, void* work,
                          int mode, void* const* bufs)
{
    const SplitPlan plan{mode, {bufs[0], bufs[1], bufs[2]}};
    potrf_blocked<float>(h, upper, n, A, lda, info, work, &plan);
}

#define DDLA_INSTANTIATE(T)                                                                                    \
    template void ppotrf<T>(const DdlaHandle_t&, char, int, T*, int, int, const int*, int&);                  \
    template void ppotrs<T>(const DdlaHandle_t&, char, char, char, int, int, const T*, const int*, T*,         \
                            const int*);                                                                       \
    template void pposv<T>(const DdlaHandle_t&, char, char, char, int, int, T*, int, int, const int*, T*, int, \
                           int, const int*, int&);                                                             \
    template void potrf<T>(const DdlaHandle_t&, char, int, T*, int, int*);                                     \
    template void diag_logdet<T>(const DdlaHandle_t&, int, const T*, int, std::complex<real_type<T>::type>&);  \
    template void pdiag_logdet<T>(const DdlaHandle_t&, int, const T*, const int*,                             \
                                  std::complex<real_type<T>::type>&);                                         \
    template void gemmVbatched<T>(const DdlaHandle_t&, char, char, int*, int*, int*, T, const T* const*, int*, \
                                  const T* const*, int*, T, T**, int*, int);                                   \
    template void gemmVbatched2s<T>(const DdlaHandle_t&, char, char, int*, int*, int*, T, const T* const*,     \
                                    int*, const T* const*, int*, T, T**, int*, char, char, int*, int*, int*, T, \
                                    const T* const*, int*, int*, T, T**, int*, bool, int, const int*, int);    \
    template void cholesky_impl::ppotrf_workspace_size<T>(const DdlaHandle_t&, char, int, T*, const int*,      \
                                                          std::size_t*, std::size_t*);                         \
    template void cholesky_impl::ppotrf_run<T>(const DdlaHandle_t&, char, int, T*, int, int, const int*, int&, \
                                               void*, std::size_t, void*, std::size_t);                        \
    template void cholesky_impl::ppotrs_workspace_size<T>(const DdlaHandle_t&, char, int, int, const int*,     \
                                                          const int*, std::size_t*, std::size_t*);             \
    template void cholesky_impl::ppotrs_run<T>(const DdlaHandle_t&, char, char, char, int, int, const T*,      \
                                               const int*, T*, const int*, void*, std::size_t, void*,          \
                                               std::size_t);                                                   \
    template void cholesky_impl::pposv_workspace_size<T>(const DdlaHandle_t&, char, char, int, int, T*,        \
                                                         const int*, const int*, std::size_t*, std::size_t*);  \
    template void cholesky_impl::pposv_run<T>(const DdlaHandle_t&, char, char, char, int, int, T*, int, int,   \
                                              const int*, T*, int, int, const int*, int&, void*, std::size_t,  \
                                              void*, std::size_t);
DDLA_INSTANTIATE(float)
DDLA_INSTANTIATE(double)
DDLA_INSTANTIATE(std::complex<float>)
DDLA_INSTANTIATE(std::complex<double>)
#undef DDLA_INSTANTIATE

#define DDLA_MIXED_INSTANTIATE(R)                                                                                \
    template void gemmVbatched<R>(const DdlaHandle_t&, char, char, int*, int*, int*, std::complex<R>,           \
                                  const R* const*, int*, const std::complex<R>* const*, int*, std::complex<R>,  \
                                  std::complex<R>**, int*, int);                                                 \
    template void gemmVbatched<R>(const DdlaHandle_t&, char, char, int*, int*, int*, std::complex<R>,           \
                                  const std::complex<R>* const*, int*, const R* const*, int*, std::complex<R>,  \
                                  std::complex<R>**, int*, int);                                                 \
    template void gemmVbatched2s<R>(const DdlaHandle_t&, char, char, int*, int*, int*, std::complex<R>,         \
                                    const R* const*, int*, const std::complex<R>* const*, int*, std::complex<R>, \
                                    std::complex<R>**, int*, char, char, int*, int*, int*, std::complex<R>,      \
                                    const R* const*, int*, int*, std::complex<R>, std::complex<R>**, int*, bool, \
                                    int, const int*, int);                                                       \
    template void gemmVbatched2s<R>(const DdlaHandle_t&, char, char, int*, int*, int*, std::complex<R>,         \
                                    const std::complex<R>* const*, int*, const R* const*, int*, std::complex<R>, \
                                    std::complex<R>**, int*, char, char, int*, int*, int*, std::complex<R>,      \
                                    const R* const*, int*, int*, std::complex<R>, std::complex<R>**, int*, bool, \
                                    int, const int*, int);                                                       \
    template void gemmVbatched2s<R>(const DdlaHandle_t&, char, char, int*, int*, int*, std::complex<R>,         \
                                    const R* const*, int*, const std::complex<R>* const*, int*, std::complex<R>, \
                                    std::complex<R>**, int*, char, char, int*, int*, int*, std::complex<R>,      \
                                    const std::complex<R>* const*, int*, int*, std::complex<R>,                 \
                                    std::complex<R>**, int*, bool, int, const int*, int);                       \
    template void gemmVbatched2s<R>(const DdlaHandle_t&, char, char, int*, int*, int*, std::complex<R>,         \
                                    const std::complex<R>* const*, int*, const R* const*, int*, std::complex<R>, \
                                    std::complex<R>**, int*, char, char, int*, int*, int*, std::complex<R>,      \
                                    const std::complex<R>* const*, int*, int*, std::complex<R>,                 \
                                    std::complex<R>**, int*, bool, int, const int*, int);
DDLA_MIXED_INSTANTIATE(float)
DDLA_MIXED_INSTANTIATE(double)
#undef DDLA_MIXED_INSTANTIATE

} // namespace ddla
