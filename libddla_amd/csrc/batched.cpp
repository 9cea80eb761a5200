// Batched SPD solves of uneven order: fused potrf + potrs + log-determinant
// per matrix in one persistent launch (kernels in batched_tile.h), its
// workspace query / run split, the autotuner that picks the kernel variant and
// processing order, and a captured hipGraph of the whole call for replay.
//
// There is no batched Cholesky in the reference; the per-matrix semantics are
// those of its pposv (src/Cholesky/pposv.cpp:15-63: solve only when info ==
// 0, B untouched otherwise) and of 2 * Re diag_logdet of the factor
// (src/diag_logdet_impl.h:22-113), for every member independently.
#include "internal.h"
#include "kernels.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace ddla {

namespace {

constexpr int kBlk = 64;
// untuned default: members of more than 4 blocks (order > 256) take the blocked path
constexpr int kDefaultSplitBlocks = 4;

int max_blocks_for(int max_n) { return std::max(1, (max_n + kBlk - 1) / kBlk); }

int grid_for_batch(const DdlaHandle_t& h, int batch, int occupancy)
{
    const int cus = h->num_cus > 0 ? h->num_cus : 256;
    return std::max(1, std::min(batch, occupancy * cus));
}

template <class T>
DdlaStream::BatchedTuning& tuning_of(const DdlaHandle_t& h)
{
    return h->batched_tuning[sizeof(T) == sizeof(double) ? 1 : 0];
}

// the blocked path can apply (some member may have more than one block)
bool blocked_possible(int max_n) { return max_blocks_for(max_n) > 1; }

template <class T>
void launch_run(const DdlaHandle_t& h, hipStream_t s, char uplo, const int* d_n, T* const* d_A, const int* d_lda,
                int nrhs, T* const* d_B, const int* d_ldb, int* d_info, double* d_logdet, int batch, int max_n,
                int occupancy, bool ordered, int split_blocks, bool left_looking, int step_blocks, void* work)
{
    kern::batched_pposv<T>(s, uplo == 'U', batch, nrhs, d_n, d_A, d_lda, d_B, d_ldb, d_info, d_logdet,
                           grid_for_batch(h, batch, occupancy), max_blocks_for(max_n), ordered, occupancy,
                           split_blocks, work, left_looking, step_blocks);
}

} // namespace

namespace cholesky_impl {

void pposv_vbatched_check(const DdlaHandle_t& h, char uplo, int batch, int nrhs, int max_n, const void* d_n,
                          const void* d_A, const void* d_lda, const void* d_B, const void* d_ldb, const void* d_info,
                          const void* d_logdet)
{
    require_gpu(h, "pposvVbatched");
    if (uplo != 'L' && uplo != 'U') throw std::invalid_argument("pposvVbatched: uplo must be 'L' or 'U'");
    if (batch < 0 || nrhs < 0) throw std::invalid_argument("pposvVbatched: negative batch count or nrhs");
    if (max_n < 0 || max_n > kern::kBatchedMaxBlocks * kBlk)
        throw std::invalid_argument("pposvVbatched: max_n must be in [0, 1024]");
    if (batch == 0) return;
    if (!d_n || !d_A || !d_lda || !d_info || !d_logdet || (nrhs > 0 && (!d_B || !d_ldb)))
        throw std::invalid_argument("pposvVbatched: null device array");
}

template <class T>
void pposv_vbatched_workspace_size(const DdlaHandle_t& h, int batch, int max_n, std::size_t* backend_bytes)
{
    require_gpu(h, "pposvVbatched");
    *backend_bytes = 0;
    if (batch <= 0) return;
    // sized for the largest grid any variant uses (two workgroups per CU) and for the blocked path
    *backend_bytes = kern::batched_work_bytes(sizeof(T), batch, grid_for_batch(h, batch, 2), max_blocks_for(max_n),
                                              blocked_possible(max_n));
}

template <class T>
void pposv_vbatched_run(const DdlaHandle_t& h, char uplo, const int* d_n, T* const* d_A, const int* d_lda, int nrhs,
                        T* const* d_B, const int* d_ldb, int* d_info, double* d_logdet, int batch, int max_n,
                        void* backend_work, std::size_t backend_bytes)
{
    if (batch == 0) return;
    std::size_t need = 0;
    pposv_vbatched_workspace_size<T>(h, batch, max_n, &need);
    if (backend_work == nullptr || backend_bytes < need)
        throw std::invalid_argument("pposvVbatched: backend workspace smaller than the workspace query");
    const auto& t = tuning_of<T>(h);
    launch_run<T>(h, h->stream, uplo, d_n, d_A, d_lda, nrhs, d_B, d_ldb, d_info, d_logdet, batch, max_n,
                  t.tuned ? t.occupancy : 1, t.tuned ? t.ordered : true,
                  t.tuned ? t.split_blocks : kDefaultSplitBlocks, t.tuned && t.left_looking,
                  t.tuned ? t.step_blocks : 1, backend_work);
}

} // namespace cholesky_impl

template <class T>
void pposvVbatched(const DdlaHandle_t& h, char uplo, const int* d_n, T* const* d_A, const int* d_lda, int nrhs,
                   T* const* d_B, const int* d_ldb, int* d_info, double* d_logdet, int batch, int max_n)
{
    cholesky_impl::pposv_vbatched_check(h, uplo, batch, nrhs, max_n, d_n, d_A, d_lda, d_B, d_ldb, d_info, d_logdet);
    if (batch == 0) return;
    std::size_t bytes = 0;
    cholesky_impl::pposv_vbatched_workspace_size<T>(h, batch, max_n, &bytes);
    void* work = h->ensure_scratch(bytes);
    cholesky_impl::pposv_vbatched_run<T>(h, uplo, d_n, d_A, d_lda, nrhs, d_B, d_ldb, d_info, d_logdet, batch, max_n,
                                         work, bytes);
}

// Times every schedule (kernel variant for 1 or 2 workgroups per CU, largest
// first or index order, and -- largest first only -- the blocked path for the
// members of more than 2, 3, 4 or 6 blocks in steps of one or two block
// columns, or the left-looking kernel for every member) on copies of the caller's
// matrices, so the inputs are left unchanged, and keeps the fastest in the
// handle for later calls.
template <class T>
void pposvVbatchedTune(const DdlaHandle_t& h, char uplo, const int* d_n, T* const* d_A, const int* d_lda, int nrhs,
                       T* const* d_B, const int* d_ldb, int batch, int max_n, int* occupancy, bool* ordered)
{
    // the tuner writes its own info / logdet scratch: only the inputs are checked
    cholesky_impl::pposv_vbatched_check(h, uplo, batch, nrhs, max_n, d_n, d_A, d_lda, d_B, d_ldb, d_n, d_n);
    auto& tuning = tuning_of<T>(h);
    if (batch == 0) {
        tuning = DdlaStream::BatchedTuning{true, 1, true, kern::kBatchedMaxBlocks, false, 1};
        if (occupancy) *occupancy = 1;
        if (ordered) *ordered = true;
        return;
    }
    hipStream_t s = h->stream;
    std::vector<int> n(batch), lda(batch), ldb(batch, 1);
    DDLA_HIP(hipMemcpyAsync(n.data(), d_n, batch * sizeof(int), hipMemcpyDeviceToHost, s));
    DDLA_HIP(hipMemcpyAsync(lda.data(), d_lda, batch * sizeof(int), hipMemcpyDeviceToHost, s));
    if (nrhs > 0) DDLA_HIP(hipMemcpyAsync(ldb.data(), d_ldb, batch * sizeof(int), hipMemcpyDeviceToHost, s));
    DDLA_HIP(hipStreamSynchronize(s));
    // members the kernel rejects (order out of range, short leading
    // dimension) are timed as empty and never read
    for (int i = 0; i < batch; ++i) {
        const int need = std::max(1, n[i]);
        if (n[i] < 0 || n[i] > max_n || lda[i] < need || (nrhs > 0 && ldb[i] < need)) n[i] = 0;
    }
    // compact copies: A_i with leading dimension max(1, n_i), B_i likewise
    std::vector<long long> offA(batch), offB(batch);
    std::vector<int> ld(batch);
    long long totA = 0, totB = 0;
    int mx = 0;
    for (int i = 0; i < batch; ++i) {
        const int m = n[i];
        ld[i] = std::max(1, m);
        offA[i] = totA;
        offB[i] = totB;
        totA += (long long)ld[i] * m;
        totB += (long long)ld[i] * nrhs;
        mx = std::max(mx, m);
    }
    std::size_t work_bytes = 0;
    cholesky_impl::pposv_vbatched_workspace_size<T>(h, batch, max_n, &work_bytes);
    const std::size_t arena = detail::padded_bytes<T>(totA + 1) + detail::padded_bytes<T>(totB + 1) +
                              2 * detail::padded_bytes<T*>(batch) + detail::padded_bytes<int>(batch) * 3 +
                              detail::padded_bytes<double>(batch) + work_bytes;
    char* base = nullptr;
    DDLA_HIP(hipMalloc(&base, arena));
    char* w = base;
    T* copyA = detail::take<T>(w, totA + 1);
    T* copyB = detail::take<T>(w, totB + 1);
    T** pA = detail::take<T*>(w, batch);
    T** pB = detail::take<T*>(w, batch);
    int* dld = detail::take<int>(w, batch);
    int* dn = detail::take<int>(w, batch);
    int* info = detail::take<int>(w, batch);
    double* logdet = detail::take<double>(w, batch);
    void* work = w;
    std::vector<T*> hA(batch), hB(batch);
    std::vector<int> nclamp(batch);
    for (int i = 0; i < batch; ++i) {
        hA[i] = copyA + offA[i];
        hB[i] = copyB + offB[i];
        nclamp[i] = n[i];
    }
    DDLA_HIP(hipMemcpyAsync(pA, hA.data(), batch * sizeof(T*), hipMemcpyHostToDevice, s));
    DDLA_HIP(hipMemcpyAsync(pB, hB.data(), batch * sizeof(T*), hipMemcpyHostToDevice, s));
    DDLA_HIP(hipMemcpyAsync(dld, ld.data(), batch * sizeof(int), hipMemcpyHostToDevice, s));
    DDLA_HIP(hipMemcpyAsync(dn, nclamp.data(), batch * sizeof(int), hipMemcpyHostToDevice, s));
    hipEvent_t e0, e1;
    DDLA_HIP(hipEventCreate(&e0));
    DDLA_HIP(hipEventCreate(&e1));
    struct Candidate {
        int occupancy;
        bool ordered;
        int split_blocks;
        bool left_looking;
        int step_blocks;
        float ms;
    };
    constexpr int kNone = kern::kBatchedMaxBlocks;
    std::vector<Candidate> cands{{1, true, kNone, false, 1, 0}, {2, true, kNone, false, 1, 0},
                                {1, false, kNone, false, 1, 0}, {2, false, kNone, false, 1, 0}};
    for (int sb : {2, 3, 4, 6})
        if (sb < max_blocks_for(max_n))
            for (int occ : {1, 2})
                for (int step : {1, 2}) cands.push_back({occ, true, sb, false, step, 0});
    for (int occ : {1, 2}) cands.push_back({occ, true, kNone, true, 1, 0});
    try {
        for (auto& c : cands) {
            float best = 1e30f;
            for (int rep = 0; rep < 2; ++rep) {
                kern::batched_copy<T>(s, batch, dn, -1, const_cast<T* const*>(d_A), d_lda, pA, dld, mx);
                if (nrhs > 0) kern::batched_copy<T>(s, batch, dn, nrhs, const_cast<T* const*>(d_B), d_ldb, pB, dld, mx);
                DDLA_HIP(hipEventRecord(e0, s));
                launch_run<T>(h, s, uplo, dn, pA, dld, nrhs, pB, dld, info, logdet, batch, max_n, c.occupancy,
                              c.ordered, c.split_blocks, c.left_looking, c.step_blocks, work);
                DDLA_HIP(hipEventRecord(e1, s));
                DDLA_HIP(hipEventSynchronize(e1));
                float ms = 0;
                DDLA_HIP(hipEventElapsedTime(&ms, e0, e1));
                best = std::min(best, ms);
            }
            c.ms = best;
        }
    } catch (...) {
        (void)hipStreamSynchronize(s);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipFree(base);
        throw;
    }
    DDLA_HIP(hipEventDestroy(e0));
    DDLA_HIP(hipEventDestroy(e1));
    DDLA_HIP(hipFree(base));
    const auto best = std::min_element(cands.begin(), cands.end(),
                                       [](const Candidate& a, const Candidate& b) { return a.ms < b.ms; });
    tuning = DdlaStream::BatchedTuning{true,          best->occupancy,    best->ordered,
                                       best->split_blocks, best->left_looking, best->step_blocks};
    if (occupancy) *occupancy = best->occupancy;
    if (ordered) *ordered = best->ordered;
}

// A captured, instantiated graph of one batched call with its own workspace.
struct BatchedGraph {
    hipGraphExec_t exec = nullptr;
    void* work = nullptr;
};

template <class T>
void* pposvVbatchedGraph(const DdlaHandle_t& h, char uplo, const int* d_n, T* const* d_A, const int* d_lda, int nrhs,
                         T* const* d_B, const int* d_ldb, int* d_info, double* d_logdet, int batch, int max_n)
{
    cholesky_impl::pposv_vbatched_check(h, uplo, batch, nrhs, max_n, d_n, d_A, d_lda, d_B, d_ldb, d_info, d_logdet);
    if (batch == 0) throw std::invalid_argument("pposvVbatchedGraph: empty batch");
    auto& t = tuning_of<T>(h);
    if (!t.tuned) pposvVbatchedTune<T>(h, uplo, d_n, d_A, d_lda, nrhs, d_B, d_ldb, batch, max_n, nullptr, nullptr);
    std::size_t bytes = 0;
    cholesky_impl::pposv_vbatched_workspace_size<T>(h, batch, max_n, &bytes);
    auto* g = new BatchedGraph;
    hipStream_t cs = nullptr;
    hipGraph_t graph = nullptr;
    try {
        DDLA_HIP(hipMalloc(&g->work, bytes));
        DDLA_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        DDLA_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        launch_run<T>(h, cs, uplo, d_n, d_A, d_lda, nrhs, d_B, d_ldb, d_info, d_logdet, batch, max_n, t.occupancy,
                      t.ordered, t.split_blocks, t.left_looking, t.step_blocks, g->work);
        DDLA_HIP(hipStreamEndCapture(cs, &graph));
        DDLA_HIP(hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0));
        DDLA_HIP(hipGraphDestroy(graph));
        DDLA_HIP(hipStreamDestroy(cs));
    } catch (...) {
        if (cs) {
            hipGraph_t partial = nullptr;
            (void)hipStreamEndCapture(cs, &partial);
            if (partial) (void)hipGraphDestroy(partial);
            (void)hipStreamDestroy(cs);
        }
        if (graph) (void)hipGraphDestroy(graph);
        if (g->work) (void)hipFree(g->work);
        delete g;
        throw;
    }
    return g;
}

void batchedGraphLaunch(void* graph, const DdlaHandle_t& h)
{
    require_gpu(h, "batchedGraphLaunch");
    if (graph == nullptr) throw std::invalid_argument("batchedGraphLaunch: null graph");
    DDLA_HIP(hipGraphLaunch(static_cast<BatchedGraph*>(graph)->exec, h->stream));
}

void batchedGraphDestroy(void* graph)
{
    if (graph == nullptr) return;
    auto* g = static_cast<BatchedGraph*>(graph);
    if (g->exec) DDLA_HIP(hipGraphExecDestroy(g->exec));
    if (g->work) {
        DDLA_HIP(hipDeviceSynchronize());
        DDLA_HIP(hipFree(g->work));
    }
    delete g;
}

#define DDLA_BATCHED_INSTANTIATE(T)                                                                               \
    template void pposvVbatched<T>(const DdlaHandle_t&, char, const int*, T* const*, const int*, int, T* const*, \
                                   const int*, int*, double*, int, int);                                         \
    template void pposvVbatchedTune<T>(const DdlaHandle_t&, char, const int*, T* const*, const int*, int,        \
                                       T* const*, const int*, int, int, int*, bool*);                            \
    template void* pposvVbatchedGraph<T>(const DdlaHandle_t&, char, const int*, T* const*, const int*, int,      \
                                         T* const*, const int*, int*, double*, int, int);                        \
    template void cholesky_impl::pposv_vbatched_workspace_size<T>(const DdlaHandle_t&, int, int, std::size_t*);  \
    template void cholesky_impl::pposv_vbatched_run<T>(const DdlaHandle_t&, char, const int*, T* const*,         \
                                                       const int*, int, T* const*, const int*, int*, double*,    \
                                                       int, int, void*, std::size_t);
DDLA_BATCHED_INSTANTIATE(float)
DDLA_BATCHED_INSTANTIATE(double)
#undef DDLA_BATCHED_INSTANTIATE

} // namespace ddla
