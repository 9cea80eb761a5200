// Launchers of the hand-written CDNA4 (gfx950) kernels.  Host code calls only
// these; every launcher enqueues on the given stream and returns immediately.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace ddla {
namespace kern {

// Spin limit of every bounded wait in the persistent kernels (panel
// factorization, triangular-solve sweeps): a wait that polls more often sets
// the launch's error word and gives up, and the host reports the error.
// Process-wide; kDefaultWaitLimit unless changed (tests force the error path
// with a tiny limit).
constexpr int kDefaultWaitLimit = 1 << 23;
// Device info of a factorization abandoned at the spin limit (the factor is
// then not valid); negative, so it is never a pivot index.
constexpr int kInfoWaitLimit = -1000;
int wait_limit();
int set_wait_limit(int spins);  // returns the previous limit; spins <= 0 restores the default

// Triangle mask applied by the GEMM epilogue, in the coordinates of the
// matrix C that is written: element (i, j) of C is written only if
//   kTriNone, or kTriLower and i + diag >= j, or kTriUpper and i + diag <= j.
enum Tri { kTriNone = 0, kTriLower = 1, kTriUpper = 2 };

// Independent problems in one launch.  Problem b's buffers sit b * stride
// bytes after problem 0's: every pointer of the call (operands, result, info,
// work) moves by the same amount, so the caller lays out each problem's data
// at one common stride.  A product may also repeat over `inner` equally
// spaced sub-blocks of each problem, inner_stride bytes apart (operands and
// result alike).  The default is one problem.
struct Batch {
    int count = 1;
    long long stride = 0;
    int inner = 1;
    long long inner_stride = 0;
};

// C := alpha * op(A) * op(B) + beta * C, op(X) = X or X^T ('N' / 'T' / 'C').
// alpha == 0 or k == 0 only scales C; beta == 0 never reads C.  If abort is
// non-null and *abort != 0 when a workgroup starts, the workgroup does nothing
// (the factorization stops at the first failed pivot).  In-place use with C
// aliasing op(B) is safe when m <= 64 (one workgroup row covers all of m),
// and with C aliasing op(A) when n <= 64.  Triangle-masked launches run under
// the symbol syrk_update_kernel, or panel_update_kernel when lookahead (the
// same code; kernel traces tell the factorization's two update kinds apart).
template <class T>
void gemm(hipStream_t s, char ta, char tb, int m, int n, int k, T alpha, const T* A, long long lda,
          const T* B, long long ldb, T beta, T* C, long long ldc, Tri tri = kTriNone, long long diag = 0,
          const int* abort = nullptr, bool lookahead = false, const Batch& batch = Batch());

// Variable-size batched GEMM over device arrays of sizes, leading dimensions
// and pointers; max_m / max_n size the grid; tri masks every C (diagonal
// offset 0).  A member with m <= 0 or n <= 0 does nothing; k == 0 only scales
// its C.
template <class T>
void gemm_vbatched(hipStream_t s, char ta, char tb, const int* m, const int* n, const int* k, T alpha,
                   const T* const* A, const int* lda, const T* const* B, const int* ldb, T beta, T* const* C,
                   const int* ldc, int batch, int max_m, int max_n, Tri tri = kTriNone);

// Validation + max reduction of vbatched dimensions: summary[0..2] = max m, n,
// k; summary[3] = 1 + first batch index with a negative size or a too-small
// leading dimension (0 if none).  summary must be zeroed beforehand.
void vbatched_validate(hipStream_t s, int na, int nb, const int* m, const int* n, const int* k, const int* lda,
                       const int* ldb, const int* ldc, int batch, int* summary);

// Panel factorization of the blocked Cholesky (one persistent launch per
// panel).  The factor is addressed as an "L view": element (i, j), i >= j, at
// A[i + j*lda], or A[j + i*lda] when upper.  panel_factor factors columns
// [k, k + kb) (kb <= 768, every earlier trailing update applied): the
// diagonal blocks, their inverses and the rows below.  On a non-positive or
// NaN pivot in global column p (0-based) it sets *info = p + 1, and
// kInfoWaitLimit if one of its waits gave up (unless *info was set first); it
// does nothing if *info != 0 on entry.  work holds panel_work_bytes<T>(n) bytes,
// prepared once per factorization by panel_reset (one fill kernel); *error
// (the last int of work, see panel_error_word) is set if an internal wait
// ever exceeded its limit.
template <class T>
size_t panel_work_bytes(int n);
template <class T>
void panel_reset(hipStream_t s, int n, void* work, const Batch& batch = Batch());
template <class T>
int* panel_error_word(int n, void* work);
template <class T>
void panel_factor(hipStream_t s, bool upper, int n, T* A, long long lda, int k, int kb, int* info, void* work,
                  const Batch& batch = Batch());

// Split-product trailing update of the FP32 factorization (L view as for
// panel_factor); mode = 1 (bf16x3: two bfloat16 planes, three products), 2
// (bf16x6: three planes, six products) or 3 (fp16x3: two half-precision
// planes of row-scaled values, three products).  split_panel writes rows
// [0, m) and columns [0, kb) (kb <= 512) of the L view starting at X -- the
// stored address of L-view element (r, k) -- into a fragment-ordered buffer
// zero-padded to split_rows(m) rows (with the row exponents after it for
// fp16x3); split_buffer_bytes(n, mode) bytes hold it for any m <= n.  split_update
// then applies, with the buffer of a split of m_split rows,
//   C_L(i, j) -= sum_k P(a_row + i, k) P(b_row + j, k),  i < m, j < n2,
// accumulated in FP32, writing only i + diag >= j when tri; C is the stored
// address of L-view element C_L(0, 0), and upper must be false (the split
// products update the lower view only).  a_row and b_row must be multiples of
// 16 with every row the launch reads inside split_rows(m_split) (checked).
int split_rows(int m);
size_t split_buffer_bytes(int n, int mode);
void split_panel(hipStream_t s, int mode, bool upper, int m, int kb, const float* X, long long ldx, void* out);
void split_update(hipStream_t s, int mode, bool upper, int m, int n2, int kb, const void* P, int m_split,
                  int a_row, int b_row, float* C, long long ldc, bool tri, long long diag, const int* abort);

// Inverts every 64 x 64 diagonal block of the order-n factor (L view as for
// panel_factor) into inv[b * 4096 ...], identity-padded for a ragged last
// block; unit: the diagonal is taken as ones (the stored one is not used).
template <class T>
void trtri_diag_blocks(hipStream_t s, bool upper, int n, const T* A, long long lda, T* inv, bool unit = false,
                       const Batch& batch = Batch());
// M := alpha M (m x n, leading dimension ld); alpha == 0 writes zeros
// without reading M.
template <class T>
void scale_matrix(hipStream_t s, int m, int n, T alpha, T* M, long long ld);

// Forward (L Y = B) then backward (L^T X = Y) substitution with the order-n
// factor (L view as for panel_factor) and its inverted 64 x 64 diagonal blocks
// dinv (from trtri_diag_blocks), in place on B (n x nrhs), one launch per
// sweep.  work must hold solve_work_bytes(nrhs) bytes; *error (device) is set
// to 1 if a sweep's internal wait ever exceeded its limit (the result is then
// not valid).
size_t solve_work_bytes(int nrhs);
template <class T>
void potrs_sweeps(hipStream_t s, bool upper, int n, int nrhs, const T* L, long long ldl, const T* dinv, T* B,
                  long long ldb, void* work, int* error);

// Staged solve (stage_tile.h): the same forward and backward sweeps over
// 512-row stages, for orders from kStageMinOrder up with fewer than
// kTrsmMinColumns right-hand sides (stage_solve).  stage_prepare forms, once
// per factor, the inverse of every stage's diagonal block and the two sweeps'
// coupling blocks from the factor and dinv; stage_sweeps then runs both
// sweeps in place on B, as often as needed.  work holds
// stage_work_bytes(n, nrhs, sizeof(T)) bytes (prepared data, per-slab
// accumulators and counters); *error as for potrs_sweeps.
constexpr int kStageMinOrder = 8192;
bool stage_solve(int n, int nrhs);
size_t stage_work_bytes(int n, int nrhs, size_t elem);
template <class T>
void stage_prepare(hipStream_t s, bool upper, int n, const T* L, long long ldl, const T* dinv, void* work);
template <class T>
void stage_sweeps(hipStream_t s, bool upper, int n, int nrhs, const T* L, long long ldl, T* B, long long ldb,
                  void* work, int* error);

// Batched fused potrf + potrs + logdet of independent matrices (order n[i] <=
// 64 * max_blocks <= 1024, stored triangle uplo, leading dimension lda[i]; B[i]
// n[i] x nrhs with ldb[i]) in one persistent launch of `grid` workgroups, after
// a one-workgroup ordering kernel (largest first) and a counter reset; all of
// it is stream-ordered and graph-capturable.  info[i] / logdet[i] as in
// batched_tile.h.  ordered: largest matrices first (else index order);
// occupancy 1 or 2: the kernel variant sized for that many workgroups per CU.
// split_blocks (with ordered, 1 <= split_blocks < max_blocks): the members of
// more than split_blocks 64-row blocks are factored on the blocked path
// instead -- one launch per 64-column block column, the trailing updates of all
// of them in one gemm_vbatched launch per step, then their solves -- and the
// persistent launch takes the rest.  work holds batched_work_bytes(sizeof(T),
// batch, grid, max_blocks, blocked) bytes, blocked = whether split_blocks may be
// used; step_blocks (1 or 2): 64-column block columns per step of the blocked
// path (with 2 its trailing updates take k = 128, half as many of them).
// left_looking: every member on the left-looking persistent kernel
// instead (batched_ll_kernel: block columns formed once by one MFMA product
// each; split_blocks is then ignored).
constexpr int kBatchedMaxBlocks = 16;
size_t batched_work_bytes(size_t elem, int batch, int grid, int max_blocks, bool blocked);
template <class T>
void batched_pposv(hipStream_t s, bool upper, int batch, int nrhs, const int* n, T* const* A, const int* lda,
                   T* const* B, const int* ldb, int* info, double* logdet, int grid, int max_blocks, bool ordered,
                   int occupancy, int split_blocks, void* work, bool left_looking = false, int step_blocks = 1);
// Batched copy of matrix i (rows n[i], columns cols or n[i] when cols < 0)
// from src[i] (leading dimension lds[i]) to dst[i] (leading dimension ldd[i]).
template <class T>
void batched_copy(hipStream_t s, int batch, const int* n, int cols, T* const* src, const int* lds, T* const* dst,
                  const int* ldd, int max_n);

// One triangular-solve sweep with the order-n factor (L view) and its inverted
// diagonal blocks: L Y = B (backward = false) or L^T X = B (backward = true),
// in place on B (n x nrhs).  work / error as for potrs_sweeps.
template <class T>
void trsm_sweep(hipStream_t s, bool upper, bool backward, int n, int nrhs, const T* L, long long ldl, const T* dinv,
                T* B, long long ldb, void* work, int* error);

// The same solve for many right-hand sides: while trsm_splits(n, nrhs) the
// order is split in two (the first part a multiple of 64 rows) and the
// off-diagonal block is applied by one GEMM, so that the GEMM carries most of
// the n^2 nrhs flops and the sweep kernel only the leaves of order <=
// kTrsmLeaf; otherwise one sweep.  potrs_sweeps takes the same route.
constexpr int kTrsmLeaf = 1024;
constexpr int kTrsmMinColumns = 256;
bool trsm_splits(int n, int nrhs);
template <class T>
void trsm(hipStream_t s, bool upper, bool backward, int n, int nrhs, const T* L, long long ldl, const T* dinv, T* B,
          long long ldb, void* work, int* error);

// Elementwise helpers of the adjoints (order n, column-major):
//   tril_copy:    out = tril(in) (strict upper triangle zero)
//   mirror_lower: out(i, j) = scale * in(max(i, j), min(i, j)), the symmetric
//                 matrix of in's lower triangle (out must not alias in)
//   diag_adjoint: Lbar(i, i) += 2 g / L(i, i)
template <class T>
void tril_copy(hipStream_t s, int n, const T* in, long long ldi, T* out, long long ldo);
template <class T>
void mirror_lower(hipStream_t s, int n, const T* in, long long ldi, T scale, T* out, long long ldo);
// out = the order-n identity
template <class T>
void set_identity(hipStream_t s, int n, T* out, long long ldo);
// Y = tril(L)^T W for W with c <= kTmvColumns columns (L lower storage, its
// strict upper triangle never read): the few-column products of the solve
// adjoint, one wavefront per row of Y.
constexpr int kTmvColumns = 16;
template <class T>
void lower_tmv(hipStream_t s, int n, int c, const T* L, long long ldl, const T* W, long long ldw, T* Y,
               long long ldy);
template <class T>
void diag_adjoint(hipStream_t s, int n, const T* L, long long ldl, double g, T* Lbar, long long ldlb);

// diag_logdet reduction over A[i * (lda + 1)], i < n: work must hold
// logdet_work_bytes(); out (device) receives {sum log|d|, flags} where flags
// = 1 nonfinite | 2 zero | 4 odd number of negative entries.
// Batched Gaussian-process likelihood kernels (gp_kernels.hip).  Problem b's working
// region sits b * stride bytes after problem 0's (pointers below are problem
// 0's); the caller's K, y, Kbar, ybar, loss, info and g have their own
// strides (elements) or one entry per problem.
//   gp_load:       L_b = the lower triangle of K_b (n x n, leading dimension n),
//                  y copied, info_b = 0;
//   gp_place_diag: X_b = the block-diagonal matrix of the inverted 64 x 64
//                  diagonal blocks dinv_b, zero elsewhere;
//   gp_loss:       loss[b] = z_b^T z_b / 2 + sum_i log L_b(i, i) (also kept at
//                  loss_slot), info[b] = info_b;
//   gp_grad:       Kbar_b(i, j) = g[b] (M_b(max, min) - a_i a_j) / 2 from the
//                  lower triangle M_b, ybar_b = g[b] a_b (when ybar is not null).
void gp_load(hipStream_t s, int batch, int n, const double* K, long long k_stride, const double* y, long long y_stride,
             double* L, double* ybuf, int* info, long long stride);
void gp_place_diag(hipStream_t s, int batch, int n, const double* dinv, double* X, long long stride);
void gp_loss(hipStream_t s, int batch, int n, const double* L, const double* z, const int* info, double* loss_slot,
             long long stride, double* loss_out, int* info_out);
void gp_grad(hipStream_t s, int batch, int n, const double* M, const double* a, long long stride, const double* g,
             double* Kbar, long long kbar_stride, double* ybar, long long ybar_stride);

struct LogdetResult { double magnitude; int flags; int pad; };
size_t logdet_work_bytes();
template <class T>
void diag_logdet(hipStream_t s, int n, const T* A, long long lda, void* work, LogdetResult* out);

// Mixed-precision refinement helpers.
// Converts the stored triangle (lower or upper) of the order-n symmetric S to
// D and computes its infinity norm into *rowmax (device); work must hold
// sym_norm_work_bytes(n) bytes.  Deterministic summation order.
size_t sym_norm_work_bytes(int n);
template <class D, class S>
void convert_sym_with_norm(hipStream_t s, int n, bool lower, const S* in, long long ldi, D* out, long long ldo,
                           void* work, double* rowmax);
// out[j] = max_i |M(i, j)| (device, double), NaN propagating.
template <class T>
void column_max(hipStream_t s, int m, int n, const T* M, long long ld, double* out);
// X += (double) W
template <class S>
void add_converted(hipStream_t s, int m, int n, const S* W, long long ldw, double* X, long long ldx);
// R = B - A X, A symmetric given by the stored triangle of S (lower or upper).
template <class T>
void symm_residual(hipStream_t s, bool lower, int n, int nrhs, const T* S, long long lds, const T* X, long long ldx,
                   const T* B, long long ldb, T* R, long long ldr);

// Setup helpers (not on the timed path).
void fill_normal(hipStream_t s, uint64_t seed, uint64_t stream, uint64_t first, long long count, double* out);
template <class T>
void add_diagonal(hipStream_t s, int n, T* A, long long lda, T delta);
template <class T>
void transpose(hipStream_t s, int m, int n, const T* in, long long ldi, T* out, long long ldo);
template <class D, class S>
void convert(hipStream_t s, int m, int n, const S* in, long long ldi, D* out, long long ldo);

// Order reversal of a symmetric matrix given by one triangle (order n, J the
// order reversal): out's triangle out_upper receives J M J,
//   transpose: out(i, j) = in(n-1-j, n-1-i), in's stored triangle = out's;
//   rotate:    out(i, j) = in(n-1-i, n-1-j), in's stored triangle = the other.
// zero_other: out's other triangle is set to zero, else it is not written;
// in's unstored triangle is never read.  in and out must not overlap.
template <class T>
void reverse_triangle(hipStream_t s, bool out_upper, bool transpose, int n, const T* in, long long ldi, T* out,
                      long long ldo, bool zero_other);
// One triangle of a column-major m x n matrix of elem-byte elements (upper:
// i <= j, lower: i >= j; the diagonal only when diag): every byte set to
// value, or copied from src (leading dimensions in elements; the regions must
// not overlap).
void memset_tri_bytes(hipStream_t s, bool upper, bool diag, int m, int n, void* p, long long ld, std::size_t elem,
                      unsigned char value);
void memcpy_tri_bytes(hipStream_t s, bool upper, bool diag, int m, int n, void* dst, long long ldd, const void* src,
                      long long lds, std::size_t elem);

// In place: exchanges the strict lower and strict upper triangles of the
// order-n A (A(i, j) <-> A(j, i), i > j); applied twice it is the identity.
template <class T>
void swap_triangles(hipStream_t s, int n, T* A, long long lda);

} // namespace kern
} // namespace ddla
