/*
 * ddla_mi355x.h -- C ABI of the MI355X-native LibDDLA Cholesky path.
 *
 * The entry points keep the names, argument order and status convention of
 * the reference's generated C ABI (fortran/ddla_c_api.cpp), so a Fortran /
 * ctypes / cgo binding written for the reference binds this library
 * unchanged.  Plain pointers and sizes only; matrices are device pointers.
 *
 * Status codes (fortran/ddla_c_api.cpp:43-51): 0 success, 1 invalid handle,
 * 2 invalid value, 3 allocation failed, 4 backend unavailable, 5 internal
 * (device) error, 6 a C++ exception escaped -- its what() text is returned by
 * ddla_c_error_string(6) on the same thread (fortran/ddla_c_api.cpp:55-98).
 *
 * One process, one GPU, a 1 x 1 grid: ddla_c_set accepts nprows = npcols = 1
 * (or <= 0 for the default) and ignores f_comm; only backend 1 (GPU) exists.
 */
#ifndef DDLA_MI355X_H
#define DDLA_MI355X_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    DDLA_C_STATUS_SUCCESS = 0,
    DDLA_C_STATUS_INVALID_HANDLE = 1,
    DDLA_C_STATUS_INVALID_VALUE = 2,
    DDLA_C_STATUS_ALLOC_FAILED = 3,
    DDLA_C_STATUS_BACKEND_UNAVAILABLE = 4,
    DDLA_C_STATUS_INTERNAL_ERROR = 5,
    DDLA_C_STATUS_EXCEPTION = 6
};

/* ---- runtime: replaces fortran/ddla_c_api.cpp:91-271 ------------------- */
const char* ddla_c_error_string(int status);                        /* :91-98   */
int ddla_c_init_backend(void** handle, int backend);                /* :100-113 */
int ddla_c_set(void* handle, int f_comm, int nprows, int npcols, char major); /* :115-130 */
int ddla_c_destroy(void** handle);                                  /* :165-176 */
int ddla_c_backend_available(int backend);                          /* :178-182 */
int ddla_c_get_backend(void* handle);                               /* :184-188 */
int ddla_c_grid_dims(void* handle, int* nprows, int* npcols);       /* :190-195 */
int ddla_c_grid_coords(void* handle, int* myprow, int* mypcol);     /* :197-202 */
int ddla_c_rank(void* handle);                                      /* :204-208 */
int ddla_c_size(void* handle);                                      /* :210-214 */
int ddla_c_descinit(int* desc, void* handle, int m, int n, int mb, int nb,
                    int irsrc, int icsrc, int lld);                 /* :216-222 */
int ddla_c_numloc(int n, int nb, int iproc, int srcproc, int nprocs); /* :224-227 */
int ddla_c_malloc(void* handle, void** ptr, long long bytes);       /* :229-234 */
int ddla_c_free(void* handle, void* ptr);                           /* :236-240 */
int ddla_c_memcpy(void* handle, void* dst, const void* src, long long bytes); /* :242-247 */
int ddla_c_memcpy2d(void* handle, void* dst, long long dpitch, const void* src, long long spitch,
                    long long width, long long height);             /* :249-258 */
int ddla_c_memset(void* handle, void* ptr, int value, long long bytes); /* :260-265 */
/* Communication mode 0 MPI, 1 host-staged MPI, 2 CCL: every valid mode is
 * accepted and recorded (a 1 x 1 grid exchanges nothing). */
int ddla_c_set_communication(void* handle, int communication);      /* :134-142 */
int ddla_c_get_communication(void* handle, int* communication);     /* :144-154 */
int ddla_c_communication_available(void* handle, int communication); /* :156-163 */
/* Communicator helpers of a single process: default communicator 0, rank 0,
 * size 1; the dimension broadcast leaves its arguments as they are. */
int ddla_c_default_comm(void);                                      /* :304-308 */
int ddla_c_comm_rank(int f_comm);                                   /* :310-317 */
int ddla_c_comm_size(int f_comm);                                   /* :319-326 */
void ddla_c_finalize(void);                                         /* :328-337 */
int ddla_c_bcast_dims(int f_comm, int* m, int* n, int* type_char, int root); /* :339-349 */
int ddla_c_synchronize(void* handle);                               /* :267-271 */
/* Spin limit of the bounded waits inside the persistent kernels (panel
 * factorization, triangular-solve sweeps); no counterpart in the reference
 * (its vendor routines have no such waits).  Process-wide: returns the previous
 * limit; spins <= 0 restores the default (2^23 polls).  A wait that exceeds it
 * makes the call fail with status 6 ("... wait limit ..."); a diagnostic for
 * hung-looking runs and for tests of that error path. */
int ddla_c_set_wait_limit(int spins);
/* Panel width of the blocked factorization (the depth of its trailing
 * updates); no counterpart in the reference (its vendor potrf picks its own
 * blocking).  Process-wide: cols <= 0 restores the built-in rule (768 columns
 * for FP64 from order 16384, else 512); other values are rounded down to a
 * multiple of 64 within [64, 768] (the bf16 split-product factorization caps
 * it at 512).  Returns the previous override (0 = none).  The factor does not
 * depend on it beyond rounding; it exists so that tests can run the wide
 * panels of large orders on small matrices. */
int ddla_c_set_panel_width(int cols);
/* The stream schedule of the blocked factorization for order n and panel
 * width panel_width (<= 0: the FP64 rule), without a GPU: *count steps, the
 * first min(*count, capacity) written to steps as 5 ints each, (op, stream,
 * a, b, c) in the order the host issues them: op 0 resets the panel
 * counters, 1 factors panel a, 2 applies panel a to columns [b, c) (rows from
 * b), all three launched on stream 0 (the handle's stream), 1 (side) or 2
 * (look-ahead); op 3 records event a on the stream, op 4 makes the stream
 * wait for event a (0 panel, 1 look-ahead, 2 rest).  No counterpart in the
 * reference; it lets a caller check the schedule's ordering on the host. */
int ddla_c_potrf_schedule(int n, int panel_width, int* steps, int capacity, int* count);
/* ddlaSetStream / ddlaGetStream (include/ddla/ddla_handle_t.h:300, :337);
 * no C entry in the reference. */
int ddla_c_set_stream(void* handle, void* stream);
void* ddla_c_get_stream(void* handle);

/* ---- Cholesky drivers: replace fortran/ddla_c_api.cpp:1155-1335 ---------- */
int ddla_s_ppotrf(void* handle, char uplo, int n, void* A, int ia, int ja, const int* descA, int* info); /* :1155 */
int ddla_d_ppotrf(void* handle, char uplo, int n, void* A, int ia, int ja, const int* descA, int* info); /* :1164 */
int ddla_s_ppotrs(void* handle, char side, char uplo, char trans, int n, int nrhs, const void* A,
                  const int* descA, void* B, const int* descB);                                      /* :1263 */
int ddla_d_ppotrs(void* handle, char side, char uplo, char trans, int n, int nrhs, const void* A,
                  const int* descA, void* B, const int* descB);                                      /* :1272 */
int ddla_s_pposv(void* handle, char side, char uplo, char trans, int n, int nrhs, void* A, int ia, int ja,
                 const int* descA, void* B, int ib, int jb, const int* descB, int* info);            /* :1299 */
int ddla_d_pposv(void* handle, char side, char uplo, char trans, int n, int nrhs, void* A, int ia, int ja,
                 const int* descA, void* B, int ib, int jb, const int* descB, int* info);            /* :1308 */

/* ---- complex Hermitian drivers: replace fortran/ddla_c_api.cpp:1173-1190
 * (ppotrf), :1281-1299 (ppotrs), :1317-1335 (pposv).  A = L L^H (uplo 'L')
 * or U^H U ('U'); matrices are std::complex<float> ('c') / <double> ('z')
 * arrays, real part first.  side 'R' solves X A = B; trans 'N' and 'C' solve
 * the same system.  Same statuses, info and synchronization as the real ones. */
int ddla_c_ppotrf(void* handle, char uplo, int n, void* A, int ia, int ja, const int* descA, int* info); /* :1173 */
int ddla_z_ppotrf(void* handle, char uplo, int n, void* A, int ia, int ja, const int* descA, int* info); /* :1182 */
int ddla_c_ppotrs(void* handle, char side, char uplo, char trans, int n, int nrhs, const void* A,
                  const int* descA, void* B, const int* descB);                                      /* :1281 */
int ddla_z_ppotrs(void* handle, char side, char uplo, char trans, int n, int nrhs, const void* A,
                  const int* descA, void* B, const int* descB);                                      /* :1290 */
int ddla_c_pposv(void* handle, char side, char uplo, char trans, int n, int nrhs, void* A, int ia, int ja,
                 const int* descA, void* B, int ib, int jb, const int* descB, int* info);            /* :1317 */
int ddla_z_pposv(void* handle, char side, char uplo, char trans, int n, int nrhs, void* A, int ia, int ja,
                 const int* descA, void* B, int ib, int jb, const int* descB, int* info);            /* :1326 */

/* ---- triangular solves: replace fortran/ddla_c_api.cpp:351-367 (ptrtrs)
 * and :1451-1468 (trsm).  B := op(A)^-1 B (side 'L', A m x m) or B op(A)^-1
 * (side 'R', A n x n); uplo 'L' / 'U', trans 'N' / 'T' / 'C', diag 'N' /
 * 'U' (the stored diagonal is then not read).  ptrtrs: descriptors, 1 x 1
 * grid, upper-case flags (include/ddla/ddla.h:96 of the reference); trsm:
 * leading dimensions and *alpha, case-insensitive flags, alpha == 0 zeroes B
 * without reading A (include/ddla/ddla.h:1620).  Both return synchronized. */
int ddla_s_ptrtrs(void* handle, char side, char uplo, char trans, char diag, int m, int n, void* A,
                  const int* descA, void* B, const int* descB);                                      /* :351 */
int ddla_d_ptrtrs(void* handle, char side, char uplo, char trans, char diag, int m, int n, void* A,
                  const int* descA, void* B, const int* descB);                                      /* :360 */
int ddla_s_trsm(void* handle, char side, char uplo, char trans, char diag, int m, int n, const float* alpha,
                void* A, int lda, void* B, int ldb);                                                 /* :1451 */
int ddla_d_trsm(void* handle, char side, char uplo, char trans, char diag, int m, int n, const double* alpha,
                void* A, int lda, void* B, int ldb);                                                 /* :1460 */

/* ---- three-phase API (no C entry in the reference): the C++ components
 * cholesky_impl::{ppotrf,ppotrs,pposv}_{workspace_size,run} of
 * src/Cholesky/cholesky_workspace.h:133-249.  The size query allocates
 * nothing; a run validates nothing beyond the workspace size and allocates
 * nothing.  Every run returns synchronized: ppotrf/pposv run read info, and
 * all three read the device error word of the persistent kernels' bounded
 * waits and fail with status 6 ("... wait limit ...") if it is set. ------ */
int ddla_s_ppotrf_workspace_size(void* handle, char uplo, int n, const int* descA, long long* backend_bytes,
                                 long long* host_bytes);
int ddla_d_ppotrf_workspace_size(void* handle, char uplo, int n, const int* descA, long long* backend_bytes,
                                 long long* host_bytes);
int ddla_s_ppotrf_run(void* handle, char uplo, int n, void* A, int ia, int ja, const int* descA, int* info,
                      void* backend_work, long long backend_bytes, void* host_work, long long host_bytes);
int ddla_d_ppotrf_run(void* handle, char uplo, int n, void* A, int ia, int ja, const int* descA, int* info,
                      void* backend_work, long long backend_bytes, void* host_work, long long host_bytes);
int ddla_s_ppotrs_workspace_size(void* handle, char side, int n, int nrhs, const int* descA, const int* descB,
                                 long long* backend_bytes, long long* host_bytes);
int ddla_d_ppotrs_workspace_size(void* handle, char side, int n, int nrhs, const int* descA, const int* descB,
                                 long long* backend_bytes, long long* host_bytes);
int ddla_s_ppotrs_run(void* handle, char side, char uplo, char trans, int n, int nrhs, const void* A,
                      const int* descA, void* B, const int* descB, void* backend_work, long long backend_bytes,
                      void* host_work, long long host_bytes);
int ddla_d_ppotrs_run(void* handle, char side, char uplo, char trans, int n, int nrhs, const void* A,
                      const int* descA, void* B, const int* descB, void* backend_work, long long backend_bytes,
                      void* host_work, long long host_bytes);
int ddla_s_pposv_workspace_size(void* handle, char side, char uplo, int n, int nrhs, const int* descA,
                                const int* descB, long long* backend_bytes, long long* host_bytes);
int ddla_d_pposv_workspace_size(void* handle, char side, char uplo, int n, int nrhs, const int* descA,
                                const int* descB, long long* backend_bytes, long long* host_bytes);
int ddla_s_pposv_run(void* handle, char side, char uplo, char trans, int n, int nrhs, void* A, int ia, int ja,
                     const int* descA, void* B, int ib, int jb, const int* descB, int* info, void* backend_work,
                     long long backend_bytes, void* host_work, long long host_bytes);
int ddla_d_pposv_run(void* handle, char side, char uplo, char trans, int n, int nrhs, void* A, int ia, int ja,
                     const int* descA, void* B, int ib, int jb, const int* descB, int* info, void* backend_work,
                     long long backend_bytes, void* host_work, long long host_bytes);

/* ---- FP64 pposv that allocates nothing.  No counterpart in the reference:
 * for callers that cannot call hipMalloc (code under stream capture, tight
 * serving loops, harnesses that forbid allocation).  Same arguments, checks,
 * info and results as ddla_d_pposv, but the scratch is a static 64 MiB device
 * array of the library (one per GPU, held until the process ends) instead of
 * the handle's scratch.  A call that needs more -- the pposv workspace size:
 * about 512 bytes per row of A, plus 8 n nrhs bytes for side 'R' -- fails
 * with status 6 and a message naming both sizes.  Every call on one GPU
 * shares the array: calls that overlap on one GPU (two handles, two threads)
 * race.  Returns synchronized. */
int ddla_d_pposv_noalloc(void* handle, char side, char uplo, char trans, int n, int nrhs, void* A, int ia, int ja,
                         const int* descA, void* B, int ib, int jb, const int* descB, int* info);

/* Complex three-phase entries (same contracts as the real ones above). */
int ddla_c_ppotrf_workspace_size(void* handle, char uplo, int n, const int* descA, long long* backend_bytes,
                                 long long* host_bytes);
int ddla_z_ppotrf_workspace_size(void* handle, char uplo, int n, const int* descA, long long* backend_bytes,
                                 long long* host_bytes);
int ddla_c_ppotrf_run(void* handle, char uplo, int n, void* A, int ia, int ja, const int* descA, int* info,
                      void* backend_work, long long backend_bytes, void* host_work, long long host_bytes);
int ddla_z_ppotrf_run(void* handle, char uplo, int n, void* A, int ia, int ja, const int* descA, int* info,
                      void* backend_work, long long backend_bytes, void* host_work, long long host_bytes);
int ddla_c_ppotrs_workspace_size(void* handle, char side, int n, int nrhs, const int* descA, const int* descB,
                                 long long* backend_bytes, long long* host_bytes);
int ddla_z_ppotrs_workspace_size(void* handle, char side, int n, int nrhs, const int* descA, const int* descB,
                                 long long* backend_bytes, long long* host_bytes);
int ddla_c_ppotrs_run(void* handle, char side, char uplo, char trans, int n, int nrhs, const void* A,
                      const int* descA, void* B, const int* descB, void* backend_work, long long backend_bytes,
                      void* host_work, long long host_bytes);
int ddla_z_ppotrs_run(void* handle, char side, char uplo, char trans, int n, int nrhs, const void* A,
                      const int* descA, void* B, const int* descB, void* backend_work, long long backend_bytes,
                      void* host_work, long long host_bytes);
int ddla_c_pposv_workspace_size(void* handle, char side, char uplo, int n, int nrhs, const int* descA,
                                const int* descB, long long* backend_bytes, long long* host_bytes);
int ddla_z_pposv_workspace_size(void* handle, char side, char uplo, int n, int nrhs, const int* descA,
                                const int* descB, long long* backend_bytes, long long* host_bytes);
int ddla_c_pposv_run(void* handle, char side, char uplo, char trans, int n, int nrhs, void* A, int ia, int ja,
                     const int* descA, void* B, int ib, int jb, const int* descB, int* info, void* backend_work,
                     long long backend_bytes, void* host_work, long long host_bytes);
int ddla_z_pposv_run(void* handle, char side, char uplo, char trans, int n, int nrhs, void* A, int ia, int ja,
                     const int* descA, void* B, int ib, int jb, const int* descB, int* info, void* backend_work,
                     long long backend_bytes, void* host_work, long long host_bytes);

/* ---- mixed-precision pposv (no counterpart in the reference; LAPACK dsposv
 * semantics): FP64 A X = B, side 'L', factored in FP32 on the matrix cores
 * (factor_precision 0: FP32 trailing updates; 1: bf16x3 and 2: bf16x6 split
 * bfloat16 products with FP32 accumulation; 3: fp16x3, three half-precision
 * split products of panels whose rows are scaled by powers of two, FP32
 * accumulation) and refined in FP64 on the GPU
 * until every column has
 * ||r||_inf <= tol ||A||_inf ||x||_inf (at most max_iter corrections), else
 * the FP64 pposv runs.  *iter = corrections applied, -3 (FP32 factor failed)
 * or -(max_iter + 1) (no convergence; fallback ran).  A is unchanged unless
 * the fallback ran; B receives X. -------------------------------------- */
int ddla_d_pposv_mixed(void* handle, char uplo, int n, int nrhs, void* A, int ia, int ja, const int* descA, void* B,
                       int ib, int jb, const int* descB, int factor_precision, double tol, int max_iter, int* iter,
                       int* info);
int ddla_d_pposv_mixed_workspace_size(void* handle, char uplo, int n, int nrhs, const int* descA, const int* descB,
                                      long long* backend_bytes, long long* host_bytes);
int ddla_d_pposv_mixed_run(void* handle, char uplo, int n, int nrhs, void* A, const int* descA, void* B,
                           const int* descB, int factor_precision, double tol, int max_iter, int* iter, int* info,
                           void* backend_work, long long backend_bytes);

/* ---- local potrf, device info (C++ ddla::potrf, include/ddla/ddla.h:2160).
 * Asynchronous on the handle's stream; *d_info (device) receives 0, the
 * 1-based LAPACK index of the first failed pivot, or
 * DDLA_POTRF_INFO_WAIT_LIMIT when a wait inside the factorization exceeded
 * the spin limit (ddla_c_set_wait_limit; the factor is then not valid). --- */
#define DDLA_POTRF_INFO_WAIT_LIMIT (-1000)
int ddla_s_potrf(void* handle, char uplo, int n, void* A, int lda, int* d_info);
int ddla_d_potrf(void* handle, char uplo, int n, void* A, int lda, int* d_info);

int ddla_c_potrf(void* handle, char uplo, int n, void* A, int lda, int* d_info);
int ddla_z_potrf(void* handle, char uplo, int n, void* A, int lda, int* d_info);

/* ---- Cholesky from the bottom-right corner, host info (C++
 * ddla::potrf_bottom_right / ppotrf_bottom_right, include/ddla/ddla.h:1090,
 * 1115; C entries fortran/ddla_c_api.cpp:1191-1258): uplo 'U' A = U U^T,
 * 'L' A = L^T L; info = first row of the failing trailing submatrix. ---- */
int ddla_s_potrf_bottom_right(void* handle, char uplo, int n, void* A, int lda, int* info); /* :1191 */
int ddla_d_potrf_bottom_right(void* handle, char uplo, int n, void* A, int lda, int* info); /* :1200 */
int ddla_s_ppotrf_bottom_right(void* handle, char uplo, int n, void* A, const int* descA, int* info); /* :1227 */
int ddla_d_ppotrf_bottom_right(void* handle, char uplo, int n, void* A, const int* descA, int* info); /* :1236 */

/* ---- diag_logdet (C++ ddla::diag_logdet, include/ddla/ddla.h:2901; no C
 * entry in the reference): result[0] = real part, result[1] = phase. ----- */
int ddla_s_diag_logdet(void* handle, int n, const void* A, int lda, float* result);
int ddla_d_diag_logdet(void* handle, int n, const void* A, int lda, double* result);
/* complex diagonals: result[0] = sum log |d_i|, result[1] = principal phase in (-pi, pi]. */
int ddla_c_diag_logdet(void* handle, int n, const void* A, int lda, float* result);
int ddla_z_diag_logdet(void* handle, int n, const void* A, int lda, double* result);
/* pdiag_logdet (C++ ddla::pdiag_logdet, include/ddla/ddla.h:2914; no C entry
 * in the reference): the same on the leading n x n block of a descriptor. */
int ddla_s_pdiag_logdet(void* handle, int n, const void* A, const int* descA, float* result);
int ddla_d_pdiag_logdet(void* handle, int n, const void* A, const int* descA, double* result);
int ddla_c_pdiag_logdet(void* handle, int n, const void* A, const int* descA, float* result);
int ddla_z_pdiag_logdet(void* handle, int n, const void* A, const int* descA, double* result);

/* ---- gemmVbatched (C++ ddla::gemmVbatched, include/ddla/ddla.h:1836; the
 * reference binds no C entry, README.md:967-970).  All arrays are device
 * resident; alpha / beta are host scalars passed by reference. ---------- */
int ddla_s_gemm_vbatched(void* handle, char transA, char transB, int* d_m, int* d_n, int* d_k, const float* alpha,
                         const void* const* d_A, int* d_lda, const void* const* d_B, int* d_ldb, const float* beta,
                         void** d_C, int* d_ldc, int batch_count);
int ddla_d_gemm_vbatched(void* handle, char transA, char transB, int* d_m, int* d_n, int* d_k, const double* alpha,
                         const void* const* d_A, int* d_lda, const void* const* d_B, int* d_ldb, const double* beta,
                         void** d_C, int* d_ldc, int batch_count);

/* complex types: alpha / beta point to (real, imaginary) pairs */
int ddla_c_gemm_vbatched(void* handle, char transA, char transB, int* d_m, int* d_n, int* d_k, const void* alpha,
                         const void* const* d_A, int* d_lda, const void* const* d_B, int* d_ldb, const void* beta,
                         void** d_C, int* d_ldc, int batch_count);
int ddla_z_gemm_vbatched(void* handle, char transA, char transB, int* d_m, int* d_n, int* d_k, const void* alpha,
                         const void* const* d_A, int* d_lda, const void* const* d_B, int* d_ldb, const void* beta,
                         void** d_C, int* d_ldc, int batch_count);
/* Mixed real/complex (C++ overloads of include/ddla/ddla.h:1878-1960): complex
 * output and scalars, real_operand 0 makes A the real operand, 1 makes B real;
 * 'C' on the real operand is 'T'. */
int ddla_c_gemm_vbatched_mixed(void* handle, int real_operand, char transA, char transB, int* d_m, int* d_n,
                               int* d_k, const void* alpha, const void* const* d_A, int* d_lda,
                               const void* const* d_B, int* d_ldb, const void* beta, void** d_C, int* d_ldc,
                               int batch_count);
int ddla_z_gemm_vbatched_mixed(void* handle, int real_operand, char transA, char transB, int* d_m, int* d_n,
                               int* d_k, const void* alpha, const void* const* d_A, int* d_lda,
                               const void* const* d_B, int* d_ldb, const void* beta, void** d_C, int* d_ldc,
                               int batch_count);

/* ---- gemmVbatched2s (C++ ddla::gemmVbatched2s, include/ddla/ddla.h:1857-1877;
 * no C entry in the reference): per host segment (segment_sizes is host
 * memory), stage 0 C0 = alpha0 op(A0) op(B0) (beta0 must be zero), then
 * stage 1 C1 = alpha1 op(C0) op(AB1) + beta1 C1 (C0_left != 0) or
 * alpha1 op(AB1) op(C0) + beta1 C1; lda1 / ldb1 are the stage-1 left /
 * right leading dimensions. ----------------------------------------------- */
int ddla_s_gemm_vbatched2s(void* handle, char transA_0, char transB_0, int* d_m_0, int* d_n_0, int* d_k_0,
                           const float* alpha_0, const void* const* d_A_0, int* d_lda_0, const void* const* d_B_0,
                           int* d_ldb_0, const float* beta_0, void** d_C_0, int* d_ldc_0, char transA_1,
                           char transB_1, int* d_m_1, int* d_n_1, int* d_k_1, const float* alpha_1,
                           const void* const* d_AB_1, int* d_lda_1, int* d_ldb_1, const float* beta_1, void** d_C_1,
                           int* d_ldc_1, int C0_left, int batch_count, const int* segment_sizes, int segment_count);
int ddla_d_gemm_vbatched2s(void* handle, char transA_0, char transB_0, int* d_m_0, int* d_n_0, int* d_k_0,
                           const double* alpha_0, const void* const* d_A_0, int* d_lda_0, const void* const* d_B_0,
                           int* d_ldb_0, const double* beta_0, void** d_C_0, int* d_ldc_0, char transA_1,
                           char transB_1, int* d_m_1, int* d_n_1, int* d_k_1, const double* alpha_1,
                           const void* const* d_AB_1, int* d_lda_1, int* d_ldb_1, const double* beta_1,
                           void** d_C_1, int* d_ldc_1, int C0_left, int batch_count, const int* segment_sizes,
                           int segment_count);

int ddla_c_gemm_vbatched2s(void* handle, char transA_0, char transB_0, int* d_m_0, int* d_n_0, int* d_k_0,
                           const void* alpha_0, const void* const* d_A_0, int* d_lda_0, const void* const* d_B_0,
                           int* d_ldb_0, const void* beta_0, void** d_C_0, int* d_ldc_0, char transA_1,
                           char transB_1, int* d_m_1, int* d_n_1, int* d_k_1, const void* alpha_1,
                           const void* const* d_AB_1, int* d_lda_1, int* d_ldb_1, const void* beta_1, void** d_C_1,
                           int* d_ldc_1, int C0_left, int batch_count, const int* segment_sizes, int segment_count);
int ddla_z_gemm_vbatched2s(void* handle, char transA_0, char transB_0, int* d_m_0, int* d_n_0, int* d_k_0,
                           const void* alpha_0, const void* const* d_A_0, int* d_lda_0, const void* const* d_B_0,
                           int* d_ldb_0, const void* beta_0, void** d_C_0, int* d_ldc_0, char transA_1,
                           char transB_1, int* d_m_1, int* d_n_1, int* d_k_1, const void* alpha_1,
                           const void* const* d_AB_1, int* d_lda_1, int* d_ldb_1, const void* beta_1, void** d_C_1,
                           int* d_ldc_1, int C0_left, int batch_count, const int* segment_sizes, int segment_count);
/* Mixed two-stage forms: real_operand selects the real stage-0 operand (0: A_0,
 * 1: B_0); ab1_real != 0 makes the external stage-1 operand real. */
int ddla_c_gemm_vbatched2s_mixed(void* handle, int real_operand, int ab1_real, char transA_0, char transB_0,
                                 int* d_m_0, int* d_n_0, int* d_k_0, const void* alpha_0, const void* const* d_A_0,
                                 int* d_lda_0, const void* const* d_B_0, int* d_ldb_0, const void* beta_0,
                                 void** d_C_0, int* d_ldc_0, char transA_1, char transB_1, int* d_m_1, int* d_n_1,
                                 int* d_k_1, const void* alpha_1, const void* const* d_AB_1, int* d_lda_1,
                                 int* d_ldb_1, const void* beta_1, void** d_C_1, int* d_ldc_1, int C0_left,
                                 int batch_count, const int* segment_sizes, int segment_count);
int ddla_z_gemm_vbatched2s_mixed(void* handle, int real_operand, int ab1_real, char transA_0, char transB_0,
                                 int* d_m_0, int* d_n_0, int* d_k_0, const void* alpha_0, const void* const* d_A_0,
                                 int* d_lda_0, const void* const* d_B_0, int* d_ldb_0, const void* beta_0,
                                 void** d_C_0, int* d_ldc_0, char transA_1, char transB_1, int* d_m_1, int* d_n_1,
                                 int* d_k_1, const void* alpha_1, const void* const* d_AB_1, int* d_lda_1,
                                 int* d_ldb_1, const void* beta_1, void** d_C_1, int* d_ldc_1, int C0_left,
                                 int batch_count, const int* segment_sizes, int segment_count);

/* ---- batched SPD solves of uneven order (no counterpart in the reference;
 * per member the semantics of pposv and 2 * Re diag_logdet of the factor):
 * factor A_i (order d_n[i] <= max_n <= 1024, triangle uplo), solve A_i X_i =
 * B_i (d_n[i] x nrhs) in place when the factorization succeeds, and write
 * d_logdet[i] = log det A_i; d_info[i] = 0, the failed pivot (B_i untouched,
 * logdet NaN) or -1 (invalid order / leading dimension).  Arrays are device
 * resident; asynchronous on the handle's stream.  _run allocates and
 * synchronizes nothing (graph-capturable); _tune times the schedules on
 * copies of the inputs and keeps the fastest in the handle; _graph captures
 * one call into a replayable hipGraph owning its workspace.  Schedules: the
 * persistent fused kernel for every member, or -- for the members of more
 * than split_blocks 64-row blocks -- a blocked factorization whose trailing
 * updates run as one gemmVbatched launch per 64-column step. ------------ */
int ddla_s_pposv_vbatched(void* handle, char uplo, const int* d_n, void* const* d_A, const int* d_lda, int nrhs,
                          void* const* d_B, const int* d_ldb, int* d_info, double* d_logdet, int batch_count,
                          int max_n);
int ddla_d_pposv_vbatched(void* handle, char uplo, const int* d_n, void* const* d_A, const int* d_lda, int nrhs,
                          void* const* d_B, const int* d_ldb, int* d_info, double* d_logdet, int batch_count,
                          int max_n);
int ddla_s_pposv_vbatched_workspace_size(void* handle, int batch_count, int max_n, long long* backend_bytes);
int ddla_d_pposv_vbatched_workspace_size(void* handle, int batch_count, int max_n, long long* backend_bytes);
int ddla_s_pposv_vbatched_run(void* handle, char uplo, const int* d_n, void* const* d_A, const int* d_lda, int nrhs,
                              void* const* d_B, const int* d_ldb, int* d_info, double* d_logdet, int batch_count,
                              int max_n, void* backend_work, long long backend_bytes);
int ddla_d_pposv_vbatched_run(void* handle, char uplo, const int* d_n, void* const* d_A, const int* d_lda, int nrhs,
                              void* const* d_B, const int* d_ldb, int* d_info, double* d_logdet, int batch_count,
                              int max_n, void* backend_work, long long backend_bytes);
int ddla_s_pposv_vbatched_tune(void* handle, char uplo, const int* d_n, void* const* d_A, const int* d_lda, int nrhs,
                               void* const* d_B, const int* d_ldb, int batch_count, int max_n, int* occupancy,
                               int* ordered);
int ddla_d_pposv_vbatched_tune(void* handle, char uplo, const int* d_n, void* const* d_A, const int* d_lda, int nrhs,
                               void* const* d_B, const int* d_ldb, int batch_count, int max_n, int* occupancy,
                               int* ordered);
int ddla_s_pposv_vbatched_graph(void* handle, char uplo, const int* d_n, void* const* d_A, const int* d_lda,
                                int nrhs, void* const* d_B, const int* d_ldb, int* d_info, double* d_logdet,
                                int batch_count, int max_n, void** graph);
int ddla_d_pposv_vbatched_graph(void* handle, char uplo, const int* d_n, void* const* d_A, const int* d_lda,
                                int nrhs, void* const* d_B, const int* d_ldb, int* d_info, double* d_logdet,
                                int batch_count, int max_n, void** graph);
int ddla_c_graph_launch(void* graph, void* handle);
int ddla_c_graph_destroy(void* graph);
/* The batched schedule the handle holds for FP32 (element_bytes 4) or FP64
 * (8): *tuned = 1 once _tune (or _graph) chose it; workgroups per CU of the
 * persistent kernel, largest-first order, and split_blocks (16: no blocked
 * path). */
int ddla_c_vbatched_schedule(void* handle, int element_bytes, int* tuned, int* occupancy, int* ordered,
                             int* split_blocks);
/* Pins that schedule without timing (occupancy 1 or 2; split_blocks in
 * [1, 16], used only with ordered != 0) on the right-looking kernels; marks it
 * tuned. */
int ddla_c_vbatched_set_schedule(void* handle, int element_bytes, int occupancy, int ordered, int split_blocks);
/* Which kernels the schedule runs: *kernel = 0 for the right-looking
 * persistent kernel with the optional blocked path in steps of one 64-column
 * block column, 2 for the same with steps of two (its trailing updates take
 * k = 128, half as many), 1 for the left-looking persistent kernel (every
 * member in one launch, each block column formed once by one MFMA product;
 * split_blocks unused).  No counterpart in the reference (it has no batched
 * Cholesky). */
int ddla_c_vbatched_kernel(void* handle, int element_bytes, int* kernel);
/* Selects that choice (0, 1 or 2) for the schedule held (an untuned handle
 * first takes 1 workgroup per CU, largest first); marks it tuned. */
int ddla_c_vbatched_set_kernel(void* handle, int element_bytes, int kernel);

/* ---- reverse-mode adjoints (no counterpart in the reference), factor of
 * A = L L^T in the lower triangle (uplo 'L'); each returns synchronized:
 * potrf_backward  Abar = sym(L^-T Phi(L^T Lbar) L^-1), Phi = tril with the
 *                 diagonal halved, sym(S) = (S + S^T)/2;
 * potrs_backward  for X = A^-1 B: Bbar = A^-1 Xbar and
 *                 Lbar -= tril(Bbar (L^T X)^T + X (L^T Bbar)^T);
 * logdet_backward for logdet A = 2 sum log L_ii: Lbar_ii += 2 g / L_ii. ---- */
int ddla_s_potrf_backward(void* handle, char uplo, int n, const void* L, int ldl, const void* Lbar, int ldlb,
                          void* Abar, int ldab);
int ddla_d_potrf_backward(void* handle, char uplo, int n, const void* L, int ldl, const void* Lbar, int ldlb,
                          void* Abar, int ldab);
int ddla_s_potrs_backward(void* handle, char uplo, int n, int nrhs, const void* L, int ldl, const void* X, int ldx,
                          const void* Xbar, int ldxb, void* Bbar, int ldbb, void* Lbar, int ldlb);
int ddla_d_potrs_backward(void* handle, char uplo, int n, int nrhs, const void* L, int ldl, const void* X, int ldx,
                          const void* Xbar, int ldxb, void* Bbar, int ldbb, void* Lbar, int ldlb);
int ddla_s_logdet_backward(void* handle, int n, const void* L, int ldl, double g, void* Lbar, int ldlb);
int ddla_d_logdet_backward(void* handle, int n, const void* L, int ldl, double g, void* Lbar, int ldlb);
/* A^-1 of A = L L^T from its factor (uplo 'L'), written as the full
 * symmetric matrix at Ainv (LAPACK ?potri returns one triangle; no
 * counterpart in the reference).  Formed as tril(L^-1) by recursive
 * triangular inversion and then tril(L^-T L^-1), about n^3 flops; it gives
 * the closed-form Gaussian-process gradient (K^-1 - alpha alpha^T) / 2.
 * Returns synchronized. */
int ddla_s_potri(void* handle, char uplo, int n, const void* L, int ldl, void* Ainv, int lda);
int ddla_d_potri(void* handle, char uplo, int n, const void* L, int ldl, void* Ainv, int lda);

/* ---- batched Gaussian-process likelihood (no counterpart in the reference;
 * callers that evaluate many independent GP likelihoods of one order at once).
 * For problem b < batch (FP64, device memory): K_b column-major n x n (leading
 * dimension n, lower triangle read) at K + b * k_stride, y_b at y + b *
 * y_stride; loss[b] = y_b^T K_b^-1 y_b / 2 + log det K_b / 2.  Each launch
 * serves every problem.  work: batch * (*bytes_per_problem) bytes, 256-byte
 * aligned, kept between the forward and the backward call (it carries the
 * factor, L^-1 and L^-1 y).  forward writes loss[b] and info[b] (potrf's
 * info; the loss is meaningless where it is nonzero) and returns
 * synchronized; backward writes Kbar_b = g[b] (K_b^-1 - a a^T) / 2 (full
 * symmetric, at Kbar + b * kbar_stride) and, unless ybar is null, ybar_b =
 * g[b] a (a = K_b^-1 y_b) for the device upstream gradients g, asynchronously
 * on the handle's stream.  n >= 1, batch >= 0 and batch * (n / 128) <= 65535. */
int ddla_d_gp_batched_workspace_size(int n, long long* bytes_per_problem);
int ddla_d_gp_batched_forward(void* handle, int batch, int n, const void* K, long long k_stride, const void* y,
                              long long y_stride, void* work, void* loss, int* info);
int ddla_d_gp_batched_backward(void* handle, int batch, int n, void* work, const void* g, void* Kbar,
                               long long kbar_stride, void* ybar, long long ybar_stride);

/* ---- harness helpers (no counterpart in the reference) ------------------
 * Philox4x64-10 + Box-Muller normals of the project's seed convention
 * (element e of stream s: counter {e/4, s, 0, 0}, key {seed, 0}), and the
 * synthetic SPD matrix A = G G^T / n + delta I built from them (G from
 * stream 0; `work` must hold n*n doubles).  Test / benchmark setup only. */
int ddla_c_fill_normal(void* handle, unsigned long long seed, unsigned long long stream,
                       unsigned long long first, long long count, double* d_out);
int ddla_d_spd_generate(void* handle, int n, unsigned long long seed, double delta, double* A, int lda,
                        double* work);
/* Batch of the same convention: member i (host arrays n, A, lda, B, ldb of
 * device pointers) gets A_i = G_i G_i^T / n_i + delta I, G_i the normals of
 * Philox stream 2i, and B_i (n_i x nrhs) the normals of stream 2i + 1, formed
 * in FP64 and stored as float when fp32 != 0.  work: 2 max(n_i)^2 doubles. */
int ddla_c_spd_generate_batch(void* handle, int batch, const int* n, unsigned long long seed, double delta, int fp32,
                              void* const* A, const int* lda, int nrhs, void* const* B, const int* ldb, double* work);

/* ---- per-launch event profiler (no counterpart in the reference) --------
 * enable = 1 starts recording an event pair around every kernel launch of
 * the handle (and clears earlier records); ddla_c_profile_read synchronizes
 * and returns, for kernel class cls, the number of launches, their summed
 * duration in ms and their summed algorithmic flops and bytes.  Classes: 0
 * trailing update (the columns beyond the look-ahead panels), 1 panel
 * factorization, 2 triangular-solve sweeps, 3 diagonal-block inverses, 4
 * refinement residual, 5 other, 6 look-ahead updates of the next two panels'
 * columns. */
int ddla_c_profile(void* handle, int enable);
int ddla_c_profile_read(void* handle, int cls, int* count, double* ms, double* flops, double* bytes);

#ifdef __cplusplus
}
#endif
#endif
