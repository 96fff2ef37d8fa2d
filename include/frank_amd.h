/*
 * frank_amd: C ABI of the MI355X (gfx950) BLR factorization engine.
 *
 * Every entry point returns an int status: 0 on success, < 0 on a bad
 * argument / HIP failure / capacity overflow, > 0 for LAPACK-style numerical
 * info (e.g. a zero pivot).  frk_last_error() gives the message of the last
 * failure on the calling thread.  No entry point aborts the process; the
 * reference aborts with "name(Types) undefined!" (src/util/omm_error_handler.cpp:15-30)
 * where this ABI returns FRK_ERR_UNSUPPORTED with the same text.
 *
 * Layouts follow FRANK's Dense (include/FRANK/classes/dense.h:43-66): row-major
 * with a leading dimension.  Device pointers are HIP device memory; host
 * pointers are plain host memory.  All work is ordered on the context's stream.
 * A context (with the matrices and factors it owns) is used by one thread at a
 * time; separate contexts may be driven from separate threads concurrently.
 */
#ifndef FRANK_AMD_H
#define FRANK_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FRK_OK 0
#define FRK_ERR_ARG (-1)
#define FRK_ERR_HIP (-2)
#define FRK_ERR_CAPACITY (-3)
#define FRK_ERR_UNSUPPORTED (-4)

/* Tile kinds (FRANK Dense / LowRank / Empty) */
#define FRK_TILE_DENSE 0
#define FRK_TILE_LOWRANK 1
#define FRK_TILE_EMPTY 2

/* Low-rank addition algorithm: FRANK_LRA (src/operations/arithmetic/addition.cpp:191-203) */
#define FRK_LRA_NAIVE 0
#define FRK_LRA_ROUNDED 1
#define FRK_LRA_FAST 2

/* Kernel generators (src/functions.cpp:92-109 laplacend, :25-34 identity;
 * gaussian has no reference counterpart) */
#define FRK_KERNEL_LAPLACE 0
#define FRK_KERNEL_GAUSSIAN 1
#define FRK_KERNEL_IDENTITY 2
/* d/d ell of the gaussian kernel, exp(-r^2 / (2 ell^2)) r^2 / ell^3 (adjoint; no reference counterpart) */
#define FRK_KERNEL_GAUSSIAN_DELL 3

/* Admissibility (include/FRANK/definitions.h:25 AdmisType) */
#define FRK_ADMIS_POSITION 0
#define FRK_ADMIS_GEOMETRY 1

/* Low-rank storage precision */
#define FRK_PREC_F64 0
#define FRK_PREC_F32 1

typedef struct frk_ctx_s* frk_ctx;
typedef struct frk_blr_s* frk_blr;
typedef struct frk_qrf_s* frk_qrf;

const char* frk_last_error(void);
int frk_version(void);

/* Execution context: one HIP stream, staging ring, workspace.
 * Replaces FRANK::initialize() (src/util/initialize.cpp:9-17). */
int frk_ctx_create(frk_ctx* out);
int frk_ctx_destroy(frk_ctx ctx);
int frk_ctx_synchronize(frk_ctx ctx);
/* The hipStream_t every call on this context is ordered on (for events / interop). */
int frk_ctx_stream(frk_ctx ctx, void** stream);
/* Device event counters of this context (no reference counterpart; the
 * reference's timers in src/util/timer.cpp cover host time only).  Copies
 * min(n, 16) slots after draining the stream, then clears them if reset != 0.
 * Slot 0: small SVD problems solved, 1: Jacobi sweeps summed over them,
 * 2: largest sweep count of one problem. */
int frk_ctx_counters(frk_ctx ctx, int64_t* out, int n, int reset);
/* Op log (no reference counterpart).  Every launch is accounted under its
 * kernel class ("gemm", "getrf_small", "getrf_gemm", "trsm_gemm",
 * "schur_gemm", "solve_gemm", "lr_qr", "lr_svd", "lr_apply_q") with its
 * algorithmic flops and bytes; launches of the one class named by
 * frk_ctx_time_class (NULL or "" = none) are bracketed by HIP events on the
 * context stream.  frk_ctx_op_classes writes the class names, comma
 * separated, into buf (NUL-terminated, truncated to len). */
int frk_ctx_time_class(frk_ctx ctx, const char* cls);

/* Rank policy at the low-rank slot capacity for matrices built on ctx from
 * now on (no reference counterpart: the reference's ranks are unbounded; the
 * capped-rank workloads, e.g. tiled QR with ranks capped at 64, need it).
 * truncate = 0 (default): an eps-rank above kcap fails with FRK_ERR_CAPACITY.
 * truncate = 1: compressions and recompressions keep the leading kcap terms. */
int frk_ctx_set_rank_cap(frk_ctx ctx, int truncate);
int frk_ctx_op_classes(frk_ctx ctx, char* buf, int len);
int frk_ctx_op_stats(frk_ctx ctx, const char* cls, int64_t* launches, double* flops, double* bytes, double* ms,
                     int64_t* timed_launches);
/* Roofline time of one op-log class (no reference counterpart): the sum over
 * its launches of max(flops / peak, bytes / 8 TB/s), with MI355X's published
 * peak of the precision each launch computes in (FP64 78.6, FP32 157.3 TFLOP/s),
 * in milliseconds. */
int frk_ctx_op_roofline(frk_ctx ctx, const char* cls, double* roofline_ms);
int frk_ctx_reset_op_stats(frk_ctx ctx);

/* ---------------------------------------------------------------------------
 * BLR matrices (one-level Hierarchical, tiles resident in HBM)
 * ------------------------------------------------------------------------- */

/* Hierarchical(kernel, params, N, N, nleaf, eps, admis, nb, nb, admis_type)
 * (src/classes/hierarchical.cpp:89-109) when eps != 0, or the fixed-rank
 * constructor (hierarchical.cpp:67-87) with `rank` when eps == 0.
 * pts: host, dim x n structure of arrays.  shift is added on the diagonal. */
int frk_blr_create_kernel(frk_ctx ctx, int kernel_kind, int dim, int64_t n, const double* pts, double ell,
                          double shift, int nb, double eps, int rank, double admis, int admis_type, int kcap,
                          int precision, frk_blr* out);
/* Rectangular form: Hierarchical(kernel, params, n_rows, n_cols, nleaf, eps, admis, nb_r, nb_c, admis_type)
 * (hierarchical.cpp:89-109); rows and columns index the same n_pts points. */
int frk_blr_create_kernel_rect(frk_ctx ctx, int kernel_kind, int dim, int64_t n_pts, const double* pts,
                               int64_t n_rows, int64_t n_cols, int nb_r, int nb_c, double ell, double shift,
                               double eps, int rank, double admis, int admis_type, int kcap, int precision,
                               frk_blr* out);
/* Empty grid with given block offsets (nb+1 entries each) and tile kinds; fill with set_* below. */
int frk_blr_create(frk_ctx ctx, int nb_r, int nb_c, const int64_t* roff, const int64_t* coff, const int32_t* kinds,
                   int kcap, double eps, int rank, int precision, frk_blr* out);
int frk_blr_destroy(frk_blr a);
/* FRANK_LRA selection for this matrix (default FRK_LRA_FAST, as the reference). */
int frk_blr_set_lra(frk_blr a, int lra);
int frk_blr_shape(frk_blr a, int* nb_r, int* nb_c);
int frk_blr_tile_info(frk_blr a, int i, int j, int32_t* kind, int32_t* rank, int64_t* rows, int64_t* cols);
/* Host copies of tiles: dense rows x cols row-major; low-rank U rows x r, S r x r, V r x cols (row-major). */
int frk_blr_get_dense(frk_blr a, int i, int j, double* host);
int frk_blr_set_dense(frk_blr a, int i, int j, const double* host);
int frk_blr_get_lowrank(frk_blr a, int i, int j, double* U, double* S, double* V);
int frk_blr_set_lowrank(frk_blr a, int i, int j, int rank, const double* U, const double* S, const double* V);

/* Deep copy (the reference's Hierarchical copy constructor clones every block,
 * src/classes/matrix_proxy.cpp:68-76): new slots with the same structure. */
int frk_blr_clone(frk_blr src, frk_blr* out);
/* Restore dst's contents from src (same structure): one device-to-device copy. */
int frk_blr_copy(frk_blr dst, frk_blr src);
/* Same structure in another storage precision (FRK_PREC_F64 / FRK_PREC_F32),
 * tiles converted element-wise (no reference counterpart). */
int frk_blr_convert(frk_blr src, int precision, frk_blr* out);

/* getrf(Hierarchical&) (src/operations/LAPACK/getrf.cpp:45-65): tile LU without
 * pivoting, in place.  Afterwards tiles below the diagonal hold L, tiles above
 * hold U, diagonal tiles hold the packed unit-L \ U factors.  Returns > 0
 * (1-based global row) when a zero pivot is met. */
int frk_blr_getrf(frk_blr a);
/* The first `steps` elimination steps of getrf.cpp:45-65 only (steps < 0 or
 * >= the block count: all of them).  A partial factorization for parity checks
 * against a CPU restatement stopped at the same step; the diagonal tile of
 * step `steps` is factorized as well.  No reference counterpart. */
int frk_blr_getrf_steps(frk_blr a, int steps);

/* tiled_householder_blr_qr(A, T) (include/FRANK/operations/LAPACK.h:450,
 * src/operations/LAPACK/qr.cpp:213-228): FP64, in place.  Afterwards the
 * diagonal tiles hold R above / reflectors below the diagonal (geqrt layout),
 * the tiles above the diagonal R, the tiles below the diagonal the reflector
 * blocks of tpqrt (a low-rank block keeps U and gets S = I).  *out receives
 * the triangular factors T(i, k), i >= k. */
int frk_blr_qr(frk_blr a, frk_qrf* out);
int frk_qrf_destroy(frk_qrf f);
/* T(i, k) (b x b row-major) and the reflector scalars tau(i, k) (b) to host;
 * either output may be NULL. */
int frk_qrf_get_t(frk_qrf f, int i, int k, double* t_out, double* tau_out);
/* left_multiply_tiled_reflector(Y, T, C, trans) (LAPACK.h:470-472, qr.cpp:230-257):
 * C <- Q C (trans = 0) or Q^T C, with Y the factored matrix of frk_blr_qr. */
int frk_blr_apply_q(frk_blr y, frk_qrf f, frk_blr c, int trans);

/* trsm(L, b, Lower); trsm(U, b, Upper) (src/operations/BLAS/trsm.cpp:40-90,117-125)
 * on device b (N x nrhs row-major, FP64).  trans != 0 solves A^T x = b with
 * U^T then L^T sweeps (adjoint; no reference counterpart). */
int frk_blr_solve(frk_blr f, double* b_dev, int nrhs, int trans);
/* One sweep: trsm(L, b, Mode::Lower) (upper = 0: unit-lower forward substitution)
 * or trsm(U, b, Mode::Upper) (upper = 1: backward), trsm.cpp:40-90; trans = 1
 * uses L^T / U^T (no reference counterpart). */
int frk_blr_trsm(frk_blr f, double* b_dev, int nrhs, int upper, int trans);

/* gemm(Hierarchical, Dense, Dense, 1, 1): y += A x (device, N x nrhs row-major). */
int frk_blr_matvec(frk_blr a, const double* x_dev, double* y_dev, int nrhs);

/* Matrix-free FP64 residual r = b - A x for the kernel matrix the BLR matrix
 * was built from (exact, uncompressed).  Device vectors, N x nrhs row-major. */
int frk_blr_kernel_residual(frk_blr a, const double* x_dev, const double* b_dev, double* r_dev, int nrhs);
/* y = K x for the exact kernel matrix K of n host points (dim x n, structure of
 * arrays), matrix-free, FP64 device vectors (n x nrhs row-major, nrhs <= 16).
 * With FRK_KERNEL_GAUSSIAN_DELL this is (dA/d ell) x, the contraction the
 * adjoint of the solve needs (no reference counterpart). */
int frk_kernel_apply(frk_ctx ctx, int kernel_kind, int dim, int64_t n, const double* pts, double ell, double shift,
                     const double* x_dev, double* y_dev, int nrhs);

/* Solve with iterative refinement (no reference counterpart): x = F^-1 b, then
 * x += F^-1 (b - A x) with the exact FP64 kernel matrix, until the normwise
 * backward error max_c ||b_c - A x_c||_inf / (||A||_inf ||x_c||_inf + ||b_c||_inf)
 * <= tol or max_iter corrections.  F: factorized matrix built from a kernel
 * (FP64 or FP32 storage).  b, x: device, N x nrhs row-major FP64. */
int frk_blr_refine(frk_blr f, const double* b_dev, double* x_dev, int nrhs, double tol, int max_iter, int* iters,
                   double* berr);

/* ---------------------------------------------------------------------------
 * Tile kernels (FP64, device pointers, row-major with leading dimension)
 * ------------------------------------------------------------------------- */

/* getrf_omm(Dense) (src/operations/LAPACK/getrf.cpp:67-84) without the pivot
 * search: A (n x n) overwritten by unit-L \ U.  info_host[b] = first zero pivot (1-based). */
int frk_dgetrf_nopiv_batched(frk_ctx ctx, double* const* tiles_host, int batch, int n, int64_t ld,
                             int32_t* info_host);

/* The batched no-pivot tile LU without any allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * stream capture, tight serving loops).  Same operation, status and info as
 * frk_dgetrf_nopiv_batched; the tiles are factorized one after another with
 * the library's fixed per-GPU scratch (the inverses of both factors, the block
 * flags and the info words: about 4 MiB), so n <= 512 and batch <= 1024
 * (FRK_ERR_CAPACITY beyond).  The scratch keeps the last call's contents and
 * is shared by every call on one GPU: calls that overlap on one GPU race. */
int frk_dgetrf_nopiv_batched_noalloc(frk_ctx ctx, double* const* tiles_host, int batch, int n, int64_t ld,
                                     int32_t* info_host);

/* trsm_omm(Dense, Dense) (src/operations/BLAS/trsm.cpp:92-104): uplo lower =>
 * unit diagonal, upper => non-unit.  side 0 = left (A X = B), 1 = right (X A = B).
 * trans != 0 uses op(A) = A^T (adjoint sweeps; no reference counterpart). */
int frk_dtrsm(frk_ctx ctx, int side, int lower, int trans, int m, int n, const double* A, int64_t lda, double* B,
              int64_t ldb);

/* gemm_omm(Dense, Dense, Dense) (src/operations/BLAS/gemm.cpp:126-146):
 * C = alpha op(A) op(B) + beta C, C m x n. */
int frk_dgemm(frk_ctx ctx, int trans_a, int trans_b, int m, int n, int k, double alpha, const double* A,
              int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc);

/* LowRank += LowRank (src/operations/arithmetic/addition.cpp:93-203), batched.
 * Every pair b: A_b = Ua Sa Va (rank ka, capacity cap: Ua m x cap, Sa cap x cap,
 * Va cap x n row-major, device), B_b = Ub Sb Vb (rank kb).  A is updated in
 * place; ranks_host[b] holds ka on entry and the new rank on exit.
 * eps != 0 selects the rounded addition with the truncation rule of
 * src/operations/misc/misc.cpp:133-146; eps == 0 keeps ka (lra selects rounded or fast);
 * lra == FRK_LRA_NAIVE concatenates (addition.cpp:68-89). */
int frk_dlr_add_batched(frk_ctx ctx, int lra, int batch, const int32_t* m, const int32_t* n, double eps, int cap,
                        double* const* Ua, double* const* Sa, double* const* Va, int32_t* ranks_host,
                        const double* const* Ub, const double* const* Sb, const double* const* Vb,
                        const int32_t* kb);
/* The same in FP32 storage and arithmetic (no reference counterpart: the
 * reference is FP64 only). */
int frk_slr_add_batched(frk_ctx ctx, int lra, int batch, const int32_t* m, const int32_t* n, double eps, int cap,
                        float* const* Ua, float* const* Sa, float* const* Va, int32_t* ranks_host,
                        const float* const* Ub, const float* const* Sb, const float* const* Vb,
                        const int32_t* kb);

/* ---------------------------------------------------------------------------
 * Tile-level operators on caller-owned tiles (FP64, device memory).
 *
 * A tile is FRANK's Dense or LowRank (include/FRANK/classes/dense.h,
 * low_rank.h).  Dense: d, rows x cols row-major with leading dimension ld.
 * LowRank: slots of capacity cap, U rows x cap, S cap x cap and V cap x cols,
 * all row-major (as frk_dlr_add_batched takes them), rank <= cap; eps is
 * the LowRank's accuracy (LowRank::eps: 0 keeps the rank fixed, else the
 * rounded addition truncates at eps).  Operations that change a low-rank
 * tile write its new rank back into the descriptor.
 * ------------------------------------------------------------------------- */
typedef struct frk_dtile {
  int32_t kind; /* FRK_TILE_DENSE or FRK_TILE_LOWRANK */
  int32_t rows, cols;
  int32_t rank, cap;
  int32_t reserved;
  double eps;
  double* d;
  int64_t ld;
  double* u;
  double* s;
  double* v;
} frk_dtile;

/* gemm(A, B, C, alpha, beta, TransA, TransB) (include/FRANK/operations/BLAS.h:49-53,
 * src/operations/BLAS/gemm.cpp:126-272), every Dense / LowRank combination of
 * A, B and C, batched over independent items: C = alpha op(A) op(B) + beta C.
 * A low-rank C is scaled first (C.S *= beta, src/operations/arithmetic/
 * multiplication.cpp:39-44) and receives the product through the addition
 * selected by lra (gemm(LowRank, LowRank, LowRank), gemm.cpp:253-272, is the
 * core product S_A V_A U_B S_B and one addition). */
int frk_dgemm_tiles_batched(frk_ctx ctx, int batch, const frk_dtile* A, const frk_dtile* B, frk_dtile* C,
                            double alpha, double beta, int trans_a, int trans_b, int lra);
/* geqrt(A, T) (include/FRANK/operations/LAPACK.h:345, src/operations/LAPACK/geqrt.cpp:34-43):
 * n x n tiles (n <= 512); A <- R above / Householder vectors below the
 * diagonal (LAPACK dgeqrt3 layout), T <- the n x n upper triangular factor. */
int frk_dgeqrt_batched(frk_ctx ctx, int batch, int n, double* const* A, int64_t lda, double* const* T, int64_t ldt);
/* larfb(V, T, C, trans) (LAPACK.h:374, src/operations/LAPACK/larfb.cpp:37-56):
 * C <- (I - Y op(T) Y^T) C with Y the unit-lower part of V (a geqrt output),
 * op(T) = T^T when trans; a low-rank C is updated through its U. */
int frk_dlarfb_batched(frk_ctx ctx, int batch, int n, const double* const* V, int64_t ldv, const double* const* T,
                       int64_t ldt, frk_dtile* C, int trans);
/* tpqrt(A, B, T) (LAPACK.h:399, src/operations/LAPACK/tpqrt.cpp:35-53): QR of
 * [R; B] with R the upper triangle of the n x n tile A (updated in place) and
 * B a dense m x n tile (overwritten by its reflector parts, m <= 512) or a
 * low-rank tile (V <- reflector parts of S V, S <- I); T <- n x n factor. */
int frk_dtpqrt_batched(frk_ctx ctx, int batch, int n, double* const* A, int64_t lda, frk_dtile* B,
                       double* const* T, int64_t ldt);
/* tpmqrt(V, T, A, B, trans) (LAPACK.h:430, src/operations/LAPACK/tpmqrt.cpp:45-170):
 * [A; B] <- H^T [A; B] (trans) or H [A; B] with H = I - [I; V] op(T) [I; V]^T:
 * C = A + V^T B, C = op(T) C, A -= C, B -= V C, every Dense / LowRank
 * combination of V, A and B (A n x nc, B m x nc, V m x n, T n x n). */
int frk_dtpmqrt_batched(frk_ctx ctx, int batch, int n, const frk_dtile* V, const double* const* T, int64_t ldt,
                        frk_dtile* A, frk_dtile* B, int trans, int lra);

/* ---------------------------------------------------------------------------
 * Batched small dense factorizations (operator surface).  Device pointers,
 * row-major with leading dimensions; one entry per problem.
 * ------------------------------------------------------------------------- */
/* qr(A, Q, R) (include/FRANK/operations/LAPACK.h:232, src/operations/LAPACK/qr.cpp:65-90):
 * Q m x k, R k x n, k = min(m, n); m <= 512, n <= 128. */
int frk_dqr_batched(frk_ctx ctx, int batch, const int32_t* m, const int32_t* n, const double* const* A,
                    const int64_t* lda, double* const* Q, const int64_t* ldq, double* const* R, const int64_t* ldr);
/* rq(A, R, Q) (LAPACK.h:488, rq.cpp:43-70): R m x m upper, Q m x n; m <= min(n, 128), n <= 512. */
int frk_drq_batched(frk_ctx ctx, int batch, const int32_t* m, const int32_t* n, const double* const* A,
                    const int64_t* lda, double* const* R, const int64_t* ldr, double* const* Q, const int64_t* ldq);
/* svd(A) (LAPACK.h:99, svd.cpp:23-44): U m x K, S K x K diagonal descending, V K x n;
 * K = min(m, n) <= 64, max(m, n) <= 512. */
int frk_dsvd_batched(frk_ctx ctx, int batch, const int32_t* m, const int32_t* n, const double* const* A,
                     const int64_t* lda, double* const* U, const int64_t* ldu, double* const* S, const int64_t* lds,
                     double* const* V, const int64_t* ldv);
/* truncated_geqp3(A, eps) (LAPACK.h:170, geqp3.cpp:71-191): Q m x r, R r x n (columns in the
 * original order), r <= cap written to ranks_host[b]. */
int frk_dtruncated_geqp3_batched(frk_ctx ctx, int batch, const int32_t* m, const int32_t* n, const double* const* A,
                                 const int64_t* lda, double eps, int cap, double* const* Q, const int64_t* ldq,
                                 double* const* R, const int64_t* ldr, int32_t* ranks_host);
/* rsvd(A, p) (include/FRANK/operations/randomized_factorizations.h:110, rsvd.cpp:17-29),
 * Omega from std::mt19937(0): U m x p, S p x p, V p x n; p <= min(m, n, 64). */
int frk_drsvd(frk_ctx ctx, int m, int n, const double* A, int64_t lda, int p, double* U, int64_t ldu, double* S,
              int64_t lds, double* V, int64_t ldv);
/* one_sided_id(A, k) (LAPACK.h:61, id.cpp:47-72): V k x n with A ~ A[:, cols] V, the k
 * selected columns (0-based) to cols_host. */
int frk_did(frk_ctx ctx, int m, int n, const double* A, int64_t lda, int k, double* V, int64_t ldv,
            int32_t* cols_host);

/* LowRank(Dense, eps) (src/classes/low_rank.cpp:50-59, truncated_geqp3 + rq)
 * when eps != 0, else LowRank(Dense, rank) (low_rank.cpp:40-48, rsvd).
 * A (m x n row-major, device) is overwritten.  Outputs U (m x cap), S (cap x cap),
 * V (cap x n) row-major, device; *rank_out receives the rank. */
int frk_dcompress(frk_ctx ctx, int m, int n, double* A, int64_t lda, double eps, int rank, int cap, double* U,
                  double* S, double* V, int32_t* rank_out);

#ifdef __cplusplus
}
#endif

#endif /* FRANK_AMD_H */
