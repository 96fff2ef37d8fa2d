// Host-side launchers for the gfx950 kernels (kernels.hip).  All launchers
// are asynchronous on `stream`; item arrays are device pointers.
#pragma once
#include "common.hpp"

namespace frk {

// Strided batched GEMM (one 64x64 output tile per workgroup).
template <typename T>
void launch_gemm(const GemmItem* d_items, int n_items, int max_m, int max_n, hipStream_t s);

// Batched triangular solve, left side, in place (strips of 16 right-hand sides),
// blocked by 64 x 64 diagonal blocks with precomputed inverses.
template <typename T>
void launch_trsm(const TrsmItem* d_items, int n_items, int max_n, int max_nrhs, hipStream_t s);

// Inverses of the 64 x 64 diagonal blocks of triangles.
template <typename T>
void launch_tri_inv(const TriInvItem* d_items, int n_items, hipStream_t s);

// No-pivot LU of <= 64 x 64 blocks (one wave each) plus the inverses of the
// unit-lower and upper factors; info[i] = first zero pivot (1-based) or 0.
template <typename T>
void launch_getrf_small(const LuItem* d_items, int n_items, int32_t* d_info, hipStream_t s);

// Batched strided copy C <- A (GemmItem: A, a_rs, a_cs, C, c_rs, c_cs, m, n).
template <typename T>
void launch_copy(const GemmItem* d_items, int n_items, int64_t max_elems, hipStream_t s);

// Householder QR of [A1 S1 | A2 S2] (optionally projected), m <= 512.
// Dataflow LU + full triangular inverses of one FP64 tile (n <= 512), nbk^2 workgroups.
void launch_tile_lu(const TileLuItem& it, hipStream_t s);
template <typename T>
void launch_qr(const QrItem* d_items, int n_items, int max_k, hipStream_t s);
template <typename T>
void launch_tpqrt(const TpqrtItem* d_items, int n_items, int max_m, int max_k, hipStream_t s);
template <typename T>
void launch_larft(const LarftItem* d_items, int n_items, hipStream_t s);
template <typename T>
void launch_tri_extract(const T* src, int64_t rs, int64_t cs, T* dst, int64_t ld, int n, int mode, hipStream_t s);

// Apply Householder products to zero-padded blocks, m <= 512.
template <typename T>
void launch_apply_q(const ApplyQItem* d_items, int n_items, int max_ncol, hipStream_t s);

// One-sided Jacobi SVD with truncation, K <= 64.
template <typename T>
void launch_svd(const SvdItem* d_items, int n_items, int max_k, hipStream_t s);

// Truncated column-pivoted QR (LAWN-176 downdating), one workgroup per tile.
template <typename T>
void launch_geqp3(const Geqp3Item* d_items, int n_items, hipStream_t s);

// Gather the first `rank` rows of the triangular factor in original column order.
struct PermuteItem {
  const void* A;   // geqp3-processed tile (row-major, ld)
  const int32_t* piv;
  const int32_t* rank;
  void* out;       // rank x n row-major, ld = ldo
  int32_t n, lda, ldo, pad;
};
template <typename T>
void launch_permute_r(const PermuteItem* d_items, int n_items, int max_n, int max_rows, hipStream_t s);

// S = flip(Rz^T): S[i][j] = Rz[r-1-j][r-1-i] (upper triangular), r from *rank.
struct FlipItem {
  const void* R;
  void* S;
  const int32_t* rank;
  int32_t ldr, lds;
};
template <typename T>
void launch_flip_r(const FlipItem* d_items, int n_items, int max_r, hipStream_t s);

// Kernel-matrix tile generation.
struct GenItem {
  void* out;
  int64_t ld;
  int32_t r0, nr, c0, nc;
};
struct KernelSpec {
  const double* pts;  // dim arrays of N points, structure of arrays
  int64_t n_pts;
  int32_t dim;
  int32_t kind;       // 0 laplace, 1 gaussian, 2 identity, 3 d/d ell of the gaussian
  double ell;         // gaussian length scale
  double shift;       // added on the global diagonal
};
template <typename T>
void launch_gen(const KernelSpec& ks, const GenItem* d_items, int n_items, int max_nr, int max_nc,
                hipStream_t s);

// *items[i].p = items[i].v (rank updates written from the host plan).
struct SetIntItem {
  int32_t* p;
  int32_t v;
  int32_t pad;
};
void launch_set_int(const SetIntItem* d_items, int n, hipStream_t s);

// Matrix-free residual R = B - A X for a kernel matrix (FP64), X/B/R row-major N x nrhs;
// with B == nullptr, R = A X.
// w = op(Tinv) b for one diagonal block (n <= 512 rows, nrhs <= 16), op = transpose when trans.
template <typename T>
void launch_diag_apply(const T* Tinv, int64_t ld, bool trans, const T* b, T* w, int n, int nrhs, hipStream_t s);
// One workgroup per SolveUpdItem (n <= 512, r <= 64, nrhs <= 16).
template <typename T>
void launch_solve_update(const SolveUpdItem* d_items, int n_items, int nrhs, hipStream_t s);
// Stacked-row product W (R x N, row-major) = [rows[0].a; ...; rows[R-1].a] (R x K) B (K x N, strided),
// then (scatter) each row of W copied to rows[r].c.  K <= 512.
// d_count (optional): row count read on the device, at most R (the table's length).
template <typename T>
void launch_rows_gemm(const RowRef* d_rows, int R, const T* B, int64_t b_rs, int64_t b_cs, T* W, int N, int K,
                      hipStream_t s, const int32_t* d_count = nullptr);
template <typename T>
void launch_rows_scatter(const RowRef* d_rows, int R, const T* W, int N, hipStream_t s,
                         const int32_t* d_count = nullptr);
// The three launches of a batch of low-rank additions (QR of both bases, SVD of
// the core, basis update), items derived on the device from the descriptors.
template <typename T>
void launch_lr_add_qr(const LrAddDesc<T>* d, int nd, const LrBatch<T>& b, int max_k, hipStream_t s);
template <typename T>
void launch_lr_add_core(const LrAddDesc<T>* d, int nd, const LrBatch<T>& b, hipStream_t s);
template <typename T>
// SVDs of items [t0, t0 + n) of a batch, in the kernel for cores of up to max_k
void launch_lr_add_svd(const LrAddDesc<T>* d, int t0, int n, int max_k, const LrBatch<T>& b, hipStream_t s);
template <typename T>
void launch_lr_add_apply(const LrAddDesc<T>* d, int nd, const LrBatch<T>& b, int max_ncol, hipStream_t s);
// One fused block step of a substitution sweep (see StepDiag).
template <typename T>
void launch_solve_step(const SolveUpdItem* d_items, int n_upd, const StepDiag& sd, int nrhs, hipStream_t s);
// Dataflow block substitution: clears the nflags flag words, then one persistent
// launch of `grid` workgroups over the sweep's tasks (solve_df.hip).
template <typename T>
void launch_solve_df(const DfSweep& sw, int nflags, int nrhs, int grid, hipStream_t s);
// Two strided stacked-row products in one launch (see RowsPair).
template <typename T>
void launch_rows_pair(const RowsPair& p, hipStream_t s);
template <typename T>
void launch_rows_gemm2(const RowsList2& p, hipStream_t s);
template <typename T>
void launch_panel_finish(const PanelFin* d_items, int n_items, int max_n, hipStream_t s);
// Low-rank x low-rank Schur update with device ranks (kcap <= 64):
// core = alpha S_l (V_l U_b) S_b, then C += U_l core V_b on 64 x 64 tiles.
template <typename T>
void launch_lr_core(const LrCoreItem& it, hipStream_t s);
template <typename T>
void launch_lr_update(const LrUpdItem& it, hipStream_t s);
// R = B - A X (B null: R = A X) with the kernel matrix generated on the fly; part
// holds kernel_residual_scratch(n, nrhs) doubles of partial sums, stream-ordered
// with s (the caller's context owns it)
size_t kernel_residual_scratch(int64_t n, int nrhs);
void launch_kernel_residual(const KernelSpec& ks, const double* X, const double* B, double* R,
                            int64_t n, int nrhs, double* part, hipStream_t s);

// Per-column max |a| over row blocks: out[b * nrhs + c] for b < kColmaxBlocks.
constexpr int kColmaxBlocks = 64;
void launch_colmax_abs(const double* a, int64_t n, int nrhs, double* out, hipStream_t s);
void launch_fill_value(double* p, int64_t n, double v, hipStream_t s);
// x += (double) d
template <typename T>
void launch_axpy_convert(double* x, const T* d, int64_t n, hipStream_t s);

// Elementwise helpers.
template <typename T>
void launch_fill_identity(T* p, int64_t rows, int64_t cols, int64_t ld, hipStream_t s);
template <typename TD, typename TS>
void launch_convert(TD* dst, const TS* src, int64_t n, hipStream_t s);
// Copy n16 16-byte words from pinned host memory (a device-visible pointer) to
// device memory with a kernel on stream s.
void launch_stage_copy(void* dst, const void* src_host, size_t n16, hipStream_t s);
// dst_host[0, n) = src[0, n), dst_host[n] = *tail: device words into pinned host
// memory (a device-visible pointer) with a kernel on stream s.
void launch_snapshot(int32_t* dst_host, const int32_t* src, int n, const int32_t* tail, hipStream_t s);

}  // namespace frk
