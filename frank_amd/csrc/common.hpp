// Shared device/host definitions for the gfx950 BLR kernels.
//
// Matrices are addressed through strided views: element (i, j) of a view
// lives at p[i * rs + j * cs].  Row-major FRANK Dense tiles use (rs = ld,
// cs = 1); the engine's low-rank bases use column-major slots (rs = 1,
// cs = ld) so every basis vector is one contiguous, coalesced stream.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace frk {

#define FRK_HIP_CHECK(expr)                                                        \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess) {                                                        \
      throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(_e) + \
                               " at " + __FILE__ + ":" + std::to_string(__LINE__)); \
    }                                                                              \
  } while (0)

constexpr int kWave = 64;

// The device keeps kCounterCopies copies of the counters, kCounterStride ints
// apart (one 128-byte line each), so that the workgroups of one launch update
// different lines (copy blockIdx % kCounterCopies); a read sums (maxes) them.
constexpr int kCounterCopies = 64, kCounterStride = 32, kCounterWords = kCounterCopies * kCounterStride;

// ---------------------------------------------------------------------------
// Batched GEMM item: C = alpha * A * B + beta * C, A m x k, B k x n.
// ---------------------------------------------------------------------------
struct GemmItem {
  const void* A;
  const void* B;
  void* C;
  int64_t a_rs, a_cs, b_rs, b_cs, c_rs, c_cs;
  int32_t m, n, k, pad;
  double alpha, beta;
};

// ---------------------------------------------------------------------------
// Batched triangular solve item (left side, op(A) already folded into the
// strides): solve T X = B in place, T n x n, B n x nrhs.  lower: forward
// substitution; upper: backward.  dinv holds the inverses of T's 64 x 64
// diagonal blocks (block k at dinv + k * 4096, element (i, j) at
// [i * d_rs + j * d_cs]); the unit flag is already folded into them.
// ---------------------------------------------------------------------------
constexpr int kTriBlk = 64;

struct TrsmItem {
  const void* A;
  const void* dinv;
  void* B;
  int64_t a_rs, a_cs, b_rs, b_cs, d_rs, d_cs;
  int32_t n, nrhs;
  int32_t lower, unit;
};

// No-pivot LU of one <= 64 x 64 diagonal block plus the inverses of its
// unit-lower and upper factors (row-major 64 x 64 outputs).
struct LuItem {
  void* a;
  void* linv;  // inverse of the unit-lower factor, row stride ld_inv (may be nullptr)
  void* uinv;  // inverse of the upper factor, row stride ld_inv (may be nullptr)
  int64_t rs, cs, ld_inv;
  int32_t n, pad;
};

// Inverse of one <= 64 x 64 diagonal block of a triangle (strided).
struct TriInvItem {
  const void* a;
  void* out;  // row-major 64 x 64
  int64_t rs, cs;
  int32_t n, lower, unit, pad;
};

// ---------------------------------------------------------------------------
// Householder QR item.  The m x K input is the column concatenation
// [A1 S1 | A2 S2] formed on load (S == nullptr: identity; A1 has k1a columns
// and S1 is k1a x k1, likewise for block 2).  Optional projection first:
// every column x becomes x - P (P^T x) (modified Gram-Schmidt against an
// orthonormal P, m x kp) with the coefficients in rows [0, kp) of the R
// buffer.  Outputs: reflectors Y (element (row, c) at Y[row*y_rs + c*y_cs],
// rows below the diagonal only; unit diagonal implicit), tau[K], R (rows
// [kp, kp + K) of the R buffer, row-major with ld = ldr; with r_upper_only
// only the upper triangle is written, so R and Y may share one tile).
// ---------------------------------------------------------------------------
struct QrItem {
  const void* A1;
  const void* S1;
  const void* A2;
  const void* S2;
  const void* P;
  void* Y;
  void* tau;
  void* R;
  int64_t a1_rs, a1_cs, a2_rs, a2_cs, p_rs, p_cs;
  int32_t s1_ld, s2_ld;
  int32_t m, k1a, k1, k2a, k2, kp;
  int64_t y_rs, y_cs;
  int32_t ldr, r_upper_only;
};

// ---------------------------------------------------------------------------
// Pentagonal QR (LAPACK tpqrt, l = 0) of [Rt; B]: Rt is the K x K upper
// triangle on top (strided, upper part read and updated in place), B the
// m x K block below (overwritten by the reflectors' lower parts).  Reflector
// j is [e_j; v_j]; tau[K] receives the scalars.
// ---------------------------------------------------------------------------
struct TpqrtItem {
  void* R;
  void* B;
  void* tau;
  int64_t r_rs, r_cs, b_rs, b_cs;
  int32_t m, K;
};

// ---------------------------------------------------------------------------
// Triangular factor of a block of K <= 64 reflectors (LAPACK larft, forward,
// columnwise): T[j][j] = tau_j, T[0:j, j] = -tau_j T[0:j, 0:j] G[0:j, j] with
// G[c][j] = w_c . w_j (strided, only c < j read).  T written upper, zeros below.
// ---------------------------------------------------------------------------
struct LarftItem {
  const void* G;
  const void* tau;
  void* T;
  int64_t g_rs, g_cs, t_rs, t_cs;
  int32_t K, pad;
};

// ---------------------------------------------------------------------------
// Apply the Householder product Q = H_0 ... H_{nref-1} (or Q^T) to the
// zero-padded m x ncol matrix [X; 0] (X is kx x ncol), optionally adding
// Base (m x kb) * Xb (kb x ncol).  Result written through a strided view.
// ---------------------------------------------------------------------------
struct ApplyQItem {
  const void* Y;     // reflector j is column j of the strided view (y_rs, y_cs)
  const void* tau;
  const void* X;
  const void* Base;
  const void* Xb;
  void* out;
  int64_t y_rs, y_cs, x_rs, x_cs, base_rs, base_cs, xb_rs, xb_cs, out_rs, out_cs;
  const int32_t* ncol_ptr;  // when set, the column count is read on the device
  int32_t m, nref, kx, ncol, kb;
  int32_t trans;
};

// ---------------------------------------------------------------------------
// SVD + truncation item.  The K x K core M is built in the kernel:
//   mode 0: M given (row-major, ld = ldm)
//   mode 1: M = Ru * Rv^T               (rounded addition)
//   mode 2: fast rounded addition block core
//           [[Sa + Zu Zv^T, Zu Rv^T], [Ru Zv^T, Ru Rv^T]] with ka = k_a
// Outputs: rank (fixed when fixed_rank >= 0, else the H2 rule at eps), the
// r x r diagonal S (ld = lds), left vectors Um (K x r, row-major, ld = ldu),
// right vectors Vm (K x r, ld = ldv).
// ---------------------------------------------------------------------------
struct SvdItem {
  const void* M;
  const void* Ru;
  const void* Rv;
  const void* Sa;
  const void* Zu;
  const void* Zv;
  void* S;
  void* Um;
  void* Vm;
  void* vwork;         // K x K scratch for the right vectors when they do not fit LDS (FP64, K > 64)
  int32_t* rank_out;
  int32_t* stats;      // optional: [0] += 1 problem, [1] += sweeps, [2] = max sweeps
  int64_t m_rs, m_cs;  // mode 0 input strides
  int32_t K, ldr, lds_in, kdim;  // kdim > 0: min-dimension used by the truncation rule
  int32_t ka, kx, fixed_rank, mode;
  int32_t lds, ldu, ldv, cap;
  int32_t cap_trunc;  // a rank above cap keeps the leading cap terms (else the rank is reported negated)
  int32_t pad2;
  double eps;
};

// ---------------------------------------------------------------------------
// One low-rank addition C += X (rounded / fast rounded): target C = U S V in its
// slots (rank kc known at launch), addend X = Ux Sx Vx with Ux m x kx1 (strided),
// Sx kx1 x kx2 (nullptr = identity), Vx kx2 x n (strided).  Element (i, t) of U
// at cu[i*cu_rs + t*cu_cs]; (t, j) of V likewise.  The QR, SVD and basis-update
// kernels of a batch derive their items from these on the device (LrBatch).
// ---------------------------------------------------------------------------
template <typename T>
struct LrAddDesc {
  // target C = U S V (slots, rank kc known on the host at launch time)
  T* cu;
  T* cs;
  T* cv;
  int32_t* d_rank;
  int kc;
  int64_t cu_rs, cu_cs, cv_rs, cv_cs;  // element (i, t) of U at cu[i*cu_rs + t*cu_cs]; (t, j) of V likewise
  int lds;
  int m, n;
  double eps;  // != 0: fixed accuracy (H2 rule); 0: keep kc
  int cap;
  bool cap_trunc = false;  // keep the leading cap terms when the eps-rank exceeds cap
  // addend X = Ux Sx Vx, Ux m x kx1 (strided), Sx kx1 x kx2 (nullptr = identity),
  // Vx kx2 x n (strided)
  const T* xu;
  int64_t xu_rs, xu_cs;
  int kx1;
  const T* xs;
  int xs_ld;
  int kx2;
  const T* xv;
  int64_t xv_rs, xv_cs;
};

// Per-batch workspace of a batched low-rank addition: addition t owns
// Yu/Yv + t*mmax*K / t*nmax*K (reflectors), tu/tv + t*K, Ru/Rv/Um/Vm + t*K*K
// (and Vw, Mw + t*K*K when set).  lra: 1 rounded, 2 fast (when eps == 0).
// An addition whose K = kc + kx2 exceeds a tile dimension has a core of
// Ke = max(min(m, K), min(n, K)) < K: its M = Ru Rv^T (Ke x Ke) is formed into
// Mw first and the SVD runs on it (Kmax_svd: the largest core of the batch).
template <typename T>
struct LrBatch {
  T *Yu, *Yv, *tu, *tv, *Ru, *Rv, *Um, *Vm, *Vw, *Mw;
  int32_t* stats;
  int32_t mmax, nmax, K, lra;
  int32_t Kmax_svd, pad;
};

// ---------------------------------------------------------------------------
// Truncated column-pivoted QR item (construction).  A (m x n, row-major,
// ld = lda) is overwritten.  Outputs rank, tau, pivot order.
// ---------------------------------------------------------------------------
struct Geqp3Item {
  void* A;
  void* tau;
  int32_t* piv;
  int32_t* rank_out;
  int32_t m, n, lda, cap;
  int32_t cap_trunc;  // stop after cap pivots instead of reporting the overflow
  int32_t pad;
  double eps;
};

// ---------------------------------------------------------------------------
// Dataflow LU of one square tile (n <= 512) without pivoting: one workgroup
// per 64 x 64 block, ordered by per-block flags in global memory.  Outputs
// the packed unit-L \ U factors in place and the full inverses of both
// factors (row stride ldi).  Flags (2 nbk^2 ints, zero at allocation) are
// compared against targets derived from epoch, so no reset is needed between
// launches as long as epoch grows.  A wait that exceeds spin_limit polls sets
// *err and abandons the factorization (every wave still exits).
// ---------------------------------------------------------------------------
struct TileLuItem {
  void* a;
  void* linv;
  void* uinv;
  int32_t* flags;
  int32_t* info;  // first zero pivot (1-based row), 0 if none
  int32_t* err;   // set when a wait timed out
  int64_t ld, ldi;
  int32_t n, epoch;
  int64_t spin_limit;
  int64_t* trace;  // optional: per workgroup 4 wall-clock stamps (start, updates done, factor published,
                   // inverse published)
  // optional FP32 input (leading dimension ld) and FP32 copies of the results
  // (L\U: ld, inverses: ldi); the FP64 a / linv / uinv then serve as the
  // workgroups' exchange buffers and need no initial contents
  const float* a32_in;
  float* a32_out;
  float* linv32;
  float* uinv32;
  // optional low-rank update applied to the tile before its factorization:
  // A += U core V with U (n x r1, element (i, a) at pre_u[i + a * pre_u_cs]),
  // core (r1 x r2, row stride pre_ldk), V (r2 x n, row stride pre_v_rs), the
  // ranks read on the device and clamped to pre_kcap; FP32 operands when pre_f32
  const void* pre_u;
  const void* pre_core;
  const void* pre_v;
  const int32_t* pre_r1;
  const int32_t* pre_r2;
  int64_t pre_u_cs, pre_v_rs;
  int32_t pre_ldk, pre_kcap, pre_f32, pre_pad;
};

// ---------------------------------------------------------------------------
// Block substitution update of one target block: b -= U (S (V x)) (low-rank,
// r columns) or b -= D x (dense, D != nullptr).  All operands through strided
// views, x (n x nrhs) and b (m x nrhs) row-major with row stride nrhs.
// ---------------------------------------------------------------------------
struct SolveUpdItem {
  const void* U;
  const void* S;
  const void* V;
  const void* D;
  const void* x;
  void* b;
  int64_t u_rs, u_cs, s_rs, s_cs, v_rs, v_cs, d_rs, d_cs;
  int32_t m, n, r, pad;
};

// ---------------------------------------------------------------------------
// Dataflow block substitution (one launch per sweep, solve_df.hip).  Blocks
// and tasks are listed in sweep order ("positions").
// ---------------------------------------------------------------------------
// Task geometry of the dataflow block substitution: rows of x per D task,
// rows of bt per A task, updates per Y task (one per wave).
constexpr int kDfRows = 16;
constexpr int kDfArows = 64;
constexpr int kDfYtiles = 4;
// completion flags per block: one per D task (512 / kDfRows) and per A task (512 / kDfArows)
constexpr int kDfXflags = 512 / kDfRows;
constexpr int kDfAflags = 512 / kDfArows;
struct DfBlock {
  const void* tinv;  // n x n, row stride ld; used as op(Tinv) = Tinv^T when trans
  int64_t ld;
  const void* b;     // n x nrhs row-major (this sweep's input)
  void* x;           // n x nrhs row-major (this sweep's output)
  void* bt;          // n x nrhs scratch: b minus the updates from two or more positions back
  int32_t n, trans;
  int32_t nd;        // D tasks of this block (rows of 16)
  int32_t has_bt;    // D tasks read bt (after A) instead of b
  int32_t tile0, ntile;  // updates into this block, by source position ascending
  int32_t last;          // the update from the previous position, or -1
  int32_t src0, nsrc;    // updates out of this block (by_src), by target position ascending
  int32_t na;            // A tasks of this block (rows of kDfArows)
  int32_t near;          // the update from two positions back (applied by the A tasks from x), or -1
  int32_t pad;
};
struct DfTile {  // b_p -= U (S (V x_q)), U n_p x r, S r x r, V r x n_q (strided)
  const void* U;
  const void* S;
  const void* V;
  int64_t u_rs, u_cs, s_rs, s_cs, v_rs, v_cs;
  int32_t r, p, q, yoff;  // rank, target / source positions, offset of y (r x NR) in the y buffer
};
struct DfSweep {
  const void* blocks;   // DfBlock[nb]
  const void* tiles;    // DfTile[ntiles], grouped by target
  const int32_t* by_src;
  const void* tasks;    // int4 {type (0 D, 1 Y, 2 A), position, arg0, arg1}
  void* y;
  int32_t* flags;       // ticket, error, D-task flags [nb][kDfXflags], A-task flags [nb][kDfAflags], ydone[ntiles]
  int32_t nb, ntiles, ntasks, nrhs;
  int64_t spin_limit;
};

// Diagonal part of one fused substitution step: x (n x nrhs) = op(Tinv) b after
// the update items[n_upd] (when has_last) is applied to a copy of b.
struct StepDiag {
  const void* tinv;
  int64_t ld;
  const void* b;
  void* x;
  int32_t n, trans, has_last, pad;
};

// One row of a stacked-row product: a is a contiguous K-vector (input row),
// c where the finished row goes (contiguous N-vector).
// Low-rank x low-rank core with device ranks: core = alpha S_l (V_l U_b) S_b.
// V_l: rl x k (strides v_rs, v_cs), U_b: k x rb, S_l / S_b square with leading
// dimension lds; core rl x rb with leading dimension ldk; ranks clamped to kcap.
struct LrCoreItem {
  const void* vl;
  int64_t v_rs, v_cs;
  const void* ub;
  int64_t u_rs, u_cs;
  const void* sl;
  const void* sb;
  void* core;
  const int32_t* rl;
  const int32_t* rb;
  int32_t lds, ldk, k, kcap;
  double alpha;
  // optional: rows r < rl of V_l written to vl_dst (row stride vl_dst_ld, nv
  // columns) and columns c < rb of U_b to ub_dst (column stride ub_dst_ld, nu
  // rows) -- the panel results, when vl / ub point at workspace
  void* vl_dst;
  void* ub_dst;
  int64_t vl_dst_ld, ub_dst_ld;
  int32_t nv, nu;
};

// Two stacked-row products in one launch (blockIdx.z selects the job), rows
// r < R of base + r * stride (R capped on the device by d_count when given).
struct RowsJob {
  const void* base;
  int64_t stride;
  const void* b;
  int64_t b_rs, b_cs;
  void* w;
  const int32_t* d_count;
  int32_t R, N, K, pad;
};
struct RowsPair {
  RowsJob job[2];
};

// One low-rank panel tile after its solve: row sigma < r of the solved factor
// at src + sigma * src_ld (n entries) is written back to dst + sigma * dst_ld
// (dst null: already in place), and output o < r of the stacked factor,
// alpha sum_sigma coef(o, sigma) row_sigma with coef(o, sigma) at
// s[o * s_rs + sigma * s_cs], to pq + o * n.
struct PanelFin {
  const void* src;
  void* dst;
  const void* s;
  void* pq;
  int64_t src_ld, dst_ld, s_rs, s_cs;
  int32_t r, n;
  double alpha;
};

// C (m x n, leading dimension ldc) += U (m x r1) core (r1 x r2) V (r2 x n),
// r1 / r2 read on the device and clamped to kcap.
struct LrUpdItem {
  const void* u;
  int64_t u_rs, u_cs;
  const void* core;
  const void* v;
  int64_t v_rs, v_cs;
  void* c;
  int64_t ldc;
  const int32_t* r1;
  const int32_t* r2;
  int32_t ldk, m, n, kcap;
};

struct RowRef {
  const void* a;
  void* c;
};

// Two stacked-row products over row lists in one launch (blockIdx.z selects
// the job): row r < R of job j at rows[r].a.
struct RowsListJob {
  const RowRef* rows;
  const void* b;
  int64_t b_rs, b_cs;
  void* w;
  int32_t R, N, K, pad;
};
struct RowsList2 {
  RowsListJob job[2];
};

}  // namespace frk
