// BLR engine: tile storage resident in HBM and the batched BLR-LU / solve /
// construction drivers built on the kernels in kernels.hip.
//
// Storage (per tile (i, j), rows m = roff[i+1]-roff[i], cols n):
//   dense    : m x n row-major, ld = n
//   low-rank : fixed-capacity slots (capacity kcap, ranks in a device table)
//              U  column-major m x kcap (ld = m)  -> every basis vector contiguous
//              S  kcap x kcap row-major (ld = kcap)
//              V  kcap x n row-major (ld = n)     -> every basis vector contiguous
// Updates rewrite the slots in place, so a factorization never reallocates.
#pragma once
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

namespace frk {

// Host waits with a deadline: poll the event (or an event recorded on the
// stream) until it completes, and throw if it has not after wait_timeout_s()
// seconds (FRK_WAIT_TIMEOUT_S, default 60), so that a device-side stall
// surfaces as an error of the call instead of a hang.
double wait_timeout_s();
void host_wait_event(hipEvent_t e, const char* what);
void host_wait_stream(hipStream_t s, const char* what);

class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t bytes) { ensure(bytes); }
  ~DevBuf();
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), bytes_(o.bytes_) {
    o.p_ = nullptr;
    o.bytes_ = 0;
  }
  void ensure(size_t bytes);  // grows; contents are not preserved
  template <class X>
  X* as() const {
    return static_cast<X*>(p_);
  }
  size_t bytes() const { return bytes_; }

 private:
  void* p_ = nullptr;
  size_t bytes_ = 0;
};

// Stream-ordered upload of small descriptor arrays through a pinned ring.
class Staging {
 public:
  Staging() = default;
  ~Staging();
  template <class Item>
  const Item* push(const std::vector<Item>& v, hipStream_t s) {
    return static_cast<const Item*>(push_bytes(v.data(), v.size() * sizeof(Item), s));
  }
  const void* push_bytes(const void* src, size_t bytes, hipStream_t s);
  void reset() { off_ = 0; }

 private:
  char* h_ = nullptr;
  char* hd_ = nullptr;  // h_ as the device sees it
  char* d_ = nullptr;
  size_t cap_ = 0, off_ = 0;
};

// Execution context: stream, staging ring, workspace arena, constants.
// Per kernel-class op log: launches, algorithmic flops and bytes (computed on
// the host from the launch descriptors) and, for the one class selected with
// Context::time_class(), device time from HIP events around each launch.
struct OpStat {
  int64_t launches = 0;
  double flops = 0, bytes = 0, ms = 0;
  int64_t timed = 0;
  double roof_s = 0;  // sum over launches of max(flops / peak, bytes / HBM bandwidth)
};

// MI355X published peaks used by the op log's roofline time (dense MFMA; the
// FP64 and FP32 vector rates equal the matrix rates on this part).
constexpr double kPeakF64 = 78.6e12, kPeakF32 = 157.3e12, kPeakBf16 = 2.5e15, kHbmBytesPerS = 8.0e12;
template <typename T>
constexpr double mfma_peak() {
  return sizeof(T) == 8 ? kPeakF64 : (sizeof(T) == 4 ? kPeakF32 : kPeakBf16);
}

class Context {
 public:
  explicit Context(hipStream_t s);
  ~Context();
  hipStream_t stream() const { return s_; }
  Staging& staging() { return stage_; }
  // Workspace: reserve() after a sync, then bump-allocate with take().
  void reserve(size_t bytes);
  template <class X>
  X* take(size_t n) {
    return static_cast<X*>(take_bytes(n * sizeof(X)));
  }
  void* take_bytes(size_t bytes);
  void reset_ws() { ws_off_ = 0; }
  void sync();  // drain the stream, reset staging and workspace
  // identity matrix (ident_n x ident_n) in both precisions
  template <class X>
  const X* identity(int n);
  int ident_ld() const { return ident_n_; }
  // rsvd sample matrix Omega = random_uniform(n, p) (mt19937 seed 0)
  template <class X>
  const X* omega(int n, int p);
  int32_t* info() { return info_.as<int32_t>(); }
  // Rank policy of the matrices built on this context (BlrMatrix::cap_trunc).
  void set_cap_trunc(bool on) { cap_trunc_ = on; }
  bool cap_trunc() const { return cap_trunc_; }
  // Per-block flags of the dataflow tile LU (zero at allocation, kTileFlags
  // ints; the last one is the timeout marker) and the launch epochs that
  // index them: every launch on this context takes the next epoch.
  static constexpr int kTileFlags = 160;
  int32_t* tile_flags();
  int32_t next_tile_epoch() { return ++tile_epoch_; }
  int32_t tile_epoch() const { return tile_epoch_; }
  // Throw if a tile LU launched at an epoch > since timed out (drains the stream).
  void check_tile_lu(int32_t since);
  // Device event counters (kCounter* slots), zero at creation.
  int32_t* counters() { return owner_->counters_.as<int32_t>(); }
  // Persistent auxiliary streams (created on first use, one hardware queue each
  // as long as the process holds few streams): 0 = lookahead factorizations,
  // 1 = second chunk of a BLR-LU step.
  hipStream_t aux_stream(int k);
  // op log
  // (launches: how many launches the entry stands for; 0 adds the cost of
  // launches already counted when they were issued)
  void account(const char* cls, double flops, double bytes, double peak = 0, int64_t launches = 1);
  // peak arithmetic rate of launches that name none (the precision the caller computes in)
  void set_default_peak(double p) { owner_->peak_ = p; }
  // Device buffers kept between calls, grown on demand: a call that reuses a
  // size allocates nothing.  Slots: 0, 1 refinement residual and correction;
  // 2, 3 tpmqrt's C copies and their ranks; 4 ranks of a typed batch's dense
  // compressions; 5 inverses of the batched tile LU; 6 tile-level tpqrt T;
  // 7 partial sums of the kernel residual.
  void* persistent(int slot, size_t bytes) {
    DevBuf& b = persist_[slot];
    if (b.bytes() < bytes) {
      host_wait_stream(s_, "persistent buffer growth");
      b.ensure(bytes);
    }
    return b.as<void>();
  }
  double default_peak() const { return owner_->peak_; }
  // host-side wall time attributed to a class of the op log (counted as timed launches)
  void host_time(const char* cls, double ms);
  void time_class(const std::string& cls) { timed_cls_ = cls; }
  bool is_timed(const char* cls) const {
    return cls && !owner_->timed_cls_.empty() && owner_->timed_cls_ == cls;
  }
  void event_begin();
  void event_end();
  // Route this context's op log (and event timing) into another context's.
  void set_stats_owner(Context* owner) { owner_ = owner ? owner : this; }
  const std::map<std::string, OpStat>& op_stats();  // resolves pending events (drains the stream)
  void reset_op_stats();

 private:
  hipStream_t s_;
  Staging stage_;
  DevBuf ws_;
  size_t ws_off_ = 0;
  DevBuf ident64_, ident32_;
  int ident_n_ = 0;
  DevBuf omega64_, omega32_;
  int omega_n_ = 0, omega_p_ = 0;
  DevBuf info_;
  DevBuf counters_;
  DevBuf tile_flags_;
  int32_t tile_epoch_ = 0;
  bool cap_trunc_ = false;
  std::map<std::string, OpStat> stats_;
  std::string timed_cls_;
  std::vector<hipEvent_t> ev_pool_;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pending_;
  hipEvent_t ev_open_ = nullptr;
  Context* owner_ = this;
  double peak_ = kPeakF64;
  DevBuf persist_[8];
  hipStream_t aux_[2] = {nullptr, nullptr};
  void resolve_events();
};

// Slots of Context::counters().
enum Counter : int32_t { kCounterSvdProblems = 0, kCounterSvdSweeps = 1, kCounterSvdMaxSweeps = 2, kNumCounters = 16 };

// A batch of launches whose descriptor arrays travel to the device in one
// copy: add() item arrays, launch() closures that receive the device base,
// run() uploads once and issues every launch in order on the context stream.
class Plan {
 public:
  explicit Plan(Context& ctx) : ctx_(ctx) {}
  template <class Item>
  size_t add(const std::vector<Item>& v) {
    const size_t off = (bytes_.size() + 255) & ~size_t(255);
    bytes_.resize(off + v.size() * sizeof(Item));
    if (!v.empty()) std::memcpy(bytes_.data() + off, v.data(), v.size() * sizeof(Item));
    return off;
  }
  // cls (a string literal) names the kernel class for the op log
  // peak: arithmetic rate the launch's flops run at (0: the context's default)
  void launch(std::function<void(const char*)> f, const char* cls = nullptr, double flops = 0, double bytes = 0,
              double peak = 0) {
    launches_.push_back(std::move(f));
    meta_.push_back({cls, flops, bytes, peak});
  }
  void run();
  bool empty() const { return launches_.empty(); }
  Context& ctx() { return ctx_; }

 private:
  Context& ctx_;
  std::vector<char> bytes_;
  std::vector<std::function<void(const char*)>> launches_;
  struct Meta {
    const char* cls;
    double flops, bytes, peak;
  };
  std::vector<Meta> meta_;
};

enum TileKind : int32_t { kDense = 0, kLowRank = 1, kEmpty = 2 };
enum Lra : int32_t { kLraNaive = 0, kLraRounded = 1, kLraFast = 2 };

template <typename T>
class BlrMatrix {
 public:
  BlrMatrix(std::vector<int64_t> roff, std::vector<int64_t> coff, std::vector<int32_t> kinds, int kcap,
            double eps, int fixed_rank, Context& ctx);
  int nb_r() const { return (int)roff.size() - 1; }
  int nb_c() const { return (int)coff.size() - 1; }
  int64_t rows(int i) const { return roff[i + 1] - roff[i]; }
  int64_t cols(int j) const { return coff[j + 1] - coff[j]; }
  int idx(int i, int j) const { return i * nb_c() + j; }
  int32_t* d_rank(int i, int j) const { return d_ranks_.as<int32_t>() + idx(i, j); }
  void pull_ranks(hipStream_t s);  // synchronous device -> host copy of the rank table
  void push_ranks(hipStream_t s);
  // Same tile structure, fresh slots; copy_from() restores contents (stream-ordered).
  std::unique_ptr<BlrMatrix<T>> clone_structure(Context& ctx) const;
  void copy_from(const BlrMatrix<T>& src, hipStream_t s);
  size_t arena_bytes() const { return arena_bytes_; }
  // raw slot storage (element-wise conversion between precisions)
  T* arena_data() const { return arena_.as<T>(); }
  T* dinv_data() const { return dinv_.as<T>(); }
  size_t dinv_count() const { return dinv_.bytes() / sizeof(T); }
  void set_ranks_from_device(const int32_t* src, hipStream_t s);

  std::vector<int64_t> roff, coff;
  std::vector<int32_t> kind;
  std::vector<int32_t> rank;  // host mirror (valid after pull_ranks)
  int kcap;
  double eps;      // fixed accuracy when != 0
  int fixed_rank;  // used when eps == 0
  int lra = kLraFast;
  // Rank policy at the slot capacity: false reports an overflow as an error,
  // true keeps the best kcap terms (the capped-rank workloads).
  bool cap_trunc = false;
  std::vector<T*> D, U, S, V;  // per tile slots (nullptr when unused)
  // inverses of the 64 x 64 diagonal blocks of the L and U factors of every
  // square dense diagonal tile (written by getrf, read by the blocked trsm)
  std::vector<T*> Linv, Uinv;

 private:
  DevBuf arena_;
  size_t arena_bytes_ = 0;
  DevBuf dinv_;
  DevBuf d_ranks_;
};

// ----------------------------------------------------------------------------
// Batched building blocks (all asynchronous on ctx.stream()).
// ----------------------------------------------------------------------------

// C += X for every descriptor (reference: rounded_addition / fast_rounded_addition).
// Workspace comes from the plan's context (reserve lr_add_workspace() first).
template <typename T>
void plan_lr_add(Plan& pl, const std::vector<LrAddDesc<T>>& d, int lra);
template <typename T>
size_t lr_add_workspace(const std::vector<LrAddDesc<T>>& d);
template <typename T>
void lr_add_batched(const std::vector<LrAddDesc<T>>& d, int lra, Context& ctx);
// naive_addition: concatenation, dense recompression past min(m, n).  Synchronous.
template <typename T>
void lr_add_naive(const std::vector<LrAddDesc<T>>& d, Context& ctx);

template <typename T>
struct CompressDesc {
  T* a;  // dense m x n row-major (overwritten), ld = lda
  int m, n;
  int64_t lda;
  double eps;  // != 0: truncated_geqp3 + RQ; 0: rsvd at fixed rank
  int rank;
  T* u;
  T* s;
  T* v;
  int64_t u_rs, u_cs, v_rs, v_cs;  // strided outputs: U (m x r), V (r x n)
  int lds;
  int32_t* d_rank;
  int cap;
  bool cap_trunc = false;
};

// LowRank(Dense, eps) / LowRank(Dense, rank) for every descriptor.  Drains
// the stream (ranks are needed on the host to size the follow-up launches).
template <typename T>
void compress_batched(const std::vector<CompressDesc<T>>& d, Context& ctx);

// No-pivot LU of square row-major tiles (n <= 512) in place; blocked with
// 64-wide panels whose diagonal-block inverses go to linv / uinv (nblk x 64 x 64
// each).  info_dev[p * ntiles + t] receives the 1-based first zero pivot of
// panel p (0 if none).
template <typename T>
void plan_getrf(Plan& pl, const std::vector<T*>& tiles, const std::vector<T*>& linv, const std::vector<T*>& uinv,
                int n, int64_t ld, int32_t* info_dev);
template <typename T>
void getrf_tiles(const std::vector<T*>& tiles, int n, int64_t ld, int32_t* info_dev, Context& ctx);
// Allocation-free batched tile LU (noalloc.hip): scratch in a file-scope device
// array; status as frk_dgetrf_nopiv_batched, message in msg.
int getrf_nopiv_noalloc(hipStream_t s, double* const* tiles, int batch, int n, int64_t ld, int32_t* info_host,
                        std::string& msg);
// Triangular solve T X = B (left, in place) for a triangle without precomputed
// block inverses (computes them first).  Drains the stream.
template <typename T>
void trsm_generic(const T* A, int64_t a_rs, int64_t a_cs, bool lower, bool unit, T* B, int64_t b_rs, int64_t b_cs,
                  int n, int nrhs, Context& ctx);

// ----------------------------------------------------------------------------
// BLR drivers.
// ----------------------------------------------------------------------------
struct KernelDesc {
  int kind;  // 0 laplace, 1 gaussian, 2 identity
  int dim;
  int64_t n;               // number of points
  const double* host_pts;  // dim x n, structure of arrays
  double ell, shift;
  int64_t n_rows = 0, n_cols = 0;  // matrix size (0: n); rows and columns index the same points
};

// Hierarchical(kernel, ..., eps | rank, admis, nb_r, nb_c) for a one-level tree.
template <typename T>
std::unique_ptr<BlrMatrix<T>> build_blr_kernel(const KernelDesc& kd, int nb_r, int nb_c, double eps, int rank,
                                               double admis, int admis_type, int kcap, Context& ctx);

// getrf(Hierarchical&): right-looking tile LU without pivoting, in place
// (lower tiles become L, upper tiles U, diagonal tiles packed L\U).
template <typename T>
void blr_getrf(BlrMatrix<T>& A, Context& ctx, int max_steps = -1);

// Forward/backward block substitution on device vectors b (N x nrhs row-major).
// trans = false: L then U (solves A x = b); trans = true: U^T then L^T.
// part: 0 both sweeps, 1 the L (or L^T) sweep only, 2 the U (or U^T) sweep only.
template <typename T>
void blr_solve(BlrMatrix<T>& F, T* b, int nrhs, bool trans, Context& ctx, int part = 0);

// y += A x on an unfactored BLR matrix.
template <typename T>
void blr_matvec(BlrMatrix<T>& A, const T* x, T* y, int nrhs, Context& ctx);

// Iterative refinement: x = F^-1 b, then x += F^-1 (b - A x) with the exact
// FP64 kernel matrix (matrix-free) until the normwise backward error
// max_c ||b_c - A x_c||_inf / (||A||_inf ||x_c||_inf + ||b_c||_inf) <= tol.
struct RefineResult {
  int iters;
  double berr;
};
template <typename T>
RefineResult blr_refine(BlrMatrix<T>& F, const KernelSpec& ks, double anorm_inf, const double* b, double* x,
                        int nrhs, double tol, int max_iter, Context& ctx);
// ||A||_inf of a kernel matrix with non-negative entries (max row sum).
double kernel_norm_inf(const KernelSpec& ks, int64_t n, Context& ctx);

// Same tile structure in another storage precision, contents converted element-wise.
template <typename TD, typename TS>
std::unique_ptr<BlrMatrix<TD>> convert_blr(const BlrMatrix<TS>& src, Context& ctx);

// Device kernel spec with uploaded points (owned by the returned buffer).
KernelSpec upload_kernel(const KernelDesc& kd, DevBuf& pts, hipStream_t s);

// ----------------------------------------------------------------------------
// Typed (Dense / LowRank) operands and batched typed gemm
// (src/operations/BLAS/gemm.cpp:126-272).
// ----------------------------------------------------------------------------
template <typename T>
struct Op {  // read-only view, possibly transposed through its strides
  int kind;
  int m, n;
  const T* d;
  int64_t d_rs, d_cs;
  const T* u;
  int64_t u_rs, u_cs;
  const T* s;
  int64_t s_rs, s_cs;
  const T* v;
  int64_t v_rs, v_cs;
  int r;
};

template <typename T>
struct Target {  // dense view, or low-rank slots in the BlrMatrix layout
  int kind;
  int m, n;
  T* d;
  int64_t d_rs, d_cs;
  T* u;  // m x cap, column-major unless u_rs / u_cs say otherwise
  T* s;  // cap x cap row-major
  T* v;  // cap x n row-major
  int cap, rank;
  int64_t u_rs = 1, u_cs = 0;  // row / column strides of U (u_cs = 0: m, column-major)
  int32_t* d_rank;
  double eps;
  int fixed_rank;
  bool cap_trunc = false;
};

template <typename T>
struct TypedBatch {
  std::vector<GemmItem> cp;    // copies, run first
  std::vector<GemmItem> g[5];  // dependent product stages
  std::vector<LrAddDesc<T>> adds;
  std::vector<CompressDesc<T>> comp;
  std::vector<int> comp_add;
};

template <typename T>
Op<T> tile_op(const BlrMatrix<T>& A, int i, int j);
template <typename T>
Op<T> transpose_op(const Op<T>& a);
template <typename T>
Target<T> tile_target(BlrMatrix<T>& A, int i, int j);
template <typename T>
size_t typed_gemm_workspace(const Op<T>& A, const Op<T>& B, const Target<T>& C);
// C = alpha A B + beta C (beta = 1 for low-rank targets); products go to the
// batch's stages, additions to its list; the ranks of low-rank operands and
// targets must be current on the host.
template <typename T>
void typed_gemm(TypedBatch<T>& b, Context& ctx, const Op<T>& A, const Op<T>& B, const Target<T>& C, double alpha,
                double beta);
template <typename T>
void run_typed_batch(TypedBatch<T>& b, Context& ctx, int lra);

// Workspace bound of plan_lr_add for nd additions of at most m x n with K columns.
template <typename T>
size_t lr_add_bound(int nd, int m, int n, int K) {
  return (size_t)nd * sizeof(T) * ((size_t)(m + n) * K + 2 * (size_t)K + 6 * (size_t)K * K) + 10 * 256 + 4096;
}

// ----------------------------------------------------------------------------
// Batched small factorizations of the operator surface (small.cpp).
// ----------------------------------------------------------------------------
template <typename T>
struct SmallJob {  // A (m x n, strided) -> o1, o2 (strided)
  const T* a;
  int64_t a_rs, a_cs;
  T* o1;
  int64_t o1_rs, o1_cs;
  T* o2;
  int64_t o2_rs, o2_cs;
  int m, n;
};
template <typename T>
struct SvdJob {
  const T* a;
  int64_t a_rs, a_cs;
  T* u;
  int64_t u_rs, u_cs;
  T* s;
  int64_t s_ld;
  T* v;
  int64_t v_rs, v_cs;
  int m, n;
};
template <typename T>
struct Geqp3Job {
  const T* a;
  int64_t a_rs, a_cs;
  int m, n;
  double eps;
  int cap;
  T* q;
  int64_t q_rs, q_cs;
  T* r;
  int64_t r_ld;
  int32_t* rank_host;
};
// qr(A, Q, R): o1 = Q (m x k), o2 = R (k x n), k = min(m, n); m <= 512, n <= 128.
template <typename T>
void qr_small_batched(const std::vector<SmallJob<T>>& jobs, Context& ctx);
// rq(A, R, Q): o1 = R (m x m), o2 = Q (m x n); m <= n, m <= 128, n <= 512.
template <typename T>
void rq_small_batched(const std::vector<SmallJob<T>>& jobs, Context& ctx);
// svd(A): U (m x K), S (K x K diagonal, descending), V (K x n); K = min(m, n) <= 64.
template <typename T>
void svd_small_batched(const std::vector<SvdJob<T>>& jobs, Context& ctx);
// truncated_geqp3(A, eps): Q (m x r), R (r x n), r <= cap.
template <typename T>
void geqp3_small_batched(const std::vector<Geqp3Job<T>>& jobs, Context& ctx);
// rsvd(A, p): U (m x p), S (p x p), V (p x n), all row-major.
template <typename T>
void rsvd_small(const T* A, int64_t lda, int m, int n, int p, T* U, int64_t ldu, T* S, int64_t lds, T* V, int64_t ldv,
                Context& ctx);
// one_sided_id(A, k): V (k x n, ld ldv), the k selected columns to cols_host.
template <typename T>
void id_small(const T* A, int64_t a_rs, int64_t a_cs, int m, int n, int k, T* V, int64_t ldv, int32_t* cols_host,
              Context& ctx);

// ----------------------------------------------------------------------------
// Tiled Householder BLR-QR (src/operations/LAPACK/qr.cpp:213-257).
// ----------------------------------------------------------------------------
template <typename T>
struct QrFactors {
  explicit QrFactors(const BlrMatrix<T>& A);
  int nb_r, nb_c;
  std::vector<T*> Tt;   // T(i, k), i >= k: b x b row-major
  std::vector<T*> tau;  // tau(i, k): b
  std::vector<T*> Yx;   // explicit unit-lower reflectors of diagonal tile k (b x b)
  std::vector<int> bsize;
  T* tmat(int i, int k) const { return Tt[i * nb_c + k]; }
  T* tau_of(int i, int k) const { return tau[i * nb_c + k]; }

 private:
  DevBuf buf_;
};

// tiled_householder_blr_qr(A, T): in place, R in the upper tiles, reflectors
// in the diagonal tiles' lower parts and in the tiles below the diagonal.
// A caller-owned tile of the tile-level operators: Dense (rows x cols, row
// stride ld) or LowRank (U rows x cap with strides u_rs / u_cs, S cap x cap and
// V cap x cols row-major), with its rank and accuracy (0: fixed rank).
template <typename T>
struct RawTile {
  int kind = kDense;
  int rows = 0, cols = 0;
  int rank = 0, cap = 0;
  double eps = 0.0;
  T* d = nullptr;
  int64_t ld = 0;
  T* u = nullptr;
  T* s = nullptr;
  T* v = nullptr;
  int64_t u_rs = 1, u_cs = 0;  // U strides (u_cs = 0: rows, column-major)
};
template <typename T>
void tiles_geqrt(const std::vector<T*>& A, int64_t lda, const std::vector<T*>& Tout, int64_t ldt, int n, Context& ctx);
template <typename T>
void tiles_larfb(const std::vector<const T*>& V, int64_t ldv, const std::vector<const T*>& Tin, int64_t ldt,
                 std::vector<RawTile<T>>& C, int n, bool trans, Context& ctx);
template <typename T>
void tiles_tpqrt(const std::vector<T*>& R, int64_t ldr, std::vector<RawTile<T>>& B, const std::vector<T*>& Tout,
                 int64_t ldt, int n, Context& ctx);
template <typename T>
void tiles_tpmqrt(const std::vector<RawTile<T>>& V, const std::vector<const T*>& Tin, int64_t ldt,
                  std::vector<RawTile<T>>& A, std::vector<RawTile<T>>& B, int n, bool trans, int lra, Context& ctx);
template <typename T>
void tiles_gemm(const std::vector<RawTile<T>>& A, const std::vector<RawTile<T>>& B, std::vector<RawTile<T>>& C,
                double alpha, double beta, bool ta, bool tb, int lra, Context& ctx);

template <typename T>
void blr_tiled_qr(BlrMatrix<T>& A, QrFactors<T>& F, Context& ctx);
// left_multiply_tiled_reflector(Y, T, C, trans): C <- Q C (trans = false) or Q^T C.
template <typename T>
void blr_apply_q(const BlrMatrix<T>& Y, QrFactors<T>& F, BlrMatrix<T>& C, bool trans, Context& ctx);

}  // namespace frk
