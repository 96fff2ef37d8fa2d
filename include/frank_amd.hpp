// C++17 operator surface over the C ABI (frank_amd.h), mirroring FRANK's
// classes and free functions so that drivers and tests written against
// include/FRANK/FRANK.h port by changing the include and the namespace:
//
//   Dense, LowRank, Hierarchical, MatrixProxy  (include/FRANK/classes/*.h)
//   initialize, setGlobalValue                 (src/util/initialize.cpp, global_key_value.cpp)
//   getrf, trsm, gemm, qr, rq, svd, truncated_geqp3, rsvd, id, one_sided_id, rid,
//   tiled_householder_blr_qr, left_multiply_tiled_reflector, operator+=, l2_error
//                                              (include/FRANK/operations/*.h)
//
// Dispatch on operand types is a std::visit over MatrixProxy (Dense | LowRank |
// Hierarchical); an unmatched combination throws frank_amd::Error with the
// reference's "name(T1, T2) undefined!" wording instead of aborting.
//
// Storage: a Dense keeps a host copy (element access as in the reference) and
// a device copy (what the kernels use); each side is refreshed on demand.
// A Hierarchical is a BLR handle whose tiles stay in HBM.  Everything is FP64.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <variant>
#include <vector>

#include "frank_amd.h"

namespace frank_amd {

enum class Side { Left, Right };
enum class Mode { Upper, Lower };
enum class AdmisType { PositionBased, GeometryBased };

class Error : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

namespace detail {
inline void check(int rc) {
  if (rc != FRK_OK) throw Error(frk_last_error());
}
inline void hip(hipError_t e) {
  if (e != hipSuccess) throw Error(std::string("HIP error: ") + hipGetErrorString(e));
}
struct Runtime {
  frk_ctx ctx = nullptr;
  std::map<std::string, std::string> globals;
  ~Runtime() {
    if (ctx) frk_ctx_destroy(ctx);
  }
};
inline Runtime& runtime() {
  static Runtime r;
  return r;
}
inline frk_ctx ctx() {
  Runtime& r = runtime();
  if (!r.ctx) check(frk_ctx_create(&r.ctx));
  return r.ctx;
}
inline int lra() {
  const auto& g = runtime().globals;
  const auto it = g.find("FRANK_LRA");
  if (it == g.end() || it->second == "fast_rounded_addition") return FRK_LRA_FAST;
  if (it->second == "rounded_addition") return FRK_LRA_ROUNDED;
  if (it->second == "naive") return FRK_LRA_NAIVE;
  throw Error("unknown FRANK_LRA value " + it->second);
}
inline std::shared_ptr<double> dalloc(size_t n) {
  double* p = nullptr;
  hip(hipMalloc(&p, sizeof(double) * std::max<size_t>(n, 1)));
  hip(hipMemset(p, 0, sizeof(double) * std::max<size_t>(n, 1)));
  return std::shared_ptr<double>(p, [](double* q) { (void)hipFree(q); });
}
}  // namespace detail

// FRANK::initialize() (src/util/initialize.cpp:9-17): creates the execution context.
inline void initialize() { (void)detail::ctx(); }
// setGlobalValue("FRANK_LRA", "rounded_addition" | "fast_rounded_addition" | "naive")
inline void setGlobalValue(const std::string& key, const std::string& value) { detail::runtime().globals[key] = value; }
inline std::string getGlobalValue(const std::string& key) { return detail::runtime().globals[key]; }

// ---------------------------------------------------------------------------
// Kernel functions with the reference's signature (src/functions.cpp).
// ---------------------------------------------------------------------------
using KernelFn = void (*)(double* A, const uint64_t A_rows, const uint64_t A_cols, const uint64_t A_stride,
                          const std::vector<std::vector<double>>& params, const int64_t row_start,
                          const int64_t col_start);

inline void identity(double* A, const uint64_t rows, const uint64_t cols, const uint64_t stride,
                     const std::vector<std::vector<double>>&, const int64_t r0, const int64_t c0) {
  for (uint64_t i = 0; i < rows; ++i)
    for (uint64_t j = 0; j < cols; ++j) A[i * stride + j] = (r0 + (int64_t)i == c0 + (int64_t)j) ? 1.0 : 0.0;
}

inline void laplacend(double* A, const uint64_t rows, const uint64_t cols, const uint64_t stride,
                      const std::vector<std::vector<double>>& x, const int64_t r0, const int64_t c0) {
  for (uint64_t i = 0; i < rows; ++i)
    for (uint64_t j = 0; j < cols; ++j) {
      double r = 0.0;
      for (const auto& xk : x) {
        const double d = xk[r0 + i] - xk[c0 + j];
        r += d * d;
      }
      A[i * stride + j] = 1.0 / (std::sqrt(r) + 1e-3);
    }
}

// get_sorted_random_vector (src/util/experiment_setup.cpp:12-23)
inline std::vector<double> get_sorted_random_vector(int64_t n) {
  std::mt19937 gen(0);
  std::uniform_real_distribution<double> dist(0.0, 1.0);
  std::vector<double> v(n);
  for (auto& x : v) x = dist(gen);
  std::sort(v.begin(), v.end());
  return v;
}

// ---------------------------------------------------------------------------
// Dense: row-major FP64, host and device copies.
// ---------------------------------------------------------------------------
class Dense {
 public:
  int64_t dim[2] = {0, 0};

  Dense() = default;
  Dense(const int64_t n_rows, const int64_t n_cols = 1) : dim{n_rows, n_cols}, h_((size_t)n_rows * n_cols, 0.0) {}
  // Dense(kernel, params, n_rows, n_cols, row_start, col_start) (dense.cpp)
  Dense(KernelFn kernel, const std::vector<std::vector<double>>& params, const int64_t n_rows, const int64_t n_cols,
        const int64_t row_start = 0, const int64_t col_start = 0)
      : Dense(n_rows, n_cols) {
    kernel(h_.data(), n_rows, n_cols, n_cols, params, row_start, col_start);
  }
  Dense(const Dense& A) : dim{A.dim[0], A.dim[1]} {
    A.sync_host();
    h_ = A.h_;
  }
  Dense& operator=(const Dense& A) {
    if (this != &A) {
      A.sync_host();
      dim[0] = A.dim[0];
      dim[1] = A.dim[1];
      h_ = A.h_;
      d_.reset();
      host_ok_ = true;
      dev_ok_ = false;
    }
    return *this;
  }
  Dense(Dense&&) = default;
  Dense& operator=(Dense&&) = default;

  double& operator()(const int64_t i, const int64_t j) {
    sync_host();
    dev_ok_ = false;
    return h_[(size_t)i * dim[1] + j];
  }
  const double& operator()(const int64_t i, const int64_t j) const {
    sync_host();
    return h_[(size_t)i * dim[1] + j];
  }
  double& operator[](const int64_t i) { return (*this)(dim[1] == 1 ? i : i / dim[1], dim[1] == 1 ? 0 : i % dim[1]); }
  const double* host() const {
    sync_host();
    return h_.data();
  }
  // device buffer for reading / for writing (the host copy becomes stale)
  const double* device() const {
    sync_device();
    return d_.get();
  }
  double* device_mut() {
    sync_device();
    host_ok_ = false;
    return d_.get();
  }
  size_t size() const { return (size_t)dim[0] * dim[1]; }

 private:
  mutable std::vector<double> h_;
  mutable std::shared_ptr<double> d_;
  mutable bool host_ok_ = true, dev_ok_ = false;
  void sync_host() const {
    if (host_ok_) return;
    h_.resize(size());
    detail::check(frk_ctx_synchronize(detail::ctx()));
    detail::hip(hipMemcpy(h_.data(), d_.get(), sizeof(double) * size(), hipMemcpyDeviceToHost));
    host_ok_ = true;
  }
  void sync_device() const {
    if (dev_ok_) return;
    if (!d_) d_ = detail::dalloc(size());
    detail::hip(hipMemcpy(d_.get(), h_.data(), sizeof(double) * size(), hipMemcpyHostToDevice));
    dev_ok_ = true;
  }
};

// ---------------------------------------------------------------------------
// LowRank: A ~ U S V (low_rank.h:29-163).
// ---------------------------------------------------------------------------
class LowRank {
 public:
  int64_t dim[2] = {0, 0};
  int64_t rank = 0;
  double eps = 0.0;
  Dense U, S, V;

  LowRank() = default;
  LowRank(const Dense& U_, const Dense& S_, const Dense& V_, double eps_ = 0.0)
      : dim{U_.dim[0], V_.dim[1]}, rank(S_.dim[0]), eps(eps_), U(U_), S(S_), V(V_) {}
  // LowRank(Dense, rank): rsvd (low_rank.cpp:40-48)
  LowRank(const Dense& A, const int64_t k) : dim{A.dim[0], A.dim[1]}, rank(k) { compress(A, 0.0, (int)k, (int)k); }
  // LowRank(Dense, eps): truncated_geqp3 + rq (low_rank.cpp:50-59)
  LowRank(const Dense& A, const double eps_, const int cap = 64) : dim{A.dim[0], A.dim[1]}, eps(eps_) {
    compress(A, eps_, 0, cap);
  }

 private:
  void compress(const Dense& A, double e, int k, int cap) {
    Dense work(A);
    Dense Uc(A.dim[0], cap), Sc(cap, cap), Vc(cap, A.dim[1]);
    int32_t r = 0;
    detail::check(frk_dcompress(detail::ctx(), (int)A.dim[0], (int)A.dim[1], work.device_mut(), A.dim[1], e, k, cap,
                                Uc.device_mut(), Sc.device_mut(), Vc.device_mut(), &r));
    rank = r;
    U = Dense(A.dim[0], r);
    S = Dense(r, r);
    V = Dense(r, A.dim[1]);
    for (int64_t i = 0; i < A.dim[0]; ++i)
      for (int64_t t = 0; t < r; ++t) U(i, t) = Uc(i, t);
    for (int64_t i = 0; i < r; ++i)
      for (int64_t t = 0; t < r; ++t) S(i, t) = Sc(i, t);
    for (int64_t t = 0; t < r; ++t)
      for (int64_t j = 0; j < A.dim[1]; ++j) V(t, j) = Vc(t, j);
  }
};

// ---------------------------------------------------------------------------
// Hierarchical: one-level BLR grid in HBM.  getrf() returns (L, U) views of
// the same factored handle, as the reference returns two Hierarchicals.
// ---------------------------------------------------------------------------
class Hierarchical {
 public:
  int64_t dim[2] = {0, 0};
  enum Part { Full, LowerFactor, UpperFactor };

  Hierarchical() = default;
  // Hierarchical(kernel, params, n_rows, n_cols, rank, nleaf, admis, nb_r, nb_c, admis_type) (hierarchical.cpp:67-87)
  Hierarchical(KernelFn kernel, const std::vector<std::vector<double>>& params, int64_t n_rows, int64_t n_cols,
               int64_t rank, int64_t nleaf, double admis = 0, int64_t nb_r = 2, int64_t nb_c = 2,
               AdmisType admis_type = AdmisType::PositionBased)
      : Hierarchical(kernel, params, n_rows, n_cols, nleaf, 0.0, (int)rank, admis, nb_r, nb_c, admis_type) {}
  // Hierarchical(kernel, params, n_rows, n_cols, nleaf, eps, admis, nb_r, nb_c, admis_type) (hierarchical.cpp:89-109)
  Hierarchical(KernelFn kernel, const std::vector<std::vector<double>>& params, int64_t n_rows, int64_t n_cols,
               int64_t nleaf, double eps, double admis = 0, int64_t nb_r = 2, int64_t nb_c = 2,
               AdmisType admis_type = AdmisType::PositionBased)
      : Hierarchical(kernel, params, n_rows, n_cols, nleaf, eps, 0, admis, nb_r, nb_c, admis_type) {}

  frk_blr handle() const { return h_.get(); }
  Part part() const { return part_; }
  Hierarchical view(Part p) const {
    Hierarchical v(*this, 0);
    v.part_ = p;
    return v;
  }
  // Deep copy (matrix_proxy.cpp:68-76 clones every block)
  Hierarchical(const Hierarchical& A) : dim{A.dim[0], A.dim[1]}, part_(A.part_) {
    frk_blr c = nullptr;
    detail::check(frk_blr_clone(A.handle(), &c));
    h_ = std::shared_ptr<frk_blr_s>(c, [](frk_blr q) { frk_blr_destroy(q); });
  }
  Hierarchical& operator=(const Hierarchical& A) {
    Hierarchical t(A);
    std::swap(*this, t);
    return *this;
  }
  Hierarchical(Hierarchical&&) = default;
  Hierarchical& operator=(Hierarchical&&) = default;

  int tile_kind(int i, int j) const {
    int32_t k;
    detail::check(frk_blr_tile_info(handle(), i, j, &k, nullptr, nullptr, nullptr));
    return k;
  }
  int tile_rank(int i, int j) const {
    int32_t r;
    detail::check(frk_blr_tile_info(handle(), i, j, nullptr, &r, nullptr, nullptr));
    return r;
  }
  Dense dense_tile(int i, int j) const {
    int64_t m, n;
    detail::check(frk_blr_tile_info(handle(), i, j, nullptr, nullptr, &m, &n));
    Dense d(m, n);
    std::vector<double> h((size_t)m * n);
    detail::check(frk_blr_get_dense(handle(), i, j, h.data()));
    for (int64_t a = 0; a < m; ++a)
      for (int64_t b = 0; b < n; ++b) d(a, b) = h[(size_t)a * n + b];
    return d;
  }
  LowRank lowrank_tile(int i, int j) const {
    int64_t m, n;
    int32_t r;
    detail::check(frk_blr_tile_info(handle(), i, j, nullptr, &r, &m, &n));
    std::vector<double> u((size_t)m * r), s((size_t)r * r), v((size_t)r * n);
    detail::check(frk_blr_get_lowrank(handle(), i, j, u.data(), s.data(), v.data()));
    Dense U(m, r), S(r, r), V(r, n);
    for (int64_t a = 0; a < m; ++a)
      for (int t = 0; t < r; ++t) U(a, t) = u[(size_t)a * r + t];
    for (int a = 0; a < r; ++a)
      for (int t = 0; t < r; ++t) S(a, t) = s[(size_t)a * r + t];
    for (int t = 0; t < r; ++t)
      for (int64_t b = 0; b < n; ++b) V(t, b) = v[(size_t)t * n + b];
    return LowRank(U, S, V);
  }
  // Dense(Hierarchical): the full matrix (host)
  Dense to_dense() const {
    int64_t rows = 0, cols = 0;
    std::vector<int64_t> ro(1, 0), co(1, 0);
    for (int i = 0; i < dim[0]; ++i) {
      int64_t m;
      detail::check(frk_blr_tile_info(handle(), i, 0, nullptr, nullptr, &m, nullptr));
      ro.push_back(ro.back() + m);
    }
    for (int j = 0; j < dim[1]; ++j) {
      int64_t n;
      detail::check(frk_blr_tile_info(handle(), 0, j, nullptr, nullptr, nullptr, &n));
      co.push_back(co.back() + n);
    }
    rows = ro.back();
    cols = co.back();
    Dense out(rows, cols);
    for (int i = 0; i < dim[0]; ++i)
      for (int j = 0; j < dim[1]; ++j) {
        const int k = tile_kind(i, j);
        if (k == 2) continue;
        Dense blk;
        if (k == 0) {
          blk = dense_tile(i, j);
        } else {
          LowRank L = lowrank_tile(i, j);
          blk = Dense(L.dim[0], L.dim[1]);
          for (int64_t a = 0; a < L.dim[0]; ++a)
            for (int64_t b = 0; b < L.dim[1]; ++b) {
              double acc = 0.0;
              for (int64_t t = 0; t < L.rank; ++t) {
                double us = 0.0;
                for (int64_t q = 0; q < L.rank; ++q) us += L.U(a, q) * L.S(q, t);
                acc += us * L.V(t, b);
              }
              blk(a, b) = acc;
            }
        }
        for (int64_t a = 0; a < blk.dim[0]; ++a)
          for (int64_t b = 0; b < blk.dim[1]; ++b) out(ro[i] + a, co[j] + b) = blk(a, b);
      }
    return out;
  }

 private:
  std::shared_ptr<frk_blr_s> h_;
  Part part_ = Full;
  Hierarchical(const Hierarchical& A, int) : dim{A.dim[0], A.dim[1]}, h_(A.h_), part_(A.part_) {}
  Hierarchical(KernelFn kernel, const std::vector<std::vector<double>>& params, int64_t n_rows, int64_t n_cols,
               int64_t nleaf, double eps, int rank, double admis, int64_t nb_r, int64_t nb_c, AdmisType admis_type)
      : dim{nb_r, nb_c} {
    (void)nleaf;
    int kind = -1;
    if (kernel == &laplacend) kind = FRK_KERNEL_LAPLACE;
    if (kernel == &identity) kind = FRK_KERNEL_IDENTITY;
    if (kind < 0) throw Error("Hierarchical(kernel, ...): only laplacend and identity are generated on the device");
    const int dimx = (int)std::max<size_t>(params.size(), 1);
    const int64_t npts = params.empty() ? std::max(n_rows, n_cols) : (int64_t)params[0].size();
    std::vector<double> pts((size_t)dimx * npts, 0.0);
    for (size_t d = 0; d < params.size(); ++d)
      for (int64_t i = 0; i < npts; ++i) pts[d * npts + i] = params[d][i];
    frk_blr c = nullptr;
    detail::check(frk_blr_create_kernel_rect(detail::ctx(), kind, dimx, npts, pts.data(), n_rows, n_cols, (int)nb_r,
                                             (int)nb_c, 0.0, 0.0, eps, rank, admis,
                                             admis_type == AdmisType::PositionBased ? FRK_ADMIS_POSITION
                                                                                    : FRK_ADMIS_GEOMETRY,
                                             eps != 0.0 ? 64 : rank, FRK_PREC_F64, &c));
    h_ = std::shared_ptr<frk_blr_s>(c, [](frk_blr q) { frk_blr_destroy(q); });
  }
};

using MatrixProxy = std::variant<Dense, LowRank, Hierarchical>;

inline const char* type_name(const Dense&) { return "Dense"; }
inline const char* type_name(const LowRank&) { return "LowRank"; }
inline const char* type_name(const Hierarchical&) { return "Hierarchical"; }

[[noreturn]] inline void undefined(const char* op, const char* a, const char* b, const char* c = nullptr) {
  std::string m = std::string(op) + "(" + a + ", " + b + (c ? std::string(", ") + c : std::string()) + ") undefined!";
  throw Error(m);
}

// ---------------------------------------------------------------------------
// getrf / trsm (LAPACK.h:40, BLAS.h:201)
// ---------------------------------------------------------------------------
// getrf(Hierarchical&): tile LU in place; returns (L, U) views of the factors.
inline std::tuple<Hierarchical, Hierarchical> getrf(Hierarchical& A) {
  detail::check(frk_blr_set_lra(A.handle(), detail::lra()));
  detail::check(frk_blr_getrf(A.handle()));
  return {A.view(Hierarchical::LowerFactor), A.view(Hierarchical::UpperFactor)};
}

// trsm(A, B, uplo, side): A a factor view of getrf (B Dense, left) or a Dense
// triangle (B Dense; Lower => unit diagonal, H4).
inline void trsm(const Hierarchical& A, Dense& B, const Mode uplo, const Side side = Side::Left) {
  if (side != Side::Left) undefined("trsm", "Hierarchical", "Dense (right side)");
  detail::check(frk_blr_trsm(A.handle(), B.device_mut(), (int)B.dim[1], uplo == Mode::Upper ? 1 : 0, 0));
}
inline void trsm(const Dense& A, Dense& B, const Mode uplo, const Side side = Side::Left) {
  detail::check(frk_dtrsm(detail::ctx(), side == Side::Left ? 0 : 1, uplo == Mode::Lower ? 1 : 0, 0, (int)B.dim[0],
                          (int)B.dim[1], A.device(), A.dim[1], B.device_mut(), B.dim[1]));
}
inline void trsm(const Dense& A, LowRank& B, const Mode uplo, const Side side = Side::Left) {
  if (side == Side::Left)
    trsm(A, B.U, uplo, side);
  else
    trsm(A, B.V, uplo, side);
}
inline void trsm(const MatrixProxy& A, MatrixProxy& B, const Mode uplo, const Side side = Side::Left) {
  std::visit(
      [&](const auto& a, auto& b) {
        using TA = std::decay_t<decltype(a)>;
        using TB = std::decay_t<decltype(b)>;
        if constexpr (std::is_same_v<TA, Hierarchical> && std::is_same_v<TB, Dense>)
          trsm(a, b, uplo, side);
        else if constexpr (std::is_same_v<TA, Dense> && (std::is_same_v<TB, Dense> || std::is_same_v<TB, LowRank>))
          trsm(a, b, uplo, side);
        else
          undefined("trsm", type_name(a), type_name(b));
      },
      A, B);
}

// ---------------------------------------------------------------------------
// gemm (BLAS.h:49-53): C = alpha A B + beta C for the Dense / LowRank /
// Hierarchical combinations the reference defines on this path.
// ---------------------------------------------------------------------------
inline void gemm(const Dense& A, const Dense& B, Dense& C, double alpha = 1, double beta = 1, bool tA = false,
                 bool tB = false) {
  const int m = (int)C.dim[0], n = (int)C.dim[1], k = (int)(tA ? A.dim[0] : A.dim[1]);
  detail::check(frk_dgemm(detail::ctx(), tA, tB, m, n, k, alpha, A.device(), A.dim[1], B.device(), B.dim[1], beta,
                          C.device_mut(), C.dim[1]));
}
inline Dense gemm(const Dense& A, const Dense& B, double alpha = 1, bool tA = false, bool tB = false) {
  Dense C(tA ? A.dim[1] : A.dim[0], tB ? B.dim[0] : B.dim[1]);
  gemm(A, B, C, alpha, 0.0, tA, tB);
  return C;
}
// LR x D -> D (gemm.cpp:148-162), D x LR -> D (:164-178), LR x LR -> D (:180-200)
inline void gemm(const LowRank& A, const Dense& B, Dense& C, double alpha = 1, double beta = 1) {
  Dense inner = gemm(A.V, B, alpha);
  Dense asb = gemm(A.S, inner);
  gemm(A.U, asb, C, 1.0, beta);
}
inline void gemm(const Dense& A, const LowRank& B, Dense& C, double alpha = 1, double beta = 1) {
  Dense ab = gemm(A, B.U, alpha);
  Dense abs = gemm(ab, B.S);
  gemm(abs, B.V, C, 1.0, beta);
}
inline void gemm(const LowRank& A, const LowRank& B, Dense& C, double alpha = 1, double beta = 1) {
  Dense inner = gemm(A.V, B.U, alpha);
  Dense x = gemm(A.S, inner);
  Dense y = gemm(x, B.S);
  Dense z = gemm(A.U, y);
  gemm(z, B.V, C, 1.0, beta);
}
// Hierarchical x Dense -> Dense: C = alpha A B + beta C (y += A x on the tiles)
inline void gemm(const Hierarchical& A, const Dense& B, Dense& C, double alpha = 1, double beta = 1) {
  if (A.part() != Hierarchical::Full) undefined("gemm", "Hierarchical (factor view)", "Dense", "Dense");
  Dense y(C.dim[0], C.dim[1]);
  detail::check(frk_blr_matvec(A.handle(), B.device(), y.device_mut(), (int)B.dim[1]));
  for (int64_t i = 0; i < C.dim[0]; ++i)
    for (int64_t j = 0; j < C.dim[1]; ++j) C(i, j) = alpha * y(i, j) + beta * C(i, j);
}
// gemm into a LowRank target (defined with the tile-level operators below)
inline void gemm(const MatrixProxy& A, const MatrixProxy& B, LowRank& C, double alpha, double beta, bool tA, bool tB);
inline void gemm(const MatrixProxy& A, const MatrixProxy& B, MatrixProxy& C, double alpha = 1, double beta = 1) {
  std::visit(
      [&](const auto& a, const auto& b, auto& c) {
        using TA = std::decay_t<decltype(a)>;
        using TB = std::decay_t<decltype(b)>;
        using TC = std::decay_t<decltype(c)>;
        if constexpr (std::is_same_v<TC, Dense> && !std::is_same_v<TB, Hierarchical> &&
                      (!std::is_same_v<TA, Hierarchical> || std::is_same_v<TB, Dense>))
          gemm(a, b, c, alpha, beta);
        else if constexpr (std::is_same_v<TC, LowRank> && !std::is_same_v<TA, Hierarchical> &&
                           !std::is_same_v<TB, Hierarchical>)
          gemm(MatrixProxy(a), MatrixProxy(b), c, alpha, beta, false, false);
        else
          undefined("gemm", type_name(a), type_name(b), type_name(c));
      },
      A, B, C);
}

// operator+=(LowRank&, const LowRank&) (arithmetic.h:31, addition.cpp:191-203), FRANK_LRA selects the method
inline LowRank& operator+=(LowRank& A, const LowRank& B) {
  const int cap = (int)std::max<int64_t>(64, A.rank + B.rank);
  const int m = (int)A.dim[0], n = (int)A.dim[1];
  Dense Ua(m, cap), Sa(cap, cap), Va(cap, n);
  for (int i = 0; i < m; ++i)
    for (int t = 0; t < A.rank; ++t) Ua(i, t) = A.U(i, t);
  for (int i = 0; i < A.rank; ++i)
    for (int t = 0; t < A.rank; ++t) Sa(i, t) = A.S(i, t);
  for (int t = 0; t < A.rank; ++t)
    for (int j = 0; j < n; ++j) Va(t, j) = A.V(t, j);
  double* ua = Ua.device_mut();
  double* sa = Sa.device_mut();
  double* va = Va.device_mut();
  const double* ub = B.U.device();
  const double* sb = B.S.device();
  const double* vb = B.V.device();
  int32_t mm = m, nn = n, ka = (int32_t)A.rank, kb = (int32_t)B.rank;
  detail::check(frk_dlr_add_batched(detail::ctx(), detail::lra(), 1, &mm, &nn, A.eps, cap, &ua, &sa, &va, &ka, &ub,
                                    &sb, &vb, &kb));
  Dense U(m, ka), S(ka, ka), V(ka, n);
  for (int i = 0; i < m; ++i)
    for (int t = 0; t < ka; ++t) U(i, t) = Ua(i, t);
  for (int i = 0; i < ka; ++i)
    for (int t = 0; t < ka; ++t) S(i, t) = Sa(i, t);
  for (int t = 0; t < ka; ++t)
    for (int j = 0; j < n; ++j) V(t, j) = Va(t, j);
  A = LowRank(U, S, V, A.eps);
  return A;
}

// ---------------------------------------------------------------------------
// Dense factorizations (LAPACK.h, randomized_factorizations.h)
// ---------------------------------------------------------------------------
inline void qr(Dense& A, Dense& Q, Dense& R) {
  const int32_t m = (int32_t)A.dim[0], n = (int32_t)A.dim[1];
  const double* a = A.device();
  double* q = Q.device_mut();
  double* r = R.device_mut();
  const int64_t lda = n, ldq = Q.dim[1], ldr = R.dim[1];
  detail::check(frk_dqr_batched(detail::ctx(), 1, &m, &n, &a, &lda, &q, &ldq, &r, &ldr));
}
inline void rq(Dense& A, Dense& R, Dense& Q) {
  const int32_t m = (int32_t)A.dim[0], n = (int32_t)A.dim[1];
  const double* a = A.device();
  double* r = R.device_mut();
  double* q = Q.device_mut();
  const int64_t lda = n, ldr = R.dim[1], ldq = Q.dim[1];
  detail::check(frk_drq_batched(detail::ctx(), 1, &m, &n, &a, &lda, &r, &ldr, &q, &ldq));
}
inline std::tuple<Dense, Dense, Dense> svd(Dense& A) {
  const int32_t m = (int32_t)A.dim[0], n = (int32_t)A.dim[1], K = std::min(m, n);
  Dense U(m, K), S(K, K), V(K, n);
  const double* a = A.device();
  double* u = U.device_mut();
  double* s = S.device_mut();
  double* v = V.device_mut();
  const int64_t lda = n, ldu = K, lds = K, ldv = n;
  detail::check(frk_dsvd_batched(detail::ctx(), 1, &m, &n, &a, &lda, &u, &ldu, &s, &lds, &v, &ldv));
  return {U, S, V};
}
inline std::tuple<Dense, Dense> truncated_geqp3(const Dense& A, const double eps, const int cap = 64) {
  const int32_t m = (int32_t)A.dim[0], n = (int32_t)A.dim[1];
  Dense Qc(m, cap), Rc(cap, n);
  const double* a = A.device();
  double* q = Qc.device_mut();
  double* r = Rc.device_mut();
  const int64_t lda = n, ldq = cap, ldr = n;
  int32_t k = 0;
  detail::check(frk_dtruncated_geqp3_batched(detail::ctx(), 1, &m, &n, &a, &lda, eps, cap, &q, &ldq, &r, &ldr, &k));
  Dense Q(m, k), R(k, n);
  for (int i = 0; i < m; ++i)
    for (int t = 0; t < k; ++t) Q(i, t) = Qc(i, t);
  for (int t = 0; t < k; ++t)
    for (int j = 0; j < n; ++j) R(t, j) = Rc(t, j);
  return {Q, R};
}
inline std::tuple<Dense, Dense, Dense> rsvd(const Dense& A, const int64_t p) {
  const int m = (int)A.dim[0], n = (int)A.dim[1];
  Dense U(m, p), S(p, p), V(p, n);
  detail::check(frk_drsvd(detail::ctx(), m, n, A.device(), n, (int)p, U.device_mut(), p, S.device_mut(), p,
                          V.device_mut(), n));
  return {U, S, V};
}
inline std::tuple<Dense, std::vector<int64_t>> one_sided_id(Dense& A, const int64_t k) {
  const int m = (int)A.dim[0], n = (int)A.dim[1];
  Dense V(k, n);
  std::vector<int32_t> c(k);
  detail::check(frk_did(detail::ctx(), m, n, A.device(), n, (int)k, V.device_mut(), n, c.data()));
  return {V, std::vector<int64_t>(c.begin(), c.end())};
}
inline Dense get_cols(const Dense& A, const std::vector<int64_t>& cols) {
  Dense out(A.dim[0], (int64_t)cols.size());
  for (int64_t i = 0; i < A.dim[0]; ++i)
    for (size_t j = 0; j < cols.size(); ++j) out(i, j) = A(i, cols[j]);
  return out;
}
inline Dense transpose(const Dense& A) {
  Dense out(A.dim[1], A.dim[0]);
  for (int64_t i = 0; i < A.dim[0]; ++i)
    for (int64_t j = 0; j < A.dim[1]; ++j) out(j, i) = A(i, j);
  return out;
}
// id(A, k) (id.cpp:100-110): A ~ U A[rows, cols] V
inline std::tuple<Dense, Dense, Dense> id(Dense& A, const int64_t k) {
  auto [V, cols] = one_sided_id(A, k);
  Dense Ac = get_cols(A, cols);
  Dense AcT = transpose(Ac);
  auto [Ut, rows] = one_sided_id(AcT, k);
  Dense S(k, k);
  for (int64_t i = 0; i < k; ++i)
    for (int64_t j = 0; j < k; ++j) S(i, j) = Ac(rows[i], j);
  return {transpose(Ut), S, V};
}
// rid(A, p, k) (rid.cpp:18-31)
inline std::tuple<Dense, Dense, Dense> rid(const Dense& A, const int64_t p, const int64_t k) {
  const int64_t n = A.dim[1];
  Dense omega(n, p);
  std::mt19937 gen(0);
  std::uniform_real_distribution<double> dist(0.0, 1.0);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < p; ++j) omega(i, j) = dist(gen);
  Dense Y = gemm(A, omega);
  Dense Q(A.dim[0], p), R(p, p);
  qr(Y, Q, R);
  Dense QtA = gemm(Q, A, 1.0, true, false);
  auto [Ub, S, V] = id(QtA, k);
  return {gemm(Q, Ub), S, V};
}

// ---------------------------------------------------------------------------
// Tile-level QR operators and the typed gemm into low-rank targets
// (LAPACK.h:345, :374, :399, :430; gemm.cpp:202-272) over the batched C ABI.
// A LowRank operand travels in slots of capacity max(64, 2 rank) and comes
// back at its new rank.
// ---------------------------------------------------------------------------
namespace detail {
struct TileArg {
  frk_dtile t{};
  std::shared_ptr<double> u, s, v;
  explicit TileArg(const Dense& D) {
    t.kind = FRK_TILE_DENSE;
    t.rows = (int32_t)D.dim[0];
    t.cols = (int32_t)D.dim[1];
    t.d = const_cast<double*>(D.device());
    t.ld = D.dim[1];
  }
  explicit TileArg(Dense& D, bool) : TileArg(static_cast<const Dense&>(D)) { t.d = D.device_mut(); }
  explicit TileArg(const LowRank& A) {
    const int64_t m = A.dim[0], n = A.dim[1], r = A.rank;
    const int cap = (int)std::max<int64_t>(64, 2 * r);
    u = dalloc((size_t)m * cap);
    s = dalloc((size_t)cap * cap);
    v = dalloc((size_t)cap * n);
    if (r > 0) {
      hip(hipMemcpy2D(u.get(), sizeof(double) * cap, A.U.host(), sizeof(double) * r, sizeof(double) * r, m,
                      hipMemcpyHostToDevice));
      hip(hipMemcpy2D(s.get(), sizeof(double) * cap, A.S.host(), sizeof(double) * r, sizeof(double) * r, r,
                      hipMemcpyHostToDevice));
      hip(hipMemcpy(v.get(), A.V.host(), sizeof(double) * r * n, hipMemcpyHostToDevice));
    }
    t.kind = FRK_TILE_LOWRANK;
    t.rows = (int32_t)m;
    t.cols = (int32_t)n;
    t.rank = (int32_t)r;
    t.cap = cap;
    t.eps = A.eps;
    t.u = u.get();
    t.s = s.get();
    t.v = v.get();
  }
  // the low-rank tile at its current rank, back on the host
  LowRank lowrank(double eps) const {
    const int64_t m = t.rows, n = t.cols, r = t.rank, cap = t.cap;
    check(frk_ctx_synchronize(ctx()));
    std::vector<double> hu((size_t)m * r), hs((size_t)r * r), hv((size_t)r * n);
    if (r > 0) {
      hip(hipMemcpy2D(hu.data(), sizeof(double) * r, u.get(), sizeof(double) * cap, sizeof(double) * r, m,
                      hipMemcpyDeviceToHost));
      hip(hipMemcpy2D(hs.data(), sizeof(double) * r, s.get(), sizeof(double) * cap, sizeof(double) * r, r,
                      hipMemcpyDeviceToHost));
      hip(hipMemcpy(hv.data(), v.get(), sizeof(double) * r * n, hipMemcpyDeviceToHost));
    }
    Dense U(m, r), S(r, r), V(r, n);
    for (int64_t i = 0; i < m; ++i)
      for (int64_t c = 0; c < r; ++c) U(i, c) = hu[(size_t)i * r + c];
    for (int64_t i = 0; i < r; ++i)
      for (int64_t c = 0; c < r; ++c) S(i, c) = hs[(size_t)i * r + c];
    for (int64_t i = 0; i < r; ++i)
      for (int64_t c = 0; c < n; ++c) V(i, c) = hv[(size_t)i * n + c];
    return LowRank(U, S, V, eps);
  }
};
inline TileArg tile_in(const MatrixProxy& A) {
  if (const auto* d = std::get_if<Dense>(&A)) return TileArg(*d);
  if (const auto* l = std::get_if<LowRank>(&A)) return TileArg(*l);
  throw Error("tile operand must be Dense or LowRank");
}
inline TileArg tile_out(MatrixProxy& A) {
  if (auto* d = std::get_if<Dense>(&A)) return TileArg(*d, true);
  if (const auto* l = std::get_if<LowRank>(&A)) return TileArg(*l);
  throw Error("tile operand must be Dense or LowRank");
}
inline void tile_back(MatrixProxy& A, const TileArg& t) {
  if (auto* l = std::get_if<LowRank>(&A)) *l = t.lowrank(l->eps);
}
}  // namespace detail

// geqrt(A, T) (LAPACK.h:345, geqrt.cpp:34-43)
inline void geqrt(Dense& A, Dense& T) {
  double* a = A.device_mut();
  double* t = T.device_mut();
  detail::check(frk_dgeqrt_batched(detail::ctx(), 1, (int)A.dim[0], &a, A.dim[1], &t, T.dim[1]));
}
// larfb(V, T, C, trans) (LAPACK.h:374, larfb.cpp:37-56): Dense C, or LowRank C through its U
inline void larfb(const Dense& V, const Dense& T, MatrixProxy& C, const bool trans) {
  detail::TileArg c = detail::tile_out(C);
  const double* v = V.device();
  const double* t = T.device();
  detail::check(frk_dlarfb_batched(detail::ctx(), 1, (int)V.dim[0], &v, V.dim[1], &t, T.dim[1], &c.t, trans));
  detail::tile_back(C, c);
}
// tpqrt(A, B, T) (LAPACK.h:399, tpqrt.cpp:35-53): Dense or LowRank B
inline void tpqrt(Dense& A, MatrixProxy& B, Dense& T) {
  detail::TileArg b = detail::tile_out(B);
  double* a = A.device_mut();
  double* t = T.device_mut();
  detail::check(frk_dtpqrt_batched(detail::ctx(), 1, (int)A.dim[0], &a, A.dim[1], &b.t, &t, T.dim[1]));
  detail::tile_back(B, b);
}
// tpmqrt(V, T, A, B, trans) (LAPACK.h:430, tpmqrt.cpp:45-170): every Dense / LowRank combination
inline void tpmqrt(const MatrixProxy& V, const Dense& T, MatrixProxy& A, MatrixProxy& B, const bool trans) {
  detail::TileArg v = detail::tile_in(V), a = detail::tile_out(A), b = detail::tile_out(B);
  const double* t = T.device();
  detail::check(frk_dtpmqrt_batched(detail::ctx(), 1, (int)T.dim[0], &v.t, &t, T.dim[1], &a.t, &b.t, trans,
                                    detail::lra() == FRK_LRA_FAST ? FRK_LRA_FAST : FRK_LRA_ROUNDED));
  detail::tile_back(A, a);
  detail::tile_back(B, b);
}
// gemm into a LowRank target (gemm.cpp:202-272: D D LR, LR D LR, D LR LR, LR LR LR)
inline void gemm(const MatrixProxy& A, const MatrixProxy& B, LowRank& C, double alpha, double beta, bool tA,
                 bool tB) {
  detail::TileArg a = detail::tile_in(A), b = detail::tile_in(B), c(C);
  detail::check(frk_dgemm_tiles_batched(detail::ctx(), 1, &a.t, &b.t, &c.t, alpha, beta, tA, tB,
                                        detail::lra() == FRK_LRA_FAST ? FRK_LRA_FAST : FRK_LRA_ROUNDED));
  C = c.lowrank(C.eps);
}

// ---------------------------------------------------------------------------
// Tiled Householder BLR-QR (LAPACK.h:450, :470-472)
// ---------------------------------------------------------------------------
class QrFactors {
 public:
  explicit QrFactors(frk_qrf f) : f_(f, [](frk_qrf q) { frk_qrf_destroy(q); }) {}
  frk_qrf handle() const { return f_.get(); }
  Dense T(int i, int k, int b) const {
    std::vector<double> t((size_t)b * b);
    detail::check(frk_qrf_get_t(handle(), i, k, t.data(), nullptr));
    Dense d(b, b);
    for (int a = 0; a < b; ++a)
      for (int c = 0; c < b; ++c) d(a, c) = t[(size_t)a * b + c];
    return d;
  }

 private:
  std::shared_ptr<frk_qrf_s> f_;
};
inline QrFactors tiled_householder_blr_qr(Hierarchical& A) {
  detail::check(frk_blr_set_lra(A.handle(), detail::lra()));
  frk_qrf f = nullptr;
  detail::check(frk_blr_qr(A.handle(), &f));
  return QrFactors(f);
}
inline void left_multiply_tiled_reflector(const Hierarchical& Y, const QrFactors& T, Hierarchical& C, bool trans) {
  detail::check(frk_blr_set_lra(C.handle(), detail::lra()));
  detail::check(frk_blr_apply_q(Y.handle(), T.handle(), C.handle(), trans ? 1 : 0));
}

// l2_error (src/util/l2_error.cpp:35-101): sqrt(sum (A - B)^2 / sum A^2)
inline double l2_error(const Dense& A, const Dense& B) {
  double num = 0.0, den = 0.0;
  for (int64_t i = 0; i < A.dim[0]; ++i)
    for (int64_t j = 0; j < A.dim[1]; ++j) {
      const double d = A(i, j) - B(i, j);
      num += d * d;
      den += A(i, j) * A(i, j);
    }
  return std::sqrt(num / den);
}
inline double l2_error(const Dense& A, const Hierarchical& B) { return l2_error(A, B.to_dense()); }

}  // namespace frank_amd
