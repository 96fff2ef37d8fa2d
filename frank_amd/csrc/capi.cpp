// extern "C" boundary over the engine (declarations in include/frank_amd.h).
#include <atomic>
#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/frank_amd.h"
#include "engine.hpp"

using namespace frk;

// Handles are reference counted: a matrix holds its context and a QR factor
// holds its matrix, so destroying them in any order (e.g. a garbage collector
// finalizing a cycle) never leaves a dangling owner.
struct frk_ctx_s {
  hipStream_t stream = nullptr;
  std::unique_ptr<Context> ctx;
  std::atomic<int> refs{1};
};

struct frk_blr_s {
  frk_ctx_s* owner = nullptr;
  std::atomic<int> refs{1};
  int precision = FRK_PREC_F64;
  std::unique_ptr<BlrMatrix<double>> a64;
  std::unique_ptr<BlrMatrix<float>> a32;
  bool has_kernel = false;
  KernelSpec ks{};
  DevBuf pts;
  std::vector<double> host_pts;
  KernelDesc kd{};
  double anorm = -1.0;
};

struct frk_qrf_s {
  frk_blr_s* a = nullptr;
  std::unique_ptr<QrFactors<double>> f64;
};

namespace {
thread_local std::string g_err;

void ctx_release(frk_ctx_s* c) {
  if (--c->refs > 0) return;
  (void)hipStreamSynchronize(c->stream);
  c->ctx.reset();
  (void)hipStreamDestroy(c->stream);
  delete c;
}

void blr_release(frk_blr_s* a) {
  if (--a->refs > 0) return;
  frk_ctx_s* o = a->owner;
  (void)hipStreamSynchronize(o->stream);
  delete a;
  ctx_release(o);
}

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

template <class F>
int guard(F&& f) {
  try {
    return f();
  } catch (const std::invalid_argument& e) {
    return fail(FRK_ERR_ARG, e.what());
  } catch (const std::runtime_error& e) {
    const std::string m = e.what();
    if (m.find("capacity") != std::string::npos) return fail(FRK_ERR_CAPACITY, m);
    if (m.find("zero pivot") != std::string::npos) return fail(1, m);
    if (m.find("HIP error") != std::string::npos) return fail(FRK_ERR_HIP, m);
    if (m.find("undefined") != std::string::npos) return fail(FRK_ERR_UNSUPPORTED, m);
    return fail(FRK_ERR_ARG, m);
  } catch (const std::exception& e) {
    return fail(FRK_ERR_ARG, e.what());
  }
}

template <class T>
BlrMatrix<T>& mat(frk_blr a);
template <>
BlrMatrix<double>& mat<double>(frk_blr a) {
  return *a->a64;
}
template <>
BlrMatrix<float>& mat<float>(frk_blr a) {
  return *a->a32;
}

void check_tile(frk_blr a, int i, int j) {
  int nr, nc;
  if (a->a64) {
    nr = a->a64->nb_r();
    nc = a->a64->nb_c();
  } else {
    nr = a->a32->nb_r();
    nc = a->a32->nb_c();
  }
  if (i < 0 || j < 0 || i >= nr || j >= nc) throw std::invalid_argument("tile index out of range");
}

template <class T>
void to_dev(T* dst, const double* src, size_t n, hipStream_t s) {
  if constexpr (sizeof(T) == 8) {
    FRK_HIP_CHECK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, s));
    host_wait_stream(s, __func__);
  } else {
    std::vector<float> tmp(src, src + n);
    FRK_HIP_CHECK(hipMemcpyAsync(dst, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice, s));
    host_wait_stream(s, __func__);
  }
}

template <class T>
void from_dev(double* dst, const T* src, size_t n, hipStream_t s) {
  if constexpr (sizeof(T) == 8) {
    FRK_HIP_CHECK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, s));
    host_wait_stream(s, __func__);
  } else {
    std::vector<float> tmp(n);
    FRK_HIP_CHECK(hipMemcpyAsync(tmp.data(), src, n * sizeof(float), hipMemcpyDeviceToHost, s));
    host_wait_stream(s, __func__);
    for (size_t k = 0; k < n; ++k) dst[k] = tmp[k];
  }
}

template <class T>
int get_dense_t(frk_blr a, int i, int j, double* host) {
  auto& A = mat<T>(a);
  const int t = A.idx(i, j);
  if (A.kind[t] != kDense) throw std::invalid_argument("tile is not dense");
  from_dev<T>(host, A.D[t], (size_t)A.rows(i) * A.cols(j), a->owner->stream);
  return FRK_OK;
}

template <class T>
int set_dense_t(frk_blr a, int i, int j, const double* host) {
  auto& A = mat<T>(a);
  const int t = A.idx(i, j);
  if (A.kind[t] != kDense) throw std::invalid_argument("tile is not dense");
  to_dev<T>(A.D[t], host, (size_t)A.rows(i) * A.cols(j), a->owner->stream);
  return FRK_OK;
}

template <class T>
int get_lowrank_t(frk_blr a, int i, int j, double* U, double* S, double* V) {
  auto& A = mat<T>(a);
  const hipStream_t s = a->owner->stream;
  A.pull_ranks(s);
  const int t = A.idx(i, j);
  if (A.kind[t] != kLowRank) throw std::invalid_argument("tile is not low-rank");
  const int r = A.rank[t];
  const int64_t m = A.rows(i), n = A.cols(j);
  std::vector<double> u((size_t)m * A.kcap), sv((size_t)A.kcap * A.kcap), v((size_t)A.kcap * n);
  from_dev<T>(u.data(), A.U[t], u.size(), s);
  from_dev<T>(sv.data(), A.S[t], sv.size(), s);
  from_dev<T>(v.data(), A.V[t], v.size(), s);
  for (int64_t row = 0; row < m; ++row)
    for (int c = 0; c < r; ++c) U[row * r + c] = u[row + (size_t)c * m];  // column-major slot
  for (int p = 0; p < r; ++p)
    for (int q = 0; q < r; ++q) S[p * r + q] = sv[(size_t)p * A.kcap + q];
  for (int p = 0; p < r; ++p) std::memcpy(V + (size_t)p * n, v.data() + (size_t)p * n, sizeof(double) * n);
  return FRK_OK;
}

template <class T>
int set_lowrank_t(frk_blr a, int i, int j, int r, const double* U, const double* S, const double* V) {
  auto& A = mat<T>(a);
  const hipStream_t s = a->owner->stream;
  const int t = A.idx(i, j);
  if (A.kind[t] != kLowRank) throw std::invalid_argument("tile is not low-rank");
  if (r < 0 || r > A.kcap) throw std::runtime_error("rank exceeds the low-rank capacity");
  const int64_t m = A.rows(i), n = A.cols(j);
  std::vector<double> u((size_t)m * A.kcap, 0.0), sv((size_t)A.kcap * A.kcap, 0.0), v((size_t)A.kcap * n, 0.0);
  for (int64_t row = 0; row < m; ++row)
    for (int c = 0; c < r; ++c) u[row + (size_t)c * m] = U[row * r + c];
  for (int p = 0; p < r; ++p)
    for (int q = 0; q < r; ++q) sv[(size_t)p * A.kcap + q] = S[p * r + q];
  for (int p = 0; p < r; ++p) std::memcpy(v.data() + (size_t)p * n, V + (size_t)p * n, sizeof(double) * n);
  to_dev<T>(A.U[t], u.data(), u.size(), s);
  to_dev<T>(A.S[t], sv.data(), sv.size(), s);
  to_dev<T>(A.V[t], v.data(), v.size(), s);
  A.pull_ranks(s);
  A.rank[t] = r;
  A.push_ranks(s);
  return FRK_OK;
}

}  // namespace

extern "C" {

const char* frk_last_error(void) { return g_err.c_str(); }

int frk_version(void) { return 1; }

int frk_ctx_create(frk_ctx* out) {
  return guard([&] {
    if (!out) throw std::invalid_argument("out is null");
    auto c = std::make_unique<frk_ctx_s>();
    FRK_HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->ctx = std::make_unique<Context>(c->stream);
    *out = c.release();
    return FRK_OK;
  });
}

int frk_ctx_destroy(frk_ctx ctx) {
  return guard([&] {
    if (!ctx) return FRK_OK;
    host_wait_stream(ctx->stream, __func__);
    ctx_release(ctx);  // freed once the last matrix built on it is destroyed
    return FRK_OK;
  });
}

int frk_ctx_synchronize(frk_ctx ctx) {
  return guard([&] {
    if (!ctx) throw std::invalid_argument("ctx is null");
    host_wait_stream(ctx->stream, __func__);
    return FRK_OK;
  });
}

int frk_ctx_stream(frk_ctx ctx, void** stream) {
  return guard([&] {
    if (!ctx || !stream) throw std::invalid_argument("null argument");
    *stream = (void*)ctx->stream;
    return FRK_OK;
  });
}

int frk_ctx_counters(frk_ctx ctx, int64_t* out, int n, int reset) {
  return guard([&] {
    if (!ctx || (!out && n > 0) || n < 0) throw std::invalid_argument("null or negative argument");
    std::vector<int32_t> c(kCounterWords);
    host_wait_stream(ctx->stream, __func__);
    FRK_HIP_CHECK(hipMemcpy(c.data(), ctx->ctx->counters(), sizeof(int32_t) * c.size(), hipMemcpyDeviceToHost));
    int64_t sum[kNumCounters] = {};
    for (int k = 0; k < kCounterCopies; ++k)
      for (int i = 0; i < kNumCounters; ++i) {
        const int64_t v = c[(size_t)k * kCounterStride + i];
        sum[i] = (i == kCounterSvdMaxSweeps) ? std::max(sum[i], v) : sum[i] + v;
      }
    for (int i = 0; i < n && i < kNumCounters; ++i) out[i] = sum[i];
    if (reset) {
      FRK_HIP_CHECK(hipMemsetAsync(ctx->ctx->counters(), 0, sizeof(int32_t) * c.size(), ctx->stream));
      host_wait_stream(ctx->stream, __func__);
    }
    return FRK_OK;
  });
}

int frk_ctx_time_class(frk_ctx ctx, const char* cls) {
  return guard([&] {
    if (!ctx) throw std::invalid_argument("ctx is null");
    ctx->ctx->sync();
    ctx->ctx->time_class(cls ? std::string(cls) : std::string());
    return FRK_OK;
  });
}

int frk_ctx_set_rank_cap(frk_ctx ctx, int truncate) {
  return guard([&] {
    if (!ctx) throw std::invalid_argument("ctx is null");
    ctx->ctx->set_cap_trunc(truncate != 0);
    return FRK_OK;
  });
}

int frk_ctx_op_classes(frk_ctx ctx, char* buf, int len) {
  return guard([&] {
    if (!ctx || !buf || len <= 0) throw std::invalid_argument("null argument or empty buffer");
    std::string all;
    for (const auto& kv : ctx->ctx->op_stats()) {
      if (!all.empty()) all += ",";
      all += kv.first;
    }
    const size_t n = std::min(all.size(), (size_t)len - 1);
    std::memcpy(buf, all.data(), n);
    buf[n] = 0;
    return FRK_OK;
  });
}

int frk_ctx_op_stats(frk_ctx ctx, const char* cls, int64_t* launches, double* flops, double* bytes, double* ms,
                     int64_t* timed_launches) {
  return guard([&] {
    if (!ctx || !cls) throw std::invalid_argument("null argument");
    const auto& st = ctx->ctx->op_stats();
    const auto it = st.find(cls);
    OpStat z;
    const OpStat& o = (it == st.end()) ? z : it->second;
    if (launches) *launches = o.launches;
    if (flops) *flops = o.flops;
    if (bytes) *bytes = o.bytes;
    if (ms) *ms = o.ms;
    if (timed_launches) *timed_launches = o.timed;
    return FRK_OK;
  });
}

int frk_ctx_op_roofline(frk_ctx ctx, const char* cls, double* roofline_ms) {
  return guard([&] {
    if (!ctx || !cls || !roofline_ms) throw std::invalid_argument("null argument");
    const auto& st = ctx->ctx->op_stats();
    const auto it = st.find(cls);
    *roofline_ms = (it == st.end()) ? 0.0 : 1e3 * it->second.roof_s;
    return FRK_OK;
  });
}

int frk_ctx_reset_op_stats(frk_ctx ctx) {
  return guard([&] {
    if (!ctx) throw std::invalid_argument("ctx is null");
    ctx->ctx->reset_op_stats();
    return FRK_OK;
  });
}

int frk_blr_create_kernel_rect(frk_ctx ctx, int kernel_kind, int dim, int64_t n_pts, const double* pts,
                               int64_t n_rows, int64_t n_cols, int nb_r, int nb_c, double ell, double shift,
                               double eps, int rank, double admis, int admis_type, int kcap, int precision,
                               frk_blr* out) {
  return guard([&] {
    if (!ctx || !out || !pts) throw std::invalid_argument("null argument");
    if (dim < 1 || dim > 3) throw std::invalid_argument("dim must be 1, 2 or 3");
    if (kernel_kind < 0 || kernel_kind > 2) throw std::invalid_argument("unknown kernel");
    if (n_rows <= 0 || n_cols <= 0 || n_rows > n_pts || n_cols > n_pts)
      throw std::invalid_argument("matrix size must be positive and at most the point count");
    if (nb_r < 1 || nb_c < 1 || n_rows < nb_r || n_cols < nb_c) throw std::invalid_argument("bad block count");
    if ((n_rows + nb_r - 1) / nb_r > 512 || (n_cols + nb_c - 1) / nb_c > 512)
      throw std::invalid_argument("tile size above 512");
    if (eps == 0.0 && rank <= 0) throw std::invalid_argument("need eps != 0 or rank > 0");
    if (kcap <= 0) kcap = (eps == 0.0) ? rank : 64;
    if (eps == 0.0 && rank > kcap) throw std::invalid_argument("rank exceeds kcap");
    KernelDesc kd{kernel_kind, dim, n_pts, pts, ell, shift};
    kd.n_rows = n_rows;
    kd.n_cols = n_cols;
    auto h = std::make_unique<frk_blr_s>();
    h->owner = ctx;
    h->precision = precision;
    if (precision == FRK_PREC_F64)
      h->a64 = build_blr_kernel<double>(kd, nb_r, nb_c, eps, rank, admis, admis_type, kcap, *ctx->ctx);
    else if (precision == FRK_PREC_F32)
      h->a32 = build_blr_kernel<float>(kd, nb_r, nb_c, eps, rank, admis, admis_type, kcap, *ctx->ctx);
    else
      throw std::invalid_argument("unknown precision");
    h->host_pts.assign(pts, pts + (size_t)dim * n_pts);
    h->kd = kd;
    h->kd.host_pts = h->host_pts.data();
    h->ks = upload_kernel(h->kd, h->pts, ctx->stream);
    h->has_kernel = n_rows == n_cols && n_rows == n_pts;
    host_wait_stream(ctx->stream, __func__);
    ++h->owner->refs;  // the matrix holds its context
    *out = h.release();
    return FRK_OK;
  });
}

int frk_blr_create_kernel(frk_ctx ctx, int kernel_kind, int dim, int64_t n, const double* pts, double ell,
                          double shift, int nb, double eps, int rank, double admis, int admis_type, int kcap,
                          int precision, frk_blr* out) {
  return frk_blr_create_kernel_rect(ctx, kernel_kind, dim, n, pts, n, n, nb, nb, ell, shift, eps, rank, admis,
                                    admis_type, kcap, precision, out);
}

int frk_blr_create(frk_ctx ctx, int nb_r, int nb_c, const int64_t* roff, const int64_t* coff, const int32_t* kinds,
                   int kcap, double eps, int rank, int precision, frk_blr* out) {
  return guard([&] {
    if (!ctx || !out || !roff || !coff || !kinds) throw std::invalid_argument("null argument");
    if (nb_r < 1 || nb_c < 1) throw std::invalid_argument("bad block counts");
    std::vector<int64_t> ro(roff, roff + nb_r + 1), co(coff, coff + nb_c + 1);
    for (int i = 0; i < nb_r; ++i)
      if (ro[i + 1] - ro[i] <= 0 || ro[i + 1] - ro[i] > 512) throw std::invalid_argument("bad row block size");
    for (int j = 0; j < nb_c; ++j)
      if (co[j + 1] - co[j] <= 0 || co[j + 1] - co[j] > 512) throw std::invalid_argument("bad column block size");
    std::vector<int32_t> kd(kinds, kinds + (size_t)nb_r * nb_c);
    if (kcap <= 0) throw std::invalid_argument("kcap must be positive");
    auto h = std::make_unique<frk_blr_s>();
    h->owner = ctx;
    h->precision = precision;
    if (precision == FRK_PREC_F64)
      h->a64 = std::make_unique<BlrMatrix<double>>(ro, co, kd, kcap, eps, rank, *ctx->ctx);
    else if (precision == FRK_PREC_F32)
      h->a32 = std::make_unique<BlrMatrix<float>>(ro, co, kd, kcap, eps, rank, *ctx->ctx);
    else
      throw std::invalid_argument("unknown precision");
    host_wait_stream(ctx->stream, __func__);
    ++h->owner->refs;  // the matrix holds its context
    *out = h.release();
    return FRK_OK;
  });
}

int frk_blr_clone(frk_blr src, frk_blr* out) {
  return guard([&] {
    if (!src || !out) throw std::invalid_argument("null argument");
    auto h = std::make_unique<frk_blr_s>();
    h->owner = src->owner;
    h->precision = src->precision;
    Context& c = *src->owner->ctx;
    if (src->a64) {
      h->a64 = src->a64->clone_structure(c);
      h->a64->copy_from(*src->a64, src->owner->stream);
    } else {
      h->a32 = src->a32->clone_structure(c);
      h->a32->copy_from(*src->a32, src->owner->stream);
    }
    if (src->has_kernel) {
      h->host_pts = src->host_pts;
      h->kd = src->kd;
      h->kd.host_pts = h->host_pts.data();
      h->ks = upload_kernel(h->kd, h->pts, src->owner->stream);
      h->has_kernel = true;
      h->anorm = src->anorm;
    }
    host_wait_stream(src->owner->stream, __func__);
    ++h->owner->refs;  // the matrix holds its context
    *out = h.release();
    return FRK_OK;
  });
}

int frk_blr_convert(frk_blr src, int precision, frk_blr* out) {
  return guard([&] {
    if (!src || !out) throw std::invalid_argument("null argument");
    auto h = std::make_unique<frk_blr_s>();
    h->owner = src->owner;
    h->precision = precision;
    Context& c = *src->owner->ctx;
    if (precision == FRK_PREC_F32)
      h->a32 = src->a64 ? convert_blr<float, double>(*src->a64, c) : std::unique_ptr<BlrMatrix<float>>();
    else
      h->a64 = src->a32 ? convert_blr<double, float>(*src->a32, c) : std::unique_ptr<BlrMatrix<double>>();
    if (!h->a32 && !h->a64) throw std::invalid_argument("source already has that precision (use frk_blr_clone)");
    if (src->has_kernel) {
      h->host_pts = src->host_pts;
      h->kd = src->kd;
      h->kd.host_pts = h->host_pts.data();
      h->ks = upload_kernel(h->kd, h->pts, src->owner->stream);
      h->has_kernel = true;
      h->anorm = src->anorm;
    }
    host_wait_stream(src->owner->stream, __func__);
    ++h->owner->refs;  // the matrix holds its context
    *out = h.release();
    return FRK_OK;
  });
}

int frk_blr_copy(frk_blr dst, frk_blr src) {
  return guard([&] {
    if (!dst || !src) throw std::invalid_argument("null argument");
    if ((dst->a64 != nullptr) != (src->a64 != nullptr)) throw std::invalid_argument("precisions differ");
    Context& c = *dst->owner->ctx;
    auto run = [&](auto& d, auto& s) {
      d.copy_from(s, dst->owner->stream);
      // op log: one read and one write of the tile arena and the diagonal inverses
      using E = std::remove_reference_t<decltype(*d.dinv_data())>;
      c.account("restore_copy", 0.0, 2.0 * ((double)d.arena_bytes() + sizeof(E) * (double)d.dinv_count()));
    };
    if (dst->a64)
      run(*dst->a64, *src->a64);
    else
      run(*dst->a32, *src->a32);
    return FRK_OK;
  });
}

int frk_blr_refine(frk_blr f, const double* b_dev, double* x_dev, int nrhs, double tol, int max_iter, int* iters,
                   double* berr) {
  return guard([&] {
    if (!f || !b_dev || !x_dev || nrhs <= 0) throw std::invalid_argument("bad argument");
    if (!f->has_kernel) throw std::invalid_argument("refinement needs a matrix built from a kernel");
    Context& c = *f->owner->ctx;
    const int64_t n = f->a64 ? f->a64->roff.back() : f->a32->roff.back();
    if (f->anorm < 0) f->anorm = kernel_norm_inf(f->ks, n, c);
    RefineResult r = f->a64 ? blr_refine<double>(*f->a64, f->ks, f->anorm, b_dev, x_dev, nrhs, tol, max_iter, c)
                            : blr_refine<float>(*f->a32, f->ks, f->anorm, b_dev, x_dev, nrhs, tol, max_iter, c);
    if (iters) *iters = r.iters;
    if (berr) *berr = r.berr;
    return FRK_OK;
  });
}

int frk_blr_destroy(frk_blr a) {
  return guard([&] {
    if (!a) return FRK_OK;
    host_wait_stream(a->owner->stream, __func__);
    blr_release(a);
    return FRK_OK;
  });
}

int frk_blr_set_lra(frk_blr a, int lra) {
  return guard([&] {
    if (!a) throw std::invalid_argument("null matrix");
    if (lra != FRK_LRA_NAIVE && lra != FRK_LRA_ROUNDED && lra != FRK_LRA_FAST)
      throw std::invalid_argument("unknown FRANK_LRA value");
    if (a->a64) a->a64->lra = lra;
    if (a->a32) a->a32->lra = lra;
    return FRK_OK;
  });
}

int frk_blr_shape(frk_blr a, int* nb_r, int* nb_c) {
  return guard([&] {
    if (!a || !nb_r || !nb_c) throw std::invalid_argument("null argument");
    *nb_r = a->a64 ? a->a64->nb_r() : a->a32->nb_r();
    *nb_c = a->a64 ? a->a64->nb_c() : a->a32->nb_c();
    return FRK_OK;
  });
}

int frk_blr_tile_info(frk_blr a, int i, int j, int32_t* kind, int32_t* rank, int64_t* rows, int64_t* cols) {
  return guard([&] {
    if (!a) throw std::invalid_argument("null matrix");
    check_tile(a, i, j);
    auto fill = [&](auto& A) {
      A.pull_ranks(a->owner->stream);
      const int t = A.idx(i, j);
      if (kind) *kind = A.kind[t];
      if (rank) *rank = A.rank[t];
      if (rows) *rows = A.rows(i);
      if (cols) *cols = A.cols(j);
    };
    if (a->a64)
      fill(*a->a64);
    else
      fill(*a->a32);
    return FRK_OK;
  });
}

int frk_blr_get_dense(frk_blr a, int i, int j, double* host) {
  return guard([&] {
    if (!a || !host) throw std::invalid_argument("null argument");
    check_tile(a, i, j);
    return a->a64 ? get_dense_t<double>(a, i, j, host) : get_dense_t<float>(a, i, j, host);
  });
}

int frk_blr_set_dense(frk_blr a, int i, int j, const double* host) {
  return guard([&] {
    if (!a || !host) throw std::invalid_argument("null argument");
    check_tile(a, i, j);
    return a->a64 ? set_dense_t<double>(a, i, j, host) : set_dense_t<float>(a, i, j, host);
  });
}

int frk_blr_get_lowrank(frk_blr a, int i, int j, double* U, double* S, double* V) {
  return guard([&] {
    if (!a || !U || !S || !V) throw std::invalid_argument("null argument");
    check_tile(a, i, j);
    return a->a64 ? get_lowrank_t<double>(a, i, j, U, S, V) : get_lowrank_t<float>(a, i, j, U, S, V);
  });
}

int frk_blr_set_lowrank(frk_blr a, int i, int j, int rank, const double* U, const double* S, const double* V) {
  return guard([&] {
    if (!a || !U || !S || !V) throw std::invalid_argument("null argument");
    check_tile(a, i, j);
    return a->a64 ? set_lowrank_t<double>(a, i, j, rank, U, S, V) : set_lowrank_t<float>(a, i, j, rank, U, S, V);
  });
}

int frk_blr_getrf(frk_blr a) { return frk_blr_getrf_steps(a, -1); }

int frk_blr_getrf_steps(frk_blr a, int steps) {
  return guard([&] {
    if (!a) throw std::invalid_argument("null matrix");
    if (a->a64)
      blr_getrf<double>(*a->a64, *a->owner->ctx, steps);
    else
      blr_getrf<float>(*a->a32, *a->owner->ctx, steps);
    return FRK_OK;
  });
}

int frk_blr_qr(frk_blr a, frk_qrf* out) {
  return guard([&] {
    if (!a || !out) throw std::invalid_argument("null argument");
    if (!a->a64) throw std::runtime_error("tiled_householder_blr_qr(FP32 Hierarchical) undefined!");
    auto f = std::make_unique<frk_qrf_s>();
    f->a = a;
    f->f64 = std::make_unique<QrFactors<double>>(*a->a64);
    blr_tiled_qr<double>(*a->a64, *f->f64, *a->owner->ctx);
    ++a->refs;  // the factors hold their matrix
    *out = f.release();
    return FRK_OK;
  });
}

int frk_qrf_destroy(frk_qrf f) {
  return guard([&] {
    if (!f) return FRK_OK;
    host_wait_stream(f->a->owner->stream, __func__);
    frk_blr_s* a = f->a;
    delete f;
    blr_release(a);
    return FRK_OK;
  });
}

int frk_qrf_get_t(frk_qrf f, int i, int k, double* t_out, double* tau_out) {
  return guard([&] {
    if (!f) throw std::invalid_argument("null factors");
    const QrFactors<double>& F = *f->f64;
    if (i < 0 || k < 0 || i >= F.nb_r || k >= F.nb_c || i < k || k >= (int)F.bsize.size())
      throw std::invalid_argument("T(i, k) exists for i >= k only");
    const size_t b = F.bsize[k];
    host_wait_stream(f->a->owner->stream, __func__);
    if (t_out) FRK_HIP_CHECK(hipMemcpy(t_out, F.tmat(i, k), sizeof(double) * b * b, hipMemcpyDeviceToHost));
    if (tau_out) FRK_HIP_CHECK(hipMemcpy(tau_out, F.tau_of(i, k), sizeof(double) * b, hipMemcpyDeviceToHost));
    return FRK_OK;
  });
}

int frk_blr_apply_q(frk_blr y, frk_qrf f, frk_blr c, int trans) {
  return guard([&] {
    if (!y || !f || !c) throw std::invalid_argument("null argument");
    if (!y->a64 || !c->a64) throw std::runtime_error("left_multiply_tiled_reflector(FP32) undefined!");
    if (f->a != y) throw std::invalid_argument("factors belong to another matrix");
    blr_apply_q<double>(*y->a64, *f->f64, *c->a64, trans != 0, *y->owner->ctx);
    return FRK_OK;
  });
}

int frk_blr_trsm(frk_blr f, double* b_dev, int nrhs, int upper, int trans) {
  return guard([&] {
    if (!f || !b_dev || nrhs <= 0) throw std::invalid_argument("bad argument");
    if (!f->a64) throw std::runtime_error("trsm(FP32 Hierarchical, FP64 Dense) undefined!");
    blr_solve<double>(*f->a64, b_dev, nrhs, trans != 0, *f->owner->ctx, upper ? 2 : 1);
    host_wait_stream(f->owner->stream, __func__);
    return FRK_OK;
  });
}

int frk_blr_solve(frk_blr f, double* b_dev, int nrhs, int trans) {
  return guard([&] {
    if (!f || !b_dev || nrhs <= 0) throw std::invalid_argument("bad argument");
    if (!f->a64) throw std::runtime_error("frk_blr_solve(FP32 factors, FP64 vector) undefined!");
    blr_solve<double>(*f->a64, b_dev, nrhs, trans != 0, *f->owner->ctx);
    host_wait_stream(f->owner->stream, __func__);
    return FRK_OK;
  });
}

int frk_blr_matvec(frk_blr a, const double* x_dev, double* y_dev, int nrhs) {
  return guard([&] {
    if (!a || !x_dev || !y_dev || nrhs <= 0) throw std::invalid_argument("bad argument");
    if (!a->a64) throw std::runtime_error("frk_blr_matvec(FP32 tiles, FP64 vector) undefined!");
    blr_matvec<double>(*a->a64, x_dev, y_dev, nrhs, *a->owner->ctx);
    host_wait_stream(a->owner->stream, __func__);
    return FRK_OK;
  });
}

int frk_blr_kernel_residual(frk_blr a, const double* x_dev, const double* b_dev, double* r_dev, int nrhs) {
  return guard([&] {
    if (!a || !x_dev || !b_dev || !r_dev || nrhs <= 0) throw std::invalid_argument("bad argument");
    if (!a->has_kernel) throw std::invalid_argument("matrix was not built from a kernel");
    const int64_t n = a->a64 ? a->a64->roff.back() : a->a32->roff.back();
    double* part = static_cast<double*>(
        a->owner->ctx->persistent(7, sizeof(double) * kernel_residual_scratch(n, nrhs)));
    launch_kernel_residual(a->ks, x_dev, b_dev, r_dev, n, nrhs, part, a->owner->stream);
    host_wait_stream(a->owner->stream, __func__);
    return FRK_OK;
  });
}

int frk_kernel_apply(frk_ctx ctx, int kernel_kind, int dim, int64_t n, const double* pts, double ell, double shift,
                     const double* x_dev, double* y_dev, int nrhs) {
  return guard([&] {
    if (!ctx || !pts || !x_dev || !y_dev || nrhs <= 0 || n <= 0) throw std::invalid_argument("bad argument");
    if (dim < 1 || dim > 3) throw std::invalid_argument("dim must be 1, 2 or 3");
    if (kernel_kind < 0 || kernel_kind > 3) throw std::invalid_argument("unknown kernel");
    KernelDesc kd{kernel_kind, dim, n, pts, ell, shift};
    DevBuf dpts;
    const KernelSpec ks = upload_kernel(kd, dpts, ctx->stream);
    double* part = static_cast<double*>(ctx->ctx->persistent(7, sizeof(double) * kernel_residual_scratch(n, nrhs)));
    launch_kernel_residual(ks, x_dev, nullptr, y_dev, n, nrhs, part, ctx->stream);
    host_wait_stream(ctx->stream, __func__);
    return FRK_OK;
  });
}

int frk_dgetrf_nopiv_batched(frk_ctx ctx, double* const* tiles_host, int batch, int n, int64_t ld,
                             int32_t* info_host) {
  return guard([&] {
    if (!ctx || !tiles_host || batch <= 0 || n <= 0 || n > 512 || ld < n) throw std::invalid_argument("bad argument");
    Context& c = *ctx->ctx;
    c.sync();
    std::vector<double*> tiles(tiles_host, tiles_host + batch);
    const int npan = (n + 63) / 64;
    if ((size_t)batch * npan > 4096) throw std::invalid_argument("batch too large for the info words (batch * ceil(n / 64) <= 4096)");
    int32_t* info = c.info();
    FRK_HIP_CHECK(hipMemsetAsync(info, 0, sizeof(int32_t) * batch * npan, ctx->stream));
    getrf_tiles<double>(tiles, n, ld, info, c);
    std::vector<int32_t> h((size_t)batch * npan);
    FRK_HIP_CHECK(hipMemcpyAsync(h.data(), info, sizeof(int32_t) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    c.sync();
    int status = FRK_OK;
    for (int b = 0; b < batch; ++b) {
      int first = 0;
      for (int p = 0; p < npan && first == 0; ++p)
        if (h[(size_t)p * batch + b] != 0) first = p * 64 + h[(size_t)p * batch + b];
      if (info_host) info_host[b] = first;
      if (first != 0 && status == FRK_OK) status = 1;
    }
    if (status != FRK_OK) g_err = "zero pivot in dense getrf";
    return status;
  });
}

int frk_dgetrf_nopiv_batched_noalloc(frk_ctx ctx, double* const* tiles_host, int batch, int n, int64_t ld,
                                     int32_t* info_host) {
  return guard([&] {
    if (!ctx || !tiles_host || batch <= 0 || n <= 0 || ld < n) throw std::invalid_argument("bad argument");
    for (int b = 0; b < batch; ++b)
      if (!tiles_host[b]) throw std::invalid_argument("bad argument: null tile pointer");
    std::string msg;
    const int st = getrf_nopiv_noalloc(ctx->stream, tiles_host, batch, n, ld, info_host, msg);
    if (st != FRK_OK) g_err = msg;
    return st;
  });
}

int frk_dtrsm(frk_ctx ctx, int side, int lower, int trans, int m, int n, const double* A, int64_t lda, double* B,
              int64_t ldb) {
  return guard([&] {
    if (!ctx || !A || !B || m <= 0 || n <= 0) throw std::invalid_argument("bad argument");
    Context& c = *ctx->ctx;
    c.sync();
    const bool unit = lower != 0;  // trsm.cpp:92-104: Lower => Unit, Upper => NonUnit
    // op(A) is lower-triangular iff (lower XOR trans)
    const bool op_lower = (lower != 0) != (trans != 0);
    if (side == 0) {  // op(A) X = B, A m x m
      if (m > 512) throw std::invalid_argument("triangle above 512");
      trsm_generic<double>(A, trans ? 1 : lda, trans ? lda : 1, op_lower, unit, B, ldb, 1, m, n, c);
    } else {  // X op(A) = B  <=>  op(A)^T X^T = B^T, A n x n
      if (n > 512) throw std::invalid_argument("triangle above 512");
      trsm_generic<double>(A, trans ? lda : 1, trans ? 1 : lda, !op_lower, unit, B, 1, ldb, n, m, c);
    }
    c.sync();
    return FRK_OK;
  });
}

int frk_dgemm(frk_ctx ctx, int trans_a, int trans_b, int m, int n, int k, double alpha, const double* A,
              int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc) {
  return guard([&] {
    if (!ctx || !C || m < 0 || n < 0 || k < 0) throw std::invalid_argument("bad argument");
    Context& c = *ctx->ctx;
    c.sync();
    GemmItem g{};
    g.A = A;
    g.B = B;
    g.C = C;
    g.a_rs = trans_a ? 1 : lda;
    g.a_cs = trans_a ? lda : 1;
    g.b_rs = trans_b ? 1 : ldb;
    g.b_cs = trans_b ? ldb : 1;
    g.c_rs = ldc;
    g.c_cs = 1;
    g.m = m;
    g.n = n;
    g.k = k;
    g.alpha = alpha;
    g.beta = beta;
    const GemmItem* dg = static_cast<const GemmItem*>(c.staging().push_bytes(&g, sizeof(g), ctx->stream));
    launch_gemm<double>(dg, 1, m, n, ctx->stream);
    c.sync();
    return FRK_OK;
  });
}

extern "C++" {
namespace {
template <typename T>
void lr_add_batched_impl(frk_ctx ctx, int lra, int batch, const int32_t* m, const int32_t* n, double eps, int cap,
                         T* const* Ua, T* const* Sa, T* const* Va, int32_t* ranks_host, const T* const* Ub,
                         const T* const* Sb, const T* const* Vb, const int32_t* kb) {
  if (!ctx || batch <= 0 || !m || !n || !Ua || !Sa || !Va || !ranks_host || !Ub || !Sb || !Vb || !kb)
    throw std::invalid_argument("null argument");
  if (lra != FRK_LRA_NAIVE && lra != FRK_LRA_ROUNDED && lra != FRK_LRA_FAST)
    throw std::invalid_argument("unknown FRANK_LRA value");
  Context& c = *ctx->ctx;
  c.sync();
  DevBuf dr(sizeof(int32_t) * batch);
  FRK_HIP_CHECK(hipMemcpyAsync(dr.as<int32_t>(), ranks_host, sizeof(int32_t) * batch, hipMemcpyHostToDevice,
                               ctx->stream));
  std::vector<LrAddDesc<T>> d(batch);
  size_t need = 4096;
  int kmax = 0, mmax = 0, nmax = 0;
  for (int b = 0; b < batch; ++b) {
    if (m[b] <= 0 || n[b] <= 0 || m[b] > 512 || n[b] > 512) throw std::invalid_argument("tile dims must be in [1, 512]");
    if (ranks_host[b] <= 0 || ranks_host[b] > cap || kb[b] <= 0) throw std::invalid_argument("bad ranks");
    kmax = std::max(kmax, ranks_host[b] + kb[b]);
    mmax = std::max(mmax, m[b]);
    nmax = std::max(nmax, n[b]);
  }
  need += (size_t)batch * sizeof(T) * ((size_t)(mmax + nmax) * kmax + 6 * (size_t)kmax * kmax + 2 * kmax + 512);
  c.reserve(need);
  for (int b = 0; b < batch; ++b) {
    LrAddDesc<T> x{};
    x.cu = Ua[b];
    x.cs = Sa[b];
    x.cv = Va[b];
    x.d_rank = dr.as<int32_t>() + b;
    x.kc = ranks_host[b];
    x.cu_rs = cap;
    x.cu_cs = 1;
    x.cv_rs = n[b];
    x.cv_cs = 1;
    x.lds = cap;
    x.m = m[b];
    x.n = n[b];
    x.eps = eps;
    x.cap = cap;
    x.xu = Ub[b];
    x.xu_rs = kb[b];
    x.xu_cs = 1;
    x.kx1 = kb[b];
    x.xs = Sb[b];
    x.xs_ld = kb[b];
    x.kx2 = kb[b];
    x.xv = Vb[b];
    x.xv_rs = n[b];
    x.xv_cs = 1;
    d[b] = x;
  }
  lr_add_batched<T>(d, lra, c);
  FRK_HIP_CHECK(hipMemcpyAsync(ranks_host, dr.as<int32_t>(), sizeof(int32_t) * batch, hipMemcpyDeviceToHost,
                               ctx->stream));
  c.sync();
  for (int b = 0; b < batch; ++b)
    if (ranks_host[b] < 0) throw std::runtime_error("low-rank capacity exceeded in addition");
}
}  // namespace
}  // extern "C++"

int frk_dlr_add_batched(frk_ctx ctx, int lra, int batch, const int32_t* m, const int32_t* n, double eps, int cap,
                        double* const* Ua, double* const* Sa, double* const* Va, int32_t* ranks_host,
                        const double* const* Ub, const double* const* Sb, const double* const* Vb,
                        const int32_t* kb) {
  return guard([&] {
    lr_add_batched_impl<double>(ctx, lra, batch, m, n, eps, cap, Ua, Sa, Va, ranks_host, Ub, Sb, Vb, kb);
    return FRK_OK;
  });
}

int frk_slr_add_batched(frk_ctx ctx, int lra, int batch, const int32_t* m, const int32_t* n, double eps, int cap,
                        float* const* Ua, float* const* Sa, float* const* Va, int32_t* ranks_host,
                        const float* const* Ub, const float* const* Sb, const float* const* Vb,
                        const int32_t* kb) {
  return guard([&] {
    lr_add_batched_impl<float>(ctx, lra, batch, m, n, eps, cap, Ua, Sa, Va, ranks_host, Ub, Sb, Vb, kb);
    return FRK_OK;
  });
}

namespace {
RawTile<double> raw_of(const frk_dtile& t) {
  RawTile<double> r;
  if (t.kind != FRK_TILE_DENSE && t.kind != FRK_TILE_LOWRANK) throw std::invalid_argument("tile kind must be dense or low-rank");
  if (t.rows <= 0 || t.cols <= 0) throw std::invalid_argument("tile dimensions must be positive");
  r.kind = t.kind == FRK_TILE_DENSE ? kDense : kLowRank;
  r.rows = t.rows;
  r.cols = t.cols;
  if (r.kind == kDense) {
    if (!t.d || t.ld < t.cols) throw std::invalid_argument("dense tile: null pointer or ld < cols");
    r.d = t.d;
    r.ld = t.ld;
  } else {
    if (!t.u || !t.s || !t.v) throw std::invalid_argument("low-rank tile: null factor pointer");
    if (t.cap <= 0 || t.rank < 0 || t.rank > t.cap) throw std::invalid_argument("low-rank tile: need 0 <= rank <= cap");
    r.u = t.u;
    r.u_rs = t.cap;  // U rows x cap row-major, as frk_dlr_add_batched takes it
    r.u_cs = 1;
    r.s = t.s;
    r.v = t.v;
    r.rank = t.rank;
    r.cap = t.cap;
    r.eps = t.eps;
  }
  return r;
}
std::vector<RawTile<double>> raws(const frk_dtile* t, int n) {
  if (!t) throw std::invalid_argument("null tile array");
  std::vector<RawTile<double>> v;
  for (int q = 0; q < n; ++q) v.push_back(raw_of(t[q]));
  return v;
}
void ranks_back(frk_dtile* t, const std::vector<RawTile<double>>& r) {
  for (size_t q = 0; q < r.size(); ++q)
    if (r[q].kind == kLowRank) t[q].rank = r[q].rank;
}
int check_lra(int lra) {
  if (lra != FRK_LRA_ROUNDED && lra != FRK_LRA_FAST) throw std::invalid_argument("lra must be rounded or fast");
  return lra == FRK_LRA_ROUNDED ? kLraRounded : kLraFast;
}
}  // namespace

int frk_dgemm_tiles_batched(frk_ctx ctx, int batch, const frk_dtile* A, const frk_dtile* B, frk_dtile* C,
                            double alpha, double beta, int trans_a, int trans_b, int lra) {
  return guard([&] {
    if (!ctx || batch <= 0) throw std::invalid_argument("bad argument");
    const int l = check_lra(lra);
    auto a = raws(A, batch), b = raws(B, batch), c = raws(C, batch);
    for (int q = 0; q < batch; ++q) {
      const int m = trans_a ? a[q].cols : a[q].rows, k = trans_a ? a[q].rows : a[q].cols;
      const int kb = trans_b ? b[q].cols : b[q].rows, n = trans_b ? b[q].rows : b[q].cols;
      if (k != kb || m != c[q].rows || n != c[q].cols) throw std::invalid_argument("gemm: shape mismatch in item " + std::to_string(q));
    }
    tiles_gemm<double>(a, b, c, alpha, beta, trans_a != 0, trans_b != 0, l, *ctx->ctx);
    ranks_back(C, c);
    return FRK_OK;
  });
}

int frk_dgeqrt_batched(frk_ctx ctx, int batch, int n, double* const* A, int64_t lda, double* const* T, int64_t ldt) {
  return guard([&] {
    if (!ctx || batch <= 0 || n <= 0 || n > 512 || !A || !T || lda < n || ldt < n) throw std::invalid_argument("bad argument");
    std::vector<double*> a(A, A + batch), t(T, T + batch);
    tiles_geqrt<double>(a, lda, t, ldt, n, *ctx->ctx);
    return FRK_OK;
  });
}

int frk_dlarfb_batched(frk_ctx ctx, int batch, int n, const double* const* V, int64_t ldv, const double* const* T,
                       int64_t ldt, frk_dtile* C, int trans) {
  return guard([&] {
    if (!ctx || batch <= 0 || n <= 0 || n > 512 || !V || !T || ldv < n || ldt < n) throw std::invalid_argument("bad argument");
    auto c = raws(C, batch);
    for (int q = 0; q < batch; ++q)
      if (c[q].rows != n) throw std::invalid_argument("larfb: C must have n rows");
    std::vector<const double*> v(V, V + batch), t(T, T + batch);
    tiles_larfb<double>(v, ldv, t, ldt, c, n, trans != 0, *ctx->ctx);
    return FRK_OK;
  });
}

int frk_dtpqrt_batched(frk_ctx ctx, int batch, int n, double* const* A, int64_t lda, frk_dtile* B,
                       double* const* T, int64_t ldt) {
  return guard([&] {
    if (!ctx || batch <= 0 || n <= 0 || n > 512 || !A || !T || lda < n || ldt < n) throw std::invalid_argument("bad argument");
    auto b = raws(B, batch);
    for (int q = 0; q < batch; ++q)
      if (b[q].cols != n) throw std::invalid_argument("tpqrt: B must have n columns");
    std::vector<double*> a(A, A + batch), t(T, T + batch);
    tiles_tpqrt<double>(a, lda, b, t, ldt, n, *ctx->ctx);
    return FRK_OK;
  });
}

int frk_dtpmqrt_batched(frk_ctx ctx, int batch, int n, const frk_dtile* V, const double* const* T, int64_t ldt,
                        frk_dtile* A, frk_dtile* B, int trans, int lra) {
  return guard([&] {
    if (!ctx || batch <= 0 || n <= 0 || n > 512 || !T || ldt < n) throw std::invalid_argument("bad argument");
    const int l = check_lra(lra);
    auto v = raws(V, batch), a = raws(A, batch), b = raws(B, batch);
    for (int q = 0; q < batch; ++q)
      if (v[q].cols != n || a[q].rows != n || v[q].rows != b[q].rows || a[q].cols != b[q].cols)
        throw std::invalid_argument("tpmqrt: shape mismatch in item " + std::to_string(q));
    std::vector<const double*> t(T, T + batch);
    tiles_tpmqrt<double>(v, t, ldt, a, b, n, trans != 0, l, *ctx->ctx);
    ranks_back(A, a);
    ranks_back(B, b);
    return FRK_OK;
  });
}

int frk_dqr_batched(frk_ctx ctx, int batch, const int32_t* m, const int32_t* n, const double* const* A,
                    const int64_t* lda, double* const* Q, const int64_t* ldq, double* const* R, const int64_t* ldr) {
  return guard([&] {
    if (!ctx || batch <= 0 || !m || !n || !A || !lda || !Q || !ldq || !R || !ldr) throw std::invalid_argument("null argument");
    std::vector<SmallJob<double>> jobs(batch);
    for (int b = 0; b < batch; ++b) {
      const int k = std::min(m[b], n[b]);
      if (!A[b] || !Q[b] || !R[b] || lda[b] < n[b] || ldq[b] < k || ldr[b] < n[b])
        throw std::invalid_argument("qr: null pointer or leading dimension below the row length");
      jobs[b] = {A[b], lda[b], 1, Q[b], ldq[b], 1, R[b], ldr[b], 1, m[b], n[b]};
    }
    qr_small_batched<double>(jobs, *ctx->ctx);
    return FRK_OK;
  });
}

int frk_drq_batched(frk_ctx ctx, int batch, const int32_t* m, const int32_t* n, const double* const* A,
                    const int64_t* lda, double* const* R, const int64_t* ldr, double* const* Q, const int64_t* ldq) {
  return guard([&] {
    if (!ctx || batch <= 0 || !m || !n || !A || !lda || !Q || !ldq || !R || !ldr) throw std::invalid_argument("null argument");
    std::vector<SmallJob<double>> jobs(batch);
    for (int b = 0; b < batch; ++b) {
      if (m[b] < 1 || m[b] > 128 || n[b] > 512 || m[b] > n[b]) throw std::invalid_argument("rq: 1 <= m <= min(n, 128), n <= 512");
      if (!A[b] || !Q[b] || !R[b] || lda[b] < n[b] || ldr[b] < m[b] || ldq[b] < n[b])
        throw std::invalid_argument("rq: null pointer or leading dimension below the row length");
      jobs[b] = {A[b], lda[b], 1, R[b], ldr[b], 1, Q[b], ldq[b], 1, m[b], n[b]};
    }
    rq_small_batched<double>(jobs, *ctx->ctx);
    return FRK_OK;
  });
}

int frk_dsvd_batched(frk_ctx ctx, int batch, const int32_t* m, const int32_t* n, const double* const* A,
                     const int64_t* lda, double* const* U, const int64_t* ldu, double* const* S, const int64_t* lds,
                     double* const* V, const int64_t* ldv) {
  return guard([&] {
    if (!ctx || batch <= 0 || !m || !n || !A || !lda || !U || !ldu || !S || !lds || !V || !ldv)
      throw std::invalid_argument("null argument");
    std::vector<SvdJob<double>> jobs(batch);
    for (int b = 0; b < batch; ++b) {
      const int K = std::min(m[b], n[b]);
      if (m[b] < 1 || n[b] < 1 || !A[b] || !U[b] || !S[b] || !V[b] || lda[b] < n[b] || ldu[b] < K || lds[b] < K ||
          ldv[b] < n[b])
        throw std::invalid_argument("svd: bad shape, null pointer or leading dimension below the row length");
      jobs[b] = {A[b], lda[b], 1, U[b], ldu[b], 1, S[b], lds[b], V[b], ldv[b], 1, m[b], n[b]};
    }
    svd_small_batched<double>(jobs, *ctx->ctx);
    return FRK_OK;
  });
}

int frk_dtruncated_geqp3_batched(frk_ctx ctx, int batch, const int32_t* m, const int32_t* n, const double* const* A,
                                 const int64_t* lda, double eps, int cap, double* const* Q, const int64_t* ldq,
                                 double* const* R, const int64_t* ldr, int32_t* ranks_host) {
  return guard([&] {
    if (!ctx || batch <= 0 || !m || !n || !A || !lda || !Q || !ldq || !R || !ldr || !ranks_host || cap <= 0)
      throw std::invalid_argument("bad argument");
    if (eps < 0.0) throw std::invalid_argument("eps must be >= 0");
    std::vector<Geqp3Job<double>> jobs(batch);
    if (cap > 128) throw std::invalid_argument("truncated_geqp3: cap <= 128");
    for (int b = 0; b < batch; ++b) {
      if (!A[b] || !Q[b] || !R[b] || lda[b] < n[b] || ldq[b] < cap || ldr[b] < n[b])
        throw std::invalid_argument("truncated_geqp3: null pointer or leading dimension below the row length");
      jobs[b] = {A[b], lda[b], 1, m[b], n[b], eps, cap, Q[b], ldq[b], 1, R[b], ldr[b], ranks_host + b};
    }
    geqp3_small_batched<double>(jobs, *ctx->ctx);
    return FRK_OK;
  });
}

int frk_drsvd(frk_ctx ctx, int m, int n, const double* A, int64_t lda, int p, double* U, int64_t ldu, double* S,
              int64_t lds, double* V, int64_t ldv) {
  return guard([&] {
    if (!ctx || !A || !U || !S || !V) throw std::invalid_argument("null argument");
    if (lda < n || ldu < p || lds < p || ldv < n) throw std::invalid_argument("rsvd: leading dimension below the row length");
    rsvd_small<double>(A, lda, m, n, p, U, ldu, S, lds, V, ldv, *ctx->ctx);
    return FRK_OK;
  });
}

int frk_did(frk_ctx ctx, int m, int n, const double* A, int64_t lda, int k, double* V, int64_t ldv,
            int32_t* cols_host) {
  return guard([&] {
    if (!ctx || !A || !V || !cols_host) throw std::invalid_argument("null argument");
    if (lda < n || ldv < n) throw std::invalid_argument("id: leading dimension below the row length");
    id_small<double>(A, lda, 1, m, n, k, V, ldv, cols_host, *ctx->ctx);
    return FRK_OK;
  });
}

int frk_dcompress(frk_ctx ctx, int m, int n, double* A, int64_t lda, double eps, int rank, int cap, double* U,
                  double* S, double* V, int32_t* rank_out) {
  return guard([&] {
    if (!ctx || !A || !U || !S || !V || !rank_out) throw std::invalid_argument("null argument");
    if (m <= 0 || n <= 0 || m > 512 || n > 512) throw std::invalid_argument("tile dims must be in [1, 512]");
    if (eps == 0.0 && (rank <= 0 || rank > cap)) throw std::invalid_argument("bad rank");
    Context& c = *ctx->ctx;
    c.sync();
    DevBuf dr(sizeof(int32_t));
    CompressDesc<double> d{};
    d.a = A;
    d.m = m;
    d.n = n;
    d.lda = lda;
    d.eps = eps;
    d.rank = rank;
    d.u = U;
    d.s = S;
    d.v = V;
    d.u_rs = cap;
    d.u_cs = 1;
    d.v_rs = n;
    d.v_cs = 1;
    d.lds = cap;
    d.d_rank = dr.as<int32_t>();
    d.cap = cap;
    compress_batched<double>({d}, c);
    FRK_HIP_CHECK(hipMemcpy(rank_out, dr.as<int32_t>(), sizeof(int32_t), hipMemcpyDeviceToHost));
    if (*rank_out < 0) throw std::runtime_error("low-rank capacity exceeded in compression");
    return FRK_OK;
  });
}

}  // extern "C"
