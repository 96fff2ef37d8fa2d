// Host orchestration of the batched BLR operations (see engine.hpp).
#include "engine.hpp"
#include "plan.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <random>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>

namespace frk {

// ============================================================================
// Host waits with a deadline
// ============================================================================
double wait_timeout_s() {
  static const double t = [] {
    const char* e = std::getenv("FRK_WAIT_TIMEOUT_S");
    const double v = e ? std::atof(e) : 0.0;
    return v > 0.0 ? v : 60.0;
  }();
  return t;
}

void host_wait_event(hipEvent_t e, const char* what) {
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  const double limit = wait_timeout_s();
  for (int64_t spin = 0;; ++spin) {
    const hipError_t q = hipEventQuery(e);
    if (q == hipSuccess) return;
    if (q != hipErrorNotReady) FRK_HIP_CHECK(q);
    // poll hard for the first few thousand rounds (the per-step waits are short
    // and on the host's critical path), then check the clock and yield
    if (spin < 4096) continue;
    const double s = std::chrono::duration<double>(clk::now() - t0).count();
    if (s > limit)
      throw std::runtime_error(std::string("device work did not complete within ") + std::to_string(limit) +
                               " s (" + what + "; FRK_WAIT_TIMEOUT_S sets the limit)");
    std::this_thread::yield();
  }
}

void host_wait_stream(hipStream_t s, const char* what) {
  hipEvent_t e;
  FRK_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  try {
    FRK_HIP_CHECK(hipEventRecord(e, s));
    host_wait_event(e, what);
  } catch (...) {
    (void)hipEventDestroy(e);
    throw;
  }
  FRK_HIP_CHECK(hipEventDestroy(e));
  // the stream's own error state (a failed launch reports here)
  FRK_HIP_CHECK(hipStreamSynchronize(s));
}

// ============================================================================
// Buffers, staging, context
// ============================================================================
DevBuf::~DevBuf() {
  if (p_) (void)hipFree(p_);
}

void DevBuf::ensure(size_t bytes) {
  if (bytes <= bytes_) return;
  if (p_) FRK_HIP_CHECK(hipFree(p_));
  p_ = nullptr;
  bytes_ = 0;
  FRK_HIP_CHECK(hipMalloc(&p_, bytes));
  bytes_ = bytes;
  // FRK_POISON=1 fills every new allocation with NaN bit patterns, so a read of
  // memory no kernel has written shows up in the results (a debugging aid).
  static const bool poison = [] {
    const char* e = std::getenv("FRK_POISON");
    return e && e[0] == '1';
  }();
  if (poison) {
    FRK_HIP_CHECK(hipMemset(p_, 0xFF, bytes));
    FRK_HIP_CHECK(hipDeviceSynchronize());
  }
}

Staging::~Staging() {
  if (h_) (void)hipHostFree(h_);
  if (d_) (void)hipFree(d_);
}

const void* Staging::push_bytes(const void* src, size_t bytes, hipStream_t s) {
  const size_t aligned = (bytes + 255) & ~size_t(255);
  if (off_ + aligned > cap_) {
    if (off_ != 0) {
      // the ring is in use by queued launches: drain before growing
      host_wait_stream(s, __func__);
      off_ = 0;
    }
    if (aligned > cap_) {
      if (h_) FRK_HIP_CHECK(hipHostFree(h_));
      if (d_) FRK_HIP_CHECK(hipFree(d_));
      size_t c = std::max<size_t>(aligned, size_t(16) << 20);
      FRK_HIP_CHECK(hipHostMalloc((void**)&h_, c));
      FRK_HIP_CHECK(hipHostGetDevicePointer((void**)&hd_, h_, 0));
      FRK_HIP_CHECK(hipMalloc((void**)&d_, c));
      cap_ = c;
    }
  }
  char* hp = h_ + off_;
  char* dp = d_ + off_;
  if (bytes) {
    // the slot is `aligned` bytes long on both sides: copy it whole in 16-byte words
    std::memcpy(hp, src, bytes);
    launch_stage_copy(dp, hd_ + off_, aligned / 16, s);
  }
  off_ += aligned;
  return dp;
}

Context::Context(hipStream_t s) : s_(s) {
  info_.ensure(sizeof(int32_t) * 4096);
  counters_.ensure(sizeof(int32_t) * kCounterWords);
  FRK_HIP_CHECK(hipMemsetAsync(counters_.as<void>(), 0, sizeof(int32_t) * kCounterWords, s_));
}

int32_t* Context::tile_flags() {
  if (tile_flags_.bytes() == 0) {
    tile_flags_.ensure(sizeof(int32_t) * kTileFlags);
    FRK_HIP_CHECK(hipMemsetAsync(tile_flags_.as<void>(), 0, sizeof(int32_t) * kTileFlags, s_));
  }
  return tile_flags_.as<int32_t>();
}

void Context::check_tile_lu(int32_t since) {
  if (tile_flags_.bytes() == 0 || tile_epoch_ <= since) return;
  int32_t e = 0;
  FRK_HIP_CHECK(hipMemcpyAsync(&e, tile_flags_.as<int32_t>() + kTileFlags - 1, sizeof(e), hipMemcpyDeviceToHost, s_));
  host_wait_stream(s_, __func__);
  if (e > since) throw std::runtime_error("tile LU: a block wait timed out (workgroups not co-resident?)");
}

hipStream_t Context::aux_stream(int k) {
  if (k < 0 || k > 1) throw std::invalid_argument("aux_stream: index 0 or 1");
  if (!aux_[k]) FRK_HIP_CHECK(hipStreamCreateWithFlags(&aux_[k], hipStreamNonBlocking));
  return aux_[k];
}

Context::~Context() {
  (void)hipStreamSynchronize(s_);
  for (hipStream_t a : aux_)
    if (a) {
      (void)hipStreamSynchronize(a);
      (void)hipStreamDestroy(a);
    }
  for (auto& pr : ev_pending_) {
    ev_pool_.push_back(pr.first);
    ev_pool_.push_back(pr.second);
  }
  for (hipEvent_t e : ev_pool_) (void)hipEventDestroy(e);
}

void Context::reserve(size_t bytes) {
  if (bytes > ws_.bytes()) {
    host_wait_stream(s_, __func__);
    ws_.ensure(bytes + (bytes >> 2));
  }
  ws_off_ = 0;
}

void* Context::take_bytes(size_t bytes) {
  const size_t aligned = (bytes + 255) & ~size_t(255);
  if (ws_off_ + aligned > ws_.bytes())
    throw std::runtime_error("workspace exhausted (internal sizing error)");
  void* p = ws_.as<char>() + ws_off_;
  ws_off_ += aligned;
  return p;
}

void Context::sync() {
  host_wait_stream(s_, "context stream");
  stage_.reset();
  ws_off_ = 0;
  resolve_events();
}

void Context::account(const char* cls, double flops, double bytes, double peak, int64_t launches) {
  OpStat& st = owner_->stats_[cls];
  st.launches += launches;
  st.flops += flops;
  st.bytes += bytes;
  st.roof_s += std::max(flops / (peak > 0 ? peak : owner_->peak_), bytes / kHbmBytesPerS);
}

void Context::host_time(const char* cls, double ms) {
  OpStat& st = owner_->stats_[cls];
  st.ms += ms;
  st.timed += 1;
}

// Launch-timing events are read only after a full synchronize, so they carry
// no system-scope fence (which would write back the caches at every record).
void Context::event_begin() {
  Context& o = *owner_;
  if (o.ev_pool_.empty()) {
    hipEvent_t e;
    FRK_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    o.ev_pool_.push_back(e);
  }
  ev_open_ = o.ev_pool_.back();
  o.ev_pool_.pop_back();
  FRK_HIP_CHECK(hipEventRecord(ev_open_, s_));
}

void Context::event_end() {
  Context& o = *owner_;
  if (o.ev_pool_.empty()) {
    hipEvent_t e;
    FRK_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    o.ev_pool_.push_back(e);
  }
  hipEvent_t e1 = o.ev_pool_.back();
  o.ev_pool_.pop_back();
  FRK_HIP_CHECK(hipEventRecord(e1, s_));
  o.ev_pending_.emplace_back(ev_open_, e1);
  ev_open_ = nullptr;
}

void Context::resolve_events() {
  if (ev_pending_.empty()) return;
  OpStat& st = stats_[timed_cls_];
  for (auto& pr : ev_pending_) {
    FRK_HIP_CHECK(hipEventSynchronize(pr.second));
    float ms = 0.f;
    FRK_HIP_CHECK(hipEventElapsedTime(&ms, pr.first, pr.second));
    st.ms += ms;
    st.timed += 1;
    ev_pool_.push_back(pr.first);
    ev_pool_.push_back(pr.second);
  }
  ev_pending_.clear();
}

const std::map<std::string, OpStat>& Context::op_stats() {
  resolve_events();
  return stats_;
}

void Context::reset_op_stats() {
  resolve_events();
  stats_.clear();
}

template <class X>
const X* Context::identity(int n) {
  if (n > ident_n_) {
    const int nn = std::max(n, 256);
    host_wait_stream(s_, __func__);
    ident64_.ensure(sizeof(double) * nn * nn);
    ident32_.ensure(sizeof(float) * nn * nn);
    launch_fill_identity<double>(ident64_.as<double>(), nn, nn, nn, s_);
    launch_fill_identity<float>(ident32_.as<float>(), nn, nn, nn, s_);
    ident_n_ = nn;
  }
  if constexpr (sizeof(X) == 8)
    return ident64_.as<X>();
  else
    return ident32_.as<X>();
}

template <class X>
const X* Context::omega(int n, int p) {
  if (n != omega_n_ || p != omega_p_) {
    // src/functions.cpp:52-66: random_uniform re-seeds std::mt19937 to 0 on
    // every call and fills row-major with uniform_real_distribution(0, 1).
    std::mt19937 gen(0);
    gen.seed(0);
    std::uniform_real_distribution<double> dist(0.0, 1.0);
    std::vector<double> h((size_t)n * p);
    for (auto& x : h) x = dist(gen);
    std::vector<float> hf(h.begin(), h.end());
    host_wait_stream(s_, __func__);
    omega64_.ensure(sizeof(double) * h.size());
    omega32_.ensure(sizeof(float) * h.size());
    FRK_HIP_CHECK(hipMemcpy(omega64_.as<double>(), h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    FRK_HIP_CHECK(hipMemcpy(omega32_.as<float>(), hf.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
    omega_n_ = n;
    omega_p_ = p;
  }
  if constexpr (sizeof(X) == 8)
    return omega64_.as<X>();
  else
    return omega32_.as<X>();
}

template const double* Context::identity<double>(int);
template const float* Context::identity<float>(int);
template const double* Context::omega<double>(int, int);
template const float* Context::omega<float>(int, int);

// ============================================================================
// BlrMatrix
// ============================================================================
template <typename T>
BlrMatrix<T>::BlrMatrix(std::vector<int64_t> ro, std::vector<int64_t> co, std::vector<int32_t> kinds, int kcap_,
                        double eps_, int fixed_rank_, Context& ctx)
    : roff(std::move(ro)), coff(std::move(co)), kind(std::move(kinds)), kcap(kcap_), eps(eps_),
      fixed_rank(fixed_rank_) {
  const int nr = nb_r(), nc = nb_c();
  if ((int)kind.size() != nr * nc) throw std::invalid_argument("tile kind table has the wrong size");
  rank.assign(nr * nc, 0);
  D.assign(nr * nc, nullptr);
  U.assign(nr * nc, nullptr);
  S.assign(nr * nc, nullptr);
  V.assign(nr * nc, nullptr);
  size_t total = 0;
  auto al = [](size_t x) { return (x + 63) & ~size_t(63); };
  for (int i = 0; i < nr; ++i)
    for (int j = 0; j < nc; ++j) {
      const int t = idx(i, j);
      if (kind[t] == kDense)
        total += al((size_t)rows(i) * cols(j));
      else if (kind[t] == kLowRank)
        total += al((size_t)rows(i) * kcap) + al((size_t)kcap * kcap) + al((size_t)kcap * cols(j));
    }
  arena_bytes_ = std::max<size_t>(total, 1) * sizeof(T);
  arena_.ensure(arena_bytes_);
  T* p = arena_.as<T>();
  for (int i = 0; i < nr; ++i)
    for (int j = 0; j < nc; ++j) {
      const int t = idx(i, j);
      if (kind[t] == kDense) {
        D[t] = p;
        p += al((size_t)rows(i) * cols(j));
      } else if (kind[t] == kLowRank) {
        U[t] = p;
        p += al((size_t)rows(i) * kcap);
        S[t] = p;
        p += al((size_t)kcap * kcap);
        V[t] = p;
        p += al((size_t)kcap * cols(j));
      }
    }
  FRK_HIP_CHECK(hipMemsetAsync(arena_.as<T>(), 0, std::max<size_t>(total, 1) * sizeof(T), ctx.stream()));
  Linv.assign(nr * nc, nullptr);
  Uinv.assign(nr * nc, nullptr);
  size_t dtot = 0;
  for (int i = 0; i < std::min(nr, nc); ++i)
    if (kind[idx(i, i)] == kDense && rows(i) == cols(i)) dtot += 2 * al((size_t)rows(i) * rows(i));
  if (dtot) {
    dinv_.ensure(dtot * sizeof(T));
    FRK_HIP_CHECK(hipMemsetAsync(dinv_.as<T>(), 0, dtot * sizeof(T), ctx.stream()));
    T* q = dinv_.as<T>();
    for (int i = 0; i < std::min(nr, nc); ++i)
      if (kind[idx(i, i)] == kDense && rows(i) == cols(i)) {
        const size_t nn = al((size_t)rows(i) * rows(i));
        Linv[idx(i, i)] = q;
        q += nn;
        Uinv[idx(i, i)] = q;
        q += nn;
      }
  }
  d_ranks_.ensure(sizeof(int32_t) * nr * nc);
  FRK_HIP_CHECK(hipMemsetAsync(d_ranks_.as<int32_t>(), 0, sizeof(int32_t) * nr * nc, ctx.stream()));
}

template <typename T>
void BlrMatrix<T>::pull_ranks(hipStream_t s) {
  FRK_HIP_CHECK(hipMemcpyAsync(rank.data(), d_ranks_.as<int32_t>(), sizeof(int32_t) * rank.size(),
                               hipMemcpyDeviceToHost, s));
  host_wait_stream(s, __func__);
  for (size_t t = 0; t < rank.size(); ++t)
    if (kind[t] == kLowRank && rank[t] < 0)
      throw std::runtime_error("low-rank capacity exceeded: rank " + std::to_string(-rank[t]) +
                               " > kcap " + std::to_string(kcap));
}

template <typename T>
std::unique_ptr<BlrMatrix<T>> BlrMatrix<T>::clone_structure(Context& ctx) const {
  auto c = std::make_unique<BlrMatrix<T>>(roff, coff, kind, kcap, eps, fixed_rank, ctx);
  c->lra = lra;
  c->cap_trunc = cap_trunc;
  return c;
}

template <typename T>
void BlrMatrix<T>::copy_from(const BlrMatrix<T>& src, hipStream_t s) {
  if (src.arena_bytes_ != arena_bytes_ || src.kind != kind || src.roff != roff || src.coff != coff)
    throw std::invalid_argument("copy_from: tile structures differ");
  FRK_HIP_CHECK(hipMemcpyAsync(arena_.as<char>(), src.arena_.as<char>(), arena_bytes_, hipMemcpyDeviceToDevice, s));
  FRK_HIP_CHECK(hipMemcpyAsync(d_ranks_.as<int32_t>(), src.d_ranks_.as<int32_t>(), sizeof(int32_t) * rank.size(),
                               hipMemcpyDeviceToDevice, s));
  if (dinv_.bytes() && src.dinv_.bytes() == dinv_.bytes())
    FRK_HIP_CHECK(hipMemcpyAsync(dinv_.as<char>(), src.dinv_.as<char>(), dinv_.bytes(), hipMemcpyDeviceToDevice, s));
  rank = src.rank;
  lra = src.lra;
}

template <typename T>
void BlrMatrix<T>::set_ranks_from_device(const int32_t* src, hipStream_t s) {
  FRK_HIP_CHECK(hipMemcpyAsync(d_ranks_.as<int32_t>(), src, sizeof(int32_t) * rank.size(), hipMemcpyDeviceToDevice, s));
}

template <typename T>
void BlrMatrix<T>::push_ranks(hipStream_t s) {
  FRK_HIP_CHECK(hipMemcpyAsync(d_ranks_.as<int32_t>(), rank.data(), sizeof(int32_t) * rank.size(),
                               hipMemcpyHostToDevice, s));
  host_wait_stream(s, __func__);
}

// ============================================================================
// Item helpers and plans
// ============================================================================
void Plan::run() {
  if (launches_.empty()) return;
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  const char* d = static_cast<const char*>(ctx_.staging().push_bytes(bytes_.data(), bytes_.size(), ctx_.stream()));
  const auto t1 = clk::now();
  ctx_.host_time("host_stage", std::chrono::duration<double, std::milli>(t1 - t0).count());
  for (size_t i = 0; i < launches_.size(); ++i) {
    const Meta& m = meta_[i];
    const bool timed = ctx_.is_timed(m.cls);
    if (timed) ctx_.event_begin();
    launches_[i](d);
    if (timed) ctx_.event_end();
    if (m.cls) ctx_.account(m.cls, m.flops, m.bytes, m.peak);
  }
  launches_.clear();
  meta_.clear();
  bytes_.clear();
}

// ============================================================================
// Dense tile LU (blocked, no pivoting)
// ============================================================================
template <typename T>
void plan_getrf(Plan& pl, const std::vector<T*>& tiles, const std::vector<T*>& linv, const std::vector<T*>& uinv,
                int n, int64_t ld, int32_t* info_dev) {
  if (tiles.empty() || n <= 0) return;
  const hipStream_t s = pl.ctx().stream();
  const int nt = (int)tiles.size();
  int pi = 0;
  for (int p = 0; p < n; p += 64, ++pi) {
    const int nbk = std::min(64, n - p);
    const int rest = n - p - nbk;
    std::vector<LuItem> small(nt);
    for (int t = 0; t < nt; ++t) {
      LuItem x{};
      x.a = tiles[t] + p * ld + p;
      x.rs = ld;
      x.cs = 1;
      x.n = nbk;
      x.linv = linv[t] + (size_t)p * n + p;
      x.uinv = uinv[t] + (size_t)p * n + p;
      x.ld_inv = n;
      small[t] = x;
    }
    const size_t off = pl.add(small);
    int32_t* inf = info_dev ? info_dev + pi * nt : nullptr;
    const double nb3 = (double)nbk * nbk * nbk;
    pl.launch([=](const char* d) { launch_getrf_small<T>(reinterpret_cast<const LuItem*>(d + off), nt, inf, s); },
              "getrf_small", nt * (4.0 / 3.0) * nb3, nt * 4.0 * sizeof(T) * nbk * nbk);
    if (rest <= 0) continue;
    // A12 <- inv(L11) A12 ; A21 <- A21 inv(U11), in place: the 64 x 64 GEMM tiling gives A12 a single
    // output row tile and A21 a single output column tile, so every workgroup reads only the region it
    // writes, and it reads all of it before writing.
    std::vector<GemmItem> pan;
    pan.reserve(2 * nt);
    for (int t = 0; t < nt; ++t) {
      T* a = tiles[t];
      pan.push_back(gemm_item(linv[t] + (size_t)p * n + p, n, 1, a + p * ld + p + nbk, ld, 1, a + p * ld + p + nbk,
                              ld, 1, nbk, rest, nbk, 1.0, 0.0));
      pan.push_back(gemm_item(a + (p + nbk) * ld + p, ld, 1, uinv[t] + (size_t)p * n + p, n, 1,
                              a + (p + nbk) * ld + p, ld, 1, rest, nbk, nbk, 1.0, 0.0));
    }
    if (nbk != 64) throw std::runtime_error("getrf: only the last panel may be narrower than 64");
    plan_gemms_onelaunch<T>(pl, pan, "getrf_gemm");
    std::vector<GemmItem> gm(nt);
    for (int t = 0; t < nt; ++t) {
      T* a = tiles[t];
      gm[t] = gemm_item(a + (p + nbk) * ld + p, ld, 1, a + p * ld + p + nbk, ld, 1, a + (p + nbk) * ld + p + nbk,
                        ld, 1, rest, rest, nbk, -1.0, 1.0);
    }
    plan_gemms<T>(pl, gm, "getrf_gemm");
  }
  // full inverses of the unit-lower and upper factors by recursive doubling of the 64 x 64 block
  // inverses: [[A, 0], [C, D]]^-1 has off-diagonal block -D^-1 (C A^-1); the upper case mirrors it.
  Context& ctx = pl.ctx();
  for (int sz = 64; sz < n; sz *= 2) {
    std::vector<GemmItem> w1, w2;
    for (int t = 0; t < nt; ++t) {
      T* a = tiles[t];
      for (int p = 0; p + sz < n; p += 2 * sz) {
        const int d0 = p + sz, sd = std::min(sz, n - d0);
        T* wl = ctx.take<T>((size_t)sz * sd + 1);
        T* wu = ctx.take<T>((size_t)sz * sd + 1);
        w1.push_back(gemm_item(a + (int64_t)d0 * ld + p, ld, 1, linv[t] + (size_t)p * n + p, n, 1, wl, sz, 1, sd, sz,
                               sz, 1.0, 0.0));
        w1.push_back(gemm_item(a + (int64_t)p * ld + d0, ld, 1, uinv[t] + (size_t)d0 * n + d0, n, 1, wu, sd, 1, sz,
                               sd, sd, 1.0, 0.0));
        w2.push_back(gemm_item(linv[t] + (size_t)d0 * n + d0, n, 1, wl, sz, 1, linv[t] + (size_t)d0 * n + p, n, 1,
                               sd, sz, sd, -1.0, 0.0));
        w2.push_back(gemm_item(uinv[t] + (size_t)p * n + p, n, 1, wu, sd, 1, uinv[t] + (size_t)p * n + d0, n, 1,
                               sz, sd, sz, -1.0, 0.0));
      }
    }
    plan_gemms<T>(pl, w1, "getrf_gemm");
    plan_gemms<T>(pl, w2, "getrf_gemm");
  }
}

// The dataflow kernel (one launch) unless FRK_TILE_LU=0 selects the blocked
// sequence of getrf_small / panel GEMM / inverse-doubling launches.
static bool use_tile_lu() {
  static const bool on = [] {
    const char* e = std::getenv("FRK_TILE_LU");
    return !(e && e[0] == '0');
  }();
  return on;
}

// One-launch FP64 LU of an n x n tile (n <= 512) with full inverses of both
// factors (row stride n, the opposite triangles zeroed).
// f32 (optional): the FP32 tile read as input and overwritten with L\U, and the
// FP32 inverses written beside the FP64 ones (a, linv, uinv are then scratch).
struct TileLu32 {
  float* a = nullptr;
  float* linv = nullptr;
  float* uinv = nullptr;
};

static void plan_tile_lu(Plan& pl, double* a, double* linv, double* uinv, int n, int64_t ld, int32_t* info,
                         TileLu32 f32 = {}, const LrUpdItem* pre = nullptr, bool pre_f32 = false) {
  Context& c = pl.ctx();
  TileLuItem it{};
  if (pre != nullptr) {
    it.pre_u = pre->u;
    it.pre_u_cs = pre->u_cs;
    it.pre_core = pre->core;
    it.pre_ldk = pre->ldk;
    it.pre_v = pre->v;
    it.pre_v_rs = pre->v_rs;
    it.pre_r1 = pre->r1;
    it.pre_r2 = pre->r2;
    it.pre_kcap = pre->kcap;
    it.pre_f32 = pre_f32 ? 1 : 0;
  }
  it.a = a;
  it.linv = linv;
  it.uinv = uinv;
  it.a32_in = f32.a;
  it.a32_out = f32.a;
  it.linv32 = f32.linv;
  it.uinv32 = f32.uinv;
  it.flags = c.tile_flags();
  it.err = it.flags + Context::kTileFlags - 1;
  it.info = info;
  it.ld = ld;
  it.ldi = n;
  it.n = n;
  it.epoch = c.next_tile_epoch();
  it.spin_limit = int64_t(1) << 22;
  const hipStream_t s = c.stream();
  // algorithmic counts of getrf (LAWN 41): 2/3 n^3 flops, read A and write L\U
  // at the storage precision.  The two triangular inverses the kernel also forms
  // (n^3/3 flops each, two more n x n results) are logged apart, under
  // "getrf_tile_inverses", with no launch of their own.
  const double n3 = (double)n * n * n;
  const double sb = f32.a ? sizeof(float) : sizeof(double);
  pl.launch([=](const char*) { launch_tile_lu(it, s); }, "getrf_tile", (2.0 / 3.0) * n3, 2.0 * sb * n * n, kPeakF64);
  c.account("getrf_tile_inverses", (2.0 / 3.0) * n3, 2.0 * sb * n * n, kPeakF64);
}

template <typename T>
static size_t getrf_workspace(int n, int nt) {
  size_t w = 4096;
  for (int sz = 64; sz < n; sz *= 2) w += (size_t)nt * ((n + 2 * sz - 1) / (2 * sz)) * 2 * sizeof(T) * ((size_t)sz * sz + 64);
  return w;
}

template <typename T>
void getrf_tiles(const std::vector<T*>& tiles, int n, int64_t ld, int32_t* info_dev, Context& ctx) {
  const size_t nn = (size_t)n * n;
  // the inverses of both factors (the kernels exchange blocks through them);
  // kept in the context between calls
  T* inv = static_cast<T*>(ctx.persistent(5, sizeof(T) * tiles.size() * nn * 2));
  FRK_HIP_CHECK(hipMemsetAsync(inv, 0, sizeof(T) * tiles.size() * nn * 2, ctx.stream()));
  std::vector<T*> li(tiles.size()), ui(tiles.size());
  for (size_t t = 0; t < tiles.size(); ++t) {
    li[t] = inv + t * nn * 2;
    ui[t] = li[t] + nn;
  }
  ctx.sync();
  ctx.reserve(getrf_workspace<T>(n, (int)tiles.size()));
  Plan pl(ctx);
  int32_t e0 = ctx.tile_epoch();
  if constexpr (sizeof(T) == 8) {
    if (tiles.size() == 1 && use_tile_lu())
      plan_tile_lu(pl, tiles[0], li[0], ui[0], n, ld, info_dev);
    else
      plan_getrf<T>(pl, tiles, li, ui, n, ld, info_dev);
  } else {
    plan_getrf<T>(pl, tiles, li, ui, n, ld, info_dev);
  }
  pl.run();
  ctx.check_tile_lu(e0);
  host_wait_stream(ctx.stream(), __func__);
}

// Generic trsm without precomputed inverses: invert the diagonal blocks first.
template <typename T>
void trsm_generic(const T* A, int64_t a_rs, int64_t a_cs, bool lower, bool unit, T* B, int64_t b_rs, int64_t b_cs,
                  int n, int nrhs, Context& ctx) {
  const hipStream_t s = ctx.stream();
  const int nbk = (n + 63) / 64;
  DevBuf inv(sizeof(T) * nbk * 4096);
  std::vector<TriInvItem> ti(nbk);
  for (int k = 0; k < nbk; ++k) {
    TriInvItem x{};
    x.a = A + (int64_t)(64 * k) * a_rs + (int64_t)(64 * k) * a_cs;
    x.out = inv.as<T>() + (size_t)k * 4096;
    x.rs = a_rs;
    x.cs = a_cs;
    x.n = std::min(64, n - 64 * k);
    x.lower = lower ? 1 : 0;
    x.unit = unit ? 1 : 0;
    ti[k] = x;
  }
  Plan pl(ctx);
  const size_t off = pl.add(ti);
  pl.launch([=](const char* d) { launch_tri_inv<T>(reinterpret_cast<const TriInvItem*>(d + off), nbk, s); });
  plan_trsms<T>(pl, {trsm_item(A, a_rs, a_cs, inv.as<T>(), 64, 1, B, b_rs, b_cs, n, nrhs, lower)});
  pl.run();
  host_wait_stream(s, __func__);
}

// ============================================================================
// Rounded low-rank addition (batched)
// ============================================================================
template <typename T>
size_t lr_add_workspace(const std::vector<LrAddDesc<T>>& d) {
  int kmax = 0, mmax = 0, nmax = 0;
  for (const auto& x : d) {
    kmax = std::max(kmax, x.kc + x.kx2);
    mmax = std::max(mmax, x.m);
    nmax = std::max(nmax, x.n);
  }
  const size_t K = kmax;
  const size_t vw = (sizeof(T) == 8 && K > 64) ? K * K : 0;  // Jacobi right-vector scratch
  bool small_core = false;                                   // a core formed before its SVD
  for (const auto& x : d) small_core = small_core || x.kc + x.kx2 > std::min(x.m, x.n);
  const size_t mw = small_core ? K * K : 0;
  return d.size() * sizeof(T) * ((size_t)(mmax + nmax) * K + 2 * K + 4 * K * K + vw + mw) + 10 * 256 + 4096;
}

template <typename T>
void plan_lr_add(Plan& pl, const std::vector<LrAddDesc<T>>& d_in, int lra) {
  if (d_in.empty()) return;
  Context& ctx = pl.ctx();
  const hipStream_t s = ctx.stream();
  const int nd = (int)d_in.size();
  // The SVD kernel is chosen by the batch's largest core.  In a large batch
  // (throughput-bound: many waves of workgroups) the cores above 64 columns are
  // ordered last and launched apart, so that the smaller ones do not run in the
  // 128-column kernel (one workgroup per CU).  Small batches keep one launch:
  // there the launches' latencies would add up.  Every kernel of the batch
  // indexes its workspace by the item, so the order is the same for all.
  auto core_dim = [lra](const LrAddDesc<T>& x) {
    const int Kt = x.kc + x.kx2;
    const bool fast = (lra == kLraFast) && (x.eps == 0.0);
    return fast ? Kt : std::max(std::min(x.m, Kt), std::min(x.n, Kt));
  };
  constexpr int kSplitBatch = 1024;
  const bool split = nd >= kSplitBatch;
  auto svd_class = [split](int k) { return (split && k > 64) ? 1 : 0; };
  std::vector<LrAddDesc<T>> d(d_in);
  if (split)
    std::stable_sort(d.begin(), d.end(), [&](const LrAddDesc<T>& a, const LrAddDesc<T>& c) {
      return svd_class(core_dim(a)) < svd_class(core_dim(c));
    });
  int kmax = 0, mmax = 0, nmax = 0, qk = 0, capmax = 0;
  double qf = 0, qb = 0, sf = 0, sb = 0, af = 0, ab = 0;
  for (const auto& x : d) {
    const int Kt = x.kc + x.kx2;
    kmax = std::max(kmax, Kt);
    mmax = std::max(mmax, x.m);
    nmax = std::max(nmax, x.n);
    capmax = std::max(capmax, std::min(x.cap, Kt));
  }
  // op log: QR (geqrf + orgqr 4mK^2 - 4/3 K^3, panel formation, projection), SVD 24 K^3, reflector
  // application 4 m K ncol with ncol <= min(K, cap) (the kept rank is known on the device only).
  for (const auto& x : d) {
    const bool fast = (lra == kLraFast) && (x.eps == 0.0);
    const int Kt = x.kc + x.kx2;
    const double kq = fast ? x.kx2 : Kt, kp = fast ? x.kc : 0;
    qk = std::max(qk, (int)kq);
    for (const double m : {(double)x.m, (double)x.n}) {
      qf += 4.0 * m * kq * kq - 4.0 / 3.0 * kq * kq * kq + 4.0 * m * kp * kq;
      qb += sizeof(T) * (m * (kq + kp + kq) + kq * kq);
    }
    if (x.xs) qf += 2.0 * x.m * x.kx1 * x.kx2;
    if (!fast) qf += 2.0 * x.m * x.kc * x.kc;
    sf += 24.0 * (double)Kt * Kt * Kt;
    sb += sizeof(T) * 5.0 * Kt * Kt;
    const double nc = std::min(Kt, capmax);
    for (const double m : {(double)x.m, (double)x.n}) {
      const double nref = std::min(m, fast ? (double)x.kx2 : (double)Kt);
      af += 4.0 * m * nref * nc;
      ab += sizeof(T) * m * (nref + nc);
    }
  }
  const int K = kmax;
  // cores smaller than K (K above a tile dimension, not in fast mode): formed first,
  // the SVD then runs at the largest core of the batch
  int kmax_svd = 0;
  bool core_prod = false;
  int cls_end[2] = {0, 0}, cls_kmax[2] = {0, 0};
  for (int t = 0; t < nd; ++t) {
    const auto& x = d[t];
    const int Kt = x.kc + x.kx2;
    const int Ke = core_dim(x);
    core_prod = core_prod || Ke < Kt;
    kmax_svd = std::max(kmax_svd, Ke);
    const int c = svd_class(Ke);
    cls_end[c] = t + 1;
    cls_kmax[c] = std::max(cls_kmax[c], Ke);
  }
  LrBatch<T> b{};
  b.Yu = ctx.take<T>((size_t)nd * mmax * K);
  b.Yv = ctx.take<T>((size_t)nd * nmax * K);
  b.tu = ctx.take<T>((size_t)nd * K);
  b.tv = ctx.take<T>((size_t)nd * K);
  b.Ru = ctx.take<T>((size_t)nd * K * K);
  b.Rv = ctx.take<T>((size_t)nd * K * K);
  b.Um = ctx.take<T>((size_t)nd * K * K);
  b.Vm = ctx.take<T>((size_t)nd * K * K);
  // FP64 cores above 64 keep their right vectors in global scratch
  b.Vw = (sizeof(T) == 8 && kmax_svd > 64) ? ctx.take<T>((size_t)nd * K * K) : nullptr;
  b.Mw = core_prod ? ctx.take<T>((size_t)nd * K * K) : nullptr;
  b.Kmax_svd = kmax_svd;
  b.stats = ctx.counters();
  b.mmax = mmax;
  b.nmax = nmax;
  b.K = K;
  b.lra = lra;
  // only the additions' descriptors are uploaded: the QR, SVD and basis-update
  // kernels derive their items from them
  const size_t od = pl.add(d);
  pl.launch([=](const char* dd) { launch_lr_add_qr<T>(reinterpret_cast<const LrAddDesc<T>*>(dd + od), nd, b, qk, s); },
            "lr_qr", qf, qb);
  if (core_prod)
    pl.launch([=](const char* dd) { launch_lr_add_core<T>(reinterpret_cast<const LrAddDesc<T>*>(dd + od), nd, b, s); });
  int t0 = 0;
  bool first = true;
  for (int c = 0; c < 2; ++c) {
    if (cls_end[c] <= t0) continue;
    const int n = cls_end[c] - t0, kc_max = cls_kmax[c];
    // the batch's SVD flops and bytes are logged with its first launch
    pl.launch([=](const char* dd) {
      launch_lr_add_svd<T>(reinterpret_cast<const LrAddDesc<T>*>(dd + od), t0, n, kc_max, b, s);
    }, "lr_svd", first ? sf : 0.0, first ? sb : 0.0);
    first = false;
    t0 = cls_end[c];
  }
  pl.launch(
      [=](const char* dd) { launch_lr_add_apply<T>(reinterpret_cast<const LrAddDesc<T>*>(dd + od), nd, b, capmax, s); },
      "lr_apply_q", af, ab);
}

template <typename T>
void lr_add_batched(const std::vector<LrAddDesc<T>>& d, int lra, Context& ctx) {
  if (d.empty()) return;
  if (lra == kLraNaive) {
    lr_add_naive<T>(d, ctx);
    return;
  }
  Plan pl(ctx);
  plan_lr_add<T>(pl, d, lra);
  pl.run();
}

// naive_addition (addition.cpp:68-89): C <- [Uc | Ux] diag(Sc, Sx) [Vc; Vx]
// (a rectangular Sx is folded into U: [Uc | Ux Sx], diag(Sc, I)); when
// kc + kx >= min(m, n) the dense sum is recompressed at rank kc instead
// (LowRank(Dense(A) + Dense(B), A.rank)).  Drains the stream.
template <typename T>
void lr_add_naive(const std::vector<LrAddDesc<T>>& d, Context& ctx) {
  if (d.empty()) return;
  const hipStream_t s = ctx.stream();
  ctx.sync();
  std::vector<int> cat, dns;
  for (int i = 0; i < (int)d.size(); ++i)
    (d[i].kc + d[i].kx2 >= std::min(d[i].m, d[i].n) ? dns : cat).push_back(i);
  if (!cat.empty()) {
    ctx.reserve(4096);
    const T* I = ctx.identity<T>(std::max(1, d[0].cap));
    const int ild = ctx.ident_ld();
    Plan pl(ctx);
    std::vector<GemmItem> cp, g1, zero;
    std::vector<SetIntItem> sr;
    for (int i : cat) {
      const auto& x = d[i];
      const int kc = x.kc, k1 = x.kx1, k2 = x.kx2;
      if (kc + k2 > x.cap) throw std::runtime_error("naive addition: rank " + std::to_string(kc + k2) +
                                                    " exceeds the low-rank capacity " + std::to_string(x.cap));
      T* ublk = x.cu + (int64_t)kc * x.cu_cs;
      if (x.xs != nullptr && k1 == k2) {
        cp.push_back(gemm_item(x.xu, x.xu_rs, x.xu_cs, nullptr, 0, 0, ublk, x.cu_rs, x.cu_cs, x.m, k1, 0, 0.0, 0.0));
        cp.push_back(gemm_item(x.xs, x.xs_ld, 1, nullptr, 0, 0, x.cs + (int64_t)kc * x.lds + kc, x.lds, 1, k1, k2, 0,
                               0.0, 0.0));
      } else if (x.xs != nullptr) {
        g1.push_back(gemm_item(x.xu, x.xu_rs, x.xu_cs, x.xs, x.xs_ld, 1, ublk, x.cu_rs, x.cu_cs, x.m, k2, k1, 1.0, 0.0));
        cp.push_back(gemm_item(I, ild, 1, nullptr, 0, 0, x.cs + (int64_t)kc * x.lds + kc, x.lds, 1, k2, k2, 0, 0.0,
                               0.0));
      } else {
        cp.push_back(gemm_item(x.xu, x.xu_rs, x.xu_cs, nullptr, 0, 0, ublk, x.cu_rs, x.cu_cs, x.m, k2, 0, 0.0, 0.0));
        cp.push_back(gemm_item(I, ild, 1, nullptr, 0, 0, x.cs + (int64_t)kc * x.lds + kc, x.lds, 1, k2, k2, 0, 0.0,
                               0.0));
      }
      // off-diagonal blocks of S are zero (k = 0 products write beta * C = 0)
      zero.push_back(gemm_item(nullptr, 0, 0, nullptr, 0, 0, x.cs + kc, x.lds, 1, kc, k2, 0, 1.0, 0.0));
      zero.push_back(gemm_item(nullptr, 0, 0, nullptr, 0, 0, x.cs + (int64_t)kc * x.lds, x.lds, 1, k2, kc, 0, 1.0, 0.0));
      cp.push_back(gemm_item(x.xv, x.xv_rs, x.xv_cs, nullptr, 0, 0, x.cv + (int64_t)kc * x.cv_rs, x.cv_rs, x.cv_cs, k2,
                             x.n, 0, 0.0, 0.0));
      sr.push_back({x.d_rank, kc + k2, 0});
    }
    plan_copies<T>(pl, cp);
    plan_gemms<T>(pl, g1, "lr_naive");
    plan_gemms<T>(pl, zero, "lr_naive");
    const size_t osr = pl.add(sr);
    const int nsr = (int)sr.size();
    pl.launch([=](const char* dd) { launch_set_int(reinterpret_cast<const SetIntItem*>(dd + osr), nsr, s); });
    pl.run();
    ctx.sync();
  }
  if (!dns.empty()) {
    size_t need = 8192;
    for (int i : dns) need += sizeof(T) * ((size_t)d[i].m * d[i].n + (size_t)d[i].m * (d[i].kc + d[i].kx2) + 256);
    ctx.reserve(need);
    Plan pl(ctx);
    std::vector<GemmItem> g1, g2, g3;
    std::vector<CompressDesc<T>> cd;
    for (int i : dns) {
      const auto& x = d[i];
      T* W = ctx.take<T>((size_t)x.m * x.n + 1);
      T* t1 = ctx.take<T>((size_t)x.m * x.kc + 1);
      T* t2 = ctx.take<T>((size_t)x.m * x.kx2 + 1);
      g1.push_back(gemm_item(x.cu, x.cu_rs, x.cu_cs, x.cs, x.lds, 1, t1, x.kc, 1, x.m, x.kc, x.kc, 1.0, 0.0));
      if (x.xs != nullptr)
        g1.push_back(gemm_item(x.xu, x.xu_rs, x.xu_cs, x.xs, x.xs_ld, 1, t2, x.kx2, 1, x.m, x.kx2, x.kx1, 1.0, 0.0));
      else
        g1.push_back(gemm_item(x.xu, x.xu_rs, x.xu_cs, nullptr, 0, 0, t2, x.kx2, 1, x.m, x.kx2, 0, 0.0, 0.0));
      g2.push_back(gemm_item(t1, x.kc, 1, x.cv, x.cv_rs, x.cv_cs, W, x.n, 1, x.m, x.n, x.kc, 1.0, 0.0));
      g3.push_back(gemm_item(t2, x.kx2, 1, x.xv, x.xv_rs, x.xv_cs, W, x.n, 1, x.m, x.n, x.kx2, 1.0, 1.0));
      CompressDesc<T> c{};
      c.a = W;
      c.m = x.m;
      c.n = x.n;
      c.lda = x.n;
      c.eps = 0.0;
      c.rank = x.kc;
      c.u = x.cu;
      c.s = x.cs;
      c.v = x.cv;
      c.u_rs = x.cu_rs;
      c.u_cs = x.cu_cs;
      c.v_rs = x.cv_rs;
      c.v_cs = x.cv_cs;
      c.lds = x.lds;
      c.d_rank = x.d_rank;
      c.cap = x.cap;
      cd.push_back(c);
    }
    // the copies of a plain xu (no S) are GEMM items with k = 0: route them as copies
    std::vector<GemmItem> c1, p1;
    for (const auto& g : g1) (g.k == 0 && g.B == nullptr ? c1 : p1).push_back(g);
    plan_copies<T>(pl, c1);
    plan_gemms<T>(pl, p1, "lr_naive");
    plan_gemms<T>(pl, g2, "lr_naive");
    plan_gemms<T>(pl, g3, "lr_naive");
    pl.run();
    Context cside(s);
    cside.set_stats_owner(&ctx);
    compress_batched<T>(cd, cside);
    ctx.sync();
  }
}

// ============================================================================
// Compression: LowRank(Dense, eps) and LowRank(Dense, rank)
// ============================================================================
template <typename T>
void compress_batched(const std::vector<CompressDesc<T>>& d, Context& ctx) {
  if (d.empty()) return;
  const hipStream_t s = ctx.stream();
  std::vector<CompressDesc<T>> de, dr;
  for (const auto& x : d) (x.eps != 0.0 ? de : dr).push_back(x);
  if (!de.empty()) {
    // truncated_geqp3 (src/operations/LAPACK/geqp3.cpp:71-191) then rq of R (low_rank.cpp:50-59)
    const int nd = (int)de.size();
    int nmax = 0, mmax = 0, capmax = 0;
    for (const auto& x : de) {
      nmax = std::max(nmax, x.n);
      mmax = std::max(mmax, x.m);
      capmax = std::max(capmax, x.cap);
    }
    ctx.sync();
    ctx.reserve((size_t)nd * sizeof(T) * ((size_t)nmax + (size_t)mmax + (size_t)capmax * nmax * 2 + 4 * capmax +
                                            (size_t)capmax * capmax + 64) +
                (size_t)nd * sizeof(int32_t) * (nmax + 64) + (size_t)16 * 4096);
    T* tau = ctx.take<T>((size_t)nd * std::max(mmax, nmax));
    int32_t* piv = ctx.take<int32_t>((size_t)nd * nmax);
    std::vector<Geqp3Item> gi(nd);
    for (int t = 0; t < nd; ++t) {
      Geqp3Item g{};
      g.A = de[t].a;
      g.tau = tau + (size_t)t * std::max(mmax, nmax);
      g.piv = piv + (size_t)t * nmax;
      g.rank_out = de[t].d_rank;
      g.m = de[t].m;
      g.n = de[t].n;
      g.lda = (int)de[t].lda;
      g.cap = de[t].cap;
      g.cap_trunc = de[t].cap_trunc;
      g.eps = de[t].eps;
      gi[t] = g;
    }
    {
      Plan pl(ctx);
      const size_t og = pl.add(gi);
      pl.launch([=](const char* dd) { launch_geqp3<T>(reinterpret_cast<const Geqp3Item*>(dd + og), nd, s); });
      pl.run();
    }
    std::vector<int32_t> rk(nd);
    for (int t = 0; t < nd; ++t)
      FRK_HIP_CHECK(hipMemcpyAsync(&rk[t], de[t].d_rank, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    host_wait_stream(s, __func__);
    int rmax = 1;
    for (int t = 0; t < nd; ++t) {
      if (rk[t] < 0)
        throw std::runtime_error("low-rank capacity exceeded during compression: rank " + std::to_string(-rk[t]));
      rmax = std::max(rmax, rk[t]);
    }
    const T* I = ctx.identity<T>(rmax + 1);
    const int ild = ctx.ident_ld();
    T* RP = ctx.take<T>((size_t)nd * rmax * nmax);
    T* Yz = ctx.take<T>((size_t)nd * nmax * rmax);
    T* tz = ctx.take<T>((size_t)nd * rmax);
    T* Rz = ctx.take<T>((size_t)nd * rmax * rmax);
    std::vector<ApplyQItem> au, av;
    std::vector<PermuteItem> pi(nd);
    std::vector<QrItem> qz(nd);
    std::vector<FlipItem> fl(nd);
    for (int t = 0; t < nd; ++t) {
      const auto& x = de[t];
      const int r = rk[t];
      ApplyQItem a{};
      a.Y = x.a;
      a.y_rs = x.lda;
      a.y_cs = 1;
      a.tau = gi[t].tau;
      a.X = I;
      a.x_rs = ild;
      a.x_cs = 1;
      a.kx = r;
      a.ncol = r;
      a.nref = std::min(r, std::min(x.m, x.n));
      a.m = x.m;
      a.out = x.u;
      a.out_rs = x.u_rs;
      a.out_cs = x.u_cs;
      au.push_back(a);
      T* rp = RP + (size_t)t * rmax * nmax;
      PermuteItem p{};
      p.A = x.a;
      p.piv = gi[t].piv;
      p.rank = x.d_rank;
      p.out = rp;
      p.n = x.n;
      p.lda = (int)x.lda;
      p.ldo = x.n;
      pi[t] = p;
      // RQ of RP via QR of its column-reversed transpose Z(i, c) = RP[r-1-c][i]
      QrItem q{};
      q.A1 = rp + (size_t)(r - 1) * x.n;
      q.a1_rs = 1;
      q.a1_cs = -(int64_t)x.n;
      q.k1a = q.k1 = r;
      q.m = x.n;
      q.Y = Yz + (size_t)t * nmax * rmax;
      q.y_rs = 1;
    q.y_cs = x.n;
      q.tau = tz + (size_t)t * rmax;
      q.R = Rz + (size_t)t * rmax * rmax;
      q.ldr = r;
      qz[t] = q;
      FlipItem f{};
      f.R = q.R;
      f.S = x.s;
      f.rank = x.d_rank;
      f.ldr = r;
      f.lds = x.lds;
      fl[t] = f;
      ApplyQItem b{};
      b.Y = q.Y;
      b.y_rs = 1;
      b.y_cs = x.n;
      b.tau = q.tau;
      b.X = I + (r - 1);
      b.x_rs = ild;
      b.x_cs = -1;
      b.kx = r;
      b.ncol = r;
      b.nref = std::min(r, x.n);
      b.m = x.n;
      b.out = x.v;
      b.out_rs = x.v_cs;
      b.out_cs = x.v_rs;
      av.push_back(b);
    }
    Plan pl(ctx);
    const size_t oau = pl.add(au), opi = pl.add(pi), oqz = pl.add(qz), ofl = pl.add(fl), oav = pl.add(av);
    pl.launch([=](const char* dd) { launch_apply_q<T>(reinterpret_cast<const ApplyQItem*>(dd + oau), nd, rmax, s); });
    pl.launch([=](const char* dd) {
      launch_permute_r<T>(reinterpret_cast<const PermuteItem*>(dd + opi), nd, nmax, rmax, s);
    });
    pl.launch([=](const char* dd) { launch_qr<T>(reinterpret_cast<const QrItem*>(dd + oqz), nd, rmax, s); });
    pl.launch([=](const char* dd) { launch_flip_r<T>(reinterpret_cast<const FlipItem*>(dd + ofl), nd, rmax, s); });
    pl.launch([=](const char* dd) { launch_apply_q<T>(reinterpret_cast<const ApplyQItem*>(dd + oav), nd, rmax, s); });
    pl.run();
    ctx.sync();
  }
  if (!dr.empty()) {
    // rsvd (src/operations/randomized/rsvd.cpp:17-29) at p = min(rank + 5, m, n), truncated to rank.
    // Every call shares one Omega(n, p): split batches whose tiles differ in (n, p).
    auto np_of = [](const CompressDesc<T>& x) {
      return std::make_pair(x.n, std::min(std::min(x.rank + 5, x.m), x.n));
    };
    for (const auto& x : dr)
      if (np_of(x) != np_of(dr[0])) {
        std::vector<CompressDesc<T>> same, rest;
        for (const auto& y : dr) (np_of(y) == np_of(dr[0]) ? same : rest).push_back(y);
        compress_batched<T>(same, ctx);
        compress_batched<T>(rest, ctx);
        return;
      }
    const int nd = (int)dr.size();
    ctx.sync();
    int nmax = 0, mmax = 0, pmax = 0;
    for (const auto& x : dr) {
      nmax = std::max(nmax, x.n);
      mmax = std::max(mmax, x.m);
      pmax = std::max(pmax, std::min(std::min(x.rank + 5, x.m), x.n));
    }
    const T* I = ctx.identity<T>(pmax + 1);
    const int ild = ctx.ident_ld();
    ctx.reserve((size_t)nd * sizeof(T) *
                    ((size_t)mmax * pmax * 2 + (size_t)nmax * pmax + 5 * (size_t)pmax * pmax + 4 * pmax + 64) +
                (size_t)16 * 4096);
    std::vector<GemmItem> g1, g2;
    std::vector<QrItem> q1, q2;
    std::vector<ApplyQItem> a1, a2;
    std::vector<SvdItem> sv;
    T* Yb = ctx.take<T>((size_t)nd * mmax * pmax);
    T* Qe = ctx.take<T>((size_t)nd * mmax * pmax);
    T* Zb = ctx.take<T>((size_t)nd * nmax * pmax);
    T* ty = ctx.take<T>((size_t)nd * pmax);
    T* tz = ctx.take<T>((size_t)nd * pmax);
    T* Ry = ctx.take<T>((size_t)nd * pmax * pmax);
    T* Rz = ctx.take<T>((size_t)nd * pmax * pmax);
    T* Um = ctx.take<T>((size_t)nd * pmax * pmax);
    T* Vm = ctx.take<T>((size_t)nd * pmax * pmax);
    T* Vw = (sizeof(T) == 8 && pmax > 64) ? ctx.take<T>((size_t)nd * pmax * pmax) : nullptr;
    for (int t = 0; t < nd; ++t) {
      const auto& x = dr[t];
      const int p = std::min(std::min(x.rank + 5, x.m), x.n);
      const T* Om = ctx.omega<T>(x.n, p);
      T* y = Yb + (size_t)t * mmax * pmax;
      T* qe = Qe + (size_t)t * mmax * pmax;
      T* z = Zb + (size_t)t * nmax * pmax;
      g1.push_back(gemm_item(x.a, x.lda, 1, Om, p, 1, y, 1, x.m, x.m, p, x.n, 1.0, 0.0));  // Y = A Omega
      QrItem q{};
      q.A1 = y;
      q.a1_rs = 1;
      q.a1_cs = x.m;
      q.k1a = q.k1 = p;
      q.m = x.m;
      q.Y = y;
      q.y_rs = 1;
    q.y_cs = x.m;
      q.tau = ty + (size_t)t * pmax;
      q.R = Ry + (size_t)t * pmax * pmax;
      q.ldr = p;
      q1.push_back(q);
      ApplyQItem a{};
      a.Y = y;
      a.y_rs = 1;
      a.y_cs = x.m;
      a.tau = q.tau;
      a.X = I;
      a.x_rs = ild;
      a.x_cs = 1;
      a.kx = p;
      a.ncol = p;
      a.nref = p;
      a.m = x.m;
      a.out = qe;
      a.out_rs = 1;
      a.out_cs = x.m;
      a1.push_back(a);
      g2.push_back(gemm_item(x.a, 1, x.lda, qe, 1, x.m, z, 1, x.n, x.n, p, x.m, 1.0, 0.0));  // Z = A^T Q
      QrItem qq{};
      qq.A1 = z;
      qq.a1_rs = 1;
      qq.a1_cs = x.n;
      qq.k1a = qq.k1 = p;
      qq.m = x.n;
      qq.Y = z;
      qq.y_rs = 1;
    qq.y_cs = x.n;
      qq.tau = tz + (size_t)t * pmax;
      qq.R = Rz + (size_t)t * pmax * pmax;
      qq.ldr = p;
      q2.push_back(qq);
      SvdItem s0{};
      s0.mode = 0;
      s0.M = qq.R;
      s0.m_rs = 1;
      s0.m_cs = p;  // M = Rz^T = (Q^T A) Qz
      s0.K = p;
      s0.fixed_rank = x.rank;
      s0.S = x.s;
      s0.lds = x.lds;
      s0.Um = Um + (size_t)t * pmax * pmax;
      s0.ldu = p;
      s0.Vm = Vm + (size_t)t * pmax * pmax;
      s0.ldv = p;
      s0.rank_out = x.d_rank;
      s0.cap = x.cap;
      s0.stats = ctx.counters();
      s0.vwork = Vw ? Vw + (size_t)t * pmax * pmax : nullptr;
      sv.push_back(s0);
      ApplyQItem b{};
      b.Y = y;
      b.y_rs = 1;
      b.y_cs = x.m;
      b.tau = q.tau;
      b.X = s0.Um;
      b.x_rs = p;
      b.x_cs = 1;
      b.kx = p;
      b.ncol = x.rank;
      b.nref = p;
      b.m = x.m;
      b.out = x.u;
      b.out_rs = x.u_rs;
      b.out_cs = x.u_cs;
      a2.push_back(b);
      ApplyQItem c{};
      c.Y = z;
      c.y_rs = 1;
      c.y_cs = x.n;
      c.tau = qq.tau;
      c.X = s0.Vm;
      c.x_rs = p;
      c.x_cs = 1;
      c.kx = p;
      c.ncol = x.rank;
      c.nref = p;
      c.m = x.n;
      c.out = x.v;
      c.out_rs = x.v_cs;
      c.out_cs = x.v_rs;
      a2.push_back(c);
    }
    Plan pl(ctx);
    plan_gemms<T>(pl, g1);
    const size_t oq1 = pl.add(q1), oa1 = pl.add(a1);
    pl.launch([=](const char* dd) { launch_qr<T>(reinterpret_cast<const QrItem*>(dd + oq1), nd, pmax, s); });
    pl.launch([=](const char* dd) { launch_apply_q<T>(reinterpret_cast<const ApplyQItem*>(dd + oa1), nd, pmax, s); });
    plan_gemms<T>(pl, g2);
    const size_t oq2 = pl.add(q2), osv = pl.add(sv), oa2 = pl.add(a2);
    const int na2 = (int)a2.size();
    pl.launch([=](const char* dd) { launch_qr<T>(reinterpret_cast<const QrItem*>(dd + oq2), nd, pmax, s); });
    pl.launch([=](const char* dd) { launch_svd<T>(reinterpret_cast<const SvdItem*>(dd + osv), nd, pmax, s); });
    pl.launch(
        [=](const char* dd) { launch_apply_q<T>(reinterpret_cast<const ApplyQItem*>(dd + oa2), na2, pmax, s); });
    pl.run();
    ctx.sync();
  }
}

// ============================================================================
// Construction from a kernel function
// ============================================================================
static std::vector<int64_t> index_split(int64_t n, int nsplit) {
  // src/classes/initialization_helpers/index_range.cpp:15-24 (ceil split, last block smaller)
  std::vector<int64_t> off(nsplit + 1);
  const int64_t child = (n + nsplit - 1) / nsplit;
  for (int i = 0; i < nsplit; ++i) off[i] = child * i;
  off[nsplit] = n;
  return off;
}

static bool admissible(const KernelDesc& kd, int64_t r0, int64_t nr, int64_t c0, int64_t nc, int bi, int bj,
                       double admis, int admis_type) {
  // src/classes/initialization_helpers/matrix_initializer.cpp:51-91
  bool ok = nr > 1 && nc > 1;
  if (admis_type == 0) return ok && std::abs(bi - bj) > (int64_t)admis;
  const double offset = 5e-1;
  double max_len_r = 0, max_len_c = 0, dist = 0;
  for (int d = 0; d < kd.dim; ++d) {
    const double* x = kd.host_pts + (size_t)d * kd.n;
    const double mnr = *std::min_element(x + r0, x + r0 + nr) - offset;
    const double mxr = *std::max_element(x + r0, x + r0 + nr) + offset;
    const double mnc = *std::min_element(x + c0, x + c0 + nc) - offset;
    const double mxc = *std::max_element(x + c0, x + c0 + nc) + offset;
    const double cr = mnr + (mxr - mnr) / 2.0, cc = mnc + (mxc - mnc) / 2.0;
    max_len_r = std::max(max_len_r, mxr - mnr);
    max_len_c = std::max(max_len_c, mxc - mnc);
    const double dd = std::fabs(cr - cc);
    dist += dd * dd;
  }
  const double diam = std::max(max_len_r, max_len_c);
  return ok && (admis * admis * diam * diam) < dist;
}

KernelSpec upload_kernel(const KernelDesc& kd, DevBuf& pts, hipStream_t s) {
  pts.ensure(sizeof(double) * kd.dim * kd.n);
  FRK_HIP_CHECK(hipMemcpyAsync(pts.as<double>(), kd.host_pts, sizeof(double) * kd.dim * kd.n,
                               hipMemcpyHostToDevice, s));
  KernelSpec ks{};
  ks.pts = pts.as<double>();
  ks.n_pts = kd.n;
  ks.dim = kd.dim;
  ks.kind = kd.kind;
  ks.ell = kd.ell;
  ks.shift = kd.shift;
  return ks;
}

template <typename T>
std::unique_ptr<BlrMatrix<T>> build_blr_kernel(const KernelDesc& kd, int nb_r, int nb_c, double eps, int rank,
                                               double admis, int admis_type, int kcap, Context& ctx) {
  const hipStream_t s = ctx.stream();
  const int64_t n_rows = kd.n_rows > 0 ? kd.n_rows : kd.n, n_cols = kd.n_cols > 0 ? kd.n_cols : kd.n;
  if (n_rows > kd.n || n_cols > kd.n) throw std::invalid_argument("matrix larger than the point set");
  auto roff = index_split(n_rows, nb_r);
  auto coff = index_split(n_cols, nb_c);
  std::vector<int32_t> kinds(nb_r * nb_c);
  for (int i = 0; i < nb_r; ++i)
    for (int j = 0; j < nb_c; ++j)
      kinds[i * nb_c + j] = admissible(kd, roff[i], roff[i + 1] - roff[i], coff[j], coff[j + 1] - coff[j], i, j,
                                       admis, admis_type)
                                ? kLowRank
                                : kDense;
  const int nb = nb_c;
  auto A = std::make_unique<BlrMatrix<T>>(roff, coff, kinds, kcap, eps, rank, ctx);
  A->cap_trunc = ctx.cap_trunc();
  const auto& off = roff;
  DevBuf pts;
  KernelSpec ks = upload_kernel(kd, pts, s);
  ctx.sync();
  std::vector<GenItem> gd;
  int mr = 0, mc = 0;
  for (int i = 0; i < nb_r; ++i)
    for (int j = 0; j < nb_c; ++j)
      if (kinds[i * nb + j] == kDense) {
        GenItem g{};
        g.out = A->D[A->idx(i, j)];
        g.ld = A->cols(j);
        g.r0 = (int)off[i];
        g.nr = (int)A->rows(i);
        g.c0 = (int)coff[j];
        g.nc = (int)A->cols(j);
        gd.push_back(g);
        mr = std::max<int>(mr, g.nr);
        mc = std::max<int>(mc, g.nc);
      }
  {
    Plan pl(ctx);
    const size_t og = pl.add(gd);
    const int ng = (int)gd.size();
    pl.launch([=](const char* dd) { launch_gen<T>(ks, reinterpret_cast<const GenItem*>(dd + og), ng, mr, mc, s); });
    pl.run();
  }
  std::vector<std::pair<int, int>> lr;
  for (int i = 0; i < nb_r; ++i)
    for (int j = 0; j < nb_c; ++j)
      if (kinds[i * nb + j] == kLowRank) lr.emplace_back(i, j);
  int64_t bmax = 1;
  for (int i = 0; i < nb_r; ++i) bmax = std::max<int64_t>(bmax, A->rows(i));
  for (int j = 0; j < nb_c; ++j) bmax = std::max<int64_t>(bmax, A->cols(j));
  const size_t per = (size_t)bmax * bmax;
  const int chunk = (int)std::max<int64_t>(1, std::min<int64_t>(1024, (int64_t(2) << 30) / (per * sizeof(T))));
  DevBuf dense;
  dense.ensure(sizeof(T) * per * std::min<size_t>(chunk, std::max<size_t>(lr.size(), 1)));
  for (size_t o = 0; o < lr.size(); o += chunk) {
    const size_t e = std::min(lr.size(), o + chunk);
    std::vector<GenItem> gl;
    std::vector<CompressDesc<T>> cd;
    for (size_t t = o; t < e; ++t) {
      const int i = lr[t].first, j = lr[t].second;
      T* buf = dense.as<T>() + (t - o) * per;
      GenItem g{};
      g.out = buf;
      g.ld = A->cols(j);
      g.r0 = (int)off[i];
      g.nr = (int)A->rows(i);
      g.c0 = (int)coff[j];
      g.nc = (int)A->cols(j);
      gl.push_back(g);
      CompressDesc<T> c{};
      c.a = buf;
      c.m = (int)A->rows(i);
      c.n = (int)A->cols(j);
      c.lda = A->cols(j);
      c.eps = eps;
      c.rank = rank;
      c.u = A->U[A->idx(i, j)];
      c.s = A->S[A->idx(i, j)];
      c.v = A->V[A->idx(i, j)];
      c.u_rs = 1;
      c.u_cs = A->rows(i);
      c.v_rs = A->cols(j);
      c.v_cs = 1;
      c.lds = kcap;
      c.d_rank = A->d_rank(i, j);
      c.cap = kcap;
      c.cap_trunc = ctx.cap_trunc();
      cd.push_back(c);
    }
    {
      Plan pl(ctx);
      const size_t og = pl.add(gl);
      const int ng = (int)gl.size();
      const int bm = (int)bmax;
      pl.launch([=](const char* dd) { launch_gen<T>(ks, reinterpret_cast<const GenItem*>(dd + og), ng, bm, bm, s); });
      pl.run();
    }
    compress_batched<T>(cd, ctx);
  }
  ctx.sync();
  A->pull_ranks(s);
  return A;
}

// ============================================================================
// BLR-LU
// ============================================================================
// Whether the diagonal tile i is factored by the one-launch tile LU.
template <typename T>
static bool tile_lu_path(const BlrMatrix<T>& A, int i) {
  return use_tile_lu() && A.rows(i) == A.cols(i) && A.rows(i) <= 512;
}

// Factorization of the dense diagonal tile i (always in FP64): plan into p on
// context c (its stream and workspace).  pre (tile path only): a low-rank
// update of the tile applied by the tile LU kernel before it factors (the
// priority block's C += U core V).
template <typename T>
static void plan_diag_getrf(Plan& p, Context& c, BlrMatrix<T>& A, int i, int32_t* info,
                            const LrUpdItem* pre = nullptr) {
  const int ii = A.idx(i, i);
  const int m_i = (int)A.rows(i), n_i = (int)A.cols(i);
  const bool tile = tile_lu_path(A, i);
  if (pre != nullptr && !tile) throw std::logic_error("tile LU pre-update without the tile LU");
  if constexpr (sizeof(T) == 8) {
    if (tile)
      plan_tile_lu(p, A.D[ii], A.Linv[ii], A.Uinv[ii], n_i, n_i, info, {}, pre, false);
    else
      plan_getrf<T>(p, {A.D[ii]}, {A.Linv[ii]}, {A.Uinv[ii]}, m_i, n_i, info);
  } else {
    const hipStream_t s = c.stream();
    const int64_t nn = (int64_t)m_i * n_i;
    double* d64 = c.take<double>(nn);
    double* l64 = c.take<double>(nn);
    double* u64 = c.take<double>(nn);
    T* dt = A.D[ii];
    T* lt = A.Linv[ii];
    T* ut = A.Uinv[ii];
    if (tile) {
      // the tile LU reads the FP32 tile and writes the FP32 results itself
      TileLu32 f32;
      f32.a = reinterpret_cast<float*>(dt);
      f32.linv = reinterpret_cast<float*>(lt);
      f32.uinv = reinterpret_cast<float*>(ut);
      plan_tile_lu(p, d64, l64, u64, n_i, n_i, info, f32, pre, true);
      return;
    }
    p.launch([=](const char*) {
      launch_convert<double, T>(d64, dt, nn, s);
      launch_fill_value(l64, nn, 0.0, s);
      launch_fill_value(u64, nn, 0.0, s);
    });
    plan_getrf<double>(p, {d64}, {l64}, {u64}, m_i, n_i, info);
    p.launch([=](const char*) {
      launch_convert<T, double>(dt, d64, nn, s);
      launch_convert<T, double>(lt, l64, nn, s);
      launch_convert<T, double>(ut, u64, nn, s);
    });
  }
}

template <typename T>
static size_t diag_getrf_workspace(int n) {
  return (sizeof(T) == 4 ? sizeof(double) * ((size_t)3 * n * n + 768) : 0) + getrf_workspace<double>(n, 1);
}

// The lookahead streams, their events and a pinned snapshot of the rank table.
// prio (lookahead stream): the priority block of the latest step, and so the
// factorization of its diagonal tile, is done; sxa (main): a step's panels and
// stacked products are done; c0 / c1: its first / second chunk is done; rest
// (second stream): the step's ranks are in snap.
struct Lookahead {
  hipStream_t s = nullptr, s2 = nullptr;  // lookahead stream, second rest-chunk stream (the context's)
  hipEvent_t prio = nullptr, rest = nullptr, sxa = nullptr, c0 = nullptr, c1 = nullptr;
  int32_t* snap = nullptr;
  int32_t* snap_dev = nullptr;  // snap as the device sees it
  Lookahead(Context& ctx, size_t nrank) : s(ctx.aux_stream(0)), s2(ctx.aux_stream(1)) {
    // device-side ordering only: no system-scope fence (a cache write-back per
    // record); rest keeps it, since the host reads the snapshot after it
    for (hipEvent_t* e : {&prio, &sxa, &c0, &c1})
      FRK_HIP_CHECK(hipEventCreateWithFlags(e, hipEventDisableTiming | hipEventDisableSystemFence));
    FRK_HIP_CHECK(hipEventCreateWithFlags(&rest, hipEventDisableTiming));
    // nrank rank slots, then the lookahead context's tile-LU timeout marker
    FRK_HIP_CHECK(hipHostMalloc((void**)&snap, sizeof(int32_t) * (nrank + 1)));
    FRK_HIP_CHECK(hipHostGetDevicePointer((void**)&snap_dev, snap, 0));
    snap[nrank] = 0;
  }
  ~Lookahead() {
    (void)hipStreamSynchronize(s);
    (void)hipStreamSynchronize(s2);
    for (hipEvent_t e : {prio, rest, sxa, c0, c1}) (void)hipEventDestroy(e);
    (void)hipHostFree(snap);
  }
};

// Whether the priority block of step k can be planned before the ranks of its
// panel tiles are known on the host: both low rank (the kernels read the ranks
// on the device) or both dense (no rank enters).
template <typename T>
static bool prio_rank_free(const BlrMatrix<T>& A, int k) {
  const int kl = A.kind[A.idx(k + 1, k)], kb = A.kind[A.idx(k, k + 1)];
  return (kl == kLowRank && kb == kLowRank && A.kcap <= 64) || (kl == kDense && kb == kDense);
}

template <typename T>
static size_t prio_workspace(const BlrMatrix<T>& A, int64_t bm) {
  return sizeof(T) * (6 * (size_t)bm * bm + 4 * (size_t)A.kcap * bm + (size_t)A.kcap * A.kcap) + 16 * 256;
}

// Priority block of step k (getrf.cpp:50-62 restricted to what the diagonal
// tile k+1 depends on): the panel solves of tiles (k+1, k) and (k, k+1) as
// products with the explicit inverses of tile k, then the Schur update of the
// dense tile (k+1, k+1).  With device_ranks (prio_rank_free) the low-rank tiles
// are processed at capacity and their ranks read on the device; otherwise rl,
// rb are the ranks of the two panel tiles.
// Exact op-log cost of the device-rank priority block of step k with panel
// ranks rl, rb (its launches are counted when issued).
template <typename T>
static void account_prio(Context& c, const BlrMatrix<T>& A, int k, int rl, int rb) {
  const double n_k = (double)A.cols(k), m_k = (double)A.rows(k);
  const double m = (double)A.rows(k + 1), n = (double)A.cols(k + 1);
  c.account("trsm_gemm", 2.0 * (rl * n_k * n_k + rb * m_k * m_k),
            sizeof(T) * (n_k * n_k + m_k * m_k + 2.0 * (rl * n_k + rb * m_k)), 0, 0);
  c.account("schur_gemm", 2.0 * rl * rb * n_k + 2.0 * rl * (rl + rb) * rb + 2.0 * m * rl * rb + 2.0 * m * n * rb,
            sizeof(T) * (2.0 * m * n + (rl + rb) * (n_k + m + n)), 0, 0);
}

// With device ranks and both panels low rank the launches are counted here and
// their cost is added by account_prio once the panel ranks are known on the host.
template <typename T>
static void plan_prio(Plan& pl, Context& W, BlrMatrix<T>& A, int k, int rl, int rb, bool device_ranks,
                      LrUpdItem* defer_update = nullptr, bool* deferred = nullptr) {
  if (deferred) *deferred = false;
  const hipStream_t s = W.stream();
  const int kk = A.idx(k, k), tl = A.idx(k + 1, k), tb = A.idx(k, k + 1), tc = A.idx(k + 1, k + 1);
  const int m_k = (int)A.rows(k), n_k = (int)A.cols(k);
  const int m = (int)A.rows(k + 1), n = (int)A.cols(k + 1);
  const int kl = A.kind[tl], kb = A.kind[tb];
  const int32_t* dl = device_ranks ? A.d_rank(k + 1, k) : nullptr;
  const int32_t* db = device_ranks ? A.d_rank(k, k + 1) : nullptr;
  if (device_ranks) rl = rb = A.kcap;
  if (device_ranks && kl == kLowRank && kb == kLowRank) {
    // three launches, no descriptor upload: both panel products (V U^-1 and
    // (L^-1 U)^T, rows at capacity, counts read on the device), the core
    // -S_l (V_l' U_b') S_b with the panel results written back, the update of C
    T* wv = W.take<T>((size_t)A.kcap * n_k + 1);
    T* wu = W.take<T>((size_t)A.kcap * m_k + 1);
    T* core = W.take<T>((size_t)A.kcap * A.kcap + 1);
    RowsPair rp{};
    rp.job[0] = RowsJob{A.V[tl], A.cols(k), A.Uinv[kk], n_k, 1, wv, dl, A.kcap, n_k, n_k, 0};
    rp.job[1] = RowsJob{A.U[tb], A.rows(k), A.Linv[kk], 1, m_k, wu, db, A.kcap, m_k, m_k, 0};
    LrCoreItem ci{};
    ci.vl = wv;
    ci.v_rs = n_k;
    ci.v_cs = 1;
    ci.ub = wu;
    ci.u_rs = 1;
    ci.u_cs = m_k;
    ci.sl = A.S[tl];
    ci.sb = A.S[tb];
    ci.core = core;
    ci.rl = dl;
    ci.rb = db;
    ci.lds = A.kcap;
    ci.ldk = A.kcap;
    ci.k = n_k;
    ci.kcap = A.kcap;
    ci.alpha = -1.0;
    ci.vl_dst = A.V[tl];
    ci.vl_dst_ld = A.cols(k);
    ci.nv = n_k;
    ci.ub_dst = A.U[tb];
    ci.ub_dst_ld = A.rows(k);
    ci.nu = m_k;
    LrUpdItem ui{};
    ui.u = A.U[tl];
    ui.u_rs = 1;
    ui.u_cs = A.rows(k + 1);
    ui.core = core;
    ui.v = A.V[tb];
    ui.v_rs = A.cols(k + 1);
    ui.v_cs = 1;
    ui.c = A.D[tc];
    ui.ldc = n;
    ui.r1 = dl;
    ui.r2 = db;
    ui.ldk = A.kcap;
    ui.m = m;
    ui.n = n;
    ui.kcap = A.kcap;
    pl.launch([=](const char*) { launch_rows_pair<T>(rp, s); }, "trsm_gemm");
    if (defer_update != nullptr) {
      // the update of C is left to the tile LU of C, which applies it first
      *defer_update = ui;
      *deferred = true;
      pl.launch([=](const char*) { launch_lr_core<T>(ci, s); }, "schur_gemm");
    } else {
      pl.launch([=](const char*) {
        launch_lr_core<T>(ci, s);
        launch_lr_update<T>(ui, s);
      }, "schur_gemm");
    }
    return;
  }
  std::vector<GemmItem> tg, tcp;
  auto rows_solve = [&](std::vector<RowRef> rows, const T* Bm, int64_t brs, int64_t bcs, int N, const int32_t* dc) {
    const int R = (int)rows.size();
    T* Wb = W.take<T>((size_t)R * N + 1);
    const size_t off = pl.add(rows);
    pl.launch(
        [=](const char* d) {
          const RowRef* rr = reinterpret_cast<const RowRef*>(d + off);
          launch_rows_gemm<T>(rr, R, Bm, brs, bcs, Wb, N, N, s, dc);
          launch_rows_scatter<T>(rr, R, Wb, N, s, dc);
        },
        "trsm_gemm", 2.0 * R * N * N, sizeof(T) * ((double)N * N + 2.0 * R * N));
  };
  if (kl == kLowRank && rl > 0) {  // V <- V U^-1
    std::vector<RowRef> rows(rl);
    for (int r = 0; r < rl; ++r) {
      T* row = A.V[tl] + (size_t)r * A.cols(k);
      rows[r] = RowRef{row, row};
    }
    rows_solve(rows, A.Uinv[kk], n_k, 1, n_k, dl);
  } else if (kl == kDense) {
    T* w = W.take<T>((size_t)m * n_k + 1);
    tg.push_back(gemm_item(A.D[tl], A.cols(k), 1, A.Uinv[kk], n_k, 1, w, n_k, 1, m, n_k, n_k, 1.0, 0.0));
    tcp.push_back(gemm_item(w, n_k, 1, nullptr, 0, 0, A.D[tl], A.cols(k), 1, m, n_k, 0, 0.0, 0.0));
  }
  if (kb == kLowRank && rb > 0) {  // U <- L^-1 U, as U^T <- U^T L^-T on the columns of U
    std::vector<RowRef> cols(rb);
    for (int r = 0; r < rb; ++r) {
      T* col = A.U[tb] + (size_t)r * A.rows(k);
      cols[r] = RowRef{col, col};
    }
    rows_solve(cols, A.Linv[kk], 1, m_k, m_k, db);
  } else if (kb == kDense) {
    T* w = W.take<T>((size_t)m_k * n + 1);
    tg.push_back(gemm_item(A.Linv[kk], m_k, 1, A.D[tb], n, 1, w, n, 1, m_k, n, m_k, 1.0, 0.0));
    tcp.push_back(gemm_item(w, n, 1, nullptr, 0, 0, A.D[tb], n, 1, m_k, n, 0, 0.0, 0.0));
  }
  plan_gemms<T>(pl, tg, "trsm_gemm");
  plan_copies<T>(pl, tcp);
  // C(k+1, k+1) -= L(k+1, k) U(k, k+1) by operand kinds (gemm.cpp:126-272)
  std::vector<GemmItem> gA, gB, gC, gD;
  T* C = A.D[tc];
  if (kl == kLowRank && kb == kLowRank) {
    if (rl > 0 && rb > 0) {
      T* P = W.take<T>((size_t)rl * n_k + 1);
      T* Q = W.take<T>((size_t)m_k * rb + 1);
      T* Sx = W.take<T>((size_t)rl * rb + 1);
      T* z = W.take<T>((size_t)m * rb + 1);
      gA.push_back(gemm_item(A.S[tl], A.kcap, 1, A.V[tl], A.cols(k), 1, P, n_k, 1, rl, n_k, rl, -1.0, 0.0));
      gA.push_back(gemm_item(A.U[tb], 1, A.rows(k), A.S[tb], A.kcap, 1, Q, 1, m_k, m_k, rb, rb, 1.0, 0.0));
      gB.push_back(gemm_item(P, n_k, 1, Q, 1, m_k, Sx, rb, 1, rl, rb, n_k, 1.0, 0.0));
      gC.push_back(gemm_item(A.U[tl], 1, A.rows(k + 1), Sx, rb, 1, z, 1, m, m, rb, rl, 1.0, 0.0));
      gD.push_back(gemm_item(z, 1, m, A.V[tb], A.cols(k + 1), 1, C, n, 1, m, n, rb, 1.0, 1.0));
    }
  } else if (kl == kLowRank && kb == kDense) {
    if (rl > 0) {
      T* w = W.take<T>((size_t)rl * n + 1);
      T* w2 = W.take<T>((size_t)rl * n + 1);
      gA.push_back(gemm_item(A.V[tl], A.cols(k), 1, A.D[tb], n, 1, w, n, 1, rl, n, n_k, -1.0, 0.0));
      gB.push_back(gemm_item(A.S[tl], A.kcap, 1, w, n, 1, w2, n, 1, rl, n, rl, 1.0, 0.0));
      gC.push_back(gemm_item(A.U[tl], 1, A.rows(k + 1), w2, n, 1, C, n, 1, m, n, rl, 1.0, 1.0));
    }
  } else if (kl == kDense && kb == kLowRank) {
    if (rb > 0) {
      T* a = W.take<T>((size_t)m * rb + 1);
      T* a2 = W.take<T>((size_t)m * rb + 1);
      gA.push_back(gemm_item(A.D[tl], n_k, 1, A.U[tb], 1, A.rows(k), a, 1, m, m, rb, n_k, -1.0, 0.0));
      gB.push_back(gemm_item(a, 1, m, A.S[tb], A.kcap, 1, a2, 1, m, m, rb, rb, 1.0, 0.0));
      gC.push_back(gemm_item(a2, 1, m, A.V[tb], A.cols(k + 1), 1, C, n, 1, m, n, rb, 1.0, 1.0));
    }
  } else if (kl == kDense && kb == kDense) {
    gA.push_back(gemm_item(A.D[tl], n_k, 1, A.D[tb], n, 1, C, n, 1, m, n, n_k, -1.0, 1.0));
  }
  plan_gemms<T>(pl, gA, "schur_gemm");
  plan_gemms<T>(pl, gB, "schur_gemm");
  plan_gemms<T>(pl, gC, "schur_gemm");
  plan_gemms<T>(pl, gD, "schur_gemm");
}

// Row chunks of the rest of a step, each planned while the previous one runs.
// Every chunk ends in one batched recompression whose latency (QR, Jacobi SVD,
// basis update) does not shrink with the batch, so few chunks win: two measured
// best at N = 32768 (FP32 tiles), one to four chunks tried.
constexpr int kRestChunks = 2;

// The Schur-update launches of one row chunk of a step, as they are built.
template <typename T>
struct ChunkBuild {
  Context* cc = nullptr;  // stream and workspace of the chunk
  std::vector<GemmItem> gA, gB, gC, gD;
  std::vector<LrAddDesc<T>> adds;
  std::vector<CompressDesc<T>> comp;
  std::vector<int> comp_add;
};

// One host thread that runs a posted job while the caller does other work
// (an exception in the job is rethrown by wait()).
class HostWorker {
 public:
  HostWorker() : th_([this] { loop(); }) {}
  ~HostWorker() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    cv_.notify_all();
    th_.join();
  }
  void post(std::function<void()> f) {
    {
      std::lock_guard<std::mutex> lk(m_);
      job_ = std::move(f);
      has_ = true;
      done_ = false;
    }
    cv_.notify_all();
  }
  void wait() {
    std::unique_lock<std::mutex> lk(m_);
    cv_.wait(lk, [&] { return done_; });
    if (err_) {
      std::exception_ptr e = err_;
      err_ = nullptr;
      std::rethrow_exception(e);
    }
  }

 private:
  void loop() {
    for (;;) {
      std::function<void()> j;
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return has_ || stop_; });
        if (!has_) return;
        j = std::move(job_);
        has_ = false;
      }
      std::exception_ptr e;
      try {
        j();
      } catch (...) {
        e = std::current_exception();
      }
      {
        std::lock_guard<std::mutex> lk(m_);
        err_ = e;
        done_ = true;
      }
      cv_.notify_all();
    }
  }
  std::mutex m_;
  std::condition_variable cv_;
  std::function<void()> job_;
  bool has_ = false, done_ = true, stop_ = false;
  std::exception_ptr err_;
  std::thread th_;
};

template <typename T>
void blr_getrf(BlrMatrix<T>& A, Context& ctx, int max_steps) {
  const hipStream_t s = ctx.stream();
  ctx.set_default_peak(mfma_peak<T>());
  const int nb = A.nb_r();
  // a partial factorization stops after max_steps elimination steps (the
  // diagonal tile of step max_steps is factorized as well: its LU is queued
  // on the chain ahead of the step)
  const int nsteps = (max_steps < 0 || max_steps >= nb) ? nb : max_steps;
  if (A.nb_c() != nb) throw std::invalid_argument("getrf needs a square block grid");
  int32_t* info = ctx.info();
  const int npan_max = 8;
  if (nb * npan_max > 4096) throw std::invalid_argument("too many diagonal tiles for the info buffer");
  int64_t bm = 1;
  for (int i = 0; i < nb; ++i) {
    const int ii = A.idx(i, i);
    if (A.kind[ii] != kDense || A.rows(i) != A.cols(i))
      throw std::runtime_error("getrf: diagonal tile is not square dense");
    if (A.rows(i) > 512) throw std::runtime_error("getrf: tile size above 512");
    bm = std::max(bm, A.rows(i));
  }
  FRK_HIP_CHECK(hipMemsetAsync(info, 0, sizeof(int32_t) * 4096, s));
  ctx.sync();
  A.pull_ranks(s);
  const int32_t epoch0 = ctx.tile_epoch();
  // Three streams.  The lookahead stream carries the critical chain: factorization
  // of the diagonal tile k, the priority block of step k (the panels (k+1, k),
  // (k, k+1) and the Schur update of (k+1, k+1)), factorization of tile k+1, ...
  // The rest of step k runs beside the factorization of tile k+1, in two row
  // chunks on the main and the second stream.  The priority block of step k+1 is
  // queued behind the factorization of tile k+1 before the rest of step k has
  // run: its low-rank kernels read the panel ranks on the device, so its plan
  // does not need them.  The host waits only for the rank snapshot of a finished
  // step, to plan the next rest, and is never on the chain.
  Lookahead la(ctx, A.rank.size());
  HostWorker worker;
  Context side(la.s);
  side.set_stats_owner(&ctx);
  Context rest2(la.s2);
  rest2.set_stats_owner(&ctx);
  using clk = std::chrono::steady_clock;
  auto ms_since = [](clk::time_point t0) { return std::chrono::duration<double, std::milli>(clk::now() - t0).count(); };
  const size_t side_ws = diag_getrf_workspace<T>((int)bm) + prio_workspace(A, bm) + 8192;
  // factorization of the diagonal tile k on the lookahead stream (workspace taken
  // after whatever the current job took: the jobs of this stream run in order)
  auto enqueue_lu = [&](int k, const LrUpdItem* pre) {
    Plan pd(side);
    plan_diag_getrf<T>(pd, side, A, k, info + k * npan_max, pre);
    pd.run();
  };
  // priority block of step k on the lookahead stream, after the factorization of
  // tile k and (after_rest) the rest of step k-1, then the factorization of tile k+1
  // priority blocks planned on device ranks whose cost is logged once their
  // panel ranks are known on the host (at the next rank snapshot)
  std::vector<char> prio_dev(nb, 0);
  auto enqueue_prio = [&](int k, bool device_ranks, bool after_rest) {
    prio_dev[k] = device_ranks && A.kind[A.idx(k + 1, k)] == kLowRank && A.kind[A.idx(k, k + 1)] == kLowRank;
    const auto tp = clk::now();
    if (after_rest) FRK_HIP_CHECK(hipStreamWaitEvent(la.s, la.c1, 0));
    side.reserve(side_ws);
    Plan pp(side);
    LrUpdItem upd{};
    bool deferred = false;
    plan_prio<T>(pp, side, A, k, A.rank[A.idx(k + 1, k)], A.rank[A.idx(k, k + 1)], device_ranks,
                 tile_lu_path(A, k + 1) ? &upd : nullptr, &deferred);
    pp.run();
    // the rest of step k needs the solved panels, not the update of tile k+1
    FRK_HIP_CHECK(hipEventRecord(la.prio, la.s));
    enqueue_lu(k + 1, deferred ? &upd : nullptr);
    ctx.host_time("host_prio", ms_since(tp));
  };
  // workspace of one step (panel products, per-update temporaries, additions) with
  // ranks up to kmax
  auto max_rank = [&]() {
    int k = 1;
    for (size_t r = 0; r < A.rank.size(); ++r) k = std::max(k, A.rank[r]);
    return k;
  };
  auto step_need = [&](int i, int kmax) {
    const int rem = nb - 1 - i;
    size_t need = 4096;
    need += (size_t)rem * 2 * sizeof(T) * ((size_t)bm * kmax + 256);
    {
      size_t rp = 0, rq = 0;
      for (int t2 = i + 1; t2 < nb; ++t2) {
        if (A.kind[A.idx(t2, i)] == kLowRank) rp += A.rank[A.idx(t2, i)];
        if (A.kind[A.idx(i, t2)] == kLowRank) rq += A.rank[A.idx(i, t2)];
      }
      need += sizeof(T) * (rp * rq + 2 * (rp + rq) * bm) + 8 * 256;
    }
    for (int t2 = i + 1; t2 < nb; ++t2) {
      if (A.kind[A.idx(t2, i)] == kDense) need += sizeof(T) * ((size_t)bm * bm + 64);
      if (A.kind[A.idx(i, t2)] == kDense) need += sizeof(T) * ((size_t)bm * bm + 64);
    }
    for (int ic = i + 1; ic < nb; ++ic)
      for (int k = i + 1; k < nb; ++k) {
        const int kl = A.kind[A.idx(ic, i)], kb = A.kind[A.idx(i, k)], kc = A.kind[A.idx(ic, k)];
        if (kc == kDense)
          need += sizeof(T) * (2 * (size_t)bm * kmax + (size_t)kmax * kmax + 512);
        else if (kl == kDense && kb == kDense)
          need += sizeof(T) * ((size_t)bm * bm + (size_t)bm * A.kcap * 2 + (size_t)A.kcap * A.kcap + 1024);
        else
          need += sizeof(T) * ((size_t)bm * kmax + (size_t)kmax * kmax + 512);
        if (kc == kLowRank) {
          const size_t K = (size_t)kmax + ((kl == kDense && kb == kDense) ? (size_t)A.kcap : (size_t)kmax);
          need += sizeof(T) * (2 * bm * K + 2 * K + 6 * K * K) + 10 * 256;
        }
      }
    return need;
  };
  // Both rest workspaces are sized once, for the first step with room for the ranks
  // to grow: a larger reservation inside the loop frees the old buffer, and a free
  // synchronizes the whole device.
  {
    const size_t need0 = step_need(0, std::min(A.kcap, 2 * max_rank() + 4));
    ctx.reserve(need0);
    rest2.reserve(need0);
  }
  side.reserve(side_ws);
  // allocate (and zero, on the lookahead stream) the lookahead context's tile-LU
  // flags before the first factorization: every step copies their timeout marker
  (void)side.tile_flags();
  enqueue_lu(0, nullptr);
  if (nb > 1 && nsteps > 0) enqueue_prio(0, prio_rank_free(A, 0), false);
  bool prio_pending = false;  // the priority block of step i waits for the host ranks
  for (int i = 0; i + 1 < nb && i < nsteps; ++i) {
    auto th = clk::now();
    if (i > 0) {
      host_wait_event(la.rest, "getrf: rank snapshot of the previous step");
      // the tile LU's timeout marker, copied behind the snapshot: a factorization
      // of a diagonal tile whose block wait gave up ends the call here, before
      // any later step runs on its output
      if (la.snap[A.rank.size()] > 0)
        throw std::runtime_error("tile LU: a block wait timed out in step " + std::to_string(i - 1) +
                                 " (workgroups not co-resident?)");
      // every launch of the previous step's rest has run: its descriptors are consumed
      ctx.staging().reset();
      rest2.staging().reset();
      for (size_t t = 0; t < A.rank.size(); ++t) {
        A.rank[t] = la.snap[t];
        if (A.kind[t] == kLowRank && A.rank[t] < 0)
          throw std::runtime_error("low-rank capacity exceeded: rank " + std::to_string(-A.rank[t]) + " > kcap " +
                                   std::to_string(A.kcap));
      }
    }
    // the panel ranks of the priority block of step i are final (step i - 1 is done)
    if (prio_dev[i]) account_prio<T>(ctx, A, i, A.rank[A.idx(i + 1, i)], A.rank[A.idx(i, i + 1)]);
    ctx.host_time("host_sync", ms_since(th));
    th = clk::now();
    if (prio_pending) {
      enqueue_prio(i, false, true);
      prio_pending = false;
    }
    // the factors of tile i and the priority panels of step i come from the lookahead stream
    FRK_HIP_CHECK(hipStreamWaitEvent(s, la.prio, 0));
    const int m_i = (int)A.rows(i), n_i = (int)A.cols(i);
    const int ii = A.idx(i, i);
    const int rem = nb - 1 - i;
    // ---- plan the step ---------------------------------------------------
    const size_t need = step_need(i, max_rank());
    ctx.reserve(need);
    Plan pl(ctx);
    // 1. the panel solves (src/operations/LAPACK/getrf.cpp:50-57) not in the priority block,
    //    as products with the explicit triangular inverses.  Low-rank tiles go through
    //    stacked-row products (one launch per side): the rows of every V (column tiles)
    //    and every column of every U (row tiles); dense tiles through a product into
    //    workspace and a copy back.
    std::vector<GemmItem> tg, tcp;
    std::vector<RowRef> vrows, urows;
    // first row of each low-rank panel tile in the stacked rows; the largest rank
    // of the panel tiles (the fused write-back below takes ranks up to 64)
    std::vector<int> vofs(nb, -1), uofs(nb, -1);
    int panel_rmax = 0;
    for (int c = i + 1; c < nb; ++c)
      for (const int t : {A.idx(c, i), A.idx(i, c)})
        if (A.kind[t] == kLowRank) panel_rmax = std::max(panel_rmax, A.rank[t]);
    const bool fused_panels = panel_rmax <= 64;
    for (int ic = i + 2; ic < nb; ++ic) {
      const int t = A.idx(ic, i);
      if (A.kind[t] == kLowRank) {  // V <- V U^-1
        vofs[ic] = (int)vrows.size();
        for (int rho = 0; rho < A.rank[t]; ++rho) {
          T* row = A.V[t] + (size_t)rho * A.cols(i);
          vrows.push_back(RowRef{row, row});
        }
      } else if (A.kind[t] == kDense) {
        const int m = (int)A.rows(ic);
        T* w = ctx.take<T>((size_t)m * n_i + 1);
        tg.push_back(gemm_item(A.D[t], A.cols(i), 1, A.Uinv[ii], n_i, 1, w, n_i, 1, m, n_i, n_i, 1.0, 0.0));
        tcp.push_back(gemm_item(w, n_i, 1, nullptr, 0, 0, A.D[t], A.cols(i), 1, m, n_i, 0, 0.0, 0.0));
      }
    }
    for (int j = i + 2; j < nb; ++j) {
      const int t = A.idx(i, j);
      if (A.kind[t] == kLowRank) {  // U <- L^-1 U, as U^T <- U^T L^-T on the columns of U
        uofs[j] = (int)urows.size();
        for (int rho = 0; rho < A.rank[t]; ++rho) {
          T* col = A.U[t] + (size_t)rho * A.rows(i);
          urows.push_back(RowRef{col, col});
        }
      } else if (A.kind[t] == kDense) {
        const int n = (int)A.cols(j);
        T* w = ctx.take<T>((size_t)m_i * n + 1);
        tg.push_back(gemm_item(A.Linv[ii], m_i, 1, A.D[t], n, 1, w, n, 1, m_i, n, m_i, 1.0, 0.0));
        tcp.push_back(gemm_item(w, n, 1, nullptr, 0, 0, A.D[t], n, 1, m_i, n, 0, 0.0, 0.0));
      }
    }
    // both panels' stacked-row products in one launch, into workspace; the
    // write-back to the tiles comes with the stacked factors P and Q below
    T* Wv = nullptr;
    T* Wu = nullptr;
    if (fused_panels && (!vrows.empty() || !urows.empty())) {
      const int Rv = (int)vrows.size(), Ru = (int)urows.size();
      Wv = ctx.take<T>((size_t)Rv * n_i + 1);
      Wu = ctx.take<T>((size_t)Ru * m_i + 1);
      const size_t ov = pl.add(vrows), ou = pl.add(urows);
      const T* Ui = A.Uinv[ii];
      const T* Li = A.Linv[ii];
      T* wv = Wv;
      T* wu = Wu;
      pl.launch(
          [=](const char* d) {
            RowsList2 p{};
            p.job[0] = RowsListJob{reinterpret_cast<const RowRef*>(d + ov), Ui, n_i, 1, wv, Rv, n_i, n_i, 0};
            p.job[1] = RowsListJob{reinterpret_cast<const RowRef*>(d + ou), Li, 1, m_i, wu, Ru, m_i, m_i, 0};
            launch_rows_gemm2<T>(p, s);
          },
          "trsm_gemm", 2.0 * Rv * n_i * n_i + 2.0 * Ru * m_i * m_i,
          sizeof(T) * ((double)n_i * n_i + 2.0 * Rv * n_i + (double)m_i * m_i + 2.0 * Ru * m_i));
    }
    for (int side2 = 0; side2 < 2 && !fused_panels; ++side2) {
      const std::vector<RowRef>& rows = side2 == 0 ? vrows : urows;
      if (rows.empty()) continue;
      const int R = (int)rows.size();
      const int N = side2 == 0 ? n_i : m_i;
      T* W = ctx.take<T>((size_t)R * N + 1);
      const size_t off = pl.add(rows);
      const T* Bm = side2 == 0 ? A.Uinv[ii] : A.Linv[ii];
      const int64_t brs = side2 == 0 ? n_i : 1, bcs = side2 == 0 ? 1 : m_i;
      pl.launch(
          [=](const char* d) {
            const RowRef* rr = reinterpret_cast<const RowRef*>(d + off);
            launch_rows_gemm<T>(rr, R, Bm, brs, bcs, W, N, N, s);
            launch_rows_scatter<T>(rr, R, W, N, s);
          },
          "trsm_gemm", 2.0 * R * N * N, sizeof(T) * ((double)N * N + 2.0 * R * N));
    }
    plan_gemms<T>(pl, tg, "trsm_gemm");
    plan_copies<T>(pl, tcp);
    // 2. Schur updates A(ic,k) -= L(ic,i) A(i,k) (getrf.cpp:58-62, gemm.cpp:126-272) of every
    // target but (i+1, i+1).  P(ic) = S V of the low-rank tiles (ic, i) and Q(k) = U S of (i, k),
    // stacked so that all products P(ic) Q(k) of the step come from one GEMM (Sxa, row offsets
    // poff, column offsets qoff)
    std::vector<T*> P(nb, nullptr), Q(nb, nullptr);
    std::vector<int> poff(nb, -1), qoff(nb, -1);
    int RP = 0, RQ = 0;
    for (int ic = i + 1; ic < nb; ++ic)
      if (A.kind[A.idx(ic, i)] == kLowRank) {
        poff[ic] = RP;
        RP += A.rank[A.idx(ic, i)];
      }
    for (int k = i + 1; k < nb; ++k)
      if (A.kind[A.idx(i, k)] == kLowRank) {
        qoff[k] = RQ;
        RQ += A.rank[A.idx(i, k)];
      }
    T* Pst = ctx.take<T>((size_t)RP * n_i + 1);
    T* Qst = ctx.take<T>((size_t)RQ * m_i + 1);
    T* Sxa = ctx.take<T>((size_t)RP * RQ + 1);
    std::vector<GemmItem> gA, gB, gC, gD;
    if (fused_panels) {
      // one launch: the solved panel rows back into their tiles (those of the
      // priority block are in place already) and P (row-major, -S V) and Q
      // (column-major, U S) formed from them
      std::vector<PanelFin> fin;
      int fin_n = 0;
      double fin_fl = 0, fin_by = 0;
      for (int ic = i + 1; ic < nb; ++ic) {
        const int t = A.idx(ic, i);
        if (A.kind[t] != kLowRank || A.rank[t] <= 0) continue;
        P[ic] = Pst + (size_t)poff[ic] * n_i;
        const bool solved_here = ic >= i + 2;
        PanelFin f{};
        f.src = solved_here ? static_cast<const void*>(Wv + (size_t)vofs[ic] * n_i) : A.V[t];
        f.dst = solved_here ? A.V[t] : nullptr;
        f.s = A.S[t];
        f.pq = P[ic];
        f.src_ld = n_i;
        f.dst_ld = A.cols(i);
        f.s_rs = A.kcap;
        f.s_cs = 1;
        f.r = A.rank[t];
        f.n = n_i;
        f.alpha = -1.0;
        fin.push_back(f);
        fin_n = std::max(fin_n, n_i);
        fin_fl += 2.0 * f.r * f.r * n_i;
        fin_by += sizeof(T) * ((solved_here ? 3.0 : 2.0) * f.r * n_i + (double)f.r * f.r);
      }
      for (int k = i + 1; k < nb; ++k) {
        const int t = A.idx(i, k);
        if (A.kind[t] != kLowRank || A.rank[t] <= 0) continue;
        Q[k] = Qst + (size_t)qoff[k] * m_i;
        const bool solved_here = k >= i + 2;
        PanelFin f{};
        f.src = solved_here ? static_cast<const void*>(Wu + (size_t)uofs[k] * m_i) : A.U[t];
        f.dst = solved_here ? A.U[t] : nullptr;
        f.s = A.S[t];
        f.pq = Q[k];
        f.src_ld = m_i;
        f.dst_ld = A.rows(i);
        f.s_rs = 1;
        f.s_cs = A.kcap;
        f.r = A.rank[t];
        f.n = m_i;
        f.alpha = 1.0;
        fin.push_back(f);
        fin_n = std::max(fin_n, m_i);
        fin_fl += 2.0 * f.r * f.r * m_i;
        fin_by += sizeof(T) * ((solved_here ? 3.0 : 2.0) * f.r * m_i + (double)f.r * f.r);
      }
      if (!fin.empty()) {
        const size_t of = pl.add(fin);
        const int nf = (int)fin.size();
        pl.launch(
            [=](const char* d) { launch_panel_finish<T>(reinterpret_cast<const PanelFin*>(d + of), nf, fin_n, s); },
            "schur_gemm", fin_fl, fin_by);
      }
    } else {
      for (int ic = i + 1; ic < nb; ++ic) {
        const int t = A.idx(ic, i);
        if (A.kind[t] != kLowRank) continue;
        const int kl = A.rank[t];
        P[ic] = Pst + (size_t)poff[ic] * n_i;
        // P(ic) = -S V: the minus sign of the Schur update lives in P
        gA.push_back(gemm_item(A.S[t], A.kcap, 1, A.V[t], A.cols(i), 1, P[ic], n_i, 1, kl, n_i, kl, -1.0, 0.0));
      }
      for (int k = i + 1; k < nb; ++k) {
        const int t = A.idx(i, k);
        if (A.kind[t] != kLowRank) continue;
        const int kb = A.rank[t];
        Q[k] = Qst + (size_t)qoff[k] * m_i;
        gA.push_back(gemm_item(A.U[t], 1, A.rows(i), A.S[t], A.kcap, 1, Q[k], 1, m_i, m_i, kb, kb, 1.0, 0.0));
      }
      plan_gemms<T>(pl, gA, "schur_gemm");
      gA.clear();
    }
    if ((int64_t)RP * RQ >= (int64_t(1) << 20)) {
      // Sxa = Pst Qst as one tiled GEMM: with ranks near the tile size (small
      // tiles, tight eps) the stacked factors have thousands of rows, more than
      // the stacked-row product's 16-row slabs cover well
      plan_gemms<T>(pl, {gemm_item(Pst, n_i, 1, Qst, 1, m_i, Sxa, RQ, 1, RP, RQ, n_i, 1.0, 0.0)}, "schur_gemm");
    } else if (RP > 0 && RQ > 0) {
      // Sxa = Pst Qst: rows of Pst against the (column-major) stacked Q
      std::vector<RowRef> prow(RP);
      for (int r2 = 0; r2 < RP; ++r2) prow[r2] = RowRef{Pst + (size_t)r2 * n_i, nullptr};
      const size_t off = pl.add(prow);
      const int rp_ = RP, rq_ = RQ, kk = n_i;
      const int64_t qcs = m_i;
      T* sxa = Sxa;
      pl.launch([=](const char* d) {
        launch_rows_gemm<T>(reinterpret_cast<const RowRef*>(d + off), rp_, Qst, 1, qcs, sxa, rq_, kk, s);
      }, "schur_gemm", 2.0 * RP * RQ * n_i, sizeof(T) * ((double)RP * n_i + (double)RQ * m_i + (double)RP * RQ));
    }
    // the context (stream and workspace) and plan the next targets go to
    Context* cur = &ctx;
    Plan* curp = &pl;
    // Schur update of one target tile (ic, k): items into cb's gA..gD, additions into
    // cb.adds, temporaries from cb's context (builders of two chunks may run at once:
    // everything else they touch is read only)
    auto pair_body = [&](ChunkBuild<T>& cb, int ic, int k) {
      const int tl = A.idx(ic, i), tb = A.idx(i, k), tc = A.idx(ic, k);
      const int kl = A.kind[tl], kb = A.kind[tb], kc = A.kind[tc];
      const int m = (int)A.rows(ic), n = (int)A.cols(k);
      if (kl == kEmpty || kb == kEmpty || kc == kEmpty) return;
      if (kc == kDense) {
        T* C = A.D[tc];
        if (kl == kLowRank && kb == kLowRank) {
          const int rl = A.rank[tl], rb = A.rank[tb];
          T* Sx = Sxa + (size_t)poff[ic] * RQ + qoff[k];
          T* z = cb.cc->template take<T>((size_t)m * rb + 1);
          cb.gC.push_back(gemm_item(A.U[tl], 1, A.rows(ic), Sx, RQ, 1, z, 1, m, m, rb, rl, 1.0, 0.0));
          cb.gD.push_back(gemm_item(z, 1, m, A.V[tb], A.cols(k), 1, C, n, 1, m, n, rb, 1.0, 1.0));
        } else if (kl == kLowRank && kb == kDense) {
          const int rl = A.rank[tl];
          T* w = cb.cc->template take<T>((size_t)rl * n + 1);
          T* w2 = cb.cc->template take<T>((size_t)rl * n + 1);
          cb.gA.push_back(gemm_item(A.V[tl], A.cols(i), 1, A.D[tb], n, 1, w, n, 1, rl, n, n_i, -1.0, 0.0));
          cb.gB.push_back(gemm_item(A.S[tl], A.kcap, 1, w, n, 1, w2, n, 1, rl, n, rl, 1.0, 0.0));
          cb.gC.push_back(gemm_item(A.U[tl], 1, A.rows(ic), w2, n, 1, C, n, 1, m, n, rl, 1.0, 1.0));
        } else if (kl == kDense && kb == kLowRank) {
          const int rb = A.rank[tb];
          T* a = cb.cc->template take<T>((size_t)m * rb + 1);
          T* a2 = cb.cc->template take<T>((size_t)m * rb + 1);
          cb.gA.push_back(gemm_item(A.D[tl], n_i, 1, A.U[tb], 1, A.rows(i), a, 1, m, m, rb, n_i, -1.0, 0.0));
          cb.gB.push_back(gemm_item(a, 1, m, A.S[tb], A.kcap, 1, a2, 1, m, m, rb, rb, 1.0, 0.0));
          cb.gC.push_back(gemm_item(a2, 1, m, A.V[tb], A.cols(k), 1, C, n, 1, m, n, rb, 1.0, 1.0));
        } else {
          cb.gA.push_back(gemm_item(A.D[tl], n_i, 1, A.D[tb], n, 1, C, n, 1, m, n, n_i, -1.0, 1.0));
        }
        return;
      }
      LrAddDesc<T> ad{};
      ad.cu = A.U[tc];
      ad.cs = A.S[tc];
      ad.cv = A.V[tc];
      ad.d_rank = A.d_rank(ic, k);
      ad.kc = A.rank[tc];
      ad.cu_rs = 1;
      ad.cu_cs = m;
      ad.cv_rs = n;
      ad.cv_cs = 1;
      ad.lds = A.kcap;
      ad.m = m;
      ad.n = n;
      ad.eps = A.eps;
      ad.cap = A.kcap;
      ad.cap_trunc = A.cap_trunc;
      if (kl == kLowRank && kb == kLowRank) {  // gemm.cpp:253-272
        const int rl = A.rank[tl], rb = A.rank[tb];
        ad.xu = A.U[tl];
        ad.xu_rs = 1;
        ad.xu_cs = A.rows(ic);
        ad.kx1 = rl;
        ad.xs = Sxa + (size_t)poff[ic] * RQ + qoff[k];  // P(ic) Q(k) (P carries the minus sign)
        ad.xs_ld = RQ;
        ad.kx2 = rb;
        ad.xv = A.V[tb];
        ad.xv_rs = A.cols(k);
        ad.xv_cs = 1;
      } else if (kl == kLowRank && kb == kDense) {  // gemm.cpp:219-234
        const int rl = A.rank[tl];
        T* w = cb.cc->template take<T>((size_t)rl * n + 1);
        cb.gA.push_back(gemm_item(A.V[tl], A.cols(i), 1, A.D[tb], n, 1, w, n, 1, rl, n, n_i, -1.0, 0.0));
        ad.xu = A.U[tl];
        ad.xu_rs = 1;
        ad.xu_cs = A.rows(ic);
        ad.kx1 = rl;
        ad.xs = A.S[tl];
        ad.xs_ld = A.kcap;
        ad.kx2 = rl;
        ad.xv = w;
        ad.xv_rs = n;
        ad.xv_cs = 1;
      } else if (kl == kDense && kb == kLowRank) {  // gemm.cpp:236-251
        const int rb = A.rank[tb];
        T* a = cb.cc->template take<T>((size_t)m * rb + 1);
        cb.gA.push_back(gemm_item(A.D[tl], n_i, 1, A.U[tb], 1, A.rows(i), a, 1, m, m, rb, n_i, -1.0, 0.0));
        ad.xu = a;
        ad.xu_rs = 1;
        ad.xu_cs = m;
        ad.kx1 = rb;
        ad.xs = A.S[tb];
        ad.xs_ld = A.kcap;
        ad.kx2 = rb;
        ad.xv = A.V[tb];
        ad.xv_rs = A.cols(k);
        ad.xv_cs = 1;
      } else {  // D.D -> LR: compress the dense product first (gemm.cpp:202-217)
        T* prod = cb.cc->template take<T>((size_t)m * n + 1);
        cb.gA.push_back(gemm_item(A.D[tl], n_i, 1, A.D[tb], n, 1, prod, n, 1, m, n, n_i, -1.0, 0.0));
        T* xu = cb.cc->template take<T>((size_t)m * A.kcap + 1);
        T* xs = cb.cc->template take<T>((size_t)A.kcap * A.kcap + 1);
        T* xv = cb.cc->template take<T>((size_t)A.kcap * n + 1);
        CompressDesc<T> c{};
        c.a = prod;
        c.m = m;
        c.n = n;
        c.lda = n;
        c.eps = A.eps;
        c.rank = A.rank[tc];
        c.u = xu;
        c.s = xs;
        c.v = xv;
        c.u_rs = 1;
        c.u_cs = m;
        c.v_rs = n;
        c.v_cs = 1;
        c.lds = A.kcap;
        c.cap = A.kcap;
        c.cap_trunc = A.cap_trunc;
        cb.comp.push_back(c);
        cb.comp_add.push_back((int)cb.adds.size());
        ad.xu = xu;
        ad.xu_rs = 1;
        ad.xu_cs = m;
        ad.xs = xs;
        ad.xs_ld = A.kcap;
        ad.xv = xv;
        ad.xv_rs = n;
        ad.xv_cs = 1;
        ad.kx1 = ad.kx2 = -1;  // known after compression
      }
      cb.adds.push_back(ad);
    };
    // naive additions drain the stream and reset the workspace of the context they run on:
    // give them their own, so Pst, Qst and Sxa survive for the later chunks
    auto naive_adds = [&](ChunkBuild<T>& cb) {
      if (cb.adds.empty()) return;
      Context cn(cb.cc->stream());
      cn.set_stats_owner(&ctx);
      lr_add_naive<T>(cb.adds, cn);
    };
    // the launches of a chunk's Schur updates into its plan (no D.D -> LR targets)
    auto plan_chunk = [&](ChunkBuild<T>& cb, Plan& p) {
      plan_gemms<T>(p, cb.gA, "schur_gemm");
      plan_gemms<T>(p, cb.gB, "schur_gemm");
      plan_gemms<T>(p, cb.gC, "schur_gemm");
      plan_gemms<T>(p, cb.gD, "schur_gemm");
      plan_lr_add<T>(p, cb.adds, A.lra);
    };
    // plan and issue the Schur updates of a list of targets (the plan may already hold launches)
    auto issue_targets = [&](const std::vector<std::pair<int, int>>& targets, const char* plan_cls) {
      ChunkBuild<T> cb;
      cb.cc = cur;
      {
        const auto tp0 = clk::now();
        for (const auto& t : targets) pair_body(cb, t.first, t.second);
        ctx.host_time("host_pairs", ms_since(tp0));
      }
      plan_gemms<T>(*curp, cb.gA, "schur_gemm");
      plan_gemms<T>(*curp, cb.gB, "schur_gemm");
      plan_gemms<T>(*curp, cb.gC, "schur_gemm");
      plan_gemms<T>(*curp, cb.gD, "schur_gemm");
      if (!cb.comp.empty()) {
        // D.D -> LR needs the compressed ranks on the host: run what is planned,
        // compress on a side context (own workspace), then plan the additions.
        curp->run();
        DevBuf crank(sizeof(int32_t) * cb.comp.size());
        for (size_t c2 = 0; c2 < cb.comp.size(); ++c2) cb.comp[c2].d_rank = crank.as<int32_t>() + c2;
        Context cside(cur->stream());
        compress_batched<T>(cb.comp, cside);
        std::vector<int32_t> cr(cb.comp.size());
        FRK_HIP_CHECK(hipMemcpyAsync(cr.data(), crank.as<int32_t>(), sizeof(int32_t) * cr.size(),
                                     hipMemcpyDeviceToHost, cur->stream()));
        host_wait_stream(cur->stream(), __func__);
        for (size_t c2 = 0; c2 < cb.comp.size(); ++c2)
          cb.adds[cb.comp_add[c2]].kx1 = cb.adds[cb.comp_add[c2]].kx2 = cr[c2];
        if (A.lra == kLraNaive) {
          naive_adds(cb);
        } else {
          Plan p2(*cur);
          plan_lr_add<T>(p2, cb.adds, A.lra);
          p2.run();
        }
        return;
      }
      if (A.lra == kLraNaive) {
        curp->run();
        naive_adds(cb);
      } else {
        const auto tl0 = clk::now();
        plan_lr_add<T>(*curp, cb.adds, A.lra);
        ctx.host_time("host_lradd", ms_since(tl0));
        ctx.host_time(plan_cls, ms_since(th));
        th = clk::now();
        curp->run();
        ctx.host_time("host_issue", ms_since(th));
      }
    };
    // 3. every target but (i+1, i+1), in two row chunks that run concurrently: the first
    //    on the main stream, the second on the third stream with its own workspace, once
    //    the panels and stacked products above are done.  Each chunk ends in a batched
    //    recompression whose latency does not shrink with the batch, so the two overlap.
    const int nchunk = std::min(kRestChunks, rem);
    std::vector<std::vector<std::pair<int, int>>> targets(nchunk);
    bool dd_lr = false;  // a D.D -> LR target needs a host synchronization inside its chunk
    for (int c = 0; c < nchunk; ++c) {
      const int ic0 = i + 1 + (rem * c) / nchunk, ic1 = i + 1 + (rem * (c + 1)) / nchunk;
      for (int ic = ic0; ic < ic1; ++ic)
        for (int k = i + 1; k < nb; ++k) {
          if (ic == i + 1 && k == i + 1) continue;
          targets[c].emplace_back(ic, k);
          dd_lr = dd_lr || (A.kind[A.idx(ic, i)] == kDense && A.kind[A.idx(i, k)] == kDense &&
                            A.kind[A.idx(ic, k)] == kLowRank);
        }
    }
    if (nchunk == 2 && !dd_lr && A.lra != kLraNaive) {
      // the second chunk's descriptors are built on the worker thread while this
      // thread issues the panels and the stacked products and builds the first chunk
      rest2.reserve(need);
      ChunkBuild<T> cb1;
      cb1.cc = &rest2;
      Plan p2(rest2);
      worker.post([&]() {
        for (const auto& t : targets[1]) pair_body(cb1, t.first, t.second);
        plan_chunk(cb1, p2);
      });
      // the job uses this frame's cb1, p2 and targets: never leave it (an exception
      // included) while the job still runs
      struct JoinOnExit {
        HostWorker& w;
        bool armed = true;
        ~JoinOnExit() {
          if (armed) try {
              w.wait();
            } catch (...) {
            }
        }
      } join_on_exit{worker};
      pl.run();
      FRK_HIP_CHECK(hipEventRecord(la.sxa, s));
      ChunkBuild<T> cb0;
      cb0.cc = &ctx;
      for (const auto& t : targets[0]) pair_body(cb0, t.first, t.second);
      plan_chunk(cb0, pl);
      ctx.host_time("host_plan", ms_since(th));
      th = clk::now();
      pl.run();
      join_on_exit.armed = false;
      worker.wait();
      FRK_HIP_CHECK(hipStreamWaitEvent(la.s2, la.sxa, 0));
      p2.run();
      ctx.host_time("host_issue", ms_since(th));
    } else {
      pl.run();
      FRK_HIP_CHECK(hipEventRecord(la.sxa, s));
      std::unique_ptr<Plan> pl2;
      for (int c = 0; c < nchunk; ++c) {
        if (c > 0) th = clk::now();
        if (c == 1) {
          FRK_HIP_CHECK(hipStreamWaitEvent(la.s2, la.sxa, 0));
          rest2.reserve(need);
          pl2 = std::make_unique<Plan>(rest2);
          cur = &rest2;
          curp = pl2.get();
        }
        issue_targets(targets[c], c == 0 ? "host_plan" : "host_plan_rest");
        curp->run();
      }
    }
    // both chunks done: the rank snapshot (second stream), and the main stream waits for
    // the second chunk before the next priority block
    FRK_HIP_CHECK(hipEventRecord(la.c0, s));
    FRK_HIP_CHECK(hipStreamWaitEvent(la.s2, la.c0, 0));
    FRK_HIP_CHECK(hipEventRecord(la.c1, la.s2));
    FRK_HIP_CHECK(hipStreamWaitEvent(s, la.c1, 0));
    // the ranks, then the tile LU's timeout marker (the rest of step i ran after the
    // priority block of step i, hence after the factorization of tile i: the marker
    // is final by now), written into the pinned snapshot by one kernel
    launch_snapshot(la.snap_dev, A.d_rank(0, 0), (int)A.rank.size(), side.tile_flags() + Context::kTileFlags - 1,
                    la.s2);
    FRK_HIP_CHECK(hipEventRecord(la.rest, la.s2));
    cur = &ctx;
    curp = &pl;
    // 4. the next priority block
    if (i + 2 < nb && i + 1 < nsteps) {
      if (prio_rank_free(A, i + 1))
        enqueue_prio(i + 1, true, true);
      else
        prio_pending = true;
    }
  }
  ctx.sync();
  side.sync();
  ctx.check_tile_lu(epoch0);
  side.check_tile_lu(0);
  A.pull_ranks(s);
  std::vector<int32_t> hinfo(nb * npan_max);
  FRK_HIP_CHECK(hipMemcpy(hinfo.data(), info, sizeof(int32_t) * hinfo.size(), hipMemcpyDeviceToHost));
  for (int i = 0; i < nb; ++i)
    for (int p = 0; p < npan_max; ++p)
      if (hinfo[i * npan_max + p] != 0)
        throw std::runtime_error("getrf: zero pivot in diagonal tile " + std::to_string(i) + " at local row " +
                                 std::to_string(p * 64 + hinfo[i * npan_max + p]));
}

// ============================================================================
// Block substitution (column-oriented: every batched update has a distinct
// target block).  One plan (one descriptor upload) per sweep.
// ============================================================================
// Fast path: per block step one launch for the diagonal inverse and one for
// every update of the remaining blocks (one workgroup per target block); the
// sweep writes its result out of place, so no per-step copy is needed.
template <typename T>
static bool solve_fast_ok(const BlrMatrix<T>& F, int nrhs) {
  if (nrhs < 1 || nrhs > 16) return false;
  for (int i = 0; i < F.nb_r(); ++i)
    if (F.rows(i) > 512 || F.cols(i) > 512 || F.rows(i) != F.cols(i)) return false;
  for (size_t t = 0; t < F.rank.size(); ++t)
    if (F.kind[t] == kLowRank && F.rank[t] > 64) return false;
  return true;
}

template <typename T>
static void blr_solve_fast(BlrMatrix<T>& F, T* b, int nrhs, bool trans, Context& ctx, int part) {
  const int nb = F.nb_r();
  const hipStream_t s = ctx.stream();
  const int64_t N = F.roff[nb];
  ctx.reserve(sizeof(T) * ((size_t)N * nrhs + 64) + 4096);
  T* y = ctx.take<T>((size_t)N * nrhs + 1);
  Plan pl(ctx);
  // one sweep: block p of the sweep order (p = 0 .. nb-1) is solved by the fused step
  // launch of position p-1, which also applies the updates from block p-1 to the
  // blocks at positions p+1 and beyond; the update into block p itself from block
  // p-1 is applied inside the diagonal workgroups of block p
  auto sweep = [&](bool forward, bool lower_factor, bool transposed, T* src, T* dst) {
    auto blk_of = [&](int p) { return forward ? p : nb - 1 - p; };
    // the update of target block k from the solved block i
    auto upd_item = [&](int k, int i, SolveUpdItem& u, double& fl, double& by) -> bool {
      const int t = transposed ? F.idx(i, k) : F.idx(k, i);
      if (F.kind[t] == kEmpty) return false;
      const int ni = (int)F.rows(i);
      u = SolveUpdItem{};
      u.x = dst + F.roff[i] * nrhs;
      u.b = src + F.roff[k] * nrhs;
      u.m = (int)F.rows(k);
      u.n = ni;
      if (F.kind[t] == kDense) {
        u.D = F.D[t];
        if (!transposed) {
          u.d_rs = F.cols(i);
          u.d_cs = 1;
        } else {
          u.d_rs = 1;
          u.d_cs = F.cols(k);
        }
        fl += 2.0 * u.m * ni * nrhs;
        by += sizeof(T) * (double)u.m * ni;
      } else {
        const int r = F.rank[t];
        u.r = r;
        if (!transposed) {  // b_k -= U (S (V x)), tile (k, i)
          u.U = F.U[t];
          u.u_rs = 1;
          u.u_cs = F.rows(k);
          u.S = F.S[t];
          u.s_rs = F.kcap;
          u.s_cs = 1;
          u.V = F.V[t];
          u.v_rs = F.cols(i);
          u.v_cs = 1;
        } else {  // b_k -= V^T (S^T (U^T x)), tile (i, k)
          u.U = F.V[t];
          u.u_rs = 1;
          u.u_cs = F.cols(k);
          u.S = F.S[t];
          u.s_rs = 1;
          u.s_cs = F.kcap;
          u.V = F.U[t];
          u.v_rs = F.rows(i);
          u.v_cs = 1;
        }
        fl += 2.0 * r * ((double)ni + r + u.m) * nrhs;
        by += sizeof(T) * (double)r * (ni + u.m);
      }
      return true;
    };
    for (int p = -1; p + 1 < nb; ++p) {
      // launch of position p: bulk updates from block p (p >= 0), then block p+1
      const int kn = blk_of(p + 1);
      std::vector<SolveUpdItem> items;
      double fl = 0, by = 0;
      if (p >= 0) {
        const int i = blk_of(p);
        SolveUpdItem u;
        for (int q = p + 2; q < nb; ++q)
          if (upd_item(blk_of(q), i, u, fl, by)) items.push_back(u);
      }
      const int n_upd = (int)items.size();
      StepDiag sd{};
      if (p >= 0) {
        SolveUpdItem u;
        if (upd_item(kn, blk_of(p), u, fl, by)) {
          items.push_back(u);
          sd.has_last = 1;
        }
      }
      const int kk = F.idx(kn, kn);
      sd.tinv = lower_factor ? F.Linv[kk] : F.Uinv[kk];
      sd.ld = F.cols(kn);
      sd.b = src + F.roff[kn] * nrhs;
      sd.x = dst + F.roff[kn] * nrhs;
      sd.n = (int)F.rows(kn);
      sd.trans = transposed ? 1 : 0;
      fl += 2.0 * sd.n * sd.n * nrhs;
      by += sizeof(T) * ((double)sd.n * sd.n + 2.0 * sd.n * nrhs);
      const size_t off = items.empty() ? 0 : pl.add(items);
      const bool any = !items.empty();
      pl.launch(
          [=](const char* d) {
            launch_solve_step<T>(any ? reinterpret_cast<const SolveUpdItem*>(d + off) : nullptr, n_upd, sd, nrhs, s);
          },
          "solve_step", fl, by);
    }
  };
  const bool do_l = part != 2, do_u = part != 1;
  const bool first_lower = !trans;  // forward sweep: L (plain) or U^T (adjoint)
  const bool do_first = first_lower ? do_l : do_u, do_second = first_lower ? do_u : do_l;
  T* cur = b;
  T* other = y;
  if (do_first) {
    sweep(true, first_lower, trans, cur, other);
    std::swap(cur, other);
  }
  if (do_second) {
    sweep(false, !first_lower, trans, cur, other);
    std::swap(cur, other);
  }
  if (cur != b) plan_copies<T>(pl, {gemm_item(cur, nrhs, 1, nullptr, 0, 0, b, nrhs, 1, (int)N, nrhs, 0, 0.0, 0.0)});
  pl.run();
}

// The dataflow block substitution (one persistent launch per sweep) applies
// when every off-diagonal tile is low rank (rank <= 64) or empty and the
// diagonal tiles are square, at most 512; FRK_SOLVE_DF=0 selects the
// launch-per-step path instead.
template <typename T>
static bool solve_df_ok(const BlrMatrix<T>& F, int nrhs) {
  const char* e = std::getenv("FRK_SOLVE_DF");
  if ((e && e[0] == '0') || nrhs < 1 || nrhs > 16) return false;
  const int nb = F.nb_r();
  for (int i = 0; i < nb; ++i)
    if (F.rows(i) > 512 || F.rows(i) != F.cols(i)) return false;
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < nb; ++j) {
      const int t = F.idx(i, j);
      if (i == j) continue;
      if (F.kind[t] == kDense) return false;
      if (F.kind[t] == kLowRank && F.rank[t] > 64) return false;
    }
  return true;
}

template <typename T>
static void blr_solve_df(BlrMatrix<T>& F, T* b, int nrhs, bool trans, Context& ctx, int part) {
  const int nb = F.nb_r();
  const hipStream_t s = ctx.stream();
  const int64_t N = F.roff[nb];
  const int NR = nrhs > 8 ? 16 : 8;
  const int rpt = kDfRows;
  // sizes of one sweep's device scratch
  int ntiles_max = 0, rsum_max = 0;
  {
    int nt = 0, rs = 0;
    for (size_t t = 0; t < F.rank.size(); ++t)
      if (F.kind[t] == kLowRank && F.rank[t] > 0) {
        ++nt;
        rs += F.rank[t];
      }
    ntiles_max = nt;
    rsum_max = rs;
  }
  const size_t nflags = 2 + (size_t)(kDfXflags + kDfAflags) * nb + ntiles_max;
  ctx.reserve(2 * sizeof(T) * ((size_t)N * nrhs + 64) + 2 * sizeof(T) * ((size_t)rsum_max * NR + 64) +
              2 * sizeof(int32_t) * (nflags + 64) + 8192);
  T* y = ctx.take<T>((size_t)N * nrhs + 1);
  T* bt = ctx.take<T>((size_t)N * nrhs + 1);
  Plan pl(ctx);
  std::vector<int32_t*> df_err;
  auto sweep = [&](bool forward, bool lower_factor, bool transposed, T* src, T* dst) {
    auto blk_of = [&](int p) { return forward ? p : nb - 1 - p; };
    std::vector<DfBlock> blocks(nb);
    std::vector<DfTile> tiles;
    std::vector<std::vector<int>> out_of(nb);  // tile indices by source position
    int yoff = 0;
    for (int p = 0; p < nb; ++p) {
      const int kb = blk_of(p);
      DfBlock& B = blocks[p];
      B = DfBlock{};
      const int kk = F.idx(kb, kb);
      B.tinv = lower_factor ? F.Linv[kk] : F.Uinv[kk];
      B.ld = F.cols(kb);
      B.b = src + F.roff[kb] * nrhs;
      B.x = dst + F.roff[kb] * nrhs;
      B.bt = bt + F.roff[kb] * nrhs;
      B.n = (int)F.rows(kb);
      B.trans = transposed ? 1 : 0;
      B.nd = (B.n + rpt - 1) / rpt;
      B.na = (B.n + kDfArows - 1) / kDfArows;
      B.has_bt = p >= 2 ? 1 : 0;
      B.tile0 = (int)tiles.size();
      B.last = -1;
      B.near = -1;
      for (int q = 0; q < p; ++q) {
        const int ib = blk_of(q);
        const int t = transposed ? F.idx(ib, kb) : F.idx(kb, ib);
        if (F.kind[t] != kLowRank || F.rank[t] <= 0) continue;
        DfTile d{};
        d.r = F.rank[t];
        d.p = p;
        d.q = q;
        d.yoff = yoff;
        yoff += d.r * NR;
        if (!transposed) {  // b_k -= U (S (V x_i)), tile (k, i)
          d.U = F.U[t];
          d.u_rs = 1;
          d.u_cs = F.rows(kb);
          d.S = F.S[t];
          d.s_rs = F.kcap;
          d.s_cs = 1;
          d.V = F.V[t];
          d.v_rs = F.cols(ib);
          d.v_cs = 1;
        } else {  // b_k -= V^T (S^T (U^T x_i)), tile (i, k)
          d.U = F.V[t];
          d.u_rs = 1;
          d.u_cs = F.cols(kb);
          d.S = F.S[t];
          d.s_rs = 1;
          d.s_cs = F.kcap;
          d.V = F.U[t];
          d.v_rs = F.rows(ib);
          d.v_cs = 1;
        }
        if (q == p - 1) B.last = (int)tiles.size();
        if (q == p - 2) B.near = (int)tiles.size();
        out_of[q].push_back((int)tiles.size());
        tiles.push_back(d);
      }
      B.ntile = (int)tiles.size() - B.tile0;
    }
    std::vector<int32_t> by_src;
    for (int q = 0; q < nb; ++q) {
      blocks[q].src0 = (int)by_src.size();
      blocks[q].nsrc = (int)out_of[q].size();
      by_src.insert(by_src.end(), out_of[q].begin(), out_of[q].end());
    }
    if (by_src.empty()) by_src.push_back(0);
    // tasks in dependency order: D(p, *) (after D(p - 1, *)), Y(p, chunks), then A(p + 2, *)
    std::vector<int4> tasks;
    for (int p = 0; p < nb; ++p) {
      for (int g = 0; g < blocks[p].nd; ++g) tasks.push_back(make_int4(0, p, g, 0));
      // y for the updates three or more blocks ahead (the next block's D tasks
      // and the A tasks of the block after it apply theirs from x_p directly)
      const int ns = blocks[p].nsrc;
      int first = 0;
      while (first < ns && tiles[by_src[blocks[p].src0 + first]].p <= p + 2) ++first;
      for (int c = first; c < ns; c += kDfYtiles) tasks.push_back(make_int4(1, p, c, std::min(ns, c + kDfYtiles)));
      if (p + 2 < nb)
        for (int h = 0; h < blocks[p + 2].na; ++h) tasks.push_back(make_int4(2, p + 2, h, 0));
    }
    DfSweep sw{};
    const size_t ob = pl.add(blocks);
    const size_t ot = tiles.empty() ? ob : pl.add(tiles);
    const size_t os = pl.add(by_src);
    const size_t ok = pl.add(tasks);
    sw.y = ctx.take<T>((size_t)std::max(yoff, 1));
    sw.flags = ctx.take<int32_t>(nflags + 1);
    sw.nb = nb;
    sw.ntiles = (int)tiles.size();
    sw.ntasks = (int)tasks.size();
    sw.nrhs = nrhs;
    sw.spin_limit = int64_t(1) << 21;
    const int nfl = (int)(2 + (kDfXflags + kDfAflags) * nb + tiles.size());
    double fl = 0, by = 0;
    for (const auto& B : blocks) {
      fl += 2.0 * B.n * B.n * nrhs;
      by += sizeof(T) * ((double)B.n * B.n + 2.0 * B.n * nrhs);
    }
    for (const auto& d : tiles) {
      const double nq = F.rows(blk_of(d.q)), np = F.rows(blk_of(d.p));
      fl += 2.0 * d.r * (nq + d.r + np) * nrhs;
      by += sizeof(T) * (double)d.r * (nq + np);
    }
    const bool has_t = !tiles.empty();
    pl.launch(
        [=](const char* dd) mutable {
          sw.blocks = dd + ob;
          sw.tiles = has_t ? dd + ot : nullptr;
          sw.by_src = reinterpret_cast<const int32_t*>(dd + os);
          sw.tasks = dd + ok;
          launch_solve_df<T>(sw, nfl, nrhs, 256, s);
        },
        "solve_df", fl, by);
    df_err.push_back(sw.flags + 1);
  };
  const bool do_l = part != 2, do_u = part != 1;
  const bool first_lower = !trans;
  const bool do_first = first_lower ? do_l : do_u, do_second = first_lower ? do_u : do_l;
  T* cur = b;
  T* other = y;
  if (do_first) {
    sweep(true, first_lower, trans, cur, other);
    std::swap(cur, other);
  }
  if (do_second) {
    sweep(false, !first_lower, trans, cur, other);
    std::swap(cur, other);
  }
  if (cur != b) plan_copies<T>(pl, {gemm_item(cur, nrhs, 1, nullptr, 0, 0, b, nrhs, 1, (int)N, nrhs, 0, 0.0, 0.0)});
  pl.run();
  // a wait past its spin limit marks the sweep failed: report it, never return its x
  for (int32_t* e : df_err) {
    int32_t h = 0;
    FRK_HIP_CHECK(hipMemcpyAsync(&h, e, sizeof(h), hipMemcpyDeviceToHost, s));
    host_wait_stream(s, "dataflow solve");
    if (h != 0) throw std::runtime_error("dataflow solve: a wait timed out");
  }
}

template <typename T>
void blr_solve(BlrMatrix<T>& F, T* b, int nrhs, bool trans, Context& ctx, int part) {
  const int nb = F.nb_r();
  ctx.set_default_peak(mfma_peak<T>());
  ctx.sync();
  F.pull_ranks(ctx.stream());
  if (solve_df_ok(F, nrhs)) {
    blr_solve_df(F, b, nrhs, trans, ctx, part);
    return;
  }
  if (solve_fast_ok(F, nrhs)) {
    blr_solve_fast(F, b, nrhs, trans, ctx, part);
    return;
  }
  int kmax = 1;
  for (int r : F.rank) kmax = std::max(kmax, r);
  ctx.reserve((size_t)nb * nb * 2 * sizeof(T) * ((size_t)kmax * nrhs + 64) +
              (size_t)2 * nb * sizeof(T) * ((size_t)F.rows(0) * nrhs + 64) + 4096);
  auto bptr = [&](int i) { return b + F.roff[i] * nrhs; };
  auto sweep = [&](Plan& pl, bool forward, bool lower_factor, bool transposed) {
    for (int step = 0; step < nb; ++step) {
      const int i = forward ? step : nb - 1 - step;
      const int ii = F.idx(i, i);
      const int64_t ld = F.cols(i);
      const T* tinv = lower_factor ? F.Linv[ii] : F.Uinv[ii];
      const int ni_ = (int)F.rows(i);
      T* w = ctx.take<T>((size_t)ni_ * nrhs + 1);
      // b_i <- op(T)^-1 b_i with the explicit inverse (op = transpose for the adjoint sweeps)
      if (!transposed)
        plan_gemms<T>(pl, {gemm_item(tinv, ld, 1, bptr(i), nrhs, 1, w, nrhs, 1, ni_, nrhs, ni_, 1.0, 0.0)},
                      "solve_gemm");
      else
        plan_gemms<T>(pl, {gemm_item(tinv, 1, ld, bptr(i), nrhs, 1, w, nrhs, 1, ni_, nrhs, ni_, 1.0, 0.0)},
                      "solve_gemm");
      plan_copies<T>(pl, {gemm_item(w, nrhs, 1, nullptr, 0, 0, bptr(i), nrhs, 1, ni_, nrhs, 0, 0.0, 0.0)});
      std::vector<GemmItem> g1, g2, g3;
      for (int k = 0; k < nb; ++k) {
        if (forward ? (k <= i) : (k >= i)) continue;
        const int t = transposed ? F.idx(i, k) : F.idx(k, i);
        const int mk = (int)F.rows(k), ni = (int)F.rows(i);
        if (F.kind[t] == kEmpty) continue;
        if (F.kind[t] == kDense) {
          const int64_t tld = transposed ? F.cols(k) : F.cols(i);
          if (!transposed)
            g1.push_back(gemm_item(F.D[t], tld, 1, bptr(i), nrhs, 1, bptr(k), nrhs, 1, mk, nrhs, ni, -1.0, 1.0));
          else
            g1.push_back(gemm_item(F.D[t], 1, tld, bptr(i), nrhs, 1, bptr(k), nrhs, 1, mk, nrhs, ni, -1.0, 1.0));
          continue;
        }
        const int r = F.rank[t];
        T* t1 = ctx.take<T>((size_t)r * nrhs + 1);
        T* t2 = ctx.take<T>((size_t)r * nrhs + 1);
        if (!transposed) {
          // b_k -= U (S (V b_i)), tile (k, i)
          g1.push_back(gemm_item(F.V[t], F.cols(i), 1, bptr(i), nrhs, 1, t1, nrhs, 1, r, nrhs, ni, 1.0, 0.0));
          g2.push_back(gemm_item(F.S[t], F.kcap, 1, t1, nrhs, 1, t2, nrhs, 1, r, nrhs, r, 1.0, 0.0));
          g3.push_back(gemm_item(F.U[t], 1, F.rows(k), t2, nrhs, 1, bptr(k), nrhs, 1, mk, nrhs, r, -1.0, 1.0));
        } else {
          // b_k -= V^T (S^T (U^T b_i)), tile (i, k)
          g1.push_back(gemm_item(F.U[t], F.rows(i), 1, bptr(i), nrhs, 1, t1, nrhs, 1, r, nrhs, ni, 1.0, 0.0));
          g2.push_back(gemm_item(F.S[t], 1, F.kcap, t1, nrhs, 1, t2, nrhs, 1, r, nrhs, r, 1.0, 0.0));
          g3.push_back(gemm_item(F.V[t], 1, F.cols(k), t2, nrhs, 1, bptr(k), nrhs, 1, mk, nrhs, r, -1.0, 1.0));
        }
      }
      plan_gemms<T>(pl, g1, "solve_gemm");
      plan_gemms<T>(pl, g2, "solve_gemm");
      plan_gemms<T>(pl, g3, "solve_gemm");
    }
  };
  Plan pl(ctx);
  const bool do_l = part != 2, do_u = part != 1;
  if (!trans) {
    if (do_l) sweep(pl, true, true, false);    // L y = b
    if (do_u) sweep(pl, false, false, false);  // U x = y
  } else {
    if (do_u) sweep(pl, true, false, true);   // U^T y = b
    if (do_l) sweep(pl, false, true, true);   // L^T x = y
  }
  pl.run();
}

template <typename T>
void blr_matvec(BlrMatrix<T>& A, const T* x, T* y, int nrhs, Context& ctx) {
  const int nr = A.nb_r(), nc = A.nb_c();
  ctx.sync();
  A.pull_ranks(ctx.stream());
  int kmax = 1;
  for (int r : A.rank) kmax = std::max(kmax, r);
  ctx.reserve((size_t)nr * nc * 2 * sizeof(T) * ((size_t)kmax * nrhs + 64) + 4096);
  Plan pl(ctx);
  for (int j = 0; j < nc; ++j) {
    std::vector<GemmItem> g1, g2, g3;
    const T* xj = x + A.coff[j] * nrhs;
    for (int i = 0; i < nr; ++i) {
      const int t = A.idx(i, j);
      T* yi = y + A.roff[i] * nrhs;
      const int mi = (int)A.rows(i), nj = (int)A.cols(j);
      if (A.kind[t] == kDense) {
        g1.push_back(gemm_item(A.D[t], nj, 1, xj, nrhs, 1, yi, nrhs, 1, mi, nrhs, nj, 1.0, 1.0));
      } else if (A.kind[t] == kLowRank) {
        const int r = A.rank[t];
        T* t1 = ctx.take<T>((size_t)r * nrhs + 1);
        T* t2 = ctx.take<T>((size_t)r * nrhs + 1);
        g1.push_back(gemm_item(A.V[t], nj, 1, xj, nrhs, 1, t1, nrhs, 1, r, nrhs, nj, 1.0, 0.0));
        g2.push_back(gemm_item(A.S[t], A.kcap, 1, t1, nrhs, 1, t2, nrhs, 1, r, nrhs, r, 1.0, 0.0));
        g3.push_back(gemm_item(A.U[t], 1, mi, t2, nrhs, 1, yi, nrhs, 1, mi, nrhs, r, 1.0, 1.0));
      }
    }
    plan_gemms<T>(pl, g1);
    plan_gemms<T>(pl, g2);
    plan_gemms<T>(pl, g3);
  }
  pl.run();
}

// ============================================================================
// Precision conversion of a whole BLR matrix (same structure)
// ============================================================================
template <typename TD, typename TS>
std::unique_ptr<BlrMatrix<TD>> convert_blr(const BlrMatrix<TS>& src, Context& ctx) {
  auto dst = std::make_unique<BlrMatrix<TD>>(src.roff, src.coff, src.kind, src.kcap, src.eps, src.fixed_rank, ctx);
  dst->lra = src.lra;
  dst->cap_trunc = src.cap_trunc;
  const hipStream_t s = ctx.stream();
  launch_convert<TD, TS>(dst->arena_data(), src.arena_data(), (int64_t)(src.arena_bytes() / sizeof(TS)), s);
  if (src.dinv_count() && dst->dinv_count() == src.dinv_count())
    launch_convert<TD, TS>(dst->dinv_data(), src.dinv_data(), (int64_t)src.dinv_count(), s);
  dst->set_ranks_from_device(src.d_rank(0, 0), s);
  host_wait_stream(s, __func__);
  dst->rank = src.rank;
  return dst;
}
template std::unique_ptr<BlrMatrix<float>> convert_blr<float, double>(const BlrMatrix<double>&, Context&);
template std::unique_ptr<BlrMatrix<double>> convert_blr<double, float>(const BlrMatrix<float>&, Context&);

// ============================================================================
// Iterative refinement with the exact FP64 kernel matrix
// ============================================================================
static std::vector<double> colmax(const double* a, int64_t n, int nrhs, Context& ctx) {
  double* part = ctx.take<double>((size_t)kColmaxBlocks * nrhs);
  launch_colmax_abs(a, n, nrhs, part, ctx.stream());
  std::vector<double> h((size_t)kColmaxBlocks * nrhs);
  FRK_HIP_CHECK(hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx.stream()));
  host_wait_stream(ctx.stream(), __func__);
  std::vector<double> out(nrhs, 0.0);
  for (int b = 0; b < kColmaxBlocks; ++b)
    for (int c = 0; c < nrhs; ++c) out[c] = std::max(out[c], h[(size_t)b * nrhs + c]);
  return out;
}

double kernel_norm_inf(const KernelSpec& ks, int64_t n, Context& ctx) {
  ctx.sync();
  ctx.reserve(sizeof(double) * ((size_t)3 * n + kColmaxBlocks) + 4096);
  double* ones = ctx.take<double>(n);
  double* zero = ctx.take<double>(n);
  double* r = ctx.take<double>(n);
  launch_fill_value(ones, n, 1.0, ctx.stream());
  launch_fill_value(zero, n, 0.0, ctx.stream());
  double* part = static_cast<double*>(ctx.persistent(7, sizeof(double) * kernel_residual_scratch(n, 1)));
  launch_kernel_residual(ks, ones, zero, r, n, 1, part, ctx.stream());  // r = -A 1 (non-negative entries)
  return colmax(r, n, 1, ctx)[0];
}

template <typename T>
RefineResult blr_refine(BlrMatrix<T>& F, const KernelSpec& ks, double anorm, const double* b, double* x, int nrhs,
                        double tol, int max_iter, Context& ctx) {
  const int64_t n = F.roff.back();
  const hipStream_t s = ctx.stream();
  ctx.set_default_peak(mfma_peak<T>());
  // residual and correction vectors persist in the context between calls
  double* r = static_cast<double*>(ctx.persistent(0, sizeof(double) * n * nrhs));
  T* d = static_cast<T*>(ctx.persistent(1, sizeof(T) * n * nrhs));
  double* part = static_cast<double*>(ctx.persistent(7, sizeof(double) * kernel_residual_scratch(n, nrhs)));
  ctx.sync();
  ctx.reserve(sizeof(double) * kColmaxBlocks * nrhs * 4 + 4096);
  const std::vector<double> bn = colmax(b, n, nrhs, ctx);
  // x0 = F^-1 b
  launch_convert<T, double>(d, b, n * nrhs, s);
  blr_solve<T>(F, d, nrhs, false, ctx);
  launch_convert<double, T>(x, d, n * nrhs, s);
  RefineResult res{0, 0.0};
  for (int it = 0;; ++it) {
    launch_kernel_residual(ks, x, b, r, n, nrhs, part, s);
    // op log: the product with the exact matrix (2 n^2 nrhs) plus the kernel
    // evaluation (difference, absolute value, shift, quotient: 4 n^2), FP64; the
    // matrix is generated on the fly, so only x, b and r cross HBM
    ctx.account("residual", 2.0 * n * n * nrhs + 4.0 * n * n, 3.0 * sizeof(double) * n * nrhs, kPeakF64);
    ctx.sync();
    ctx.reserve(sizeof(double) * kColmaxBlocks * nrhs * 4 + 4096);
    const std::vector<double> rn = colmax(r, n, nrhs, ctx);
    const std::vector<double> xn = colmax(x, n, nrhs, ctx);
    double berr = 0.0;
    for (int c = 0; c < nrhs; ++c) berr = std::max(berr, rn[c] / (anorm * xn[c] + bn[c]));
    res.iters = it;
    res.berr = berr;
    if (berr <= tol || it >= max_iter) break;
    launch_convert<T, double>(d, r, n * nrhs, s);
    blr_solve<T>(F, d, nrhs, false, ctx);
    launch_axpy_convert<T>(x, d, n * nrhs, s);
  }
  return res;
}

// ----------------------------------------------------------------------------
#define FRK_ENGINE_INST(T)                                                                                  \
  template class BlrMatrix<T>;                                                                              \
  template void plan_lr_add<T>(Plan&, const std::vector<LrAddDesc<T>>&, int);                               \
  template size_t lr_add_workspace<T>(const std::vector<LrAddDesc<T>>&);                                    \
  template void lr_add_batched<T>(const std::vector<LrAddDesc<T>>&, int, Context&);                         \
  template void lr_add_naive<T>(const std::vector<LrAddDesc<T>>&, Context&);                                \
  template void compress_batched<T>(const std::vector<CompressDesc<T>>&, Context&);                         \
  template void plan_getrf<T>(Plan&, const std::vector<T*>&, const std::vector<T*>&, const std::vector<T*>&, \
                              int, int64_t, int32_t*);                                                      \
  template void getrf_tiles<T>(const std::vector<T*>&, int, int64_t, int32_t*, Context&);                   \
  template void trsm_generic<T>(const T*, int64_t, int64_t, bool, bool, T*, int64_t, int64_t, int, int,     \
                                Context&);                                                                  \
  template std::unique_ptr<BlrMatrix<T>> build_blr_kernel<T>(const KernelDesc&, int, int, double, int,    \
                                                             double, int, int, Context&);                   \
  template void blr_getrf<T>(BlrMatrix<T>&, Context&, int);                                                      \
  template void blr_solve<T>(BlrMatrix<T>&, T*, int, bool, Context&, int);                                  \
  template void blr_matvec<T>(BlrMatrix<T>&, const T*, T*, int, Context&);                                  \
  template RefineResult blr_refine<T>(BlrMatrix<T>&, const KernelSpec&, double, const double*, double*, int,   \
                                      double, int, Context&);

FRK_ENGINE_INST(double)
FRK_ENGINE_INST(float)

}  // namespace frk
