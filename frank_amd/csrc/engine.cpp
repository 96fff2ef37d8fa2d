// Host orchestration of the batched BLR operations (see engine.hpp).
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>

namespace frk {

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
      FRK_HIP_CHECK(hipStreamSynchronize(s));
      off_ = 0;
    }
    if (aligned > cap_) {
      if (h_) FRK_HIP_CHECK(hipHostFree(h_));
      if (d_) FRK_HIP_CHECK(hipFree(d_));
      size_t c = std::max<size_t>(aligned, size_t(16) << 20);
      FRK_HIP_CHECK(hipHostMalloc((void**)&h_, c));
      FRK_HIP_CHECK(hipMalloc((void**)&d_, c));
      cap_ = c;
    }
  }
  char* hp = h_ + off_;
  char* dp = d_ + off_;
  if (bytes) {
    std::memcpy(hp, src, bytes);
    FRK_HIP_CHECK(hipMemcpyAsync(dp, hp, bytes, hipMemcpyHostToDevice, s));
  }
  off_ += aligned;
  return dp;
}

Context::Context(hipStream_t s) : s_(s) { info_.ensure(sizeof(int32_t) * 4096); }

Context::~Context() = default;

void Context::reserve(size_t bytes) {
  if (bytes > ws_.bytes()) {
    FRK_HIP_CHECK(hipStreamSynchronize(s_));
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
  FRK_HIP_CHECK(hipStreamSynchronize(s_));
  stage_.reset();
  ws_off_ = 0;
}

template <class X>
const X* Context::identity(int n) {
  if (n > ident_n_) {
    const int nn = std::max(n, 256);
    FRK_HIP_CHECK(hipStreamSynchronize(s_));
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
    FRK_HIP_CHECK(hipStreamSynchronize(s_));
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
  arena_.ensure(std::max<size_t>(total, 1) * sizeof(T));
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
  d_ranks_.ensure(sizeof(int32_t) * nr * nc);
  FRK_HIP_CHECK(hipMemsetAsync(d_ranks_.as<int32_t>(), 0, sizeof(int32_t) * nr * nc, ctx.stream()));
}

template <typename T>
void BlrMatrix<T>::pull_ranks(hipStream_t s) {
  FRK_HIP_CHECK(hipMemcpyAsync(rank.data(), d_ranks_.as<int32_t>(), sizeof(int32_t) * rank.size(),
                               hipMemcpyDeviceToHost, s));
  FRK_HIP_CHECK(hipStreamSynchronize(s));
  for (size_t t = 0; t < rank.size(); ++t)
    if (kind[t] == kLowRank && rank[t] < 0)
      throw std::runtime_error("low-rank capacity exceeded: rank " + std::to_string(-rank[t]) +
                               " > kcap " + std::to_string(kcap));
}

template <typename T>
void BlrMatrix<T>::push_ranks(hipStream_t s) {
  FRK_HIP_CHECK(hipMemcpyAsync(d_ranks_.as<int32_t>(), rank.data(), sizeof(int32_t) * rank.size(),
                               hipMemcpyHostToDevice, s));
  FRK_HIP_CHECK(hipStreamSynchronize(s));
}

// ============================================================================
// Item helpers
// ============================================================================
static GemmItem gemm_item(const void* A, int64_t a_rs, int64_t a_cs, const void* B, int64_t b_rs, int64_t b_cs,
                          void* C, int64_t c_rs, int64_t c_cs, int m, int n, int k, double alpha, double beta) {
  GemmItem g{};
  g.A = A;
  g.B = B;
  g.C = C;
  g.a_rs = a_rs;
  g.a_cs = a_cs;
  g.b_rs = b_rs;
  g.b_cs = b_cs;
  g.c_rs = c_rs;
  g.c_cs = c_cs;
  g.m = m;
  g.n = n;
  g.k = k;
  g.alpha = alpha;
  g.beta = beta;
  return g;
}

template <typename T>
static void run_gemms(const std::vector<GemmItem>& items, Context& ctx) {
  if (items.empty()) return;
  int mm = 0, nn = 0;
  for (const auto& g : items) {
    mm = std::max(mm, g.m);
    nn = std::max(nn, g.n);
  }
  // grid.y is limited to 65535 items per launch
  const size_t chunk = 65535;
  for (size_t o = 0; o < items.size(); o += chunk) {
    std::vector<GemmItem> part(items.begin() + o, items.begin() + std::min(items.size(), o + chunk));
    launch_gemm<T>(ctx.staging().push(part, ctx.stream()), (int)part.size(), mm, nn, ctx.stream());
  }
}

template <typename T>
static void run_trsms(const std::vector<TrsmItem>& items, Context& ctx) {
  if (items.empty()) return;
  int n = 0, r = 0;
  for (const auto& t : items) {
    n = std::max(n, t.n);
    r = std::max(r, t.nrhs);
  }
  launch_trsm<T>(ctx.staging().push(items, ctx.stream()), (int)items.size(), n, r, ctx.stream());
}

// ============================================================================
// Dense tile LU (blocked, no pivoting)
// ============================================================================
template <typename T>
void getrf_tiles(const std::vector<T*>& tiles, int n, int64_t ld, int32_t* info_dev, Context& ctx) {
  if (tiles.empty() || n <= 0) return;
  const int nt = (int)tiles.size();
  const int npan = (n + 63) / 64;
  int pi = 0;
  for (int p = 0; p < n; p += 64, ++pi) {
    const int nbk = std::min(64, n - p);
    const int rest = n - p - nbk;
    std::vector<GemmItem> small(nt);
    for (int t = 0; t < nt; ++t) {
      small[t] = GemmItem{};
      small[t].C = tiles[t] + p * ld + p;
      small[t].c_rs = ld;
      small[t].c_cs = 1;
      small[t].m = nbk;
    }
    launch_getrf_small<T>(ctx.staging().push(small, ctx.stream()), nt, info_dev ? info_dev + pi * nt : nullptr,
                          ctx.stream());
    if (rest <= 0) continue;
    std::vector<TrsmItem> tr;
    tr.reserve(2 * nt);
    for (int t = 0; t < nt; ++t) {
      T* a = tiles[t];
      TrsmItem x{};
      // A12 <- L11^{-1} A12 (unit lower)
      x.A = a + p * ld + p;
      x.a_rs = ld;
      x.a_cs = 1;
      x.B = a + p * ld + p + nbk;
      x.b_rs = ld;
      x.b_cs = 1;
      x.n = nbk;
      x.nrhs = rest;
      x.lower = 1;
      x.unit = 1;
      tr.push_back(x);
      // A21 <- A21 U11^{-1}  <=>  U11^T A21^T = A21^T (lower, non-unit)
      TrsmItem y{};
      y.A = a + p * ld + p;
      y.a_rs = 1;
      y.a_cs = ld;
      y.B = a + (p + nbk) * ld + p;
      y.b_rs = 1;
      y.b_cs = ld;
      y.n = nbk;
      y.nrhs = rest;
      y.lower = 1;
      y.unit = 0;
      tr.push_back(y);
    }
    run_trsms<T>(tr, ctx);
    std::vector<GemmItem> gm(nt);
    for (int t = 0; t < nt; ++t) {
      T* a = tiles[t];
      gm[t] = gemm_item(a + (p + nbk) * ld + p, ld, 1, a + p * ld + p + nbk, ld, 1, a + (p + nbk) * ld + p + nbk,
                        ld, 1, rest, rest, nbk, -1.0, 1.0);
    }
    run_gemms<T>(gm, ctx);
  }
  (void)npan;
}

// ============================================================================
// Rounded low-rank addition (batched)
// ============================================================================
template <typename T>
void lr_add_batched(const std::vector<LrAddDesc<T>>& d, int lra, Context& ctx) {
  if (d.empty()) return;
  const int nd = (int)d.size();
  const hipStream_t s = ctx.stream();
  int kmax = 0, mmax = 0, nmax = 0;
  for (const auto& x : d) {
    kmax = std::max(kmax, x.kc + x.kx2);
    mmax = std::max(mmax, x.m);
    nmax = std::max(nmax, x.n);
  }
  const int K = kmax;
  // per-item workspace
  T* Yu = ctx.take<T>((size_t)nd * mmax * K);
  T* Yv = ctx.take<T>((size_t)nd * nmax * K);
  T* tu = ctx.take<T>((size_t)nd * K);
  T* tv = ctx.take<T>((size_t)nd * K);
  T* Ru = ctx.take<T>((size_t)nd * K * K);
  T* Rv = ctx.take<T>((size_t)nd * K * K);
  T* Um = ctx.take<T>((size_t)nd * K * K);
  T* Vm = ctx.take<T>((size_t)nd * K * K);
  std::vector<QrItem> qi;
  qi.reserve(2 * nd);
  std::vector<SvdItem> si(nd);
  std::vector<ApplyQItem> ai;
  ai.reserve(2 * nd);
  for (int t = 0; t < nd; ++t) {
    const auto& x = d[t];
    const bool fast = (lra == kLraFast) && (x.eps == 0.0);
    const int Kt = x.kc + x.kx2;
    T* yu = Yu + (size_t)t * mmax * K;
    T* yv = Yv + (size_t)t * nmax * K;
    T* ru = Ru + (size_t)t * K * K;
    T* rv = Rv + (size_t)t * K * K;
    T* um = Um + (size_t)t * K * K;
    T* vm = Vm + (size_t)t * K * K;
    QrItem q{};
    q.Y = yu;
    q.ldy = x.m;
    q.tau = tu + (size_t)t * K;
    q.R = ru;
    q.ldr = K;
    q.m = x.m;
    QrItem v = q;
    v.Y = yv;
    v.ldy = x.n;
    v.tau = tv + (size_t)t * K;
    v.R = rv;
    v.m = x.n;
    if (!fast) {
      // Uc = [Uc Sc | Ux Sx] ; Vc^T = [Vc^T | Vx^T]
      q.A1 = x.cu;
      q.a1_rs = x.cu_rs;
      q.a1_cs = x.cu_cs;
      q.S1 = x.cs;
      q.s1_ld = x.lds;
      q.k1a = q.k1 = x.kc;
      q.A2 = x.xu;
      q.a2_rs = x.xu_rs;
      q.a2_cs = x.xu_cs;
      q.S2 = x.xs;
      q.s2_ld = x.xs_ld;
      q.k2a = x.kx1;
      q.k2 = x.kx2;
      v.A1 = x.cv;
      v.a1_rs = x.cv_cs;
      v.a1_cs = x.cv_rs;
      v.S1 = nullptr;
      v.k1a = v.k1 = x.kc;
      v.A2 = x.xv;
      v.a2_rs = x.xv_cs;
      v.a2_cs = x.xv_rs;
      v.S2 = nullptr;
      v.k2a = v.k2 = x.kx2;
    } else {
      // project X's bases on the (orthonormal) bases of C, QR the remainders
      q.A1 = x.xu;
      q.a1_rs = x.xu_rs;
      q.a1_cs = x.xu_cs;
      q.S1 = x.xs;
      q.s1_ld = x.xs_ld;
      q.k1a = x.kx1;
      q.k1 = x.kx2;
      q.k2a = q.k2 = 0;
      q.P = x.cu;
      q.p_rs = x.cu_rs;
      q.p_cs = x.cu_cs;
      q.kp = x.kc;
      v.A1 = x.xv;
      v.a1_rs = x.xv_cs;
      v.a1_cs = x.xv_rs;
      v.S1 = nullptr;
      v.k1a = v.k1 = x.kx2;
      v.k2a = v.k2 = 0;
      v.P = x.cv;
      v.p_rs = x.cv_cs;
      v.p_cs = x.cv_rs;
      v.kp = x.kc;
    }
    qi.push_back(q);
    qi.push_back(v);
    SvdItem sv{};
    sv.mode = 1;
    sv.Ru = ru;
    sv.Rv = rv;
    sv.ldr = K;
    sv.K = Kt;
    if (!fast) {
      sv.kx = Kt;
      sv.ka = 0;
    } else {
      sv.kx = x.kx2;
      sv.ka = x.kc;
      sv.Sa = x.cs;
      sv.lds_in = x.lds;
    }
    sv.fixed_rank = (x.eps != 0.0) ? -1 : x.kc;
    sv.kdim = std::min(std::min(x.m, Kt), std::min(x.n, Kt));
    sv.eps = x.eps;
    sv.S = x.cs;
    sv.lds = x.lds;
    sv.Um = um;
    sv.ldu = K;
    sv.Vm = vm;
    sv.ldv = K;
    sv.rank_out = x.d_rank;
    sv.cap = x.cap;
    si[t] = sv;
    ApplyQItem a{};
    a.Y = yu;
    a.y_rs = 1;
    a.y_cs = x.m;
    a.tau = tu + (size_t)t * K;
    a.m = x.m;
    a.ncol_ptr = x.d_rank;
    a.ncol = Kt;
    a.out = x.cu;
    a.out_rs = x.cu_rs;
    a.out_cs = x.cu_cs;
    a.trans = 0;
    ApplyQItem b = a;
    b.Y = yv;
    b.y_cs = x.n;
    b.tau = tv + (size_t)t * K;
    b.m = x.n;
    b.out = x.cv;
    b.out_rs = x.cv_cs;
    b.out_cs = x.cv_rs;
    if (!fast) {
      a.nref = std::min(x.m, Kt);
      a.X = um;
      a.x_rs = K;
      a.x_cs = 1;
      a.kx = Kt;
      b.nref = std::min(x.n, Kt);
      b.X = vm;
      b.x_rs = K;
      b.x_cs = 1;
      b.kx = Kt;
    } else {
      a.nref = std::min(x.m, x.kx2);
      a.X = um + (size_t)x.kc * K;
      a.x_rs = K;
      a.x_cs = 1;
      a.kx = x.kx2;
      a.Base = x.cu;
      a.base_rs = x.cu_rs;
      a.base_cs = x.cu_cs;
      a.Xb = um;
      a.xb_rs = K;
      a.xb_cs = 1;
      a.kb = x.kc;
      b.nref = std::min(x.n, x.kx2);
      b.X = vm + (size_t)x.kc * K;
      b.x_rs = K;
      b.x_cs = 1;
      b.kx = x.kx2;
      b.Base = x.cv;
      b.base_rs = x.cv_cs;
      b.base_cs = x.cv_rs;
      b.Xb = vm;
      b.xb_rs = K;
      b.xb_cs = 1;
      b.kb = x.kc;
    }
    ai.push_back(a);
    ai.push_back(b);
  }
  int qk = 0;
  for (const auto& q : qi) qk = std::max(qk, q.k1 + q.k2);
  launch_qr<T>(ctx.staging().push(qi, s), (int)qi.size(), qk, s);
  launch_svd<T>(ctx.staging().push(si, s), nd, K, s);
  int capmax = 0;
  for (const auto& x : d) capmax = std::max(capmax, std::min(x.cap, x.kc + x.kx2));
  launch_apply_q<T>(ctx.staging().push(ai, s), (int)ai.size(), capmax, s);
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
    const int nd = (int)de.size();
    int nmax = 0, mmax = 0;
    for (const auto& x : de) {
      nmax = std::max(nmax, x.n);
      mmax = std::max(mmax, x.m);
    }
    ctx.sync();
    const int capmax = de[0].cap;
    ctx.reserve((size_t)nd * sizeof(T) * ((size_t)nmax + (size_t)mmax + (size_t)capmax * nmax * 2 + 4 * capmax +
                                            (size_t)capmax * capmax + 64) +
                (size_t)nd * sizeof(int32_t) * (nmax + 64) + (size_t)nd * 4096);
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
      g.eps = de[t].eps;
      gi[t] = g;
    }
    launch_geqp3<T>(ctx.staging().push(gi, s), nd, s);
    // ranks to host
    std::vector<int32_t> rk(nd);
    for (int t = 0; t < nd; ++t)
      FRK_HIP_CHECK(hipMemcpyAsync(&rk[t], de[t].d_rank, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FRK_HIP_CHECK(hipStreamSynchronize(s));
    int rmax = 1;
    for (int t = 0; t < nd; ++t) {
      if (rk[t] < 0)
        throw std::runtime_error("low-rank capacity exceeded during compression: rank " + std::to_string(-rk[t]));
      rmax = std::max(rmax, rk[t]);
    }
    const T* I = ctx.identity<T>(std::max(rmax, 1) + 1);
    const int ild = ctx.ident_ld();
    T* RP = ctx.take<T>((size_t)nd * rmax * nmax);
    T* Yz = ctx.take<T>((size_t)nd * nmax * rmax);
    T* tz = ctx.take<T>((size_t)nd * rmax);
    T* Rz = ctx.take<T>((size_t)nd * rmax * rmax);
    std::vector<ApplyQItem> au;
    std::vector<PermuteItem> pi(nd);
    std::vector<QrItem> qz(nd);
    std::vector<FlipItem> fl(nd);
    std::vector<ApplyQItem> av;
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
      FRK_HIP_CHECK(hipMemsetAsync(rp, 0, sizeof(T) * (size_t)r * x.n, s));
      PermuteItem p{};
      p.A = x.a;
      p.piv = gi[t].piv;
      p.rank = x.d_rank;
      p.out = rp;
      p.n = x.n;
      p.lda = (int)x.lda;
      p.ldo = x.n;
      pi[t] = p;
      QrItem q{};
      q.A1 = rp + (size_t)(r - 1) * x.n;
      q.a1_rs = 1;
      q.a1_cs = -(int64_t)x.n;
      q.k1a = q.k1 = r;
      q.m = x.n;
      q.Y = Yz + (size_t)t * nmax * rmax;
      q.ldy = x.n;
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
    launch_apply_q<T>(ctx.staging().push(au, s), nd, rmax, s);
    launch_permute_r<T>(ctx.staging().push(pi, s), nd, nmax, rmax, s);
    launch_qr<T>(ctx.staging().push(qz, s), nd, rmax, s);
    launch_flip_r<T>(ctx.staging().push(fl, s), nd, rmax, s);
    launch_apply_q<T>(ctx.staging().push(av, s), nd, rmax, s);
    ctx.sync();
  }
  if (!dr.empty()) {
    const int nd = (int)dr.size();
    ctx.sync();
    int nmax = 0, mmax = 0, pmax = 0;
    for (const auto& x : dr) {
      nmax = std::max(nmax, x.n);
      mmax = std::max(mmax, x.m);
      pmax = std::max(pmax, std::min(std::min(x.rank + 5, x.m), x.n));
    }
    ctx.reserve((size_t)nd * sizeof(T) * ((size_t)mmax * pmax * 2 + (size_t)nmax * pmax * 2 + 3 * pmax * pmax + 4 * pmax + 64) +
                (size_t)nd * 4096);
    // all tiles share one Omega only when they share n and p
    std::vector<GemmItem> g1;
    std::vector<QrItem> q1;
    std::vector<ApplyQItem> a1;
    std::vector<GemmItem> g2;
    std::vector<QrItem> q2;
    std::vector<SvdItem> sv;
    std::vector<ApplyQItem> a2;
    T* Yb = ctx.take<T>((size_t)nd * mmax * pmax);
    T* Qe = ctx.take<T>((size_t)nd * mmax * pmax);
    T* Zb = ctx.take<T>((size_t)nd * nmax * pmax);
    T* ty = ctx.take<T>((size_t)nd * pmax);
    T* tz = ctx.take<T>((size_t)nd * pmax);
    T* Ry = ctx.take<T>((size_t)nd * pmax * pmax);
    T* Rz = ctx.take<T>((size_t)nd * pmax * pmax);
    T* Um = ctx.take<T>((size_t)nd * pmax * pmax);
    T* Vm = ctx.take<T>((size_t)nd * pmax * pmax);
    const T* I = ctx.identity<T>(pmax + 1);
    const int ild = ctx.ident_ld();
    for (int t = 0; t < nd; ++t) {
      const auto& x = dr[t];
      const int p = std::min(std::min(x.rank + 5, x.m), x.n);
      const T* Om = ctx.omega<T>(x.n, p);
      T* y = Yb + (size_t)t * mmax * pmax;
      T* qe = Qe + (size_t)t * mmax * pmax;
      T* z = Zb + (size_t)t * nmax * pmax;
      // Y = A * Omega (m x p, column-major)
      g1.push_back(gemm_item(x.a, x.lda, 1, Om, p, 1, y, 1, x.m, x.m, p, x.n, 1.0, 0.0));
      QrItem q{};
      q.A1 = y;
      q.a1_rs = 1;
      q.a1_cs = x.m;
      q.k1a = q.k1 = p;
      q.m = x.m;
      q.Y = y;
      q.ldy = x.m;
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
      // Z = A^T Qe (n x p, column-major) = (Q^T A)^T
      g2.push_back(gemm_item(x.a, 1, x.lda, qe, 1, x.m, z, 1, x.n, x.n, p, x.m, 1.0, 0.0));
      QrItem qq{};
      qq.A1 = z;
      qq.a1_rs = 1;
      qq.a1_cs = x.n;
      qq.k1a = qq.k1 = p;
      qq.m = x.n;
      qq.Y = z;
      qq.ldy = x.n;
      qq.tau = tz + (size_t)t * pmax;
      qq.R = Rz + (size_t)t * pmax * pmax;
      qq.ldr = p;
      q2.push_back(qq);
      SvdItem s0{};
      s0.mode = 0;
      s0.M = qq.R;
      s0.m_rs = 1;
      s0.m_cs = p;  // M = Rz^T
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
    run_gemms<T>(g1, ctx);
    launch_qr<T>(ctx.staging().push(q1, s), nd, pmax, s);
    launch_apply_q<T>(ctx.staging().push(a1, s), nd, pmax, s);
    run_gemms<T>(g2, ctx);
    launch_qr<T>(ctx.staging().push(q2, s), nd, pmax, s);
    launch_svd<T>(ctx.staging().push(sv, s), nd, pmax, s);
    launch_apply_q<T>(ctx.staging().push(a2, s), (int)a2.size(), pmax, s);
    ctx.sync();
  }
}

// ============================================================================
// Construction from a kernel function
// ============================================================================
static std::vector<int64_t> index_split(int64_t n, int nsplit) {
  // src/classes/initialization_helpers/index_range.cpp:15-24
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
std::unique_ptr<BlrMatrix<T>> build_blr_kernel(const KernelDesc& kd, int nb, double eps, int rank, double admis,
                                               int admis_type, int kcap, Context& ctx) {
  const hipStream_t s = ctx.stream();
  auto off = index_split(kd.n, nb);
  std::vector<int32_t> kinds(nb * nb);
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < nb; ++j)
      kinds[i * nb + j] = admissible(kd, off[i], off[i + 1] - off[i], off[j], off[j + 1] - off[j], i, j, admis,
                                     admis_type)
                              ? kLowRank
                              : kDense;
  auto A = std::make_unique<BlrMatrix<T>>(off, off, kinds, kcap, eps, rank, ctx);
  DevBuf pts;
  KernelSpec ks = upload_kernel(kd, pts, s);
  ctx.sync();
  // dense tiles straight into their slots
  std::vector<GenItem> gd;
  int mr = 0, mc = 0;
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < nb; ++j)
      if (kinds[i * nb + j] == kDense) {
        GenItem g{};
        g.out = A->D[A->idx(i, j)];
        g.ld = A->cols(j);
        g.r0 = (int)off[i];
        g.nr = (int)A->rows(i);
        g.c0 = (int)off[j];
        g.nc = (int)A->cols(j);
        gd.push_back(g);
        mr = std::max<int>(mr, g.nr);
        mc = std::max<int>(mc, g.nc);
      }
  launch_gen<T>(ks, ctx.staging().push(gd, s), (int)gd.size(), mr, mc, s);
  // low-rank tiles: generate a chunk densely, compress into the slots
  std::vector<std::pair<int, int>> lr;
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < nb; ++j)
      if (kinds[i * nb + j] == kLowRank) lr.emplace_back(i, j);
  const int64_t bmax = off[1] - off[0];
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
      g.c0 = (int)off[j];
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
      cd.push_back(c);
    }
    launch_gen<T>(ks, ctx.staging().push(gl, s), (int)gl.size(), (int)bmax, (int)bmax, s);
    compress_batched<T>(cd, ctx);
  }
  ctx.sync();
  A->pull_ranks(s);
  return A;
}

// ============================================================================
// BLR-LU
// ============================================================================
template <typename T>
void blr_getrf(BlrMatrix<T>& A, Context& ctx) {
  const hipStream_t s = ctx.stream();
  const int nb = A.nb_r();
  if (A.nb_c() != nb) throw std::invalid_argument("getrf needs a square block grid");
  int32_t* info = ctx.info();
  const int npan_max = 8;
  if (nb * npan_max > 4096) throw std::invalid_argument("too many diagonal tiles for the info buffer");
  FRK_HIP_CHECK(hipMemsetAsync(info, 0, sizeof(int32_t) * 4096, s));
  for (int i = 0; i < nb; ++i) {
    ctx.sync();
    A.pull_ranks(s);
    const int m_i = (int)A.rows(i), n_i = (int)A.cols(i);
    const int ii = A.idx(i, i);
    if (A.kind[ii] != kDense) throw std::runtime_error("getrf: diagonal tile is not dense");
    if (m_i > 512) throw std::runtime_error("getrf: tile size above 512");
    // workspace sizing for this step
    const int rem = nb - 1 - i;
    size_t need = 4096;
    {
      size_t upd = (size_t)rem * rem;
      int kmax = 0;
      for (int r = 0; r < (int)A.rank.size(); ++r) kmax = std::max(kmax, A.rank[r]);
      const size_t K = (size_t)2 * std::max(kmax, 1) + 1;
      const size_t bm = (size_t)std::max<int64_t>(A.rows(0), A.cols(0));
      need += upd * sizeof(T) * (2 * bm * K + 6 * K * K + 2 * K + (size_t)A.kcap * A.kcap + bm * A.kcap + 64);
      need += (size_t)rem * 2 * sizeof(T) * (bm * A.kcap + 64);
      need += upd * sizeof(T) * 4 * 64;
      // dense products feeding D.D -> LR compression
      size_t ndd = 0, nld = 0;
      for (int ic = i + 1; ic < nb; ++ic)
        for (int k = i + 1; k < nb; ++k) {
          if (A.kind[A.idx(ic, i)] == kDense && A.kind[A.idx(i, k)] == kDense && A.kind[A.idx(ic, k)] == kLowRank)
            ++ndd;
          if (A.kind[A.idx(ic, k)] == kDense) ++nld;
        }
      need += (ndd + nld) * sizeof(T) * (bm * bm + 64);
    }
    ctx.reserve(need);
    // 1. diagonal tile
    getrf_tiles<T>({A.D[ii]}, m_i, n_i, info + i * npan_max, ctx);
    const T* Dii = A.D[ii];
    const int64_t ldd = n_i;
    // 2. panel solves
    std::vector<TrsmItem> tr;
    for (int ic = i + 1; ic < nb; ++ic) {
      const int t = A.idx(ic, i);
      TrsmItem x{};
      x.A = Dii;
      x.a_rs = 1;
      x.a_cs = ldd;
      x.lower = 1;
      x.unit = 0;
      x.n = n_i;
      if (A.kind[t] == kLowRank) {  // V <- V U^{-1}
        x.B = A.V[t];
        x.b_rs = 1;
        x.b_cs = A.cols(i);
        x.nrhs = A.rank[t];
      } else if (A.kind[t] == kDense) {
        x.B = A.D[t];
        x.b_rs = 1;
        x.b_cs = A.cols(i);
        x.nrhs = (int)A.rows(ic);
      } else
        continue;
      tr.push_back(x);
    }
    for (int j = i + 1; j < nb; ++j) {
      const int t = A.idx(i, j);
      TrsmItem x{};
      x.A = Dii;
      x.a_rs = ldd;
      x.a_cs = 1;
      x.lower = 1;
      x.unit = 1;
      x.n = m_i;
      if (A.kind[t] == kLowRank) {  // U <- L^{-1} U
        x.B = A.U[t];
        x.b_rs = 1;
        x.b_cs = A.rows(i);
        x.nrhs = A.rank[t];
      } else if (A.kind[t] == kDense) {
        x.B = A.D[t];
        x.b_rs = A.cols(j);
        x.b_cs = 1;
        x.nrhs = (int)A.cols(j);
      } else
        continue;
      tr.push_back(x);
    }
    run_trsms<T>(tr, ctx);
    if (rem == 0) break;
    // 3. Schur updates A(ic,k) -= L(ic,i) A(i,k)
    std::vector<T*> P(nb, nullptr), Q(nb, nullptr);
    std::vector<GemmItem> gA, gB, gC, gD;
    for (int ic = i + 1; ic < nb; ++ic) {
      const int t = A.idx(ic, i);
      if (A.kind[t] != kLowRank) continue;
      const int kl = A.rank[t];
      P[ic] = ctx.take<T>((size_t)kl * n_i + 1);
      // P = S_L V_L (kl x n_i, row-major)
      gA.push_back(gemm_item(A.S[t], A.kcap, 1, A.V[t], A.cols(i), 1, P[ic], n_i, 1, kl, n_i, kl, 1.0, 0.0));
    }
    for (int k = i + 1; k < nb; ++k) {
      const int t = A.idx(i, k);
      if (A.kind[t] != kLowRank) continue;
      const int kb = A.rank[t];
      Q[k] = ctx.take<T>((size_t)m_i * kb + 1);
      // Q = U_B S_B (m_i x kb, column-major)
      gA.push_back(gemm_item(A.U[t], 1, A.rows(i), A.S[t], A.kcap, 1, Q[k], 1, m_i, m_i, kb, kb, 1.0, 0.0));
    }
    std::vector<LrAddDesc<T>> adds;
    std::vector<CompressDesc<T>> comp;
    std::vector<std::pair<int, int>> comp_tiles;
    std::vector<std::array<T*, 3>> comp_slots;
    for (int ic = i + 1; ic < nb; ++ic)
      for (int k = i + 1; k < nb; ++k) {
        const int tl = A.idx(ic, i), tb = A.idx(i, k), tc = A.idx(ic, k);
        const int kl = A.kind[tl], kb = A.kind[tb], kc = A.kind[tc];
        const int m = (int)A.rows(ic), n = (int)A.cols(k);
        if (kl == kEmpty || kb == kEmpty || kc == kEmpty) continue;
        if (kc == kDense) {
          T* C = A.D[tc];
          if (kl == kLowRank && kb == kLowRank) {
            const int rl = A.rank[tl], rb = A.rank[tb];
            T* Sx = ctx.take<T>((size_t)rl * rb + 1);
            T* z = ctx.take<T>((size_t)m * rb + 1);
            gB.push_back(gemm_item(P[ic], n_i, 1, Q[k], 1, m_i, Sx, rb, 1, rl, rb, n_i, -1.0, 0.0));
            gC.push_back(gemm_item(A.U[tl], 1, A.rows(ic), Sx, rb, 1, z, 1, m, m, rb, rl, 1.0, 0.0));
            gD.push_back(gemm_item(z, 1, m, A.V[tb], A.cols(k), 1, C, n, 1, m, n, rb, 1.0, 1.0));
          } else if (kl == kLowRank && kb == kDense) {
            const int rl = A.rank[tl];
            T* w = ctx.take<T>((size_t)rl * n + 1);
            T* w2 = ctx.take<T>((size_t)rl * n + 1);
            gA.push_back(gemm_item(A.V[tl], A.cols(i), 1, A.D[tb], n, 1, w, n, 1, rl, n, n_i, -1.0, 0.0));
            gB.push_back(gemm_item(A.S[tl], A.kcap, 1, w, n, 1, w2, n, 1, rl, n, rl, 1.0, 0.0));
            gC.push_back(gemm_item(A.U[tl], 1, A.rows(ic), w2, n, 1, C, n, 1, m, n, rl, 1.0, 1.0));
          } else if (kl == kDense && kb == kLowRank) {
            const int rb = A.rank[tb];
            T* a = ctx.take<T>((size_t)m * rb + 1);
            T* a2 = ctx.take<T>((size_t)m * rb + 1);
            gA.push_back(gemm_item(A.D[tl], n_i, 1, A.U[tb], 1, A.rows(i), a, 1, m, m, rb, n_i, -1.0, 0.0));
            gB.push_back(gemm_item(a, 1, m, A.S[tb], A.kcap, 1, a2, 1, m, m, rb, rb, 1.0, 0.0));
            gC.push_back(gemm_item(a2, 1, m, A.V[tb], A.cols(k), 1, C, n, 1, m, n, rb, 1.0, 1.0));
          } else {
            gA.push_back(gemm_item(A.D[tl], n_i, 1, A.D[tb], n, 1, C, n, 1, m, n, n_i, -1.0, 1.0));
          }
          continue;
        }
        // low-rank target: C += X
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
        if (kl == kLowRank && kb == kLowRank) {
          const int rl = A.rank[tl], rb = A.rank[tb];
          T* Sx = ctx.take<T>((size_t)rl * rb + 1);
          gB.push_back(gemm_item(P[ic], n_i, 1, Q[k], 1, m_i, Sx, rb, 1, rl, rb, n_i, -1.0, 0.0));
          ad.xu = A.U[tl];
          ad.xu_rs = 1;
          ad.xu_cs = A.rows(ic);
          ad.kx1 = rl;
          ad.xs = Sx;
          ad.xs_ld = rb;
          ad.kx2 = rb;
          ad.xv = A.V[tb];
          ad.xv_rs = A.cols(k);
          ad.xv_cs = 1;
        } else if (kl == kLowRank && kb == kDense) {
          const int rl = A.rank[tl];
          T* w = ctx.take<T>((size_t)rl * n + 1);
          gA.push_back(gemm_item(A.V[tl], A.cols(i), 1, A.D[tb], n, 1, w, n, 1, rl, n, n_i, -1.0, 0.0));
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
        } else if (kl == kDense && kb == kLowRank) {
          const int rb = A.rank[tb];
          T* a = ctx.take<T>((size_t)m * rb + 1);
          gA.push_back(gemm_item(A.D[tl], n_i, 1, A.U[tb], 1, A.rows(i), a, 1, m, m, rb, n_i, -1.0, 0.0));
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
        } else {
          // D.D -> LR: compress the dense product first (gemm.cpp:202-217)
          T* prod = ctx.take<T>((size_t)m * n + 1);
          gA.push_back(gemm_item(A.D[tl], n_i, 1, A.D[tb], n, 1, prod, n, 1, m, n, n_i, -1.0, 0.0));
          T* xu = ctx.take<T>((size_t)m * A.kcap + 1);
          T* xs = ctx.take<T>((size_t)A.kcap * A.kcap + 1);
          T* xv = ctx.take<T>((size_t)A.kcap * n + 1);
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
          c.d_rank = nullptr;  // assigned below
          comp.push_back(c);
          comp_slots.push_back({xu, xs, xv});
          comp_tiles.emplace_back((int)adds.size(), 0);
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
        adds.push_back(ad);
      }
    run_gemms<T>(gA, ctx);
    run_gemms<T>(gB, ctx);
    run_gemms<T>(gC, ctx);
    run_gemms<T>(gD, ctx);
    if (!comp.empty()) {
      // rank table entries for the compressed products live in the info buffer tail
      DevBuf crank(sizeof(int32_t) * comp.size());
      for (size_t c = 0; c < comp.size(); ++c) comp[c].d_rank = crank.as<int32_t>() + c;
      // compression drains the stream and resets the workspace: keep the
      // already-planned buffers alive by running it on a separate arena
      Context side(s);
      compress_batched<T>(comp, side);
      std::vector<int32_t> cr(comp.size());
      FRK_HIP_CHECK(hipMemcpy(cr.data(), crank.as<int32_t>(), sizeof(int32_t) * cr.size(), hipMemcpyDeviceToHost));
      for (size_t c = 0; c < comp.size(); ++c) {
        auto& ad = adds[comp_tiles[c].first];
        ad.kx1 = ad.kx2 = cr[c];
      }
    }
    lr_add_batched<T>(adds, A.lra, ctx);
  }
  ctx.sync();
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
// Block substitution (column-oriented so every batched update has a distinct
// target block)
// ============================================================================
template <typename T>
void blr_solve(BlrMatrix<T>& F, T* b, int nrhs, bool trans, Context& ctx) {
  const int nb = F.nb_r();
  ctx.sync();
  F.pull_ranks(ctx.stream());
  int kmax = 1;
  for (int r : F.rank) kmax = std::max(kmax, r);
  auto bptr = [&](int i) { return b + F.roff[i] * nrhs; };
  // one sweep: diagonal solves in `order`, updates towards `targets(i)`
  auto sweep = [&](bool forward, bool lower_factor, bool transposed) {
    for (int step = 0; step < nb; ++step) {
      const int i = forward ? step : nb - 1 - step;
      ctx.reserve((size_t)nb * 2 * sizeof(T) * ((size_t)kmax * nrhs + 64) + 4096);
      const int ii = F.idx(i, i);
      const int64_t ld = F.cols(i);
      TrsmItem x{};
      x.A = F.D[ii];
      x.B = bptr(i);
      x.b_rs = nrhs;
      x.b_cs = 1;
      x.n = (int)F.rows(i);
      x.nrhs = nrhs;
      if (!transposed) {
        x.a_rs = ld;
        x.a_cs = 1;
        x.lower = lower_factor ? 1 : 0;
      } else {
        x.a_rs = 1;
        x.a_cs = ld;
        x.lower = lower_factor ? 0 : 1;
      }
      x.unit = lower_factor ? 1 : 0;
      run_trsms<T>({x}, ctx);
      std::vector<GemmItem> g1, g2, g3;
      for (int k = 0; k < nb; ++k) {
        if (forward ? (k <= i) : (k >= i)) continue;
        // contribution of block i to block k through tile t
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
          // b_k -= U (S (V b_i)), tile (k, i): U rows(k) x r, V r x cols(i)
          g1.push_back(gemm_item(F.V[t], F.cols(i), 1, bptr(i), nrhs, 1, t1, nrhs, 1, r, nrhs, ni, 1.0, 0.0));
          g2.push_back(gemm_item(F.S[t], F.kcap, 1, t1, nrhs, 1, t2, nrhs, 1, r, nrhs, r, 1.0, 0.0));
          g3.push_back(gemm_item(F.U[t], 1, F.rows(k), t2, nrhs, 1, bptr(k), nrhs, 1, mk, nrhs, r, -1.0, 1.0));
        } else {
          // b_k -= V^T (S^T (U^T b_i)), tile (i, k): U rows(i) x r, V r x cols(k)
          g1.push_back(gemm_item(F.U[t], F.rows(i), 1, bptr(i), nrhs, 1, t1, nrhs, 1, r, nrhs, ni, 1.0, 0.0));
          g2.push_back(gemm_item(F.S[t], 1, F.kcap, t1, nrhs, 1, t2, nrhs, 1, r, nrhs, r, 1.0, 0.0));
          g3.push_back(gemm_item(F.V[t], 1, F.cols(k), t2, nrhs, 1, bptr(k), nrhs, 1, mk, nrhs, r, -1.0, 1.0));
        }
      }
      run_gemms<T>(g1, ctx);
      run_gemms<T>(g2, ctx);
      run_gemms<T>(g3, ctx);
      ctx.staging();
    }
  };
  if (!trans) {
    sweep(true, true, false);    // L y = b
    sweep(false, false, false);  // U x = y
  } else {
    sweep(true, false, true);   // U^T y = b
    sweep(false, true, true);   // L^T x = y
  }
}

template <typename T>
void blr_matvec(BlrMatrix<T>& A, const T* x, T* y, int nrhs, Context& ctx) {
  const int nr = A.nb_r(), nc = A.nb_c();
  ctx.sync();
  A.pull_ranks(ctx.stream());
  int kmax = 1;
  for (int r : A.rank) kmax = std::max(kmax, r);
  for (int j = 0; j < nc; ++j) {
    ctx.reserve((size_t)nr * 2 * sizeof(T) * ((size_t)kmax * nrhs + 64) + 4096);
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
    run_gemms<T>(g1, ctx);
    run_gemms<T>(g2, ctx);
    run_gemms<T>(g3, ctx);
  }
}

// ----------------------------------------------------------------------------
#define FRK_ENGINE_INST(T)                                                                                  \
  template class BlrMatrix<T>;                                                                              \
  template void lr_add_batched<T>(const std::vector<LrAddDesc<T>>&, int, Context&);                         \
  template void compress_batched<T>(const std::vector<CompressDesc<T>>&, Context&);                         \
  template void getrf_tiles<T>(const std::vector<T*>&, int, int64_t, int32_t*, Context&);                   \
  template std::unique_ptr<BlrMatrix<T>> build_blr_kernel<T>(const KernelDesc&, int, double, int, double, \
                                                             int, int, Context&);                           \
  template void blr_getrf<T>(BlrMatrix<T>&, Context&);                                                      \
  template void blr_solve<T>(BlrMatrix<T>&, T*, int, bool, Context&);                                       \
  template void blr_matvec<T>(BlrMatrix<T>&, const T*, T*, int, Context&);

FRK_ENGINE_INST(double)
FRK_ENGINE_INST(float)

}  // namespace frk
