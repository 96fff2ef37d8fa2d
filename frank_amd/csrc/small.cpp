// Batched small dense factorizations of the operator surface:
//   qr (src/operations/LAPACK/qr.cpp:65-90), rq (rq.cpp:43-70), svd
//   (svd.cpp:23-44), truncated_geqp3 (geqp3.cpp:71-191), rsvd
//   (src/operations/randomized/rsvd.cpp:17-29) and one_sided_id
//   (src/operations/LAPACK/id.cpp:47-72).
// They reuse the tile kernels of the recompression path: the register-resident
// Householder QR, the reflector application, the Jacobi SVD and the
// column-pivoted QR.  Every function drains the stream before returning.
#include <algorithm>
#include <stdexcept>

#include "engine.hpp"
#include "plan.hpp"

namespace frk {

template <typename T>
static T* ws(Context& ctx, size_t n) {
  return ctx.take<T>(n + 1);
}

template <typename T>
void qr_small_batched(const std::vector<SmallJob<T>>& jobs, Context& ctx) {
  if (jobs.empty()) return;
  const hipStream_t s = ctx.stream();
  const int nj = (int)jobs.size();
  int mmax = 0, kmax = 0;
  for (const auto& j : jobs) {
    if (j.m < 1 || j.n < 1 || j.m > 512 || j.n > 128) throw std::invalid_argument("qr: m <= 512, n <= 128");
    mmax = std::max(mmax, j.m);
    kmax = std::max(kmax, j.n);
  }
  ctx.sync();
  ctx.reserve((size_t)nj * sizeof(T) * ((size_t)mmax * kmax + (size_t)kmax * kmax + kmax + 256) + 8192);
  const T* I = ctx.identity<T>(std::max(mmax, kmax) + 1);
  const int ild = ctx.ident_ld();
  std::vector<QrItem> qi(nj);
  std::vector<ApplyQItem> ai(nj);
  std::vector<GemmItem> cp;
  for (int t = 0; t < nj; ++t) {
    const auto& j = jobs[t];
    const int k = std::min(j.m, j.n);
    T* Y = ws<T>(ctx, (size_t)j.m * j.n);
    T* tau = ws<T>(ctx, j.n);
    T* R = ws<T>(ctx, (size_t)j.n * j.n);
    QrItem q{};
    q.A1 = j.a;
    q.a1_rs = j.a_rs;
    q.a1_cs = j.a_cs;
    q.k1a = q.k1 = j.n;
    q.m = j.m;
    q.Y = Y;
    q.y_rs = 1;
    q.y_cs = j.m;
    q.tau = tau;
    q.R = R;
    q.ldr = j.n;
    qi[t] = q;
    ApplyQItem a{};
    a.Y = Y;
    a.y_rs = 1;
    a.y_cs = j.m;
    a.tau = tau;
    a.m = j.m;
    a.nref = k;
    a.X = I;
    a.x_rs = ild;
    a.x_cs = 1;
    a.kx = k;
    a.ncol = k;
    a.out = j.o1;
    a.out_rs = j.o1_rs;
    a.out_cs = j.o1_cs;
    a.trans = 0;
    ai[t] = a;
    cp.push_back(gemm_item(R, j.n, 1, nullptr, 0, 0, j.o2, j.o2_rs, j.o2_cs, k, j.n, 0, 0.0, 0.0));
  }
  Plan pl(ctx);
  const size_t oq = pl.add(qi), oa = pl.add(ai);
  pl.launch([=](const char* d) { launch_qr<T>(reinterpret_cast<const QrItem*>(d + oq), nj, kmax, s); }, "small_qr");
  pl.launch([=](const char* d) { launch_apply_q<T>(reinterpret_cast<const ApplyQItem*>(d + oa), nj, kmax, s); },
            "small_qr");
  plan_copies<T>(pl, cp);
  pl.run();
  ctx.sync();
}

// A = R Q through the QR of the reversed transpose Z = J_n A^T J_m (J: index
// reversal): Z = Qz Rz gives R = J_m Rz^T J_m (upper) and Q = J_m Qz^T J_n,
// the same reflectors as LAPACK's dgerqf.
template <typename T>
void rq_small_batched(const std::vector<SmallJob<T>>& jobs, Context& ctx) {
  std::vector<SmallJob<T>> z;
  ctx.sync();
  size_t need = 8192;
  for (const auto& j : jobs) {
    if (j.m > j.n) throw std::invalid_argument("rq: needs m <= n");
    need += sizeof(T) * ((size_t)j.m * j.m + 64);
  }
  ctx.reserve(need);
  std::vector<T*> rz(jobs.size());
  for (size_t t = 0; t < jobs.size(); ++t) {
    const auto& j = jobs[t];
    SmallJob<T> q{};
    q.m = j.n;
    q.n = j.m;
    // Z(i, c) = A(m-1-c, n-1-i): Householder on Z's column c is dgerqf's on row m-1-c
    q.a = j.a + (int64_t)(j.m - 1) * j.a_rs + (int64_t)(j.n - 1) * j.a_cs;
    q.a_rs = -j.a_cs;
    q.a_cs = -j.a_rs;
    // Q(m-1-c, n-1-i) = Qz(i, c)
    q.o1 = j.o2 + (int64_t)(j.m - 1) * j.o2_rs + (int64_t)(j.n - 1) * j.o2_cs;
    q.o1_rs = -j.o2_cs;
    q.o1_cs = -j.o2_rs;
    rz[t] = ctx.take<T>((size_t)j.m * j.m + 1);
    q.o2 = rz[t];
    q.o2_rs = j.m;
    q.o2_cs = 1;
    z.push_back(q);
  }
  std::vector<GemmItem> cp;
  for (size_t t = 0; t < jobs.size(); ++t) {
    const auto& j = jobs[t];
    const int m = j.m;
    // R(i, c) = Rz(m-1-c, m-1-i)
    cp.push_back(gemm_item(rz[t] + (int64_t)(m - 1) * m + (m - 1), -1, -(int64_t)m, nullptr, 0, 0, j.o1, j.o1_rs,
                           j.o1_cs, m, m, 0, 0.0, 0.0));
  }
  // the workspace holding Rz must survive the QR call: run it on a side context
  Context side(ctx.stream());
  side.set_stats_owner(&ctx);
  qr_small_batched<T>(z, side);
  Plan pl(ctx);
  plan_copies<T>(pl, cp);
  pl.run();
  ctx.sync();
}

// svd(A) (svd.cpp:23-44), K = min(m, n) <= 64: U (m x K), S (K x K diagonal,
// descending), V (K x n) with A = U S V.
template <typename T>
void svd_small_batched(const std::vector<SvdJob<T>>& jobs, Context& ctx) {
  if (jobs.empty()) return;
  const hipStream_t s = ctx.stream();
  ctx.sync();
  size_t need = 16384;
  for (const auto& j : jobs) {
    const int K = std::min(j.m, j.n);
    if (K > 64 || std::max(j.m, j.n) > 512) throw std::invalid_argument("svd: min(m, n) <= 64, max(m, n) <= 512");
    need += sizeof(T) * ((size_t)std::max(j.m, j.n) * K + 4 * (size_t)K * K + (size_t)std::max(j.m, j.n) * K + 256);
  }
  ctx.reserve(need);
  // 1. QR of A (m >= n) or of A^T (m < n)
  std::vector<SmallJob<T>> qj;
  std::vector<T*> Qb(jobs.size()), Rb(jobs.size());
  for (size_t t = 0; t < jobs.size(); ++t) {
    const auto& j = jobs[t];
    const bool tall = j.m >= j.n;
    const int K = std::min(j.m, j.n), L = std::max(j.m, j.n);
    Qb[t] = ws<T>(ctx, (size_t)L * K);
    Rb[t] = ws<T>(ctx, (size_t)K * K);
    SmallJob<T> q{};
    q.m = L;
    q.n = K;
    q.a = j.a;
    q.a_rs = tall ? j.a_rs : j.a_cs;
    q.a_cs = tall ? j.a_cs : j.a_rs;
    q.o1 = Qb[t];
    q.o1_rs = K;
    q.o1_cs = 1;
    q.o2 = Rb[t];
    q.o2_rs = K;
    q.o2_cs = 1;
    qj.push_back(q);
  }
  {
    Context side(s);
    side.set_stats_owner(&ctx);
    qr_small_batched<T>(qj, side);
  }
  // 2. Jacobi SVD of R (tall) or R^T (wide), then the outer products
  std::vector<SvdItem> si;
  std::vector<GemmItem> g, cp;
  int kmax = 1;
  for (size_t t = 0; t < jobs.size(); ++t) {
    const auto& j = jobs[t];
    const bool tall = j.m >= j.n;
    const int K = std::min(j.m, j.n), L = std::max(j.m, j.n);
    kmax = std::max(kmax, K);
    T* Um = ws<T>(ctx, (size_t)K * K);
    T* Vm = ws<T>(ctx, (size_t)K * K);
    T* W = ws<T>(ctx, (size_t)L * K);
    SvdItem v{};
    v.mode = 0;
    v.M = Rb[t];
    v.m_rs = tall ? K : 1;
    v.m_cs = tall ? 1 : K;
    v.K = K;
    v.fixed_rank = K;
    v.kdim = K;
    v.S = j.s;
    v.lds = (int)j.s_ld;
    v.Um = Um;
    v.ldu = K;
    v.Vm = Vm;
    v.ldv = K;
    v.rank_out = nullptr;
    v.cap = K;
    v.stats = ctx.counters();
    si.push_back(v);
    if (tall) {  // A = Q R = (Q Um) S Vm^T
      g.push_back(gemm_item(Qb[t], K, 1, Um, K, 1, j.u, j.u_rs, j.u_cs, L, K, K, 1.0, 0.0));
      cp.push_back(gemm_item(Vm, 1, K, nullptr, 0, 0, j.v, j.v_rs, j.v_cs, K, K, 0, 0.0, 0.0));
    } else {  // A = R^T Q^T = Um S (Q Vm)^T
      cp.push_back(gemm_item(Um, K, 1, nullptr, 0, 0, j.u, j.u_rs, j.u_cs, K, K, 0, 0.0, 0.0));
      g.push_back(gemm_item(Qb[t], K, 1, Vm, K, 1, W, K, 1, L, K, K, 1.0, 0.0));
      cp.push_back(gemm_item(W, 1, K, nullptr, 0, 0, j.v, j.v_rs, j.v_cs, K, L, 0, 0.0, 0.0));
    }
  }
  Plan pl(ctx);
  const size_t os = pl.add(si);
  const int ns = (int)si.size();
  pl.launch([=](const char* d) { launch_svd<T>(reinterpret_cast<const SvdItem*>(d + os), ns, kmax, s); }, "small_svd");
  plan_gemms<T>(pl, g, "small_svd");
  plan_copies<T>(pl, cp);
  pl.run();
  ctx.sync();
}

// truncated_geqp3 (geqp3.cpp:71-191): Q (m x r), R (r x n, columns back in the
// original order), rank r <= cap; A is read only.
template <typename T>
void geqp3_small_batched(const std::vector<Geqp3Job<T>>& jobs, Context& ctx) {
  if (jobs.empty()) return;
  const hipStream_t s = ctx.stream();
  const int nj = (int)jobs.size();
  int mmax = 0, nmax = 0;
  for (const auto& j : jobs) {
    if (j.m < 1 || j.n < 1 || j.m > 512 || j.n > 512) throw std::invalid_argument("truncated_geqp3: tiles up to 512");
    mmax = std::max(mmax, j.m);
    nmax = std::max(nmax, j.n);
  }
  ctx.sync();
  ctx.reserve((size_t)nj * (sizeof(T) * ((size_t)mmax * nmax + (size_t)std::max(mmax, nmax) + 128) +
                            sizeof(int32_t) * ((size_t)nmax + 64)) +
              8192);
  DevBuf dr(sizeof(int32_t) * nj);
  std::vector<T*> a(nj), tau(nj);
  std::vector<int32_t*> piv(nj);
  std::vector<GemmItem> cp;
  std::vector<Geqp3Item> gi(nj);
  for (int t = 0; t < nj; ++t) {
    const auto& j = jobs[t];
    a[t] = ws<T>(ctx, (size_t)j.m * j.n);
    tau[t] = ws<T>(ctx, std::max(j.m, j.n));
    piv[t] = ctx.take<int32_t>((size_t)j.n + 1);
    cp.push_back(gemm_item(j.a, j.a_rs, j.a_cs, nullptr, 0, 0, a[t], j.n, 1, j.m, j.n, 0, 0.0, 0.0));
    Geqp3Item g{};
    g.A = a[t];
    g.tau = tau[t];
    g.piv = piv[t];
    g.rank_out = dr.as<int32_t>() + t;
    g.m = j.m;
    g.n = j.n;
    g.lda = j.n;
    g.cap = j.cap;
    g.eps = j.eps;
    gi[t] = g;
  }
  {
    Plan pl(ctx);
    plan_copies<T>(pl, cp);
    const size_t og = pl.add(gi);
    pl.launch([=](const char* d) { launch_geqp3<T>(reinterpret_cast<const Geqp3Item*>(d + og), nj, s); },
              "small_geqp3");
    pl.run();
  }
  std::vector<int32_t> rk(nj);
  FRK_HIP_CHECK(hipMemcpyAsync(rk.data(), dr.as<int32_t>(), sizeof(int32_t) * nj, hipMemcpyDeviceToHost, s));
  host_wait_stream(s, __func__);
  int rmax = 1;
  for (int t = 0; t < nj; ++t) {
    if (rk[t] < 0) throw std::runtime_error("truncated_geqp3: rank above the capacity");
    rmax = std::max(rmax, rk[t]);
    if (jobs[t].rank_host) *jobs[t].rank_host = rk[t];
  }
  const T* I = ctx.identity<T>(rmax + 1);
  const int ild = ctx.ident_ld();
  std::vector<ApplyQItem> au(nj);
  std::vector<PermuteItem> pi(nj);
  for (int t = 0; t < nj; ++t) {
    const auto& j = jobs[t];
    ApplyQItem q{};
    q.Y = a[t];
    q.y_rs = j.n;
    q.y_cs = 1;
    q.tau = tau[t];
    q.X = I;
    q.x_rs = ild;
    q.x_cs = 1;
    q.kx = rk[t];
    q.ncol = rk[t];
    q.nref = std::min(rk[t], std::min(j.m, j.n));
    q.m = j.m;
    q.out = j.q;
    q.out_rs = j.q_rs;
    q.out_cs = j.q_cs;
    au[t] = q;
    PermuteItem p{};
    p.A = a[t];
    p.piv = piv[t];
    p.rank = dr.as<int32_t>() + t;
    p.out = j.r;
    p.n = j.n;
    p.lda = j.n;
    p.ldo = (int)j.r_ld;
    pi[t] = p;
  }
  Plan pl(ctx);
  const size_t oa = pl.add(au), op = pl.add(pi);
  pl.launch([=](const char* d) { launch_apply_q<T>(reinterpret_cast<const ApplyQItem*>(d + oa), nj, rmax, s); },
            "small_geqp3");
  pl.launch([=](const char* d) { launch_permute_r<T>(reinterpret_cast<const PermuteItem*>(d + op), nj, nmax, rmax, s); });
  pl.run();
  ctx.sync();
}

// one_sided_id (id.cpp:47-72): k steps of column-pivoted QR, X = R11^-1 R12,
// V = [I | X] with its columns scattered back by the pivots; cols = the first k pivots.
template <typename T>
void id_small(const T* A, int64_t a_rs, int64_t a_cs, int m, int n, int k, T* V, int64_t ldv, int32_t* cols_host,
              Context& ctx) {
  if (m < 1 || n < 1 || m > 512 || n > 512) throw std::invalid_argument("id: tiles up to 512");
  if (k < 1 || k > std::min(m, n)) throw std::invalid_argument("id: need 1 <= k <= min(m, n)");
  const hipStream_t s = ctx.stream();
  ctx.sync();
  ctx.reserve(sizeof(T) * ((size_t)m * n + (size_t)k * n + (size_t)std::max(m, n) + 256) +
              sizeof(int32_t) * ((size_t)n + 64) + 8192);
  T* a = ws<T>(ctx, (size_t)m * n);
  T* tau = ws<T>(ctx, std::max(m, n));
  T* W = ws<T>(ctx, (size_t)k * n);
  int32_t* piv = ctx.take<int32_t>((size_t)n + 1);
  DevBuf dr(sizeof(int32_t));
  Geqp3Item g{};
  g.A = a;
  g.tau = tau;
  g.piv = piv;
  g.rank_out = dr.as<int32_t>();
  g.m = m;
  g.n = n;
  g.lda = n;
  g.cap = k;
  g.eps = -1.0;
  {
    Plan pl(ctx);
    plan_copies<T>(pl, {gemm_item(A, a_rs, a_cs, nullptr, 0, 0, a, n, 1, m, n, 0, 0.0, 0.0)});
    const size_t og = pl.add(std::vector<Geqp3Item>{g});
    pl.launch([=](const char* d) { launch_geqp3<T>(reinterpret_cast<const Geqp3Item*>(d + og), 1, s); }, "small_id");
    pl.run();
    ctx.sync();
  }
  // W = [I_k | R12] then R12 <- R11^-1 R12
  {
    const T* I = ctx.identity<T>(k + 1);
    const int ild = ctx.ident_ld();
    Plan pl(ctx);
    std::vector<GemmItem> cp{gemm_item(I, ild, 1, nullptr, 0, 0, W, n, 1, k, k, 0, 0.0, 0.0)};
    if (n > k) cp.push_back(gemm_item(a + k, n, 1, nullptr, 0, 0, W + k, n, 1, k, n - k, 0, 0.0, 0.0));
    plan_copies<T>(pl, cp);
    pl.run();
    ctx.sync();
  }
  if (n > k) {
    Context side(s);
    side.set_stats_owner(&ctx);
    trsm_generic<T>(a, n, 1, false, false, W + k, n, 1, k, n - k, side);
  }
  {
    PermuteItem p{};
    p.A = W;
    p.piv = piv;
    p.rank = dr.as<int32_t>();
    p.out = V;
    p.n = n;
    p.lda = n;
    p.ldo = (int)ldv;
    Plan pl(ctx);
    const size_t op = pl.add(std::vector<PermuteItem>{p});
    pl.launch([=](const char* d) { launch_permute_r<T>(reinterpret_cast<const PermuteItem*>(d + op), 1, n, k, s); });
    pl.run();
  }
  FRK_HIP_CHECK(hipMemcpyAsync(cols_host, piv, sizeof(int32_t) * k, hipMemcpyDeviceToHost, s));
  ctx.sync();
}

// rsvd(A, p) (rsvd.cpp:17-29): Omega = random_uniform(n, p) from std::mt19937(0),
// Y = A Omega, Q = qr(Y), [Ub, S, V] = svd(Q^T A), U = Q Ub.  U m x p, S p x p, V p x n.
template <typename T>
void rsvd_small(const T* A, int64_t lda, int m, int n, int p, T* U, int64_t ldu, T* S, int64_t lds, T* V, int64_t ldv,
                Context& ctx) {
  if (p < 1 || p > std::min(m, n) || p > 64 || m > 512 || n > 512)
    throw std::invalid_argument("rsvd: need 1 <= p <= min(m, n, 64), m, n <= 512");
  ctx.sync();
  ctx.reserve(sizeof(T) * (3 * (size_t)m * p + (size_t)p * p * 2 + (size_t)p * n + 256) + 8192);
  const T* Om = ctx.omega<T>(n, p);
  T* Y = ws<T>(ctx, (size_t)m * p);
  T* Q = ws<T>(ctx, (size_t)m * p);
  T* R = ws<T>(ctx, (size_t)p * p);
  T* B = ws<T>(ctx, (size_t)p * n);
  T* Ub = ws<T>(ctx, (size_t)p * p);
  {
    Plan pl(ctx);
    plan_gemms<T>(pl, {gemm_item(A, lda, 1, Om, p, 1, Y, p, 1, m, p, n, 1.0, 0.0)}, "small_rsvd");
    pl.run();
  }
  Context side(ctx.stream());
  side.set_stats_owner(&ctx);
  qr_small_batched<T>({SmallJob<T>{Y, p, 1, Q, p, 1, R, p, 1, m, p}}, side);
  {
    Plan pl(ctx);
    plan_gemms<T>(pl, {gemm_item(Q, 1, p, A, lda, 1, B, n, 1, p, n, m, 1.0, 0.0)}, "small_rsvd");
    pl.run();
  }
  svd_small_batched<T>({SvdJob<T>{B, n, 1, Ub, p, 1, S, lds, V, ldv, 1, p, n}}, side);
  Plan pl(ctx);
  plan_gemms<T>(pl, {gemm_item(Q, p, 1, Ub, p, 1, U, ldu, 1, m, p, p, 1.0, 0.0)}, "small_rsvd");
  pl.run();
  ctx.sync();
}

template void rsvd_small<double>(const double*, int64_t, int, int, int, double*, int64_t, double*, int64_t, double*,
                                 int64_t, Context&);
template void rsvd_small<float>(const float*, int64_t, int, int, int, float*, int64_t, float*, int64_t, float*,
                                int64_t, Context&);
template void qr_small_batched<double>(const std::vector<SmallJob<double>>&, Context&);
template void qr_small_batched<float>(const std::vector<SmallJob<float>>&, Context&);
template void rq_small_batched<double>(const std::vector<SmallJob<double>>&, Context&);
template void rq_small_batched<float>(const std::vector<SmallJob<float>>&, Context&);
template void svd_small_batched<double>(const std::vector<SvdJob<double>>&, Context&);
template void svd_small_batched<float>(const std::vector<SvdJob<float>>&, Context&);
template void geqp3_small_batched<double>(const std::vector<Geqp3Job<double>>&, Context&);
template void geqp3_small_batched<float>(const std::vector<Geqp3Job<float>>&, Context&);
template void id_small<double>(const double*, int64_t, int64_t, int, int, int, double*, int64_t, int32_t*, Context&);
template void id_small<float>(const float*, int64_t, int64_t, int, int, int, float*, int64_t, int32_t*, Context&);

}  // namespace frk
