// Tiled Householder BLR-QR (src/operations/LAPACK/qr.cpp:213-257) and its
// tile operations geqrt / larfb / tpqrt / tpmqrt (src/operations/LAPACK/
// {geqrt,larfb,tpqrt,tpmqrt}.cpp), plus the typed batched gemm they need
// (src/operations/BLAS/gemm.cpp:126-272).
//
// Device representation of the compact-WY factors: for a diagonal tile the
// reflectors stay below the diagonal of the tile (LAPACK geqrt layout) and an
// explicit unit-lower copy Yx is kept for the blocked updates; for a tile
// below the diagonal the reflectors' lower parts replace the tile (dense), or
// its V factor with S set to the identity (low rank, tpqrt.cpp:35-53).  The
// triangular factors T are formed from G = W^T W by LAPACK's larft recurrence
// on 64 x 64 diagonal blocks and merged by T_AD = -T_AA G_AD T_DD.
#include <algorithm>
#include <stdexcept>

#include "engine.hpp"
#include "plan.hpp"

namespace frk {

// ============================================================================
// Typed operand views and batched typed gemm
// ============================================================================
template <typename T>
Op<T> tile_op(const BlrMatrix<T>& A, int i, int j) {
  Op<T> o{};
  const int t = A.idx(i, j);
  o.kind = A.kind[t];
  o.m = (int)A.rows(i);
  o.n = (int)A.cols(j);
  if (o.kind == kDense) {
    o.d = A.D[t];
    o.d_rs = A.cols(j);
    o.d_cs = 1;
  } else if (o.kind == kLowRank) {
    o.r = A.rank[t];
    o.u = A.U[t];
    o.u_rs = 1;
    o.u_cs = A.rows(i);
    o.s = A.S[t];
    o.s_rs = A.kcap;
    o.s_cs = 1;
    o.v = A.V[t];
    o.v_rs = A.cols(j);
    o.v_cs = 1;
  }
  return o;
}

template <typename T>
Op<T> transpose_op(const Op<T>& a) {
  Op<T> o = a;
  o.m = a.n;
  o.n = a.m;
  if (a.kind == kDense) {
    o.d_rs = a.d_cs;
    o.d_cs = a.d_rs;
  } else if (a.kind == kLowRank) {
    o.u = a.v;
    o.u_rs = a.v_cs;
    o.u_cs = a.v_rs;
    o.s_rs = a.s_cs;
    o.s_cs = a.s_rs;
    o.v = a.u;
    o.v_rs = a.u_cs;
    o.v_cs = a.u_rs;
  }
  return o;
}

template <typename T>
Op<T> dense_op(const T* d, int m, int n, int64_t rs, int64_t cs) {
  Op<T> o{};
  o.kind = kDense;
  o.m = m;
  o.n = n;
  o.d = d;
  o.d_rs = rs;
  o.d_cs = cs;
  return o;
}

template <typename T>
Target<T> tile_target(BlrMatrix<T>& A, int i, int j) {
  Target<T> c{};
  const int t = A.idx(i, j);
  c.kind = A.kind[t];
  c.m = (int)A.rows(i);
  c.n = (int)A.cols(j);
  if (c.kind == kDense) {
    c.d = A.D[t];
    c.d_rs = A.cols(j);
    c.d_cs = 1;
  } else {
    c.u = A.U[t];
    c.s = A.S[t];
    c.v = A.V[t];
    c.cap = A.kcap;
    c.cap_trunc = A.cap_trunc;
    c.rank = A.rank[t];
    c.d_rank = A.d_rank(i, j);
    c.eps = A.eps;
    c.fixed_rank = A.fixed_rank;
  }
  return c;
}

template <typename T>
Op<T> target_op(const Target<T>& c) {
  Op<T> o{};
  o.kind = c.kind;
  o.m = c.m;
  o.n = c.n;
  if (c.kind == kDense) {
    o.d = c.d;
    o.d_rs = c.d_rs;
    o.d_cs = c.d_cs;
  } else {
    o.r = c.rank;
    o.u = c.u;
    o.u_rs = c.u_rs;
    o.u_cs = c.u_cs ? c.u_cs : c.m;
    o.s = c.s;
    o.s_rs = c.cap;
    o.s_cs = 1;
    o.v = c.v;
    o.v_rs = c.n;
    o.v_cs = 1;
  }
  return o;
}

// Upper bound of the workspace one typed_gemm may take.
template <typename T>
size_t typed_gemm_workspace(const Op<T>& A, const Op<T>& B, const Target<T>& C) {
  const size_t m = C.m, n = C.n, k = A.n;
  const size_t ra = std::max(A.r, 1), rb = std::max(B.r, 1), cap = std::max(C.cap, 64);
  size_t w = sizeof(T) * (2 * (m + n + k) * (ra + rb) + 4 * (ra + rb) * (ra + rb) + 8 * 64);
  if (C.kind == kLowRank && A.kind == kDense && B.kind == kDense)
    w += sizeof(T) * (m * n + (m + n) * cap + cap * cap + 3 * 64);
  return w;
}

template <typename T>
static T* take_ws(Context& ctx, size_t n) {
  return ctx.take<T>(n + 1);
}

// C = alpha A B + beta C for Dense/LowRank operands (gemm.cpp:126-272); LowRank
// targets only with beta = 1 (every caller here passes 1).  Appends the
// products to the batch's stages and the additions to its list.
template <typename T>
void typed_gemm(TypedBatch<T>& b, Context& ctx, const Op<T>& A, const Op<T>& B, const Target<T>& C, double alpha,
                double beta) {
  if (A.n != B.m || A.m != C.m || B.n != C.n) throw std::invalid_argument("typed gemm: shape mismatch");
  const int m = C.m, n = C.n, k = A.n;
  if (C.kind == kDense) {
    T* c = C.d;
    if (A.kind == kDense && B.kind == kDense) {  // :126-146
      b.g[0].push_back(gemm_item(A.d, A.d_rs, A.d_cs, B.d, B.d_rs, B.d_cs, c, C.d_rs, C.d_cs, m, n, k, alpha, beta));
    } else if (A.kind == kLowRank && B.kind == kDense) {  // :148-162
      const int r = A.r;
      T* inner = take_ws<T>(ctx, (size_t)r * n);
      T* asb = take_ws<T>(ctx, (size_t)r * n);
      b.g[0].push_back(gemm_item(A.v, A.v_rs, A.v_cs, B.d, B.d_rs, B.d_cs, inner, n, 1, r, n, k, alpha, 0.0));
      b.g[1].push_back(gemm_item(A.s, A.s_rs, A.s_cs, inner, n, 1, asb, n, 1, r, n, r, 1.0, 0.0));
      b.g[2].push_back(gemm_item(A.u, A.u_rs, A.u_cs, asb, n, 1, c, C.d_rs, C.d_cs, m, n, r, 1.0, beta));
    } else if (A.kind == kDense && B.kind == kLowRank) {  // :164-178
      const int r = B.r;
      T* ab = take_ws<T>(ctx, (size_t)m * r);
      T* abs = take_ws<T>(ctx, (size_t)m * r);
      b.g[0].push_back(gemm_item(A.d, A.d_rs, A.d_cs, B.u, B.u_rs, B.u_cs, ab, r, 1, m, r, k, alpha, 0.0));
      b.g[1].push_back(gemm_item(ab, r, 1, B.s, B.s_rs, B.s_cs, abs, r, 1, m, r, r, 1.0, 0.0));
      b.g[2].push_back(gemm_item(abs, r, 1, B.v, B.v_rs, B.v_cs, c, C.d_rs, C.d_cs, m, n, r, 1.0, beta));
    } else {  // LR LR D :180-200
      const int ra = A.r, rb = B.r;
      T* inner = take_ws<T>(ctx, (size_t)ra * rb);
      T* x = take_ws<T>(ctx, (size_t)ra * rb);
      T* y = take_ws<T>(ctx, (size_t)ra * rb);
      T* z = take_ws<T>(ctx, (size_t)m * rb);
      b.g[0].push_back(gemm_item(A.v, A.v_rs, A.v_cs, B.u, B.u_rs, B.u_cs, inner, rb, 1, ra, rb, k, alpha, 0.0));
      b.g[1].push_back(gemm_item(A.s, A.s_rs, A.s_cs, inner, rb, 1, x, rb, 1, ra, rb, ra, 1.0, 0.0));
      b.g[2].push_back(gemm_item(x, rb, 1, B.s, B.s_rs, B.s_cs, y, rb, 1, ra, rb, rb, 1.0, 0.0));
      b.g[3].push_back(gemm_item(A.u, A.u_rs, A.u_cs, y, rb, 1, z, rb, 1, m, rb, ra, 1.0, 0.0));
      b.g[4].push_back(gemm_item(z, rb, 1, B.v, B.v_rs, B.v_cs, c, C.d_rs, C.d_cs, m, n, rb, 1.0, beta));
    }
    return;
  }
  if (beta != 1.0) throw std::invalid_argument("typed gemm: low-rank targets support beta = 1 only");
  LrAddDesc<T> ad{};
  ad.cu = C.u;
  ad.cs = C.s;
  ad.cv = C.v;
  ad.d_rank = C.d_rank;
  ad.kc = C.rank;
  ad.cu_rs = C.u_rs;
  ad.cu_cs = C.u_cs ? C.u_cs : m;
  ad.cv_rs = n;
  ad.cv_cs = 1;
  ad.lds = C.cap;
  ad.m = m;
  ad.n = n;
  ad.eps = C.eps;
  ad.cap = C.cap;
  ad.cap_trunc = C.cap_trunc;
  // S of the addend must be row-major: materialise a transposed view
  auto s_rowmajor = [&](const Op<T>& o, int& ld) -> const T* {
    if (o.s_cs == 1) {
      ld = (int)o.s_rs;
      return o.s;
    }
    T* w = take_ws<T>(ctx, (size_t)o.r * o.r);
    b.cp.push_back(gemm_item(o.s, o.s_rs, o.s_cs, nullptr, 0, 0, w, o.r, 1, o.r, o.r, 0, 0.0, 0.0));
    ld = o.r;
    return w;
  };
  if (A.kind == kLowRank && B.kind == kDense) {  // :219-234
    const int r = A.r;
    T* w = take_ws<T>(ctx, (size_t)r * n);
    b.g[0].push_back(gemm_item(A.v, A.v_rs, A.v_cs, B.d, B.d_rs, B.d_cs, w, n, 1, r, n, k, alpha, 0.0));
    ad.xu = A.u;
    ad.xu_rs = A.u_rs;
    ad.xu_cs = A.u_cs;
    ad.kx1 = ad.kx2 = r;
    ad.xs = s_rowmajor(A, ad.xs_ld);
    ad.xv = w;
    ad.xv_rs = n;
    ad.xv_cs = 1;
  } else if (A.kind == kDense && B.kind == kLowRank) {  // :236-251
    const int r = B.r;
    T* w = take_ws<T>(ctx, (size_t)m * r);
    b.g[0].push_back(gemm_item(A.d, A.d_rs, A.d_cs, B.u, B.u_rs, B.u_cs, w, 1, m, m, r, k, alpha, 0.0));
    ad.xu = w;
    ad.xu_rs = 1;
    ad.xu_cs = m;
    ad.kx1 = ad.kx2 = r;
    ad.xs = s_rowmajor(B, ad.xs_ld);
    ad.xv = B.v;
    ad.xv_rs = B.v_rs;
    ad.xv_cs = B.v_cs;
  } else if (A.kind == kLowRank && B.kind == kLowRank) {  // :253-272
    const int ra = A.r, rb = B.r;
    T* left = take_ws<T>(ctx, (size_t)ra * k);
    T* right = take_ws<T>(ctx, (size_t)k * rb);
    T* sx = take_ws<T>(ctx, (size_t)ra * rb);
    b.g[0].push_back(gemm_item(A.s, A.s_rs, A.s_cs, A.v, A.v_rs, A.v_cs, left, k, 1, ra, k, ra, 1.0, 0.0));
    b.g[0].push_back(gemm_item(B.u, B.u_rs, B.u_cs, B.s, B.s_rs, B.s_cs, right, rb, 1, k, rb, rb, 1.0, 0.0));
    b.g[1].push_back(gemm_item(left, k, 1, right, rb, 1, sx, rb, 1, ra, rb, k, alpha, 0.0));
    ad.xu = A.u;
    ad.xu_rs = A.u_rs;
    ad.xu_cs = A.u_cs;
    ad.kx1 = ra;
    ad.xs = sx;
    ad.xs_ld = rb;
    ad.kx2 = rb;
    ad.xv = B.v;
    ad.xv_rs = B.v_rs;
    ad.xv_cs = B.v_cs;
  } else {  // D D LR :202-217: compress the dense product, then add
    T* prod = take_ws<T>(ctx, (size_t)m * n);
    b.g[0].push_back(gemm_item(A.d, A.d_rs, A.d_cs, B.d, B.d_rs, B.d_cs, prod, n, 1, m, n, k, alpha, 0.0));
    const int cap = C.cap;
    T* xu = take_ws<T>(ctx, (size_t)m * cap);
    T* xs = take_ws<T>(ctx, (size_t)cap * cap);
    T* xv = take_ws<T>(ctx, (size_t)cap * n);
    CompressDesc<T> c{};
    c.a = prod;
    c.m = m;
    c.n = n;
    c.lda = n;
    c.eps = C.eps;
    c.rank = C.eps != 0.0 ? 0 : C.rank;
    c.u = xu;
    c.s = xs;
    c.v = xv;
    c.u_rs = 1;
    c.u_cs = m;
    c.v_rs = n;
    c.v_cs = 1;
    c.lds = cap;
    c.cap = cap;
    c.cap_trunc = C.cap_trunc;
    b.comp.push_back(c);
    b.comp_add.push_back((int)b.adds.size());
    ad.xu = xu;
    ad.xu_rs = 1;
    ad.xu_cs = m;
    ad.xs = xs;
    ad.xs_ld = cap;
    ad.xv = xv;
    ad.xv_rs = n;
    ad.xv_cs = 1;
    ad.kx1 = ad.kx2 = -1;
  }
  b.adds.push_back(ad);
}

// Run every stage of the batch in order, then the compressions and additions;
// drains the stream.
template <typename T>
void run_typed_batch(TypedBatch<T>& b, Context& ctx, int lra) {
  {
    Plan pl(ctx);
    plan_copies<T>(pl, b.cp);
    for (auto& g : b.g) plan_gemms<T>(pl, g, "typed_gemm");
    pl.run();
  }
  if (!b.comp.empty()) {
    int32_t* crank = static_cast<int32_t*>(ctx.persistent(4, sizeof(int32_t) * b.comp.size()));
    for (size_t c = 0; c < b.comp.size(); ++c) b.comp[c].d_rank = crank + c;
    Context cside(ctx.stream());
    cside.set_stats_owner(&ctx);
    compress_batched<T>(b.comp, cside);
    std::vector<int32_t> cr(b.comp.size());
    FRK_HIP_CHECK(hipMemcpyAsync(cr.data(), crank, sizeof(int32_t) * cr.size(), hipMemcpyDeviceToHost, ctx.stream()));
    host_wait_stream(ctx.stream(), "typed batch compression ranks");
    for (size_t c = 0; c < b.comp.size(); ++c) {
      if (cr[c] < 0) throw std::runtime_error("low-rank capacity exceeded in a dense-product compression");
      b.adds[b.comp_add[c]].kx1 = b.adds[b.comp_add[c]].kx2 = cr[c];
    }
  }
  if (!b.adds.empty()) {
    Plan pl(ctx);
    plan_lr_add<T>(pl, b.adds, lra);
    pl.run();
  }
  ctx.sync();
}

// ============================================================================
// Compact-WY building blocks
// ============================================================================
// T (n x n, row-major) of the reflectors whose Gram matrix is G: larft on the
// 64 x 64 diagonal blocks, then T_AD = -T_AA G_AD T_DD with doubling block size.
template <typename T>
static void plan_larft_blocked(Plan& pl, const T* G, int64_t ldg, const T* tau, T* Tm, int64_t ldt, int n) {
  Context& ctx = pl.ctx();
  const hipStream_t s = ctx.stream();
  std::vector<LarftItem> items;
  for (int p = 0; p < n; p += 64) {
    LarftItem it{};
    it.G = G + (int64_t)p * ldg + p;
    it.g_rs = ldg;
    it.g_cs = 1;
    it.tau = tau + p;
    it.T = Tm + (int64_t)p * ldt + p;
    it.t_rs = ldt;
    it.t_cs = 1;
    it.K = std::min(64, n - p);
    items.push_back(it);
  }
  const size_t off = pl.add(items);
  const int ni = (int)items.size();
  double kf = 0;
  for (const auto& it : items) kf += (double)it.K * it.K * it.K / 3.0;
  pl.launch([=](const char* d) { launch_larft<T>(reinterpret_cast<const LarftItem*>(d + off), ni, s); }, "larft",
            kf, sizeof(T) * 2.0 * n * 64);
  for (int sz = 64; sz < n; sz *= 2) {
    std::vector<GemmItem> w1, w2;
    for (int p = 0; p + sz < n; p += 2 * sz) {
      const int d0 = p + sz, sd = std::min(sz, n - d0);
      T* W = take_ws<T>(ctx, (size_t)sz * sd);
      w1.push_back(gemm_item(G + (int64_t)p * ldg + d0, ldg, 1, Tm + (int64_t)d0 * ldt + d0, ldt, 1, W, sd, 1, sz, sd,
                             sd, 1.0, 0.0));
      w2.push_back(gemm_item(Tm + (int64_t)p * ldt + p, ldt, 1, W, sd, 1, Tm + (int64_t)p * ldt + d0, ldt, 1, sz, sd,
                             sz, -1.0, 0.0));
    }
    plan_gemms<T>(pl, w1, "qr_gemm");
    plan_gemms<T>(pl, w2, "qr_gemm");
  }
}

template <typename T>
static size_t larft_workspace(int n) {
  size_t w = 4096;
  for (int sz = 64; sz < n; sz *= 2) w += sizeof(T) * ((size_t)((n + 2 * sz - 1) / (2 * sz)) * sz * sz + 64);
  return w;
}

// geqrt (geqrt.cpp:34-43): R above / reflectors below the diagonal of the
// b x b tile A, tau[b], T (b x b), explicit unit-lower Yx (b x b).
template <typename T>
static void plan_geqrt(Plan& pl, T* A, int64_t ld, int b, T* tau, T* Tm, T* Yx) {
  Context& ctx = pl.ctx();
  const hipStream_t s = ctx.stream();
  if (b > 512) throw std::invalid_argument("geqrt: tile above 512");
  for (int p = 0; p < b; p += 64) {
    const int kb = std::min(64, b - p), m = b - p;
    QrItem q{};
    q.A1 = A + (int64_t)p * ld + p;
    q.a1_rs = ld;
    q.a1_cs = 1;
    q.k1a = q.k1 = kb;
    q.k2 = 0;
    q.m = m;
    q.Y = A + (int64_t)p * ld + p;
    q.y_rs = ld;
    q.y_cs = 1;
    q.tau = tau + p;
    q.R = A + (int64_t)p * ld + p;
    q.ldr = (int32_t)ld;
    q.r_upper_only = 1;
    const size_t oq = pl.add(std::vector<QrItem>{q});
    pl.launch([=](const char* d) { launch_qr<T>(reinterpret_cast<const QrItem*>(d + oq), 1, kb, s); }, "geqrt",
              4.0 * m * kb * kb - 4.0 / 3.0 * kb * kb * kb, sizeof(T) * 2.0 * m * kb);
    const int nt = b - p - kb;
    if (nt <= 0) continue;
    std::vector<ApplyQItem> ai;
    for (int c0 = 0; c0 < nt; c0 += 128) {
      ApplyQItem a{};
      a.Y = A + (int64_t)p * ld + p;
      a.y_rs = ld;
      a.y_cs = 1;
      a.tau = tau + p;
      a.m = m;
      a.nref = std::min(m, kb);
      a.X = A + (int64_t)p * ld + p + kb + c0;
      a.x_rs = ld;
      a.x_cs = 1;
      a.kx = m;
      a.ncol = std::min(128, nt - c0);
      a.out = A + (int64_t)p * ld + p + kb + c0;
      a.out_rs = ld;
      a.out_cs = 1;
      a.trans = 1;
      ai.push_back(a);
    }
    const size_t oa = pl.add(ai);
    const int na = (int)ai.size();
    const int mc = std::min(128, nt);
    pl.launch([=](const char* d) { launch_apply_q<T>(reinterpret_cast<const ApplyQItem*>(d + oa), na, mc, s); },
              "geqrt", 4.0 * m * kb * nt, sizeof(T) * 2.0 * m * nt);
  }
  pl.launch([=](const char*) { launch_tri_extract<T>(A, ld, 1, Yx, b, b, 0, s); });
  T* G = take_ws<T>(ctx, (size_t)b * b);
  plan_gemms<T>(pl, {gemm_item(Yx, 1, b, Yx, b, 1, G, b, 1, b, b, b, 1.0, 0.0)}, "qr_gemm");
  plan_larft_blocked<T>(pl, G, b, tau, Tm, b, b);
}

// larfb (larfb.cpp:37-56) batched: X <- (I - Yx op(T) Yx^T) X with op(T) = T^T
// for trans (Q^T X) and T otherwise; X is b x nx (strided).
template <typename T>
struct LarfbJob {
  T* X;
  int64_t x_rs, x_cs;
  int nx;
};

template <typename T>
static void plan_larfb(Plan& pl, const T* Yx, const T* Tm, int b, bool trans, const std::vector<LarfbJob<T>>& jobs) {
  Context& ctx = pl.ctx();
  std::vector<GemmItem> g1, g2, g3;
  for (const auto& j : jobs) {
    if (j.nx <= 0) continue;
    T* w1 = take_ws<T>(ctx, (size_t)b * j.nx);
    T* w2 = take_ws<T>(ctx, (size_t)b * j.nx);
    g1.push_back(gemm_item(Yx, 1, b, j.X, j.x_rs, j.x_cs, w1, j.nx, 1, b, j.nx, b, 1.0, 0.0));
    if (trans)
      g2.push_back(gemm_item(Tm, 1, b, w1, j.nx, 1, w2, j.nx, 1, b, j.nx, b, 1.0, 0.0));
    else
      g2.push_back(gemm_item(Tm, b, 1, w1, j.nx, 1, w2, j.nx, 1, b, j.nx, b, 1.0, 0.0));
    g3.push_back(gemm_item(Yx, b, 1, w2, j.nx, 1, j.X, j.x_rs, j.x_cs, b, j.nx, b, -1.0, 1.0));
  }
  plan_gemms<T>(pl, g1, "qr_gemm");
  plan_gemms<T>(pl, g2, "qr_gemm");
  plan_gemms<T>(pl, g3, "qr_gemm");
}

// ============================================================================
// QR factor storage
// ============================================================================
template <typename T>
QrFactors<T>::QrFactors(const BlrMatrix<T>& A) : nb_r(A.nb_r()), nb_c(A.nb_c()) {
  const int nk = std::min(nb_r, nb_c);
  size_t total = 0;
  auto al = [](size_t x) { return (x + 63) & ~size_t(63); };
  for (int k = 0; k < nk; ++k) {
    const size_t b = A.cols(k);
    total += al(b * b) + al(b);  // Yx, tau of the diagonal tile
    for (int i = k; i < nb_r; ++i) total += al(b * b) + al(b);
  }
  buf_.ensure(std::max<size_t>(total, 1) * sizeof(T));
  // the factors are written on a non-blocking stream: finish the fill first
  FRK_HIP_CHECK(hipMemset(buf_.as<T>(), 0, std::max<size_t>(total, 1) * sizeof(T)));
  FRK_HIP_CHECK(hipDeviceSynchronize());
  Tt.assign(nb_r * nb_c, nullptr);
  tau.assign(nb_r * nb_c, nullptr);
  Yx.assign(nk, nullptr);
  bsize.assign(nk, 0);
  T* p = buf_.as<T>();
  for (int k = 0; k < nk; ++k) {
    const size_t b = A.cols(k);
    bsize[k] = (int)b;
    Yx[k] = p;
    p += al(b * b);
    p += al(b);
    for (int i = k; i < nb_r; ++i) {
      Tt[i * nb_c + k] = p;
      p += al(b * b);
      tau[i * nb_c + k] = p;
      p += al(b);
    }
  }
}

// ============================================================================
// tpqrt / tpmqrt on BLR tiles
// ============================================================================
// tpqrt(A, B, T) (tpqrt.cpp:35-53) on raw operands: R (b x b, upper
// triangle, row stride ldr) is updated; B is dense (m x b, row stride ld,
// overwritten by its reflector parts) or low rank (S with leading dimension
// lds, V r x b row-major: V <- S V, S <- I, then V is overwritten by the
// reflector parts, tpqrt.cpp:35-53).  tau (b) and T (b x b, row stride b)
// receive the reflector scalars and the triangular factor.
template <typename T>
struct TpqrtB {
  int kind = kDense;
  T* d = nullptr;
  int64_t ld = 0;
  int m = 0;
  T* s = nullptr;
  T* v = nullptr;
  int r = 0, lds = 0;
};

template <typename T>
static void tpqrt_raw(T* R, int64_t ldr, int b, const TpqrtB<T>& B, T* tau, T* Tm, Context& ctx) {
  const hipStream_t s = ctx.stream();
  if (B.kind == kLowRank) {
    const int r = B.r;
    if (r > 64) throw std::runtime_error("tpqrt: low-rank block above rank 64");
    ctx.reserve(sizeof(T) * ((size_t)r * b + (size_t)b * b + 256) + larft_workspace<T>(b) + 8192);
    Plan pl(ctx);
    T* Bv = take_ws<T>(ctx, (size_t)r * b);
    plan_gemms<T>(pl, {gemm_item(B.s, B.lds, 1, B.v, b, 1, Bv, b, 1, r, b, r, 1.0, 0.0)}, "qr_gemm");
    TpqrtItem t{};
    t.R = R;
    t.r_rs = ldr;
    t.r_cs = 1;
    t.B = Bv;
    t.b_rs = b;
    t.b_cs = 1;
    t.tau = tau;
    t.m = r;
    t.K = b;
    const size_t ot = pl.add(std::vector<TpqrtItem>{t});
    pl.launch([=](const char* d) { launch_tpqrt<T>(reinterpret_cast<const TpqrtItem*>(d + ot), 1, r, b, s); },
              "tpqrt", 2.0 * r * b * b, sizeof(T) * (2.0 * r * b + b * b));
    T* V = B.v;
    T* S = B.s;
    const int lds = B.lds;
    plan_copies<T>(pl, {gemm_item(Bv, b, 1, nullptr, 0, 0, V, b, 1, r, b, 0, 0.0, 0.0)});
    pl.launch([=](const char*) { launch_fill_identity<T>(S, r, r, lds, s); });
    T* G = take_ws<T>(ctx, (size_t)b * b);
    plan_gemms<T>(pl, {gemm_item(Bv, 1, b, Bv, b, 1, G, b, 1, b, b, r, 1.0, 0.0)}, "qr_gemm");
    plan_larft_blocked<T>(pl, G, b, tau, Tm, b, b);
    pl.run();
    ctx.sync();
    return;
  }
  if (B.kind != kDense) throw std::runtime_error("tpqrt: empty block");
  const int m = B.m;
  T* Bd = B.d;
  const int64_t ldb = B.ld;
  if (m > 512) throw std::runtime_error("tpqrt: dense block above 512 rows");
  ctx.reserve(sizeof(T) * (4 * (size_t)b * b + (size_t)64 * 64 * 2 + 1024) + larft_workspace<T>(b) + 8192);
  Plan pl(ctx);
  const T* I = ctx.identity<T>(64);
  const int ild = ctx.ident_ld();
  for (int p = 0; p < b; p += 64) {
    const int kb = std::min(64, b - p);
    TpqrtItem t{};
    t.R = R + (int64_t)p * ldr + p;
    t.r_rs = ldr;
    t.r_cs = 1;
    t.B = Bd + p;
    t.b_rs = ldb;
    t.b_cs = 1;
    t.tau = tau + p;
    t.m = m;
    t.K = kb;
    const size_t ot = pl.add(std::vector<TpqrtItem>{t});
    pl.launch([=](const char* d) { launch_tpqrt<T>(reinterpret_cast<const TpqrtItem*>(d + ot), 1, m, kb, s); },
              "tpqrt", 2.0 * m * kb * kb, sizeof(T) * 2.0 * m * kb);
    const int nt = b - p - kb;
    if (nt <= 0) continue;
    // trailing columns of [R_p; B] <- H_p^T [R_p; B] in compact WY form
    T* Gp = take_ws<T>(ctx, (size_t)kb * kb);
    T* Tp = take_ws<T>(ctx, (size_t)kb * kb);
    T* W = take_ws<T>(ctx, (size_t)kb * nt);
    T* W2 = take_ws<T>(ctx, (size_t)kb * nt);
    const T* Vp = Bd + p;
    T* Rt = R + (int64_t)p * ldr + p + kb;
    T* Bt = Bd + p + kb;
    plan_gemms<T>(pl, {gemm_item(Vp, 1, ldb, Vp, ldb, 1, Gp, kb, 1, kb, kb, m, 1.0, 0.0)}, "qr_gemm");
    plan_larft_blocked<T>(pl, Gp, kb, tau + p, Tp, kb, kb);
    plan_copies<T>(pl, {gemm_item(Rt, ldr, 1, nullptr, 0, 0, W, nt, 1, kb, nt, 0, 0.0, 0.0)});
    plan_gemms<T>(pl, {gemm_item(Vp, 1, ldb, Bt, ldb, 1, W, nt, 1, kb, nt, m, 1.0, 1.0)}, "qr_gemm");
    plan_gemms<T>(pl, {gemm_item(Tp, 1, kb, W, nt, 1, W2, nt, 1, kb, nt, kb, 1.0, 0.0)}, "qr_gemm");
    plan_gemms<T>(pl, {gemm_item(I, ild, 1, W2, nt, 1, Rt, ldr, 1, kb, nt, kb, -1.0, 1.0),
                       gemm_item(Vp, ldb, 1, W2, nt, 1, Bt, ldb, 1, m, nt, kb, -1.0, 1.0)},
                  "qr_gemm");
  }
  T* G = take_ws<T>(ctx, (size_t)b * b);
  plan_gemms<T>(pl, {gemm_item(Bd, 1, ldb, Bd, ldb, 1, G, b, 1, b, b, m, 1.0, 0.0)}, "qr_gemm");
  plan_larft_blocked<T>(pl, G, b, tau, Tm, b, b);
  pl.run();
  ctx.sync();
}

// tpqrt(A_kk, A_ik, T_ik) on BLR tiles: R in A_kk's upper triangle; A_ik's
// reflector parts replace the dense tile, or V (with S <- I) of a low-rank tile.
template <typename T>
static void tpqrt_tile(BlrMatrix<T>& A, QrFactors<T>& F, int i, int k, Context& ctx) {
  const int kk = A.idx(k, k), ik = A.idx(i, k);
  TpqrtB<T> B;
  B.kind = A.kind[ik];
  if (B.kind == kLowRank) {
    B.s = A.S[ik];
    B.v = A.V[ik];
    B.r = A.rank[ik];
    B.lds = A.kcap;
  } else if (B.kind == kDense) {
    B.d = A.D[ik];
    B.ld = A.cols(k);
    B.m = (int)A.rows(i);
  }
  tpqrt_raw<T>(A.D[kk], A.cols(k), (int)A.cols(k), B, F.tau_of(i, k), F.tmat(i, k), ctx);
}

// The reflector block of tile (i, k) as an operand: dense tile, or U (I) V.
template <typename T>
static Op<T> reflector_op(const BlrMatrix<T>& Y, int i, int k) { return tile_op(Y, i, k); }

// tpmqrt(V, T, A, B, trans) (tpmqrt.cpp:45-170) for A = C(k, j), B = C(i, j)
// over all j in js, batched.  V = Y(i, k), T = T(i, k).
template <typename T>
static void tpmqrt_row(const BlrMatrix<T>& Y, QrFactors<T>& F, int i, int k, BlrMatrix<T>& Cm,
                       const std::vector<int>& js, bool trans, Context& ctx) {
  if (js.empty()) return;
  const hipStream_t s = ctx.stream();
  const int b = (int)Y.cols(k);
  const T* Tm = F.tmat(i, k);
  const Op<T> V = reflector_op(Y, i, k);
  const Op<T> Vt = transpose_op(V);
  const int lra = Cm.lra;
  ctx.sync();
  Cm.pull_ranks(s);
  // all-dense case: dtprfb ('L', T|N, 'F', 'C') directly
  std::vector<int> dense_js, typed_js;
  for (int j : js) {
    const bool all_dense = V.kind == kDense && Cm.kind[Cm.idx(k, j)] == kDense && Cm.kind[Cm.idx(i, j)] == kDense;
    (all_dense ? dense_js : typed_js).push_back(j);
  }
  if (!dense_js.empty()) {
    size_t need = 8192;
    for (int j : dense_js) need += sizeof(T) * (2 * (size_t)b * Cm.cols(j) + 128);
    ctx.reserve(need);
    Plan pl(ctx);
    const T* I = ctx.identity<T>(b);
    const int ild = ctx.ident_ld();
    std::vector<GemmItem> cp, g1, g2, g3;
    for (int j : dense_js) {
      const int n = (int)Cm.cols(j), m = (int)Cm.rows(i);
      T* Ad = Cm.D[Cm.idx(k, j)];
      T* Bd = Cm.D[Cm.idx(i, j)];
      T* W = take_ws<T>(ctx, (size_t)b * n);
      T* W2 = take_ws<T>(ctx, (size_t)b * n);
      cp.push_back(gemm_item(Ad, n, 1, nullptr, 0, 0, W, n, 1, b, n, 0, 0.0, 0.0));
      g1.push_back(gemm_item(Vt.d, Vt.d_rs, Vt.d_cs, Bd, n, 1, W, n, 1, b, n, m, 1.0, 1.0));
      g2.push_back(trans ? gemm_item(Tm, 1, b, W, n, 1, W2, n, 1, b, n, b, 1.0, 0.0)
                         : gemm_item(Tm, b, 1, W, n, 1, W2, n, 1, b, n, b, 1.0, 0.0));
      g3.push_back(gemm_item(I, ild, 1, W2, n, 1, Ad, n, 1, b, n, b, -1.0, 1.0));
      g3.push_back(gemm_item(V.d, V.d_rs, V.d_cs, W2, n, 1, Bd, n, 1, m, n, b, -1.0, 1.0));
    }
    plan_copies<T>(pl, cp);
    plan_gemms<T>(pl, g1, "qr_gemm");
    plan_gemms<T>(pl, g2, "qr_gemm");
    plan_gemms<T>(pl, g3, "qr_gemm");
    pl.run();
    ctx.sync();
  }
  if (typed_js.empty()) return;
  // C = A (a low-rank copy when A is low rank and V or B is, else dense(A))
  const int nj = (int)typed_js.size();
  const int cap = Cm.kcap;
  size_t slot = 0;
  auto al = [](size_t x) { return (x + 63) & ~size_t(63); };
  for (int j : typed_js) slot = std::max(slot, al((size_t)b * cap) + al((size_t)cap * cap) + al((size_t)cap * Cm.cols(j)) +
                                                   al((size_t)b * Cm.cols(j)));
  T* cbuf = static_cast<T*>(ctx.persistent(2, sizeof(T) * slot * nj));
  int32_t* crank = static_cast<int32_t*>(ctx.persistent(3, sizeof(int32_t) * nj));
  std::vector<Target<T>> Cj(nj);
  {
    ctx.reserve(sizeof(T) * ((size_t)b * cap * nj + 4096) + 8192);
    Plan pl(ctx);
    std::vector<GemmItem> cp;
    std::vector<GemmItem> g1, g2;
    std::vector<int32_t> r0(nj, 0);
    for (int q = 0; q < nj; ++q) {
      const int j = typed_js[q];
      const int n = (int)Cm.cols(j);
      const Op<T> Aop = tile_op(Cm, k, j);
      const Op<T> Bop = tile_op(Cm, i, j);
      Target<T>& c = Cj[q];
      T* base = cbuf + slot * q;
      c.m = b;
      c.n = n;
      if (Aop.kind == kLowRank && (V.kind == kLowRank || Bop.kind == kLowRank)) {
        c.kind = kLowRank;
        c.u = base;
        c.s = base + al((size_t)b * cap);
        c.v = c.s + al((size_t)cap * cap);
        c.cap = cap;
        c.rank = Aop.r;
        c.d_rank = crank + q;
        c.eps = Cm.eps;
        c.fixed_rank = Cm.fixed_rank;
        c.cap_trunc = Cm.cap_trunc;
        r0[q] = Aop.r;
        cp.push_back(gemm_item(Aop.u, Aop.u_rs, Aop.u_cs, nullptr, 0, 0, c.u, 1, b, b, Aop.r, 0, 0.0, 0.0));
        cp.push_back(gemm_item(Aop.s, Aop.s_rs, Aop.s_cs, nullptr, 0, 0, c.s, cap, 1, Aop.r, Aop.r, 0, 0.0, 0.0));
        cp.push_back(gemm_item(Aop.v, Aop.v_rs, Aop.v_cs, nullptr, 0, 0, c.v, n, 1, Aop.r, n, 0, 0.0, 0.0));
      } else {
        c.kind = kDense;
        c.d = base + al((size_t)b * cap) + al((size_t)cap * cap) + al((size_t)cap * n);
        c.d_rs = n;
        c.d_cs = 1;
        if (Aop.kind == kDense) {
          cp.push_back(gemm_item(Aop.d, Aop.d_rs, Aop.d_cs, nullptr, 0, 0, c.d, n, 1, b, n, 0, 0.0, 0.0));
        } else {  // dense(A) = (U S) V
          T* us = take_ws<T>(ctx, (size_t)b * Aop.r);
          g1.push_back(gemm_item(Aop.u, Aop.u_rs, Aop.u_cs, Aop.s, Aop.s_rs, Aop.s_cs, us, Aop.r, 1, b, Aop.r, Aop.r,
                                 1.0, 0.0));
          g2.push_back(gemm_item(us, Aop.r, 1, Aop.v, Aop.v_rs, Aop.v_cs, c.d, n, 1, b, n, Aop.r, 1.0, 0.0));
        }
      }
    }
    FRK_HIP_CHECK(hipMemcpyAsync(crank, r0.data(), sizeof(int32_t) * nj, hipMemcpyHostToDevice, s));
    plan_copies<T>(pl, cp);
    plan_gemms<T>(pl, g1, "qr_gemm");
    plan_gemms<T>(pl, g2, "qr_gemm");
    pl.run();
    ctx.sync();
  }
  auto pull_c_ranks = [&]() {
    std::vector<int32_t> rr(nj);
    FRK_HIP_CHECK(hipMemcpyAsync(rr.data(), crank, sizeof(int32_t) * nj, hipMemcpyDeviceToHost, s));
    host_wait_stream(s, "tpmqrt ranks");
    for (int q = 0; q < nj; ++q)
      if (Cj[q].kind == kLowRank) {
        if (rr[q] < 0) throw std::runtime_error("low-rank capacity exceeded in tpmqrt");
        Cj[q].rank = rr[q];
      }
  };
  // 1. C += V^T B
  {
    size_t need = 8192;
    for (int q = 0; q < nj; ++q) {
      const Op<T> Bop = tile_op(Cm, i, typed_js[q]);
      need += typed_gemm_workspace(Vt, Bop, Cj[q]);
    }
    need += lr_add_bound<T>(nj, b, (int)Cm.cols(0), 2 * cap);
    ctx.reserve(need);
    TypedBatch<T> tb;
    for (int q = 0; q < nj; ++q) typed_gemm(tb, ctx, Vt, tile_op(Cm, i, typed_js[q]), Cj[q], 1.0, 1.0);
    run_typed_batch(tb, ctx, lra);
    pull_c_ranks();
  }
  // 2. C = op(T) C (trmm on C's U, or on dense C)
  {
    size_t need = 8192;
    for (int q = 0; q < nj; ++q) need += sizeof(T) * ((size_t)b * (Cj[q].kind == kLowRank ? Cj[q].rank : Cj[q].n) + 64);
    ctx.reserve(need);
    Plan pl(ctx);
    std::vector<GemmItem> g1, cp;
    for (int q = 0; q < nj; ++q) {
      const Target<T>& c = Cj[q];
      const bool lr = c.kind == kLowRank;
      const int nx = lr ? c.rank : c.n;
      T* X = lr ? c.u : c.d;
      const int64_t xrs = lr ? 1 : c.d_rs, xcs = lr ? b : 1;
      T* w = take_ws<T>(ctx, (size_t)b * nx);
      g1.push_back(trans ? gemm_item(Tm, 1, b, X, xrs, xcs, w, nx, 1, b, nx, b, 1.0, 0.0)
                         : gemm_item(Tm, b, 1, X, xrs, xcs, w, nx, 1, b, nx, b, 1.0, 0.0));
      cp.push_back(gemm_item(w, nx, 1, nullptr, 0, 0, X, xrs, xcs, b, nx, 0, 0.0, 0.0));
    }
    plan_gemms<T>(pl, g1, "qr_gemm");
    plan_copies<T>(pl, cp);
    pl.run();
    ctx.sync();
  }
  // 3. A -= C ; B -= V C
  {
    Cm.pull_ranks(s);
    const T* I = ctx.identity<T>(b);
    const int ild = ctx.ident_ld();
    const Op<T> Iop = dense_op<T>(I, b, b, ild, 1);
    size_t need = 8192;
    for (int q = 0; q < nj; ++q) {
      const int j = typed_js[q];
      const Op<T> Cop = target_op(Cj[q]);
      need += typed_gemm_workspace(Iop, Cop, tile_target(Cm, k, j));
      need += typed_gemm_workspace(V, Cop, tile_target(Cm, i, j));
    }
    need += lr_add_bound<T>(2 * nj, std::max(b, (int)Cm.rows(i)), (int)Cm.cols(0), 2 * cap);
    ctx.reserve(need);
    TypedBatch<T> tb;
    for (int q = 0; q < nj; ++q) {
      const int j = typed_js[q];
      const Op<T> Cop = target_op(Cj[q]);
      typed_gemm(tb, ctx, Iop, Cop, tile_target(Cm, k, j), -1.0, 1.0);
      typed_gemm(tb, ctx, V, Cop, tile_target(Cm, i, j), -1.0, 1.0);
    }
    run_typed_batch(tb, ctx, lra);
    Cm.pull_ranks(s);
  }
}

template <typename T>
static void larfb_row(const BlrMatrix<T>& Y, QrFactors<T>& F, int k, BlrMatrix<T>& Cm, const std::vector<int>& js,
                      bool trans, Context& ctx) {
  if (js.empty()) return;
  ctx.sync();
  Cm.pull_ranks(ctx.stream());
  const int b = (int)Y.cols(k);
  size_t need = 8192;
  std::vector<LarfbJob<T>> jobs;
  for (int j : js) {
    const int t = Cm.idx(k, j);
    LarfbJob<T> jb{};
    if (Cm.kind[t] == kLowRank) {  // larfb.cpp: a low-rank C is updated through its U
      jb.X = Cm.U[t];
      jb.x_rs = 1;
      jb.x_cs = Cm.rows(k);
      jb.nx = Cm.rank[t];
    } else if (Cm.kind[t] == kDense) {
      jb.X = Cm.D[t];
      jb.x_rs = Cm.cols(j);
      jb.x_cs = 1;
      jb.nx = (int)Cm.cols(j);
    } else {
      continue;
    }
    need += sizeof(T) * (2 * (size_t)b * jb.nx + 128);
    jobs.push_back(jb);
  }
  ctx.reserve(need);
  Plan pl(ctx);
  plan_larfb<T>(pl, F.Yx[k], F.tmat(k, k), b, trans, jobs);
  pl.run();
  ctx.sync();
}

// ============================================================================
// Drivers
// ============================================================================
template <typename T>
void blr_tiled_qr(BlrMatrix<T>& A, QrFactors<T>& F, Context& ctx) {
  const int p = A.nb_r(), q = A.nb_c();
  if (p < q) throw std::invalid_argument("tiled QR needs at least as many block rows as block columns");
  for (int k = 0; k < q; ++k) {
    if (A.kind[A.idx(k, k)] != kDense || A.rows(k) != A.cols(k))
      throw std::runtime_error("tiled QR: diagonal tile is not square dense");
    for (int i = 0; i < p; ++i)
      if (A.cols(k) != A.cols(0) || A.rows(i) != A.cols(0))
        throw std::runtime_error("tiled QR: tiles must share one square size");
  }
  const hipStream_t s = ctx.stream();
  const int b = (int)A.cols(0);
  for (int k = 0; k < q; ++k) {
    ctx.sync();
    A.pull_ranks(s);
    {
      ctx.reserve(sizeof(T) * ((size_t)b * b + 256) + larft_workspace<T>(b) + 8192);
      Plan pl(ctx);
      plan_geqrt<T>(pl, A.D[A.idx(k, k)], b, b, F.tau_of(k, k), F.tmat(k, k), F.Yx[k]);
      pl.run();
      ctx.sync();
    }
    std::vector<int> js;
    for (int j = k + 1; j < q; ++j) js.push_back(j);
    larfb_row<T>(A, F, k, A, js, true, ctx);
    for (int i = k + 1; i < p; ++i) {
      A.pull_ranks(s);
      tpqrt_tile<T>(A, F, i, k, ctx);
      A.pull_ranks(s);
      tpmqrt_row<T>(A, F, i, k, A, js, true, ctx);
    }
  }
  ctx.sync();
  A.pull_ranks(s);
}

// left_multiply_tiled_reflector (qr.cpp:230-257): C <- Q C or Q^T C.
template <typename T>
void blr_apply_q(const BlrMatrix<T>& Y, QrFactors<T>& F, BlrMatrix<T>& C, bool trans, Context& ctx) {
  const int p = Y.nb_r(), q = Y.nb_c();
  if (C.nb_r() != p) throw std::invalid_argument("apply Q: block rows differ");
  std::vector<int> js;
  if (trans) {
    for (int k = 0; k < q; ++k) {
      js.clear();
      for (int j = k; j < C.nb_c(); ++j) js.push_back(j);
      larfb_row<T>(Y, F, k, C, js, true, ctx);
      for (int i = k + 1; i < p; ++i) tpmqrt_row<T>(Y, F, i, k, C, js, true, ctx);
    }
  } else {
    for (int k = q - 1; k >= 0; --k) {
      js.clear();
      for (int j = k; j < C.nb_c(); ++j) js.push_back(j);
      for (int i = p - 1; i > k; --i) tpmqrt_row<T>(Y, F, i, k, C, js, false, ctx);
      larfb_row<T>(Y, F, k, C, js, false, ctx);
    }
  }
  ctx.sync();
  C.pull_ranks(ctx.stream());
}

// ============================================================================
// Tile-level operators on caller-owned tiles (the C ABI's batched geqrt,
// larfb, tpqrt, tpmqrt and typed gemm).  Low-rank tiles use the BLR slot
// layout: U rows x cap column-major, S cap x cap and V cap x cols row-major.
// ============================================================================
template <typename T>
static Op<T> raw_op(const RawTile<T>& t) {
  Op<T> o{};
  o.kind = t.kind;
  o.m = t.rows;
  o.n = t.cols;
  if (t.kind == kDense) {
    o.d = t.d;
    o.d_rs = t.ld;
    o.d_cs = 1;
  } else {
    o.r = t.rank;
    o.u = t.u;
    o.u_rs = t.u_rs;
    o.u_cs = t.u_cs ? t.u_cs : t.rows;
    o.s = t.s;
    o.s_rs = t.cap;
    o.s_cs = 1;
    o.v = t.v;
    o.v_rs = t.cols;
    o.v_cs = 1;
  }
  return o;
}

template <typename T>
static Target<T> raw_target(const RawTile<T>& t, int32_t* d_rank) {
  Target<T> c{};
  c.kind = t.kind;
  c.m = t.rows;
  c.n = t.cols;
  if (t.kind == kDense) {
    c.d = t.d;
    c.d_rs = t.ld;
    c.d_cs = 1;
  } else {
    c.u = t.u;
    c.u_rs = t.u_rs;
    c.u_cs = t.u_cs;
    c.s = t.s;
    c.v = t.v;
    c.cap = t.cap;
    c.rank = t.rank;
    c.d_rank = d_rank;
    c.eps = t.eps;
    c.fixed_rank = t.eps == 0.0 ? t.rank : -1;
    c.cap_trunc = false;
  }
  return c;
}

// Device ranks of the low-rank tiles of a batch (one slot per tile), for the
// recompressions to write, and their return to the host descriptors.
template <typename T>
struct RankSlots {
  DevBuf buf;
  std::vector<RawTile<T>*> tiles;
  int32_t* push(std::vector<RawTile<T>*> ts, hipStream_t s) {
    tiles = std::move(ts);
    std::vector<int32_t> h(tiles.size() + 1, 0);
    for (size_t q = 0; q < tiles.size(); ++q) h[q] = tiles[q]->rank;
    buf.ensure(sizeof(int32_t) * h.size());
    FRK_HIP_CHECK(hipMemcpyAsync(buf.as<int32_t>(), h.data(), sizeof(int32_t) * h.size(), hipMemcpyHostToDevice, s));
    FRK_HIP_CHECK(hipStreamSynchronize(s));
    return buf.as<int32_t>();
  }
  void pull(hipStream_t s, const char* what) {
    std::vector<int32_t> h(tiles.size() + 1, 0);
    FRK_HIP_CHECK(hipMemcpyAsync(h.data(), buf.as<int32_t>(), sizeof(int32_t) * h.size(), hipMemcpyDeviceToHost, s));
    FRK_HIP_CHECK(hipStreamSynchronize(s));
    for (size_t q = 0; q < tiles.size(); ++q) {
      if (tiles[q]->kind != kLowRank) continue;
      if (h[q] < 0) throw std::runtime_error(std::string("low-rank capacity exceeded in ") + what);
      tiles[q]->rank = h[q];
    }
  }
};

template <typename T>
void tiles_geqrt(const std::vector<T*>& A, int64_t lda, const std::vector<T*>& Tout, int64_t ldt, int n,
                 Context& ctx) {
  const hipStream_t s = ctx.stream();
  const size_t nt = A.size();
  ctx.sync();
  ctx.reserve(nt * (sizeof(T) * (3 * (size_t)n * n + n + 4 * 64) + larft_workspace<T>(n)) + 8192);
  Plan pl(ctx);
  std::vector<GemmItem> cp;
  for (size_t t = 0; t < nt; ++t) {
    T* tau = take_ws<T>(ctx, n);
    T* Tm = take_ws<T>(ctx, (size_t)n * n);
    T* Yx = take_ws<T>(ctx, (size_t)n * n);
    FRK_HIP_CHECK(hipMemsetAsync(Tm, 0, sizeof(T) * (size_t)n * n, s));
    plan_geqrt<T>(pl, A[t], lda, n, tau, Tm, Yx);
    cp.push_back(gemm_item(Tm, n, 1, nullptr, 0, 0, Tout[t], ldt, 1, n, n, 0, 0.0, 0.0));
  }
  plan_copies<T>(pl, cp);
  pl.run();
  ctx.sync();
}

template <typename T>
void tiles_larfb(const std::vector<const T*>& V, int64_t ldv, const std::vector<const T*>& Tin, int64_t ldt,
                 std::vector<RawTile<T>>& C, int n, bool trans, Context& ctx) {
  const hipStream_t s = ctx.stream();
  ctx.sync();
  size_t need = 8192;
  for (const auto& c : C) need += sizeof(T) * (2 * (size_t)n * n + 2 * (size_t)n * std::max(c.cols, c.rank) + 4 * 64);
  ctx.reserve(need);
  Plan pl(ctx);
  std::vector<GemmItem> cp, g1, g2, g3;
  for (size_t q = 0; q < C.size(); ++q) {
    T* Yx = take_ws<T>(ctx, (size_t)n * n);
    T* Tm = take_ws<T>(ctx, (size_t)n * n);
    const T* v = V[q];
    pl.launch([=](const char*) { launch_tri_extract<T>(v, ldv, 1, Yx, n, n, 0, s); });
    cp.push_back(gemm_item(Tin[q], ldt, 1, nullptr, 0, 0, Tm, n, 1, n, n, 0, 0.0, 0.0));
    // larfb.cpp: a low-rank C is updated through its U
    const bool lr = C[q].kind == kLowRank;
    T* X = lr ? C[q].u : C[q].d;
    const int64_t xrs = lr ? C[q].u_rs : C[q].ld, xcs = lr ? (C[q].u_cs ? C[q].u_cs : C[q].rows) : 1;
    const int nx = lr ? C[q].rank : C[q].cols;
    if (nx <= 0) continue;
    T* w1 = take_ws<T>(ctx, (size_t)n * nx);
    T* w2 = take_ws<T>(ctx, (size_t)n * nx);
    g1.push_back(gemm_item(Yx, 1, n, X, xrs, xcs, w1, nx, 1, n, nx, n, 1.0, 0.0));
    g2.push_back(trans ? gemm_item(Tm, 1, n, w1, nx, 1, w2, nx, 1, n, nx, n, 1.0, 0.0)
                       : gemm_item(Tm, n, 1, w1, nx, 1, w2, nx, 1, n, nx, n, 1.0, 0.0));
    g3.push_back(gemm_item(Yx, n, 1, w2, nx, 1, X, xrs, xcs, n, nx, n, -1.0, 1.0));
  }
  plan_copies<T>(pl, cp);
  plan_gemms<T>(pl, g1, "qr_gemm");
  plan_gemms<T>(pl, g2, "qr_gemm");
  plan_gemms<T>(pl, g3, "qr_gemm");
  pl.run();
  ctx.sync();
}

template <typename T>
void tiles_tpqrt(const std::vector<T*>& R, int64_t ldr, std::vector<RawTile<T>>& B, const std::vector<T*>& Tout,
                 int64_t ldt, int n, Context& ctx) {
  const hipStream_t s = ctx.stream();
  ctx.sync();
  T* scratch = static_cast<T*>(ctx.persistent(6, sizeof(T) * ((size_t)n * n + n + 64)));
  for (size_t q = 0; q < B.size(); ++q) {
    TpqrtB<T> b;
    b.kind = B[q].kind;
    if (b.kind == kLowRank) {
      b.s = B[q].s;
      b.v = B[q].v;
      b.r = B[q].rank;
      b.lds = B[q].cap;
    } else {
      b.d = B[q].d;
      b.ld = B[q].ld;
      b.m = B[q].rows;
    }
    T* Tm = scratch;
    T* tau = Tm + (size_t)n * n;
    FRK_HIP_CHECK(hipMemsetAsync(Tm, 0, sizeof(T) * (size_t)n * n, s));
    tpqrt_raw<T>(R[q], ldr, n, b, tau, Tm, ctx);
    FRK_HIP_CHECK(hipMemcpy2DAsync(Tout[q], sizeof(T) * ldt, Tm, sizeof(T) * n, sizeof(T) * n, n,
                                   hipMemcpyDeviceToDevice, s));
    FRK_HIP_CHECK(hipStreamSynchronize(s));
  }
}

// tpmqrt(V, T, A, B, trans) (tpmqrt.cpp:45-170) for a batch of independent
// items, every operand combination through the typed gemm:
//   C = A;  C += V^T B;  C = op(T) C;  A -= C;  B -= V C
// (C low rank when A is low rank and V or B is, else dense).
template <typename T>
void tiles_tpmqrt(const std::vector<RawTile<T>>& V, const std::vector<const T*>& Tin, int64_t ldt,
                  std::vector<RawTile<T>>& A, std::vector<RawTile<T>>& B, int n, bool trans, int lra, Context& ctx) {
  const hipStream_t s = ctx.stream();
  const int nj = (int)V.size();
  ctx.sync();
  auto al = [](size_t x) { return (x + 63) & ~size_t(63); };
  size_t slot = 0;
  for (int q = 0; q < nj; ++q) {
    const size_t cap = std::max(A[q].cap, 1), nc = A[q].cols;
    slot = std::max(slot, al((size_t)n * cap) + al(cap * cap) + al(cap * nc) + al((size_t)n * nc) + al((size_t)n * n));
  }
  T* cbuf = static_cast<T*>(ctx.persistent(2, sizeof(T) * slot * nj));
  std::vector<RawTile<T>> Cr(nj);
  std::vector<const T*> Tm(nj);
  {
    size_t need = 8192;
    for (int q = 0; q < nj; ++q) need += sizeof(T) * ((size_t)n * std::max(A[q].rank, 1) + 64);
    ctx.reserve(need);
    Plan pl(ctx);
    std::vector<GemmItem> cp, g1, g2;
    for (int q = 0; q < nj; ++q) {
      const Op<T> Aop = raw_op(A[q]);
      const int nc = A[q].cols;
      T* base = cbuf + slot * q;
      const size_t cap = std::max(A[q].cap, 1);
      T* tcopy = base + al((size_t)n * cap) + al(cap * cap) + al(cap * nc) + al((size_t)n * nc);
      cp.push_back(gemm_item(Tin[q], ldt, 1, nullptr, 0, 0, tcopy, n, 1, n, n, 0, 0.0, 0.0));
      Tm[q] = tcopy;
      RawTile<T>& c = Cr[q];
      c.rows = n;
      c.cols = nc;
      if (A[q].kind == kLowRank && (V[q].kind == kLowRank || B[q].kind == kLowRank)) {
        c.kind = kLowRank;
        c.cap = (int)cap;
        c.u = base;
        c.s = base + al((size_t)n * cap);
        c.v = c.s + al(cap * cap);
        c.rank = Aop.r;
        c.eps = A[q].eps;
        cp.push_back(gemm_item(Aop.u, Aop.u_rs, Aop.u_cs, nullptr, 0, 0, c.u, 1, n, n, Aop.r, 0, 0.0, 0.0));
        cp.push_back(gemm_item(Aop.s, Aop.s_rs, Aop.s_cs, nullptr, 0, 0, c.s, c.cap, 1, Aop.r, Aop.r, 0, 0.0, 0.0));
        cp.push_back(gemm_item(Aop.v, Aop.v_rs, Aop.v_cs, nullptr, 0, 0, c.v, nc, 1, Aop.r, nc, 0, 0.0, 0.0));
      } else {
        c.kind = kDense;
        c.d = base + al((size_t)n * cap) + al(cap * cap) + al(cap * nc);
        c.ld = nc;
        if (Aop.kind == kDense) {
          cp.push_back(gemm_item(Aop.d, Aop.d_rs, Aop.d_cs, nullptr, 0, 0, c.d, nc, 1, n, nc, 0, 0.0, 0.0));
        } else {  // dense(A) = (U S) V
          T* us = take_ws<T>(ctx, (size_t)n * std::max(Aop.r, 1));
          g1.push_back(gemm_item(Aop.u, Aop.u_rs, Aop.u_cs, Aop.s, Aop.s_rs, Aop.s_cs, us, Aop.r, 1, n, Aop.r, Aop.r,
                                 1.0, 0.0));
          g2.push_back(gemm_item(us, Aop.r, 1, Aop.v, Aop.v_rs, Aop.v_cs, c.d, nc, 1, n, nc, Aop.r, 1.0, 0.0));
        }
      }
    }
    plan_copies<T>(pl, cp);
    plan_gemms<T>(pl, g1, "qr_gemm");
    plan_gemms<T>(pl, g2, "qr_gemm");
    pl.run();
    ctx.sync();
  }
  // 1. C += V^T B
  {
    RankSlots<T> rs;
    std::vector<RawTile<T>*> ts;
    for (auto& c : Cr) ts.push_back(&c);
    int32_t* dr = rs.push(ts, s);
    size_t need = 8192;
    for (int q = 0; q < nj; ++q)
      need += typed_gemm_workspace(transpose_op(raw_op(V[q])), raw_op(B[q]), raw_target(Cr[q], dr + q));
    need += lr_add_bound<T>(nj, n, n, 2 * 64 + 64);
    ctx.reserve(need);
    TypedBatch<T> tb;
    for (int q = 0; q < nj; ++q)
      typed_gemm(tb, ctx, transpose_op(raw_op(V[q])), raw_op(B[q]), raw_target(Cr[q], dr + q), 1.0, 1.0);
    run_typed_batch(tb, ctx, lra);
    rs.pull(s, "tpmqrt");
  }
  // 2. C = op(T) C (on C's U, or on dense C)
  {
    size_t need = 8192;
    for (int q = 0; q < nj; ++q)
      need += sizeof(T) * ((size_t)n * (Cr[q].kind == kLowRank ? std::max(Cr[q].rank, 1) : Cr[q].cols) + 64);
    ctx.reserve(need);
    Plan pl(ctx);
    std::vector<GemmItem> g1, cp;
    for (int q = 0; q < nj; ++q) {
      const RawTile<T>& c = Cr[q];
      const bool lr = c.kind == kLowRank;
      const int nx = lr ? c.rank : c.cols;
      if (nx <= 0) continue;
      T* X = lr ? c.u : c.d;
      const int64_t xrs = lr ? 1 : c.ld, xcs = lr ? n : 1;
      T* w = take_ws<T>(ctx, (size_t)n * nx);
      g1.push_back(trans ? gemm_item(Tm[q], 1, n, X, xrs, xcs, w, nx, 1, n, nx, n, 1.0, 0.0)
                         : gemm_item(Tm[q], n, 1, X, xrs, xcs, w, nx, 1, n, nx, n, 1.0, 0.0));
      cp.push_back(gemm_item(w, nx, 1, nullptr, 0, 0, X, xrs, xcs, n, nx, 0, 0.0, 0.0));
    }
    plan_gemms<T>(pl, g1, "qr_gemm");
    plan_copies<T>(pl, cp);
    pl.run();
    ctx.sync();
  }
  // 3. A -= C ; B -= V C
  {
    RankSlots<T> rs;
    std::vector<RawTile<T>*> ts;
    for (int q = 0; q < nj; ++q) {
      ts.push_back(&A[q]);
      ts.push_back(&B[q]);
    }
    int32_t* dr = rs.push(ts, s);
    const T* I = ctx.identity<T>(n);
    const int ild = ctx.ident_ld();
    const Op<T> Iop = dense_op<T>(I, n, n, ild, 1);
    size_t need = 8192;
    int mb = n;
    for (int q = 0; q < nj; ++q) {
      const Op<T> Cop = raw_op(Cr[q]);
      need += typed_gemm_workspace(Iop, Cop, raw_target(A[q], dr + 2 * q));
      need += typed_gemm_workspace(raw_op(V[q]), Cop, raw_target(B[q], dr + 2 * q + 1));
      mb = std::max(mb, B[q].rows);
    }
    need += lr_add_bound<T>(2 * nj, mb, n, 2 * 64 + 64);
    ctx.reserve(need);
    TypedBatch<T> tb;
    for (int q = 0; q < nj; ++q) {
      const Op<T> Cop = raw_op(Cr[q]);
      typed_gemm(tb, ctx, Iop, Cop, raw_target(A[q], dr + 2 * q), -1.0, 1.0);
      typed_gemm(tb, ctx, raw_op(V[q]), Cop, raw_target(B[q], dr + 2 * q + 1), -1.0, 1.0);
    }
    run_typed_batch(tb, ctx, lra);
    rs.pull(s, "tpmqrt");
  }
}

// gemm(A, B, C, alpha, beta, TransA, TransB) (gemm.cpp:126-272) for a batch of
// independent items, every Dense / LowRank combination; a low-rank C is
// scaled by beta first (C.S *= beta, multiplication.cpp:39-44) and then
// receives the product through the rounded addition.
template <typename T>
void tiles_gemm(const std::vector<RawTile<T>>& A, const std::vector<RawTile<T>>& B, std::vector<RawTile<T>>& C,
                double alpha, double beta, bool ta, bool tb, int lra, Context& ctx) {
  const hipStream_t s = ctx.stream();
  const int nj = (int)A.size();
  ctx.sync();
  // low-rank targets: C.S *= beta (beta = 0 empties the tile)
  if (beta != 1.0) {
    size_t need = 8192;
    for (const auto& c : C)
      if (c.kind == kLowRank) need += sizeof(T) * ((size_t)c.rank * c.rank + 64);
    ctx.reserve(need);
    Plan pl(ctx);
    std::vector<GemmItem> g, cp;
    const T* I = ctx.identity<T>(64);
    const int ild = ctx.ident_ld();
    for (auto& c : C) {
      if (c.kind != kLowRank || c.rank <= 0) continue;
      if (beta == 0.0) {
        c.rank = 0;
        continue;
      }
      T* w = take_ws<T>(ctx, (size_t)c.rank * c.rank);
      g.push_back(gemm_item(c.s, c.cap, 1, I, ild, 1, w, c.rank, 1, c.rank, c.rank, c.rank, beta, 0.0));
      cp.push_back(gemm_item(w, c.rank, 1, nullptr, 0, 0, c.s, c.cap, 1, c.rank, c.rank, 0, 0.0, 0.0));
    }
    plan_gemms<T>(pl, g, "typed_gemm");
    plan_copies<T>(pl, cp);
    pl.run();
    ctx.sync();
  }
  RankSlots<T> rs;
  std::vector<RawTile<T>*> ts;
  for (auto& c : C) ts.push_back(&c);
  int32_t* dr = rs.push(ts, s);
  size_t need = 8192;
  int mb = 1, nbx = 1, kmax = 1;
  std::vector<Op<T>> Ao(nj), Bo(nj);
  for (int q = 0; q < nj; ++q) {
    Ao[q] = ta ? transpose_op(raw_op(A[q])) : raw_op(A[q]);
    Bo[q] = tb ? transpose_op(raw_op(B[q])) : raw_op(B[q]);
    need += typed_gemm_workspace(Ao[q], Bo[q], raw_target(C[q], dr + q));
    mb = std::max(mb, C[q].rows);
    nbx = std::max(nbx, C[q].cols);
    kmax = std::max({kmax, C[q].kind == kLowRank ? C[q].cap : 0, Ao[q].r + Bo[q].r});
  }
  need += lr_add_bound<T>(nj, mb, nbx, 2 * kmax + 64);
  ctx.reserve(need);
  TypedBatch<T> tbx;
  for (int q = 0; q < nj; ++q)
    typed_gemm(tbx, ctx, Ao[q], Bo[q], raw_target(C[q], dr + q), alpha, C[q].kind == kDense ? beta : 1.0);
  run_typed_batch(tbx, ctx, lra);
  rs.pull(s, "gemm");
}

template void tiles_geqrt<double>(const std::vector<double*>&, int64_t, const std::vector<double*>&, int64_t, int,
                                  Context&);
template void tiles_larfb<double>(const std::vector<const double*>&, int64_t, const std::vector<const double*>&,
                                  int64_t, std::vector<RawTile<double>>&, int, bool, Context&);
template void tiles_tpqrt<double>(const std::vector<double*>&, int64_t, std::vector<RawTile<double>>&,
                                  const std::vector<double*>&, int64_t, int, Context&);
template void tiles_tpmqrt<double>(const std::vector<RawTile<double>>&, const std::vector<const double*>&, int64_t,
                                   std::vector<RawTile<double>>&, std::vector<RawTile<double>>&, int, bool, int,
                                   Context&);
template void tiles_gemm<double>(const std::vector<RawTile<double>>&, const std::vector<RawTile<double>>&,
                                 std::vector<RawTile<double>>&, double, double, bool, bool, int, Context&);

template struct QrFactors<double>;
template struct QrFactors<float>;
template void blr_tiled_qr<double>(BlrMatrix<double>&, QrFactors<double>&, Context&);
template void blr_tiled_qr<float>(BlrMatrix<float>&, QrFactors<float>&, Context&);
template void blr_apply_q<double>(const BlrMatrix<double>&, QrFactors<double>&, BlrMatrix<double>&, bool, Context&);
template void blr_apply_q<float>(const BlrMatrix<float>&, QrFactors<float>&, BlrMatrix<float>&, bool, Context&);
template void typed_gemm<double>(TypedBatch<double>&, Context&, const Op<double>&, const Op<double>&,
                                 const Target<double>&, double, double);
template void typed_gemm<float>(TypedBatch<float>&, Context&, const Op<float>&, const Op<float>&,
                                const Target<float>&, double, double);
template void run_typed_batch<double>(TypedBatch<double>&, Context&, int);
template void run_typed_batch<float>(TypedBatch<float>&, Context&, int);
template Op<double> tile_op<double>(const BlrMatrix<double>&, int, int);
template Op<float> tile_op<float>(const BlrMatrix<float>&, int, int);
template Op<double> transpose_op<double>(const Op<double>&);
template Op<float> transpose_op<float>(const Op<float>&);
template Target<double> tile_target<double>(BlrMatrix<double>&, int, int);
template Target<float> tile_target<float>(BlrMatrix<float>&, int, int);
template size_t typed_gemm_workspace<double>(const Op<double>&, const Op<double>&, const Target<double>&);
template size_t typed_gemm_workspace<float>(const Op<float>&, const Op<float>&, const Target<float>&);

}  // namespace frk
