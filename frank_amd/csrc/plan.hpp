// Host-side launch planning helpers shared by the engine's translation units.
#pragma once

#include <algorithm>
#include <vector>

#include "engine.hpp"

namespace frk {

inline GemmItem gemm_item(const void* A, int64_t a_rs, int64_t a_cs, const void* B, int64_t b_rs, int64_t b_cs,
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

inline TrsmItem trsm_item(const void* A, int64_t a_rs, int64_t a_cs, const void* dinv, int64_t d_rs, int64_t d_cs,
                          void* B, int64_t b_rs, int64_t b_cs, int n, int nrhs, bool lower) {
  TrsmItem x{};
  x.A = A;
  x.a_rs = a_rs;
  x.a_cs = a_cs;
  x.dinv = dinv;
  x.d_rs = d_rs;
  x.d_cs = d_cs;
  x.B = B;
  x.b_rs = b_rs;
  x.b_cs = b_cs;
  x.n = n;
  x.nrhs = nrhs;
  x.lower = lower ? 1 : 0;
  return x;
}

template <typename T>
void plan_gemm_group(Plan& pl, const std::vector<GemmItem>& items, const char* cls) {
  const hipStream_t s = pl.ctx().stream();
  const size_t chunk = 65535;  // grid.y limit
  for (size_t o = 0; o < items.size(); o += chunk) {
    std::vector<GemmItem> part(items.begin() + o, items.begin() + std::min(items.size(), o + chunk));
    int mm = 0, nn = 0;
    double fl = 0, by = 0;
    for (const auto& g : part) {
      mm = std::max(mm, g.m);
      nn = std::max(nn, g.n);
      fl += 2.0 * g.m * g.n * g.k;
      by += sizeof(T) * ((double)g.m * g.k + (double)g.k * g.n + (double)g.m * g.n * (g.beta != 0.0 ? 2 : 1));
    }
    const size_t off = pl.add(part);
    const int cnt = (int)part.size();
    pl.launch([=](const char* d) { launch_gemm<T>(reinterpret_cast<const GemmItem*>(d + off), cnt, mm, nn, s); }, cls,
              fl, by);
  }
}

// Independent GEMMs, launched per output-shape class (tiny, tall-skinny,
// short-wide, general) so every item runs in the tile shape that fits it.
template <typename T>
void plan_gemms(Plan& pl, const std::vector<GemmItem>& items, const char* cls = "gemm") {
  if (items.empty()) return;
  std::vector<GemmItem> cl[4];
  for (const auto& g : items) {
    if (g.m <= 0 || g.n <= 0) continue;
    const int c = (g.m <= 16 && g.n <= 16) ? 0 : (g.n <= 16 ? 1 : (g.m <= 16 ? 2 : 3));
    cl[c].push_back(g);
  }
  for (auto& v : cl)
    if (!v.empty()) plan_gemm_group<T>(pl, v, cls);
}

// GEMMs that must run in exactly the given launch (in-place panel updates
// rely on the 64 x 64 tiling of one combined launch).
template <typename T>
void plan_gemms_onelaunch(Plan& pl, const std::vector<GemmItem>& items, const char* cls = "gemm") {
  if (!items.empty()) plan_gemm_group<T>(pl, items, cls);
}

template <typename T>
void plan_copies(Plan& pl, const std::vector<GemmItem>& items) {
  if (items.empty()) return;
  const hipStream_t s = pl.ctx().stream();
  int64_t mx = 0;
  for (const auto& g : items) mx = std::max<int64_t>(mx, (int64_t)g.m * g.n);
  const size_t off = pl.add(items);
  const int cnt = (int)items.size();
  pl.launch([=](const char* d) { launch_copy<T>(reinterpret_cast<const GemmItem*>(d + off), cnt, mx, s); });
}

template <typename T>
void plan_trsms(Plan& pl, const std::vector<TrsmItem>& items) {
  if (items.empty()) return;
  const hipStream_t s = pl.ctx().stream();
  int n = 0, r = 0;
  for (const auto& t : items) {
    n = std::max(n, t.n);
    r = std::max(r, t.nrhs);
  }
  const size_t off = pl.add(items);
  const int cnt = (int)items.size();
  pl.launch([=](const char* d) { launch_trsm<T>(reinterpret_cast<const TrsmItem*>(d + off), cnt, n, r, s); });
}

template <typename T>
void run_gemms(const std::vector<GemmItem>& items, Context& ctx) {
  Plan pl(ctx);
  plan_gemms<T>(pl, items);
  pl.run();
}


}  // namespace frk
