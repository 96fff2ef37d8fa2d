// Driver-style test of the C++ operator surface (include/frank_amd.hpp),
// written the way the reference's drivers and GTests use FRANK:
// examples/h_lu.cpp, test/blr_fixed_acc_test.cpp, test/dense_id_test.cpp,
// test/lowrank_test.cpp.  Prints one "ok <name> <value>" line per check and
// exits non-zero on the first failure.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "frank_amd.hpp"

using namespace frank_amd;

static int failures = 0;
static void expect(bool ok, const char* name, double value) {
  std::printf("%s %s %.3e\n", ok ? "ok" : "FAIL", name, value);
  std::fflush(stdout);
  if (!ok) ++failures;
}

int main() {
  initialize();
  setGlobalValue("FRANK_LRA", "rounded_addition");
  // --- BLR-LU (examples/h_lu.cpp, blr_fixed_acc_test.cpp:46-83) ------------
  {
    const int64_t N = 1024, nleaf = 128, nb = N / nleaf;
    const double eps = 1e-8;
    const std::vector<std::vector<double>> x{get_sorted_random_vector(N)};
    Hierarchical A(laplacend, x, N, N, nleaf, eps, 0.0, nb, nb);
    const Dense D(laplacend, x, N, N);
    expect(l2_error(D, A) <= eps, "blr_compression_error", l2_error(D, A));
    Dense xt(N, 1);
    for (int64_t i = 0; i < N; ++i) xt(i, 0) = std::sin(0.1 * i);
    Dense b(N, 1);
    gemm(A, xt, b, 1.0, 0.0);
    Hierarchical L, U;
    std::tie(L, U) = getrf(A);
    trsm(L, b, Mode::Lower);
    trsm(U, b, Mode::Upper);
    expect(l2_error(xt, b) <= eps, "blr_lu_solve_error", l2_error(xt, b));
  }
  // --- multimethod dispatch: unmatched operand types are reported -------
  {
    const std::vector<std::vector<double>> x{get_sorted_random_vector(256)};
    MatrixProxy H = Hierarchical(laplacend, x, 256, 256, 64, 1e-6, 0.0, 4, 4);
    MatrixProxy D = Dense(256, 256);
    MatrixProxy C = Dense(256, 256);
    std::string msg;
    try {
      gemm(D, H, C);
    } catch (const Error& e) {
      msg = e.what();
    }
    expect(msg == "gemm(Dense, Hierarchical, Dense) undefined!", "dispatch_undefined", 0.0);
  }
  // --- dense factorizations (test/dense_qr_test.cpp, dense_svd_test.cpp, dense_id_test.cpp) -----
  {
    const int64_t n = 256;
    const std::vector<std::vector<double>> x{get_sorted_random_vector(2 * n)};
    Dense A(laplacend, x, n, n, 0, n);
    Dense Q(n, n), R(n, n);
    Dense Aq(laplacend, x, n, 128, 0, n);
    Dense Q2(n, 128), R2(128, 128);
    qr(Aq, Q2, R2);
    expect(l2_error(Aq, gemm(Q2, R2)) < 1e-13, "qr_reconstruction", l2_error(Aq, gemm(Q2, R2)));
    Dense As(laplacend, x, n, 64, 0, n);
    auto [Us, Ss, Vs] = svd(As);
    const double es = l2_error(As, gemm(gemm(Us, Ss), Vs));
    expect(es < 1e-12, "svd_reconstruction", es);
    auto [Ui, Si, Vi] = id(A, 16);
    const double ei = l2_error(A, gemm(gemm(Ui, Si), Vi));
    expect(ei < 1e-7, "id_error", ei);
    auto [Ur, Sr, Vr] = rid(A, 21, 16);
    const double er = l2_error(A, gemm(gemm(Ur, Sr), Vr));
    expect(er < 1e-7, "rid_error", er);
    auto [Uv, Sv, Vv] = rsvd(A, 16);
    const double ev = l2_error(A, gemm(gemm(Uv, Sv), Vv));
    expect(ev < 1e-8, "rsvd_error", ev);
  }
  // --- rounded addition (test/lowrank_test.cpp:81-115) -------------------
  {
    const int64_t n = 64;
    const std::vector<std::vector<double>> x{get_sorted_random_vector(4 * n)};
    const double eps = 1e-8;
    Dense D1(laplacend, x, n, n, 0, 2 * n), D2(laplacend, x, n, n, n, 3 * n);
    LowRank A(D1, eps), B(D2, eps);
    A += B;
    Dense S(n, n);
    for (int64_t i = 0; i < n; ++i)
      for (int64_t j = 0; j < n; ++j) S(i, j) = D1(i, j) + D2(i, j);
    const double ea = l2_error(S, gemm(gemm(A.U, A.S), A.V));
    expect(ea < 10 * eps, "rounded_addition_error", ea);
  }
  // --- tiled Householder BLR-QR (test/blr_fixed_acc_test.cpp:148-189) ---
  {
    const int64_t n = 256, nleaf = 32, nb = n / nleaf;
    const double eps = 1e-8;
    const std::vector<std::vector<double>> x{get_sorted_random_vector(n)};
    const Dense D(laplacend, x, n, n);
    Hierarchical A(laplacend, x, n, n, nleaf, eps, 0.0, nb, nb);
    QrFactors T = tiled_householder_blr_qr(A);
    Hierarchical Q(identity, x, n, n, nleaf, eps, 0.0, nb, nb);
    left_multiply_tiled_reflector(A, T, Q, false);
    Dense Ad = A.to_dense(), Qd = Q.to_dense();
    Dense R(n, n);
    for (int64_t i = 0; i < n; ++i)
      for (int64_t j = 0; j < n; ++j)
        if ((i / nleaf) < (j / nleaf) || ((i / nleaf) == (j / nleaf) && i <= j)) R(i, j) = Ad(i, j);
    const double res = l2_error(D, gemm(Qd, R));
    expect(res <= eps, "blr_qr_residual", res);
  }
  // --- tile-level geqrt / larfb / tpqrt / tpmqrt (LAPACK.h:345-430) ------
  // QR of [A0; B0] with B0 dense and with B0 low rank; the same reflectors
  // applied to a copy of [A0; B0] must give [R; 0].
  for (int lowrank_b = 0; lowrank_b < 2; ++lowrank_b) {
    const int64_t n = 128;
    const std::vector<std::vector<double>> x{get_sorted_random_vector(2 * n)};
    const Dense A0(laplacend, x, n, n, 0, 0);
    const Dense B0(laplacend, x, n, n, n, 0);
    Dense A(A0), T0(n, n), T1(n, n);
    MatrixProxy B = lowrank_b ? MatrixProxy(LowRank(B0, 1e-12)) : MatrixProxy(Dense(B0));
    Dense Bd0 = lowrank_b ? gemm(gemm(std::get<LowRank>(B).U, std::get<LowRank>(B).S), std::get<LowRank>(B).V)
                          : Dense(B0);
    geqrt(A, T0);
    MatrixProxy Ctop = Dense(A0), Cbot = Dense(Bd0);
    larfb(A, T0, Ctop, true);
    tpqrt(A, B, T1);
    tpmqrt(B, T1, Ctop, Cbot, true);
    const Dense& ct = std::get<Dense>(Ctop);
    const Dense& cb = std::get<Dense>(Cbot);
    double num = 0.0, bot = 0.0, den = 0.0;
    for (int64_t i = 0; i < n; ++i)
      for (int64_t j = 0; j < n; ++j) {
        const double r = (i <= j) ? A(i, j) : 0.0;
        num += (ct(i, j) - r) * (ct(i, j) - r);
        den += r * r;
        bot += cb(i, j) * cb(i, j);
      }
    const double e = std::sqrt((num + bot) / den);
    expect(e <= 1e-12, lowrank_b ? "tile_qr_ops_lowrank_b" : "tile_qr_ops_dense_b", e);
  }
  std::printf("%s\n", failures ? "FAILED" : "ALL OK");
  return failures ? 1 : 0;
}
