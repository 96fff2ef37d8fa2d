"""Pinning the oracle (CPU only): golden vectors, closed forms, brute force,
scipy/LAPACK routines, and the reference's own tolerance contracts.

The reference ships no golden data (its tests regenerate every input from
seeded generators), so the RNG is pinned by vectors written by libstdc++
itself (tests/golden/mt19937_golden.cpp) plus the C++ standard's fixed
10000th output of std::mt19937, and everything else by mathematics.  Each
planted-mistake test mutates the oracle and shows the corresponding check
fails.
"""
import json
import math
import os

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import frank_oracle as fo

GOLD = os.path.join(os.path.dirname(__file__), "golden", "mt19937_seed0.json")


# --------------------------------------------------------------------------- RNG
def test_mt19937_words_match_libstdcxx():
    g = json.load(open(GOLD))
    assert list(fo.mt19937_words(0, 16)) == g["words"]


def test_mt19937_standard_10000th_value():
    assert int(fo.mt19937_words(5489, 10000)[-1]) == json.load(open(GOLD))["default_seed_10000th"] == 4123659995


def test_uniform_real_distribution_matches_libstdcxx():
    g = json.load(open(GOLD))
    np.testing.assert_array_equal(fo.mt19937_uniform(0, 16), np.array(g["uniform"]))


def test_sorted_random_vector_and_random_uniform():
    g = json.load(open(GOLD))
    v = fo.get_sorted_random_vector(16)
    np.testing.assert_array_equal(v, np.sort(np.array(g["uniform"])))
    r = fo.random_uniform(4, 4)  # row-major fill, re-seeded every call (H6)
    np.testing.assert_array_equal(r.ravel(), np.array(g["uniform"]))
    np.testing.assert_array_equal(fo.random_uniform(4, 4), r)


def test_product_input_generator_matches_oracle():
    from frank_amd.inputs import sorted_uniform_points
    np.testing.assert_array_equal(sorted_uniform_points(1000), fo.get_sorted_random_vector(1000))


# --------------------------------------------------------------------------- kernels / norms
def test_laplacend_closed_form():
    x = np.array([0.0, 0.5, 2.0])
    a = fo.laplacend([x], 0, 3, 0, 3)
    assert a[0, 0] == 1.0 / 1e-3
    assert a[0, 2] == 1.0 / (2.0 + 1e-3)
    y = np.array([0.0, 3.0, 0.0])
    b = fo.laplacend([x, y], 0, 3, 0, 3)  # 2-D: distance sqrt(0.25 + 9)
    assert b[0, 1] == 1.0 / (math.sqrt(0.25 + 9.0) + 1e-3)


def test_norm_is_squared_frobenius():
    a = np.array([[3.0, 4.0]])
    assert fo.norm(a) == 25.0  # H1: the reference returns ||A||_F^2


def test_l2_error_formula():
    a = np.ones((2, 2))
    b = np.zeros((2, 2))
    assert fo.l2_error(a, b) == math.sqrt(4.0 / (4.0 + fo.DBL_EPSILON))


# --------------------------------------------------------------------------- truncation rank (H2)
def h2_rank_brute(s, eps):
    thr = eps * math.sqrt(float(np.sum(np.square(s))))
    k = len(s)
    for r in range(1, k):
        if math.sqrt((k - r) * s[r] ** 2) <= thr:
            return r
    return k


@pytest.mark.parametrize("seed", range(20))
def test_truncation_rank_matches_rule(seed):
    rng = np.random.default_rng(seed)
    s = np.sort(10.0 ** rng.uniform(-14, 0, size=rng.integers(1, 40)))[::-1]
    for eps in (1e-2, 1e-6, 1e-10):
        assert fo.find_svd_truncation_rank(s, eps) == h2_rank_brute(s, eps)


def test_truncation_rank_planted_mistake_is_caught(monkeypatch):
    s = 10.0 ** (-np.arange(20) / 2.0)
    good = fo.find_svd_truncation_rank(s, 1e-6)
    # off-by-one in the tail length: (min_dim - r + 1) instead of (min_dim - r)
    def bad(sd, eps):
        thr = eps * math.sqrt(fo.seq_sum(np.square(sd)))
        k = len(sd)
        for r in range(1, k):
            if math.sqrt((k - r + 1) * sd[r] ** 2) <= thr:
                return r
        return k
    assert good == h2_rank_brute(s, 1e-6)
    assert any(bad(10.0 ** (-np.arange(n) / 1.5), e) != h2_rank_brute(10.0 ** (-np.arange(n) / 1.5), e)
               for n in range(5, 30) for e in (1e-3, 1e-6, 1e-9))


# --------------------------------------------------------------------------- dense LAPACK-level ops
@pytest.mark.parametrize("n", [8, 16, 32])
def test_getrf_lu_reconstructs(n):
    # test/dense_getrf_test.cpp:12-52: diagonal boost, L U == A
    x = fo.get_sorted_random_vector(n)
    a = fo.laplacend([x], 0, n, 0, n) + 1e4 * np.eye(n)
    L, U = fo.getrf_dense(a)
    np.testing.assert_allclose(L @ U, a, rtol=1e-15, atol=1e-12)
    L2, U2 = fo.getrf_nopiv(a)
    np.testing.assert_allclose(L, L2, atol=1e-14)
    np.testing.assert_allclose(U, U2, rtol=1e-13)


def test_getrf_detects_pivoting():
    a = np.array([[1.0, 2.0], [3.0, 1.0]])
    with pytest.raises(fo.PivotingError):
        fo.getrf_dense(a)


@pytest.mark.parametrize("side", ["left", "right"])
@pytest.mark.parametrize("uplo", ["lower", "upper"])
def test_trsm_semantics(side, uplo):
    rng = np.random.default_rng(1)
    n = 12
    t = np.tril(rng.standard_normal((n, n))) + 5 * np.eye(n)
    if uplo == "upper":
        t = t.T.copy()
    b = rng.standard_normal((n, 5) if side == "left" else (5, n))
    x = fo.trsm_dense(t, b, uplo, side)
    tt = np.tril(t, -1) + np.eye(n) if uplo == "lower" else t  # H4: lower => unit diagonal
    if side == "left":
        np.testing.assert_allclose(tt @ x, b, atol=1e-12)
    else:
        np.testing.assert_allclose(x @ tt, b, atol=1e-12)
    xt = fo.trsm_dense_trans(t, b, uplo, side)
    if side == "left":
        np.testing.assert_allclose(tt.T @ xt, b, atol=1e-12)
    else:
        np.testing.assert_allclose(xt @ tt.T, b, atol=1e-12)


def test_qr_rq_svd_contracts():
    rng = np.random.default_rng(2)
    a = rng.standard_normal((40, 12))
    q, r = fo.qr(a)
    np.testing.assert_allclose(q @ r, a, atol=1e-12)  # test/dense_qr_test.cpp:27-45
    np.testing.assert_allclose(q.T @ q, np.eye(12), atol=1e-14)
    b = rng.standard_normal((12, 40))
    r2, q2 = fo.rq(b)
    np.testing.assert_allclose(r2 @ q2, b, atol=1e-12)
    np.testing.assert_allclose(q2 @ q2.T, np.eye(12), atol=1e-14)
    assert np.allclose(np.tril(r2, -1), 0.0)
    u, s, v = fo.svd(a)
    np.testing.assert_allclose(u @ s @ v, a, atol=1e-11)  # test/dense_svd_test.cpp:24-28
    assert np.all(np.diff(np.diag(s)) <= 0)


# --------------------------------------------------------------------------- truncated_geqp3 (H5)
@pytest.mark.parametrize("eps", [1e-2, 1e-4, 1e-8, 1e-12])
@pytest.mark.parametrize("shape", [(24, 24), (30, 17), (17, 30)])
def test_truncated_geqp3_matches_bruteforce(eps, shape):
    m, n = shape
    x = fo.get_sorted_random_vector(m + n)
    a = fo.laplacend([x], 0, m, m, n)
    q, r = fo.truncated_geqp3(a, eps)
    rank, piv = fo.truncated_geqp3_bruteforce(a, eps)
    assert q.shape[1] == rank
    np.testing.assert_allclose(q.T @ q, np.eye(rank), atol=1e-13)
    # residual of the truncated factorization is bounded by the stop rule
    res = a - q @ r
    assert np.max(np.linalg.norm(res, axis=0)) <= eps * math.sqrt(np.sum(a * a)) * (1 + 1e-6) + 1e-14


def test_truncated_geqp3_zero_matrix():
    q, r = fo.truncated_geqp3(np.zeros((5, 4)), 1e-8)  # test/dense_qr_test.cpp:98-128
    assert q.shape == (5, 1) and q[0, 0] == 1.0 and np.all(r == 0)


def test_truncated_geqp3_planted_mistake_is_caught(monkeypatch):
    # a wrong stop threshold (||A||_F^2 instead of ||A||_F) changes the rank
    x = fo.get_sorted_random_vector(48)
    a = 1e3 * fo.laplacend([x], 0, 24, 24, 24)
    rank, _ = fo.truncated_geqp3_bruteforce(a, 1e-6)
    orig_norm = fo.norm
    monkeypatch.setattr(fo, "norm", lambda z: orig_norm(z) ** 2)
    q, _ = fo.truncated_geqp3(a, 1e-6)
    assert q.shape[1] != rank


# --------------------------------------------------------------------------- LowRank
def test_rsvd_contract():
    # test/rsvd_test.cpp:12-32: n = 2048, k = 16, error < 1e-8 (separated Laplace block)
    n, k = 2048, 16
    x = fo.get_sorted_random_vector(2 * n)
    a = fo.laplacend([x], 0, n, n, n)
    lr = fo.LowRank.from_rank(a, k)  # rsvd with oversampling min(k + 5, m, n)
    assert lr.rank == k
    assert fo.l2_error(a, lr.dense()) < 1e-8


def test_lowrank_from_eps_contract():
    x = fo.get_sorted_random_vector(128)
    a = fo.laplacend([x], 0, 64, 64, 64)
    for eps in (1e-6, 1e-8, 1e-10):
        lr = fo.LowRank.from_eps(a, eps)
        assert abs(fo.l2_error(a, lr.dense()) - eps) <= 10 * eps  # test/lowrank_test.cpp:63-79
        assert np.allclose(np.tril(lr.S, -1), 0.0)  # S is the R factor of an RQ (H5)


def _dense_truncated_sum(A, B, eps):
    # the reference truncates the SVD of the K x K core, K = kA + kB: the sum's
    # K largest singular values with min_dim = K in the H2 rule
    s = np.linalg.svd(A.dense() + B.dense(), compute_uv=False)
    K = min(A.rank + B.rank, min(A.shape))
    return h2_rank_brute(s[:K], eps)


@pytest.mark.parametrize("eps", [1e-6, 1e-8, 1e-10, 1e-12])
def test_rounded_addition_closed_form(eps):
    # rank equals the H2 rule on the singular values of the exact sum; error
    # within eps +- 10 eps (test/lowrank_test.cpp:81-115)
    x = fo.get_sorted_random_vector(256)
    d1 = fo.laplacend([x], 0, 64, 128, 64)
    d2 = fo.laplacend([x], 64, 64, 192, 64)
    A = fo.LowRank.from_eps(d1, eps)
    B = fo.LowRank.from_eps(d2, eps)
    C = fo.rounded_addition(A, B)
    assert C.rank == _dense_truncated_sum(A, B, eps)
    assert abs(fo.l2_error(d1 + d2, C.dense()) - eps) <= 10 * eps


@pytest.mark.parametrize("lra", ["rounded_addition", "fast_rounded_addition"])
@pytest.mark.parametrize("k", [1, 2, 4, 8])
def test_fixed_rank_addition_keeps_rank(lra, k):
    # test/lowrank_test.cpp:31-61: rank preserved over three additions; result is the best rank-k
    x = fo.get_sorted_random_vector(64)
    d = fo.laplacend([x], 0, 32, 32, 32)
    A = fo.LowRank.from_rank(d, k)
    B = fo.LowRank.from_rank(d, k)
    C = A.copy()
    for _ in range(3):
        C = fo.lr_add(C, B, lra)
    assert C.rank == k
    dense = A.dense() + 3 * B.dense()
    u, s, vt = np.linalg.svd(dense)
    best = (u[:, :k] * s[:k]) @ vt[:k]
    np.testing.assert_allclose(C.dense(), best, atol=1e-10 * np.abs(dense).max())


@pytest.mark.parametrize("ta", [False, True])
@pytest.mark.parametrize("tb", [False, True])
@pytest.mark.parametrize("kinds", ["DDD", "LDD", "DLD", "LLD", "DDL", "LDL", "DLL", "LLL"])
def test_gemm_variants_against_dense(ta, tb, kinds):
    # test/gemm_test.cpp:33-46 naive gemm as the reference; compressed tolerance 10 x 1e-6
    x = fo.get_sorted_random_vector(96)
    blocks = [fo.laplacend([x], 0, 32, 64, 32), fo.laplacend([x], 32, 32, 64, 32),
              fo.laplacend([x], 0, 32, 32, 32)]
    eps = 1e-6

    def make(kch, d):
        return fo.LowRank.from_eps(d, eps) if kch == "L" else d.copy()

    A = make(kinds[0], blocks[0])
    B = make(kinds[1], blocks[1])
    C = make(kinds[2], blocks[2])
    dA = A.dense() if isinstance(A, fo.LowRank) else A
    dB = B.dense() if isinstance(B, fo.LowRank) else B
    dC = C.dense() if isinstance(C, fo.LowRank) else C.copy()
    for alpha, beta in [(1.0, 1.0), (-1.0, 1.0), (1.0, 0.0)]:
        Cw = C.copy() if isinstance(C, fo.LowRank) else C.copy()
        out = fo.gemm(A, B, Cw, alpha, beta, ta, tb, lra="rounded_addition")
        ref = alpha * ((dA.T if ta else dA) @ (dB.T if tb else dB)) + beta * dC
        got = out.dense() if isinstance(out, fo.LowRank) else out
        tol = 1e-14 if kinds == "DDD" else 10 * eps
        assert fo.l2_error(ref, got) <= tol


# --------------------------------------------------------------------------- ID / RID
def test_id_contracts():
    # test/dense_id_test.cpp:25-70: n = 2048, k = 16, error < 1e-7
    n, k = 2048, 16
    x = fo.get_sorted_random_vector(2 * n)
    a = fo.laplacend([x], 0, n, n, n)
    v, cols = fo.one_sided_id(a.copy(), k)
    assert fo.l2_error(a, a[:, cols] @ v) < 1e-7
    u, s, vv = fo.id_(a, k)
    assert fo.l2_error(a, u @ s @ vv) < 1e-7
    u, s, vv = fo.rid(a, k + 5, k)
    assert fo.l2_error(a, u @ s @ vv) < 1e-7


# --------------------------------------------------------------------------- BLR composites
@pytest.mark.parametrize("eps", [1e-6, 1e-10])
@pytest.mark.parametrize("admis", [0.0, 1.0])
def test_blr_lu_contract(eps, admis):
    # test/blr_fixed_acc_test.cpp:67-83
    N = 128
    x = fo.get_sorted_random_vector(N)
    ker = lambda r0, nr, c0, nc: fo.laplacend([x], r0, nr, c0, nc)
    A = fo.build_blr(ker, N, N, 4, 4, eps=eps, admis=admis, points=[x])
    D = fo.laplacend([x], 0, N, 0, N)
    assert fo.l2_error(D, A) <= eps
    xs = fo.random_uniform(N, 1)
    b = fo.blr_matvec(A, xs)
    L, U = fo.blr_getrf(A.copy(), lra="rounded_addition")
    y = fo.blr_trsm_upper(U, fo.blr_trsm_lower(L, b))
    assert fo.l2_error(xs, y) <= eps
    # the factorization of a dense-only grid (admissibility never met) is exact LU
    A2 = fo.build_blr(ker, N, N, 4, 4, eps=eps, admis=100.0, points=[x])
    L2, U2 = fo.blr_getrf(A2.copy())
    np.testing.assert_allclose(fo.lu_product(L2, U2), D, rtol=0, atol=1e-10 * np.abs(D).max())


def test_blr_transposed_sweeps():
    N = 96
    x = fo.get_sorted_random_vector(N)
    ker = lambda r0, nr, c0, nc: fo.laplacend([x], r0, nr, c0, nc)
    A = fo.build_blr(ker, N, N, 3, 3, eps=1e-12, points=[x])
    A0 = A.dense()
    L, U = fo.blr_getrf(A.copy(), lra="rounded_addition")
    g = np.random.default_rng(4).standard_normal((N, 2))
    lam = fo.blr_trsm_lower_trans(L, fo.blr_trsm_upper_trans(U, g))
    np.testing.assert_allclose(A0.T @ lam, g, atol=1e-8 * np.abs(g).max())


def test_index_split_and_admissibility():
    assert fo.index_split(0, 10, 3) == [(0, 4), (4, 4), (8, 2)]  # ceil split, last smaller
    assert fo.is_admissible((0, 4), (8, 4), 0, 2, 1.0, "position", [])
    assert not fo.is_admissible((0, 4), (4, 4), 0, 1, 1.0, "position", [])
    assert not fo.is_admissible((0, 1), (8, 4), 0, 2, 0.0, "position", [])  # vectors never admissible


def test_morton_sort_matches_bitwise_box_order():
    """Morton order (morton_index.cpp:75-98) against an independent bit interleave."""
    from frank_amd.inputs import morton_sort, uniform_points
    pts = uniform_points(300, 2)
    level = 4
    xs, perm = morton_sort(pts, level)
    assert sorted(perm.tolist()) == list(range(300))
    lo = pts.min(axis=1) - 0.5
    size = (pts.max(axis=1) - pts.min(axis=1)).max() + 1.0

    def box(p):
        cell = [int((p[d] - lo[d]) / size * (1 << level)) for d in range(2)]
        bits = ""
        for lev in reversed(range(level)):
            for d in reversed(range(2)):
                bits += str((cell[d] >> lev) & 1)
        return int(bits, 2)

    keys = [box(xs[:, i]) for i in range(300)]
    assert keys == sorted(keys)
    # stable: equal keys keep the input order
    for a, b in zip(range(299), range(1, 300)):
        if keys[a] == keys[b]:
            assert perm[a] < perm[b]
