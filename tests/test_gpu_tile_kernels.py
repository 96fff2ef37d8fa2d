"""Tile-kernel parity on the GPU (through the C ABI) against the oracle.

Contracts follow the reference's unit tests: dense getrf L U == A
(test/dense_getrf_test.cpp:12-52), trsm 1e-12 (test/dense_trsm_test.cpp:26-28),
gemm 1e-14 (test/gemm_test.cpp:27-28), compression error within eps +- 10 eps
(test/lowrank_test.cpp:63-79), rounded addition error (lowrank_test.cpp:81-115)
and fixed-rank preservation (lowrank_test.cpp:31-61).
"""
import numpy as np
import pytest

from oracle import frank_oracle as fo

pytestmark = pytest.mark.gpu


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


@pytest.mark.parametrize("n", [8, 16, 32, 65, 256, 333, 512])
@pytest.mark.parametrize("batch", [1, 2])
def test_getrf_nopiv_matches_lu(torch_gpu, fa, n, batch):
    """batch 1 runs the one-launch dataflow tile LU, batch 2 the blocked launch sequence."""
    torch = torch_gpu
    print(f"getrf n={n} batch={batch}", flush=True)
    x = fo.get_sorted_random_vector(2 * n)
    a = fo.laplacend([x], 0, n, 0, n) + 1000.0 * np.eye(n)  # diagonal boost as in dense_getrf_test
    scales = (1.0, 2.0)[:batch]
    tiles = [dev(torch, s * a) for s in scales]
    info = fa.getrf_nopiv_batched(tiles)
    assert list(info) == [0] * batch
    for t, scale in zip(tiles, scales):
        lu = t.cpu().numpy()
        Lf = np.tril(lu, -1) + np.eye(n)
        Uf = np.triu(lu)
        ref_l, ref_u = fo.getrf_nopiv(scale * a)
        np.testing.assert_allclose(Lf @ Uf, scale * a, rtol=0, atol=1e-12 * np.abs(a).max() * scale)
        np.testing.assert_allclose(Lf, ref_l, rtol=0, atol=1e-12)
        np.testing.assert_allclose(Uf, ref_u, rtol=0, atol=1e-12 * np.abs(ref_u).max())


def test_getrf_zero_pivot_reported(torch_gpu, fa):
    torch = torch_gpu
    a = np.eye(70)
    a[40, 40] = 0.0
    info = fa.getrf_nopiv_batched([dev(torch, a)])
    assert info[0] == 41


@pytest.mark.parametrize("side", ["left", "right"])
@pytest.mark.parametrize("uplo", ["lower", "upper"])
@pytest.mark.parametrize("trans", [False, True])
def test_trsm(torch_gpu, fa, side, uplo, trans):
    torch = torch_gpu
    n, k = 48, 37
    x = fo.get_sorted_random_vector(2 * n)
    a = fo.laplacend([x], 0, n, 0, n) + 1000.0 * np.eye(n)
    Lf, Uf = fo.getrf_nopiv(a)
    tri = Lf if uplo == "lower" else Uf
    b = fo.random_uniform(n, k) if side == "left" else fo.random_uniform(k, n)
    B = dev(torch, b)
    fa.trsm(dev(torch, tri), B, uplo, side, trans)
    if not trans:
        ref = fo.trsm_dense(tri, b, uplo, side)
    else:
        ref = fo.trsm_dense_trans(tri, b, uplo, side)
    np.testing.assert_allclose(B.cpu().numpy(), ref, rtol=0, atol=1e-12 * max(1.0, np.abs(ref).max()))


@pytest.mark.parametrize("ta", [False, True])
@pytest.mark.parametrize("tb", [False, True])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-1.0, 1.0), (0.5, -2.0)])
def test_gemm(torch_gpu, fa, ta, tb, alpha, beta):
    torch = torch_gpu
    rng = np.random.default_rng(3)
    m, n, k = 70, 131, 45
    a = rng.standard_normal((k, m) if ta else (m, k))
    b = rng.standard_normal((n, k) if tb else (k, n))
    c = rng.standard_normal((m, n))
    C_ = dev(torch, c)
    fa.gemm(dev(torch, a), dev(torch, b), C_, alpha, beta, ta, tb)
    ref = alpha * ((a.T if ta else a) @ (b.T if tb else b)) + beta * c
    np.testing.assert_allclose(C_.cpu().numpy(), ref, rtol=0, atol=1e-13 * np.abs(ref).max())


def _separated_tile(m, n):
    x = fo.get_sorted_random_vector(m + n)
    return fo.laplacend([x], 0, m, m, n)


@pytest.mark.parametrize("eps", [1e-6, 1e-8, 1e-10])
@pytest.mark.parametrize("m,n", [(128, 128), (256, 256), (200, 150)])
def test_compress_eps_matches_truncated_geqp3(torch_gpu, fa, eps, m, n):
    torch = torch_gpu
    d = _separated_tile(m, n)
    U, S, V = fa.compress(dev(torch, d), eps=eps, cap=64)
    ref = fo.LowRank.from_eps(d, eps)
    assert U.shape[1] == ref.rank, (U.shape[1], ref.rank)
    u, s, v = U.cpu().numpy(), S.cpu().numpy(), V.cpu().numpy()
    err = fo.l2_error(d, (u @ s) @ v)
    ref_err = fo.l2_error(d, ref.dense())
    assert abs(err - ref_err) <= 1e-6 * ref_err + 1e-15  # same rank, same truncation
    assert abs(err - eps) <= 10 * eps  # test/lowrank_test.cpp:63-79 EXPECT_NEAR(err, eps, 10 eps)
    np.testing.assert_allclose(u.T @ u, np.eye(u.shape[1]), atol=1e-13)
    np.testing.assert_allclose(v @ v.T, np.eye(v.shape[0]), atol=1e-13)
    assert np.allclose(np.tril(s, -1), 0.0)


def test_compress_zero_matrix(torch_gpu, fa):
    torch = torch_gpu
    U, S, V = fa.compress(dev(torch, np.zeros((64, 48))), eps=1e-8, cap=16)
    assert U.shape[1] == 1
    u = U.cpu().numpy()
    assert u[0, 0] == 1.0 and np.all(u[1:, 0] == 0.0)  # test/dense_qr_test.cpp:98-128
    assert np.all(S.cpu().numpy() == 0.0)


@pytest.mark.parametrize("rank", [4, 8, 16])
def test_compress_rank_rsvd(torch_gpu, fa, rank):
    torch = torch_gpu
    d = _separated_tile(256, 256)
    U, S, V = fa.compress(dev(torch, d), rank=rank, cap=32)
    ref = fo.LowRank.from_rank(d, rank)
    prod = (U.cpu().numpy() @ S.cpu().numpy()) @ V.cpu().numpy()
    np.testing.assert_allclose(prod, ref.dense(), rtol=0, atol=1e-11 * np.abs(d).max())
    np.testing.assert_allclose(np.diag(S.cpu().numpy()), np.diag(ref.S), rtol=0, atol=1e-12 * ref.S[0, 0])


def _lr_slot(torch, lr, cap):
    m, n = lr.shape
    U = torch.zeros((m, cap), dtype=torch.float64, device="cuda")
    S = torch.zeros((cap, cap), dtype=torch.float64, device="cuda")
    V = torch.zeros((cap, n), dtype=torch.float64, device="cuda")
    k = lr.rank
    U[:, :k] = torch.from_numpy(lr.U)
    S[:k, :k] = torch.from_numpy(lr.S)
    V[:k, :] = torch.from_numpy(lr.V)
    return (U, S, V, k)


@pytest.mark.parametrize("eps", [1e-6, 1e-8, 1e-10, 1e-12])
def test_rounded_addition_eps(torch_gpu, fa, eps):
    torch = torch_gpu
    d = _separated_tile(64, 64)
    d2 = _separated_tile(64, 64) * 0.5 + fo.laplacend([fo.get_sorted_random_vector(256)], 0, 64, 192, 64)
    A = fo.LowRank.from_eps(d, eps)
    B = fo.LowRank.from_eps(d2, eps)
    ref = fo.rounded_addition(A, B)
    cap = 48
    sa = _lr_slot(torch, A, cap)
    Bt = (dev(torch, B.U), dev(torch, B.S), dev(torch, B.V))
    ranks = fa.lr_add_batched([sa], [Bt], eps=eps, lra="rounded_addition", cap=cap)
    assert ranks[0] == ref.rank
    k = ranks[0]
    prod = (sa[0][:, :k].cpu().numpy() @ sa[1][:k, :k].cpu().numpy()) @ sa[2][:k].cpu().numpy()
    np.testing.assert_allclose(prod, ref.dense(), rtol=0, atol=1e-13 * np.abs(d).max() + 2 * eps * np.abs(d).max())
    err = fo.l2_error(d + d2, prod)
    assert abs(err - fo.l2_error(d + d2, ref.dense())) <= 10 * eps


@pytest.mark.parametrize("b,r0,c0,r1,c1", [(128, 0, 256, 0, 384), (256, 0, 768, 256, 768)])
def test_rounded_addition_large_cores(torch_gpu, fa, b, r0, c0, r1, c1):
    """FP64 cores of 77 and 92 columns (2-D Laplace blocks, the second term
    overlapping the first): the LDS Jacobi with a global K x K scratch, whose
    left vectors are formed as M v / sigma.  Rank, product and singular values
    against the oracle's rounded addition; the factors are orthonormal to the
    accuracy that formula gives, u ||M|| / sigma_min (u the unit roundoff)."""
    torch = torch_gpu
    from frank_amd.inputs import morton_sort, uniform_points
    pts, _ = morton_sort(uniform_points(1024, 2), 6)
    pts = [np.ascontiguousarray(p) for p in pts]
    eps, cap = 1e-8, 96
    d = fo.laplacend(pts, r0, b, c0, b)
    d2 = 0.5 * d + fo.laplacend(pts, r1, b, c1, b)
    A = fo.LowRank.from_eps(d, eps)
    B = fo.LowRank.from_eps(d2, eps)
    assert 64 < A.rank + B.rank <= 128
    ref = fo.rounded_addition(A, B)
    sa = _lr_slot(torch, A, cap)
    ranks = fa.lr_add_batched([sa], [(dev(torch, B.U), dev(torch, B.S), dev(torch, B.V))], eps=eps,
                              lra="rounded_addition", cap=cap)
    assert ranks[0] == ref.rank
    k = ranks[0]
    u, s, v = (sa[0][:, :k].cpu().numpy(), sa[1][:k, :k].cpu().numpy(), sa[2][:k].cpu().numpy())
    scale = np.abs(d + d2).max()
    np.testing.assert_allclose(u @ s @ v, ref.dense(), rtol=0, atol=1e-12 * scale + 2 * eps * scale)
    assert abs(fo.l2_error(d + d2, u @ s @ v) - fo.l2_error(d + d2, ref.dense())) <= 10 * eps
    np.testing.assert_allclose(np.diag(s), np.diag(ref.S), rtol=0, atol=1e-13 * ref.S[0, 0])
    sv = np.diag(ref.S)
    np.testing.assert_allclose(u.T @ u, np.eye(k), atol=64 * np.finfo(float).eps * sv[0] / sv[k - 1])
    np.testing.assert_allclose(v @ v.T, np.eye(k), atol=1e-12)


@pytest.mark.parametrize("lra", ["fast_rounded_addition", "rounded_addition"])
@pytest.mark.parametrize("rank", [1, 2, 4, 8])
def test_fixed_rank_addition_keeps_rank(torch_gpu, fa, lra, rank):
    torch = torch_gpu
    d = _separated_tile(64, 64)
    A = fo.LowRank.from_rank(d, rank)
    B = fo.LowRank.from_rank(d, rank)
    ref = fo.lr_add(A, B, lra)
    sa = _lr_slot(torch, A, 16)
    ranks = fa.lr_add_batched([sa], [(dev(torch, B.U), dev(torch, B.S), dev(torch, B.V))], eps=0.0, lra=lra,
                              cap=16)
    assert ranks == [rank]
    prod = (sa[0][:, :rank].cpu().numpy() @ sa[1][:rank, :rank].cpu().numpy()) @ sa[2][:rank].cpu().numpy()
    np.testing.assert_allclose(prod, ref.dense(), rtol=0, atol=1e-12 * np.abs(d).max())
    u = sa[0][:, :rank].cpu().numpy()
    np.testing.assert_allclose(u.T @ u, np.eye(rank), atol=1e-12)


def test_lr_add_batched_uneven(torch_gpu, fa):
    """Several independent additions with uneven shapes in one batched call."""
    torch = torch_gpu
    rng = np.random.default_rng(7)
    shapes = [(64, 128), (128, 64), (256, 256), (96, 200)]
    As, Bs, refs, dens = [], [], [], []
    for (m, n) in shapes:
        x = fo.get_sorted_random_vector(m + n)
        d = fo.laplacend([x], 0, m, m, n)
        e = 0.3 * fo.laplacend([x[::-1].copy()], 0, m, m, n)
        A = fo.LowRank.from_eps(d, 1e-8)
        B = fo.LowRank.from_eps(e, 1e-8)
        refs.append(fo.rounded_addition(A, B))
        As.append(_lr_slot(torch, A, 40))
        Bs.append((dev(torch, B.U), dev(torch, B.S), dev(torch, B.V)))
        dens.append(d + e)
    ranks = fa.lr_add_batched(As, Bs, eps=1e-8, lra="rounded_addition", cap=40)
    for a, r, ref, dd in zip(As, ranks, refs, dens):
        assert r == ref.rank
        prod = (a[0][:, :r].cpu().numpy() @ a[1][:r, :r].cpu().numpy()) @ a[2][:r].cpu().numpy()
        np.testing.assert_allclose(prod, ref.dense(), rtol=0, atol=1e-12 * np.abs(dd).max())


@pytest.mark.parametrize("ka,kb", [(4, 4), (8, 16), (40, 30)])
def test_naive_addition(torch_gpu, fa, ka, kb):
    """naive_addition (addition.cpp:68-89): concatenation below min(m, n), dense
    recompression at A's rank (rsvd) at or above it."""
    torch = torch_gpu
    d = _separated_tile(64, 64)
    A = fo.LowRank.from_rank(d, ka)
    B = fo.LowRank.from_rank(0.5 * d, kb)
    ref = fo.naive_addition(A, B)
    cap = 80
    sa = _lr_slot(torch, A, cap)
    ranks = fa.lr_add_batched([sa], [(dev(torch, B.U), dev(torch, B.S), dev(torch, B.V))], eps=0.0, lra="naive",
                              cap=cap)
    assert ranks == [ref.rank]
    k = ranks[0]
    prod = (sa[0][:, :k].cpu().numpy() @ sa[1][:k, :k].cpu().numpy()) @ sa[2][:k].cpu().numpy()
    np.testing.assert_allclose(prod, ref.dense(), rtol=0, atol=1e-10 * np.abs(d).max())


@pytest.mark.parametrize("eps", [1e-4, 1e-6])
def test_rounded_addition_fp32(torch_gpu, fa, eps):
    """FP32 rounded addition against the FP64 oracle: error within eps + FP32 rounding."""
    torch = torch_gpu
    d = _separated_tile(128, 96)
    d2 = 0.25 * _separated_tile(128, 96)[::-1, ::-1].copy()
    A = fo.LowRank.from_eps(d, eps)
    B = fo.LowRank.from_eps(d2, eps)
    ref = fo.rounded_addition(A, B)
    cap = 48
    sa = tuple(x.float() if hasattr(x, "float") else x for x in _lr_slot(torch, A, cap)[:3]) + (A.rank,)
    Bt = (dev(torch, B.U).float(), dev(torch, B.S).float(), dev(torch, B.V).float())
    ranks = fa.lr_add_batched([sa], [Bt], eps=eps, lra="rounded_addition", cap=cap)
    k = ranks[0]
    assert abs(k - ref.rank) <= 1
    prod = (sa[0][:, :k].double().cpu().numpy() @ sa[1][:k, :k].double().cpu().numpy()) @ sa[2][:k].double().cpu().numpy()
    assert fo.l2_error(d + d2, prod) <= 10 * eps + 1e-5
