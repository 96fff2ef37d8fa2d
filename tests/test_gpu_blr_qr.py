"""Tiled Householder BLR-QR parity on the GPU against the oracle.

Tile operations: geqrt / tpqrt / larft on small grids built from the same host
tiles on both sides, compared with the oracle's LAPACK-based restatement
(geqrt.cpp:34-43, tpqrt.cpp:35-53): R, the reflectors and T to 1e-11.

Full factorization: test/blr_fixed_acc_test.cpp:148-207 (n_rows 256,
n_cols in {128, 256}, nleaf 32, eps in {1e-6, 1e-8, 1e-10}, admis in
{0, 1, 4}, position/geometry admissibility): ||A - QR|| / ||A|| <= eps and
||Q^T Q - I|| / ||I|| <= eps with Q formed by left_multiply_tiled_reflector on
the identity, as the reference test does; the residual is also checked to be
within 10x of the oracle's.
"""
import numpy as np
import pytest

from oracle import frank_oracle as fo

pytestmark = pytest.mark.gpu


def _tiles_of(H):
    out = []
    for blk in H.blocks:
        out.append((blk.U, blk.S, blk.V) if isinstance(blk, fo.LowRank) else blk)
    return out


def _offsets(ranges):
    return [r[0] for r in ranges] + [ranges[-1][0] + ranges[-1][1]]


@pytest.mark.parametrize("b", [32, 64, 128, 256])
@pytest.mark.parametrize("bottom", ["dense", "lowrank"])
def test_qr_tile_ops_match_lapack(torch_gpu, fa, b, bottom):
    """2 x 1 grid: geqrt(A00), tpqrt(A00, A10) against the oracle on identical tiles."""
    print(f"qr tiles b={b} bottom={bottom}", flush=True)
    rng = np.random.default_rng(b)
    A00 = rng.standard_normal((b, b)) + 4 * np.eye(b)
    if bottom == "dense":
        A10 = rng.standard_normal((b, b))
    else:
        r = 8
        u, _ = np.linalg.qr(rng.standard_normal((b, r)))
        s = np.triu(rng.standard_normal((r, r)))
        v = rng.standard_normal((r, b))
        A10 = fo.LowRank(u, s, v, 1e-8)
    H = fo.Hier(2, 1, [A00.copy(), A10.copy() if bottom == "dense" else A10.copy()],
                [(0, b), (b, b)], [(0, b)])
    G = fa.BlrMatrix.from_tiles([0, b, 2 * b], [0, b], _tiles_of(H), eps=1e-8, kcap=32)
    G.set_lra("rounded_addition")
    F = G.qr()
    Href, Tt = fo.tiled_householder_blr_qr(H)
    # diagonal tile: R above, reflectors below the diagonal
    np.testing.assert_allclose(G.tile(0, 0), Href[0, 0], rtol=0, atol=1e-11 * np.abs(Href[0, 0]).max())
    T00, tau00 = F.t(0, 0)
    np.testing.assert_allclose(T00, Tt[0, 0], rtol=0, atol=1e-11 * max(1.0, np.abs(Tt[0, 0]).max()))
    np.testing.assert_allclose(tau00, np.diag(Tt[0, 0]), rtol=0, atol=1e-13)
    T10, _ = F.t(1, 0)
    np.testing.assert_allclose(T10, Tt[1, 0], rtol=0, atol=1e-11 * max(1.0, np.abs(Tt[1, 0]).max()))
    got = G.tile(1, 0)
    ref = Href[1, 0]
    if bottom == "dense":
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-11 * np.abs(ref).max())
    else:
        u, s, v = got
        np.testing.assert_allclose(s, np.eye(s.shape[0]), atol=0)
        np.testing.assert_allclose(v, ref.V, rtol=0, atol=1e-11 * np.abs(ref.V).max())
        np.testing.assert_allclose(u, ref.U, rtol=0, atol=0)


def _build_qr_case(n_rows, n_cols, nleaf, eps, admis, admis_type):
    x = fo.get_sorted_random_vector(max(n_rows, n_cols))
    pts = [x]
    return pts


@pytest.mark.parametrize("n_cols", [128, 256])
@pytest.mark.parametrize("eps", [1e-6, 1e-8, 1e-10])
@pytest.mark.parametrize("admis", [0.0, 1.0, 4.0])
@pytest.mark.parametrize("admis_type", ["position", "geometry"])
def test_tiled_qr_fixed_accuracy(torch_gpu, fa, n_cols, eps, admis, admis_type):
    n_rows, nleaf = 256, 32
    print(f"blr_qr {n_rows}x{n_cols} eps={eps} admis={admis} {admis_type}", flush=True)
    x = fo.get_sorted_random_vector(max(n_rows, n_cols))
    pts = [x]
    nb_r, nb_c = n_rows // nleaf, n_cols // nleaf
    A = fa.BlrMatrix.from_kernel_rect(pts, n_rows, n_cols, nb_r, nb_c, eps=eps, admis=admis, admis_type=admis_type,
                                      kcap=64)
    A.set_lra("rounded_addition")
    D = fo.laplacend(pts, 0, n_rows, 0, n_cols)
    F = A.qr()
    Q = fa.BlrMatrix.from_kernel_rect(pts, n_rows, n_cols, nb_r, nb_c, eps=eps, admis=admis, admis_type=admis_type,
                                      kcap=64, kernel="identity")
    Q.set_lra("rounded_addition")
    A.apply_q(F, Q, trans=False)
    Qd = Q.dense()
    # R from the upper tiles (upper triangle of the diagonal tiles)
    Ad = A.dense()
    R = np.zeros((n_cols, n_cols))
    for i in range(nb_c):
        for j in range(i, nb_c):
            blk = Ad[i * nleaf:(i + 1) * nleaf, j * nleaf:(j + 1) * nleaf]
            R[i * nleaf:(i + 1) * nleaf, j * nleaf:(j + 1) * nleaf] = np.triu(blk) if i == j else blk
    residual = fo.l2_error(D, Qd @ R)
    assert residual <= eps
    A.apply_q(F, Q, trans=True)
    QtQ = Q.dense()[:n_cols, :n_cols]
    orth = fo.l2_error(np.eye(n_cols), QtQ)
    assert orth <= eps
    # oracle on its own construction of the same matrix: residual within 10x
    ker = lambda r0, nr, c0, nc: fo.laplacend(pts, r0, nr, c0, nc)
    H = fo.build_blr(ker, n_rows, n_cols, nb_r, nb_c, eps=eps, admis=admis, admis_type=admis_type, points=pts)
    H, Tt = fo.tiled_householder_blr_qr(H)
    idk = lambda r0, nr, c0, nc: fo.identity(pts, r0, nr, c0, nc)
    Qr = fo.build_blr(idk, n_rows, n_cols, nb_r, nb_c, eps=eps, admis=admis, admis_type=admis_type, points=pts)
    Qr = fo.left_multiply_tiled_reflector(H, Tt, Qr, False)
    ref_residual = fo.l2_error(D, Qr.dense() @ fo.blr_r_factor(H))
    assert residual <= 10 * max(ref_residual, 1e-15)


def test_clone_keeps_the_rank_cap_policy(torch_gpu, fa):
    """A clone (and its restore by copy_from) keeps the capped-rank policy of
    the matrix it copies: the tiled QR of the clone truncates at the cap as
    the original's does, with the same ranks."""
    import numpy as np
    from frank_amd.inputs import morton_sort, uniform_points
    N, nb, eps, cap = 512, 4, 1e-12, 8
    print(f"clone cap policy N={N} cap={cap}", flush=True)
    pts, _ = morton_sort(uniform_points(N, 2), 6)
    pts = [np.ascontiguousarray(p) for p in pts]
    ctx = fa.Context()
    ctx.set_rank_cap(True)
    A = fa.BlrMatrix.from_kernel_rect(pts, N, N, nb, nb, eps=eps, kcap=cap, ctx=ctx)
    A.set_lra("rounded_addition")
    B = A.clone()
    B.copy_from(A)
    A.qr()
    B.qr()
    assert np.array_equal(A.ranks(), B.ranks())
    assert A.ranks().max() == cap
