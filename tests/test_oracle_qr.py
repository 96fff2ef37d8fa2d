"""Pinning the oracle's tiled-QR side and its naive addition (CPU only).

Every check compares the oracle with something it does not share code with:
BLAS dtrmm, LAPACK dgeqrt / dgemqrt / dtpqrt / dtpmqrt, the explicit
reflector matrix Q = I - W T W^T formed densely, numpy's QR and SVD, and the
reference's own tolerance contracts for the tiled QR
(test/blr_fixed_acc_test.cpp:148-190).  Each planted-mistake test mutates the
oracle and shows that the corresponding check fails.
"""
import numpy as np
import pytest
from scipy.linalg import blas

from oracle import frank_oracle as fo


def _lowrank(rng, m, n, k, decay=0.5):
    U = np.linalg.qr(rng.standard_normal((m, k)))[0]
    V = np.linalg.qr(rng.standard_normal((n, k)))[0].T
    S = np.diag(decay ** np.arange(k))
    return fo.LowRank(U, S, V, 0.0)


def _dense(x):
    return x.dense() if isinstance(x, fo.LowRank) else np.asarray(x)


def _reflector_q(V, T):
    """Q = I - V T V^T for unit-lower reflectors stored below the diagonal of V (geqrt layout)."""
    n = V.shape[0]
    W = np.tril(V, -1)[:, :T.shape[0]] + np.eye(n, T.shape[0])
    return np.eye(n) - W @ T @ W.T


def _pentagonal_q(Vb, T):
    """Q = I - W T W^T with W = [I; Vb] (tpqrt, l = 0)."""
    k = T.shape[0]
    W = np.vstack([np.eye(k), _dense(Vb)])
    return np.eye(W.shape[0]) - W @ T @ W.T


# --------------------------------------------------------------------------- trmm
def _trmm_err(side, uplo, trans, diag, lowrank):
    rng = np.random.default_rng(3)
    A = rng.standard_normal((12, 12))
    if lowrank:
        B = _lowrank(rng, 12, 12, 4)
        Bd = B.dense()
    else:
        B = rng.standard_normal((12, 12))
        Bd = B.copy()
    out = fo.trmm(A, B, side, uplo, trans, diag, 0.5)
    ref = blas.dtrmm(0.5, A, Bd, side=0 if side == "left" else 1, lower=1 if uplo == "lower" else 0,
                     trans_a=1 if trans == "t" else 0, diag=1 if diag == "u" else 0)
    return np.abs(_dense(out) - ref).max() / np.abs(ref).max()


@pytest.mark.parametrize("lowrank", [False, True])
@pytest.mark.parametrize("diag", ["n", "u"])
@pytest.mark.parametrize("trans", ["n", "t"])
@pytest.mark.parametrize("uplo", ["lower", "upper"])
@pytest.mark.parametrize("side", ["left", "right"])
def test_trmm_matches_blas_dtrmm(side, uplo, trans, diag, lowrank):
    assert _trmm_err(side, uplo, trans, diag, lowrank) < 1e-14


def test_trmm_planted_mistake_is_caught(monkeypatch):
    orig = fo.trmm

    def wrong(A, B, side, uplo, trans="n", diag="n", alpha=1.0):  # ignores the transpose
        return orig(A, B, side, uplo, "n", diag, alpha)
    monkeypatch.setattr(fo, "trmm", wrong)
    assert _trmm_err("left", "upper", "t", "n", False) > 1e-3


# --------------------------------------------------------------------------- geqrt / larfb
def test_geqrt_reconstructs_with_explicit_q():
    rng = np.random.default_rng(5)
    A = rng.standard_normal((16, 16))
    VR, T = fo.geqrt(A)
    Q = _reflector_q(VR, T)
    R = np.triu(VR)
    assert np.abs(Q @ R - A).max() < 1e-13 * np.abs(A).max()
    assert np.abs(Q.T @ Q - np.eye(16)).max() < 1e-14


def _larfb_err(trans, lowrank):
    rng = np.random.default_rng(9)
    VR, T = fo.geqrt(rng.standard_normal((16, 16)))
    Q = _reflector_q(VR, T)
    C = _lowrank(rng, 16, 10, 3) if lowrank else rng.standard_normal((16, 10))
    Cd = _dense(C).copy()
    out = fo.larfb(VR, T, C, trans)
    ref = (Q.T if trans else Q) @ Cd
    return np.abs(_dense(out) - ref).max() / np.abs(ref).max()


@pytest.mark.parametrize("lowrank", [False, True])
@pytest.mark.parametrize("trans", [False, True])
def test_larfb_matches_explicit_reflector(trans, lowrank):
    assert _larfb_err(trans, lowrank) < 1e-13


def test_larfb_lowrank_acts_on_u_only():
    rng = np.random.default_rng(10)
    VR, T = fo.geqrt(rng.standard_normal((16, 16)))
    C = _lowrank(rng, 16, 10, 3)
    out = fo.larfb(VR, T, C.copy(), True)
    assert np.array_equal(out.S, C.S) and np.array_equal(out.V, C.V)


def test_larfb_planted_mistake_is_caught(monkeypatch):
    orig = fo.larfb

    def wrong(V, T, C, trans):  # applies Q instead of Q^T
        return orig(V, T, C, not trans)
    monkeypatch.setattr(fo, "larfb", wrong)
    assert _larfb_err(True, False) > 1e-3


# --------------------------------------------------------------------------- tpqrt
@pytest.mark.parametrize("lowrank", [False, True])
def test_tpqrt_reconstructs(lowrank):
    rng = np.random.default_rng(11)
    A = np.triu(rng.standard_normal((16, 16))) + 4 * np.eye(16)
    B = _lowrank(rng, 16, 16, 5) if lowrank else rng.standard_normal((16, 16))
    top = np.vstack([A, _dense(B)])
    A2, B2, T = fo.tpqrt(A.copy(), B.copy() if not lowrank else B.copy())
    R = np.triu(A2)
    if lowrank:
        # the low-rank block keeps U; the reflectors live in its V (S = I)
        assert np.array_equal(B2.S, np.eye(B.rank)) and np.array_equal(B2.U, B.U)
        W = B2.U @ B2.V
    else:
        W = B2
    Q = _pentagonal_q(W, T)
    assert np.abs(Q.T @ Q - np.eye(32)).max() < 1e-13
    recon = Q @ np.vstack([R, np.zeros((16, 16))])
    assert np.abs(recon - top).max() < 1e-13 * np.abs(top).max()
    # R agrees with numpy's QR of the stacked matrix up to row signs
    Rn = np.linalg.qr(top, mode="r")
    sg = np.sign(np.diag(R)) * np.sign(np.diag(Rn))
    assert np.abs(R - sg[:, None] * Rn).max() < 1e-12 * np.abs(Rn).max()


# --------------------------------------------------------------------------- tpmqrt
def _tpmqrt_case(v_lr, a_lr, b_lr, trans, eps=1e-10):
    """Oracle tpmqrt on one combination of operand types against the explicit
    pentagonal reflector applied to the dense expansions."""
    rng = np.random.default_rng(13)
    b, nc = 16, 16
    R0 = np.triu(rng.standard_normal((b, b))) + 4 * np.eye(b)
    Bv = _lowrank(rng, b, b, 5) if v_lr else rng.standard_normal((b, b))
    _, V, T = fo.tpqrt(R0, Bv.copy() if v_lr else Bv.copy())
    Q = _pentagonal_q(V, T)
    A = _lowrank(rng, b, nc, 4) if a_lr else rng.standard_normal((b, nc))
    B = _lowrank(rng, b, nc, 3) if b_lr else rng.standard_normal((b, nc))
    if a_lr:
        A.eps = eps
    if b_lr:
        B.eps = eps
    stack = np.vstack([_dense(A), _dense(B)])
    ref = (Q.T if trans else Q) @ stack
    A2, B2 = fo.tpmqrt(V, T, A.copy(), B.copy(), trans)
    got = np.vstack([_dense(A2), _dense(B2)])
    return np.abs(got - ref).max() / np.abs(ref).max(), (A2, B2)


@pytest.mark.parametrize("trans", [False, True])
@pytest.mark.parametrize("types", ["DDD", "LDD", "DLD", "DDL", "LLD", "LDL", "DLL", "LLL"])
def test_tpmqrt_every_specialisation_matches_explicit_reflector(types, trans):
    """tpmqrt.cpp:45-170: V / A / B each Dense (D) or LowRank (L); T is dense."""
    v_lr, a_lr, b_lr = (c == "L" for c in types)
    err, (A2, B2) = _tpmqrt_case(v_lr, a_lr, b_lr, trans)
    # low-rank targets are recompressed at eps = 1e-10 (rounded addition)
    assert err < (1e-8 if (a_lr or b_lr) else 1e-13), err
    # the result types follow the reference: a LowRank A or B stays LowRank
    assert isinstance(A2, fo.LowRank) == a_lr and isinstance(B2, fo.LowRank) == b_lr


def test_tpmqrt_all_dense_matches_lapack_dtpmqrt_path():
    """The all-dense branch is LAPACK dtpmqrt: same numbers as the generic
    (typed gemm + trmm) sequence of the low-rank branches on Dense operands."""
    rng = np.random.default_rng(17)
    R0 = np.triu(rng.standard_normal((16, 16))) + 4 * np.eye(16)
    _, V, T = fo.tpqrt(R0, rng.standard_normal((16, 16)))
    A = rng.standard_normal((16, 12))
    B = rng.standard_normal((16, 12))
    a1, b1 = fo.tpmqrt(V, T, A.copy(), B.copy(), True)
    C = A + V.T @ B
    C = np.triu(T).T @ C
    np.testing.assert_allclose(a1, A - C, atol=1e-13)
    np.testing.assert_allclose(b1, B - V @ C, atol=1e-13)


def test_tpmqrt_planted_mistake_is_caught(monkeypatch):
    orig = fo.trmm

    def wrong(A, B, side, uplo, trans="n", diag="n", alpha=1.0):  # T used without its transpose
        return orig(A, B, side, uplo, "n", diag, alpha)
    monkeypatch.setattr(fo, "trmm", wrong)
    err, _ = _tpmqrt_case(True, False, False, True)
    assert err > 1e-3


# --------------------------------------------------------------------------- tiled BLR-QR
def _tiled_qr_metrics(N, nleaf, eps, admis, n_cols=None):
    n_cols = n_cols or N
    x = fo.get_sorted_random_vector(max(N, n_cols))
    pts = [x]
    ker = lambda r0, nr, c0, nc: fo.laplacend(pts, r0, nr, c0, nc)
    nb_r, nb_c = N // nleaf, n_cols // nleaf
    H = fo.build_blr(ker, N, n_cols, nb_r, nb_c, eps=eps, admis=admis, points=pts)
    A0 = H.dense()
    D = fo.laplacend(pts, 0, N, 0, n_cols)
    H, Tt = fo.tiled_householder_blr_qr(H)
    idk = lambda r0, nr, c0, nc: fo.identity(pts, r0, nr, c0, nc)
    Q = fo.build_blr(idk, N, n_cols, nb_r, nb_c, eps=eps, admis=admis, points=pts)
    Q = fo.left_multiply_tiled_reflector(H, Tt, Q, False)
    Qd = Q.dense()
    R = fo.blr_r_factor(H)
    residual = fo.l2_error(D, Qd @ R)
    Q = fo.left_multiply_tiled_reflector(H, Tt, Q, True)
    orth = fo.l2_error(np.eye(n_cols), Q.dense()[:n_cols, :n_cols])
    return residual, orth, R, A0


@pytest.mark.parametrize("n_cols", [128, 256])
@pytest.mark.parametrize("eps", [1e-6, 1e-10])
@pytest.mark.parametrize("admis", [0.0, 1.0])
def test_tiled_householder_blr_qr_contract(n_cols, eps, admis):
    """test/blr_fixed_acc_test.cpp:148-190 (256 rows, nleaf 32): residual and
    orthogonality <= eps, and R equal to numpy's R of the compressed matrix up
    to row signs."""
    residual, orth, R, A0 = _tiled_qr_metrics(256, 32, eps, admis, n_cols)
    assert residual <= eps, residual
    assert orth <= eps, orth
    Rn = np.linalg.qr(A0, mode="r")
    sg = np.sign(np.diag(R)) * np.sign(np.diag(Rn))
    assert np.linalg.norm(R - sg[:, None] * Rn) <= 10 * eps * np.linalg.norm(Rn)


def test_tiled_qr_planted_mistake_is_caught(monkeypatch):
    """A larfb applied with the wrong transpose inside the composite breaks the contract."""
    orig = fo.larfb
    monkeypatch.setattr(fo, "larfb", lambda V, T, C, trans: orig(V, T, C, False))
    residual, orth, _, _ = _tiled_qr_metrics(128, 32, 1e-8, 0.0)
    assert residual > 1e-3 or orth > 1e-3


def test_tiled_qr_planted_tpmqrt_mistake_is_caught(monkeypatch):
    orig = fo.tpmqrt

    def wrong(V, T, A, B, trans, lra="rounded_addition"):  # drops the update of B
        A2, _ = orig(V, T, A, B, trans, lra)
        return A2, B
    monkeypatch.setattr(fo, "tpmqrt", wrong)
    residual, orth, _, _ = _tiled_qr_metrics(128, 32, 1e-8, 0.0)
    assert residual > 1e-3 or orth > 1e-3


def test_blr_r_factor_is_upper_triangle():
    rng = np.random.default_rng(21)
    nb, b = 3, 8
    blocks = []
    for i in range(nb):
        for j in range(nb):
            blocks.append(_lowrank(rng, b, b, 2) if i != j else rng.standard_normal((b, b)))
    rr = [(i * b, b) for i in range(nb)]
    H = fo.Hier(nb, nb, blocks, rr, list(rr))
    np.testing.assert_array_equal(fo.blr_r_factor(H), np.triu(H.dense()))


# --------------------------------------------------------------------------- naive addition
def test_naive_addition_concatenates_exactly():
    """addition.cpp:68-89 below min(m, n): the sum is exact and the rank grows."""
    rng = np.random.default_rng(23)
    A = _lowrank(rng, 40, 32, 5)
    B = _lowrank(rng, 40, 32, 6)
    C = fo.naive_addition(A, B)
    assert C.rank == 11
    np.testing.assert_allclose(C.dense(), A.dense() + B.dense(), atol=1e-14)


def test_naive_addition_recompresses_past_min_dim():
    """rank sum >= min(m, n): LowRank(Dense(A) + Dense(B), A.rank) through rsvd,
    near the best rank-A.rank approximation of the dense sum (Eckart-Young)."""
    rng = np.random.default_rng(29)
    A = _lowrank(rng, 24, 20, 12, decay=0.3)
    B = _lowrank(rng, 24, 20, 10, decay=0.3)
    C = fo.naive_addition(A, B)
    assert C.rank == A.rank
    Sm = A.dense() + B.dense()
    s = np.linalg.svd(Sm, compute_uv=False)
    best = np.sqrt(np.sum(s[A.rank:] ** 2))
    assert np.linalg.norm(Sm - C.dense()) <= 2.0 * best + 1e-14


def test_naive_addition_planted_mistake_is_caught(monkeypatch):
    """A recompression to B's rank instead of A's changes the rank contract."""
    orig = fo.LowRank.from_rank
    monkeypatch.setattr(fo.LowRank, "from_rank", staticmethod(lambda a, rank: orig(a, rank - 2)))
    rng = np.random.default_rng(29)
    A = _lowrank(rng, 24, 20, 12, decay=0.3)
    B = _lowrank(rng, 24, 20, 10, decay=0.3)
    assert fo.naive_addition(A, B).rank != A.rank
