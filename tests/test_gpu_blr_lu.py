"""BLR-LU parity on the GPU against the oracle (the reference's BLR contracts).

test/blr_fixed_acc_test.cpp:46-83 (N in {128, 256}, nleaf 32, eps in
{1e-6, 1e-8, 1e-10}, admis in {0, 1, 4}, position/geometry admissibility):
compression error <= eps and solve error <= eps.  On top of the reference's
contract this checks the GPU path against the oracle on the same inputs:
identical ranks after construction and after the factorization, and the
factorization error ||A - LU|| / ||A|| within 10x of the oracle's.
"""
import numpy as np
import pytest

from oracle import frank_oracle as fo

pytestmark = pytest.mark.gpu


def build_both(N, nleaf, eps, admis, admis_type, rank=0):
    x = fo.get_sorted_random_vector(N)
    pts = [x]
    ker = lambda r0, nr, c0, nc: fo.laplacend(pts, r0, nr, c0, nc)
    nb = N // nleaf
    A = fo.build_blr(ker, N, N, nb, nb, eps=eps, rank=rank, admis=admis, admis_type=admis_type, points=pts)
    return pts, A


@pytest.mark.parametrize("N", [128, 256])
@pytest.mark.parametrize("eps", [1e-6, 1e-8, 1e-10])
@pytest.mark.parametrize("admis", [0.0, 1.0, 4.0])
@pytest.mark.parametrize("admis_type", ["position", "geometry"])
def test_blr_lu_fixed_accuracy(torch_gpu, fa, N, eps, admis, admis_type):
    torch = torch_gpu
    print(f"blr_lu N={N} eps={eps} admis={admis} {admis_type}", flush=True)
    pts, Aref = build_both(N, 32, eps, admis, admis_type)
    G = fa.BlrMatrix.from_kernel(pts, N // 32, eps=eps, admis=admis, admis_type=admis_type, kcap=32)
    # construction: same tile kinds and ranks, error <= eps
    assert np.array_equal(G.ranks(), Aref.ranks())
    D = fo.laplacend(pts, 0, N, 0, N)
    assert fo.l2_error(D, G.dense()) <= eps
    # solve contract: b = A x, x <- U^-1 L^-1 b
    xs = fo.random_uniform(N, 1)
    b = torch.zeros((N, 1), dtype=torch.float64, device="cuda")
    G.matvec(torch.from_numpy(xs).cuda(), b)
    G.set_lra("rounded_addition")
    G.getrf()
    G.solve(b)
    assert fo.l2_error(xs, b.cpu().numpy()) <= eps
    # oracle factorization on the same input
    Lr, Ur = fo.blr_getrf(Aref.copy(), lra="rounded_addition")
    gl = np.where(np.tril(np.ones((N // 32, N // 32)), -1) > 0, Lr.ranks(), Ur.ranks())
    assert np.array_equal(G.ranks(), gl)
    F = G.dense()
    # diagonal tiles hold packed unit-L \ U; off-diagonal tiles are whole L / U blocks
    Lg = np.zeros_like(F)
    Ug = np.zeros_like(F)
    nb = N // 32
    for i in range(nb):
        for j in range(nb):
            blk = F[i * 32:(i + 1) * 32, j * 32:(j + 1) * 32]
            if i > j:
                Lg[i * 32:(i + 1) * 32, j * 32:(j + 1) * 32] = blk
            elif i < j:
                Ug[i * 32:(i + 1) * 32, j * 32:(j + 1) * 32] = blk
            else:
                Lg[i * 32:(i + 1) * 32, j * 32:(j + 1) * 32] = np.tril(blk, -1) + np.eye(32)
                Ug[i * 32:(i + 1) * 32, j * 32:(j + 1) * 32] = np.triu(blk)
    A0 = Aref.dense()
    err_gpu = fo.l2_error(A0, Lg @ Ug)
    err_ref = fo.l2_error(A0, fo.lu_product(Lr, Ur))
    assert err_gpu <= 10 * max(err_ref, 1e-15), (err_gpu, err_ref)


def test_blr_lu_oracle_input_fixed_rank(torch_gpu, fa):
    """Fixed-rank BLR (fast rounded addition): GPU LU on the oracle's tiles."""
    torch = torch_gpu
    N, nb, k = 1024, 8, 8
    pts, Aref = build_both(N, N // nb, 0.0, 0.0, "position", rank=k)
    off = [i * (N // nb) for i in range(nb + 1)]
    tiles = []
    for blk in Aref.blocks:
        tiles.append(blk if isinstance(blk, np.ndarray) else (blk.U, blk.S, blk.V))
    G = fa.BlrMatrix.from_tiles(off, off, tiles, rank=k, kcap=k)
    G.getrf()
    Lr, Ur = fo.blr_getrf(Aref.copy())
    F = G.dense()
    LU = fo.lu_product(Lr, Ur)
    b = N // nb
    Lg = np.zeros_like(F)
    Ug = np.zeros_like(F)
    for i in range(nb):
        for j in range(nb):
            blk = F[i * b:(i + 1) * b, j * b:(j + 1) * b]
            if i > j:
                Lg[i * b:(i + 1) * b, j * b:(j + 1) * b] = blk
            elif i < j:
                Ug[i * b:(i + 1) * b, j * b:(j + 1) * b] = blk
            else:
                Lg[i * b:(i + 1) * b, j * b:(j + 1) * b] = np.tril(blk, -1) + np.eye(b)
                Ug[i * b:(i + 1) * b, j * b:(j + 1) * b] = np.triu(blk)
    A0 = Aref.dense()
    err_ref = fo.l2_error(A0, LU)
    err_gpu = fo.l2_error(A0, Lg @ Ug)
    assert err_gpu <= 10 * err_ref, (err_gpu, err_ref)
    assert np.all(G.ranks()[G.ranks() >= 0] == k)


@pytest.mark.parametrize("trans", [False, True])
def test_blr_solve_multi_rhs_and_transpose(torch_gpu, fa, trans):
    torch = torch_gpu
    N, nleaf, eps = 512, 64, 1e-10
    pts, Aref = build_both(N, nleaf, eps, 0.0, "position")
    G = fa.BlrMatrix.from_kernel(pts, N // nleaf, eps=eps, kcap=32)
    A0 = G.dense()
    G.set_lra("rounded_addition")
    G.getrf()
    rng = np.random.default_rng(11)
    xs = rng.standard_normal((N, 3))
    rhs = (A0.T if trans else A0) @ xs
    b = torch.from_numpy(rhs.copy()).cuda()
    G.solve(b, trans=trans)
    assert fo.l2_error(xs, b.cpu().numpy()) <= 1e-8


@pytest.mark.parametrize("zero_tile", ["lower", "upper"])
def test_blr_lu_rank_zero_priority_panel(torch_gpu, fa, zero_tile):
    """One of the two panels of a priority block has rank 0 (the other not).

    Nonsymmetric BLR matrix from host tiles, 4 x 4 tiles of 128, every
    off-diagonal tile low rank; tile (2, 1) (or its mirror (1, 2)) has rank 0.
    getrf + solve must match the dense FP64 solve of the same matrix.
    """
    torch = torch_gpu
    print(f"rank-zero priority panel: {zero_tile}", flush=True)
    nb, b = 4, 128
    N = nb * b
    rng = np.random.default_rng(7)
    off = [i * b for i in range(nb + 1)]
    tiles = []
    for i in range(nb):
        for j in range(nb):
            if i == j:
                tiles.append(rng.standard_normal((b, b)) + 4.0 * b * np.eye(b))
                continue
            r = 0 if (zero_tile == "lower" and (i, j) == (2, 1)) or (zero_tile == "upper" and (i, j) == (1, 2)) else 3
            U = np.linalg.qr(rng.standard_normal((b, max(r, 1))))[0][:, :r]
            V = np.linalg.qr(rng.standard_normal((b, max(r, 1))))[0][:, :r].T
            S = np.triu(rng.standard_normal((r, r))) + 5.0 * np.eye(r)
            tiles.append((U, S, V))
    G = fa.BlrMatrix.from_tiles(off, off, tiles, eps=1e-12, kcap=32)
    A0 = G.dense()
    G.set_lra("rounded_addition")
    G.getrf()
    xs = rng.standard_normal((N, 2))
    rhs = A0 @ xs
    bt = torch.from_numpy(rhs.copy()).cuda()
    G.solve(bt)
    err = np.linalg.norm(bt.cpu().numpy() - xs) / np.linalg.norm(xs)
    assert err <= 1e-9, err


@pytest.mark.parametrize("prec,trans", [("f64", False), ("f64", True), ("f32", False)])
def test_dataflow_solve_matches_launch_per_step(torch_gpu, fa, prec, trans, monkeypatch):
    """The one-launch dataflow block substitution against the launch-per-step
    sweep on the same factors (both trsm.cpp:40-90 block substitution; weak
    admissibility, every off-diagonal tile low rank): equal to 1e-12 (FP64) /
    1e-5 (FP32) relative, plain and transposed."""
    import numpy as np
    torch = torch_gpu
    N, nb, nrhs = 4096, 8, 8
    print(f"dataflow solve {prec} trans={trans}", flush=True)
    x = fo.get_sorted_random_vector(N)
    A = fa.BlrMatrix.from_kernel([x], nb, eps=1e-8, shift=1e-2, kcap=64)
    A.set_lra("rounded_addition")
    F = A.convert("f32") if prec == "f32" else A
    F.getrf()
    b = np.random.default_rng(5).standard_normal((N, nrhs))
    outs = []
    for mode in ("0", "1"):
        monkeypatch.setenv("FRK_SOLVE_DF", mode)
        v = torch.from_numpy(b).cuda().contiguous()
        if prec == "f32":  # FP32 factors solve through the refinement's first solve (no correction)
            xs = torch.zeros_like(v)
            F.refine(v, xs, tol=1.0, max_iter=0)
            v = xs
        else:
            F.solve(v, trans=trans)
        outs.append(v.cpu().numpy())
    tol = 1e-5 if prec == "f32" else 1e-12
    assert np.abs(outs[0] - outs[1]).max() <= tol * np.abs(outs[0]).max()
