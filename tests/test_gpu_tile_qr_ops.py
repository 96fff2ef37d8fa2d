"""Tile-level geqrt / larfb / tpqrt / tpmqrt and the typed gemm through the C
ABI (frk_dgeqrt_batched, frk_dlarfb_batched, frk_dtpqrt_batched,
frk_dtpmqrt_batched, frk_dgemm_tiles_batched) on identical tiles.

References: LAPACK (scipy) dgeqrt with nb = n (geqrt.cpp:34-43 calls dgeqrt3
on the whole tile), dgemqrt for larfb (larfb.cpp:37-56), dtpqrt / dtpmqrt
with l = 0 (tpqrt.cpp:35-53, tpmqrt.cpp:45-170); dense products of the
expanded operands for the typed gemm (gemm.cpp:126-272).  The same Householder
convention (dlarfg) is used, so R, V and T compare entry by entry (1e-12
relative).  Low-rank operands go through rounded additions: their results are
compared densely to the recompression accuracy (eps = 1e-12 -> 1e-9).
"""
import numpy as np
import pytest
from scipy.linalg import lapack

pytestmark = pytest.mark.gpu


def _rng(seed):
    return np.random.default_rng(seed)


def _cu(torch, x):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).cuda()


def _close(a, b, tol):
    scale = max(np.abs(b).max(), 1e-300)
    assert np.abs(a - b).max() <= tol * scale, (np.abs(a - b).max() / scale)


def _lr(rng, m, n, r):
    U, _ = np.linalg.qr(rng.standard_normal((m, r)))
    V, _ = np.linalg.qr(rng.standard_normal((n, r)))
    S = np.diag(10.0 ** (-np.arange(r) / 3.0))
    return U, S, V.T.copy()


@pytest.mark.parametrize("n", [64, 200, 256])
def test_geqrt_matches_lapack(torch_gpu, fa, n):
    from frank_amd import tile_ops as to
    torch = torch_gpu
    print(f"geqrt n={n}", flush=True)
    rng = _rng(n)
    As = [rng.standard_normal((n, n)) for _ in range(2)]
    A = [_cu(torch, a) for a in As]
    T = [torch.zeros((n, n), dtype=torch.float64, device="cuda") for _ in As]
    to.geqrt(A, T)
    for a, g, t in zip(As, A, T):
        ra, rt, info = lapack.dgeqrt(n, a)
        assert info == 0
        _close(g.cpu().numpy(), ra, 1e-12)
        _close(np.triu(t.cpu().numpy()), np.triu(rt), 1e-12)


@pytest.mark.parametrize("trans", [True, False])
def test_larfb_matches_dgemqrt_dense_and_lowrank(torch_gpu, fa, trans):
    from frank_amd import tile_ops as to
    torch = torch_gpu
    n, nc, r = 128, 96, 12
    print(f"larfb n={n} trans={trans}", flush=True)
    rng = _rng(5)
    ra, rt, _ = lapack.dgeqrt(n, rng.standard_normal((n, n)))
    Cd = rng.standard_normal((n, nc))
    U, S, V = _lr(rng, n, nc, r)
    tiles = [to.np_tile(Cd), to.Tile.lowrank(_cu(torch, U), _cu(torch, S), _cu(torch, V), cap=32)]
    to.larfb([_cu(torch, ra)] * 2, [_cu(torch, rt)] * 2, tiles, trans)
    tr = b"T" if trans else b"N"
    ref_d, info = lapack.dgemqrt(ra, rt, Cd, side=b"L", trans=tr)
    assert info == 0
    _close(tiles[0].d.cpu().numpy(), ref_d, 1e-12)
    ref_u, _ = lapack.dgemqrt(ra, rt, U, side=b"L", trans=tr)  # larfb.cpp: a low-rank C is updated through U
    Ug = tiles[1].factors()[0].cpu().numpy()
    _close(Ug, ref_u, 1e-12)
    assert tiles[1].rank == r


@pytest.mark.parametrize("m", [64, 256])
def test_tpqrt_matches_lapack_dense_and_lowrank(torch_gpu, fa, m):
    from frank_amd import tile_ops as to
    torch = torch_gpu
    n, r = 256, 10
    print(f"tpqrt n={n} m={m}", flush=True)
    rng = _rng(m)
    R0 = np.triu(rng.standard_normal((n, n))) + 4 * np.eye(n)
    B0 = rng.standard_normal((m, n))
    A = [_cu(torch, R0)]
    T = [torch.zeros((n, n), dtype=torch.float64, device="cuda")]
    Bt = [to.np_tile(B0)]
    to.tpqrt(A, Bt, T)
    ra, rb, rt, info = lapack.dtpqrt(0, n, R0, B0)
    assert info == 0
    _close(np.triu(A[0].cpu().numpy()), np.triu(ra), 1e-12)
    _close(Bt[0].d.cpu().numpy(), rb, 1e-12)
    _close(np.triu(T[0].cpu().numpy()), np.triu(rt), 1e-12)
    # low-rank B: V <- reflector parts of S V, S <- I (tpqrt.cpp:35-53)
    U, S, V = _lr(rng, m, n, r)
    A = [_cu(torch, R0)]
    T = [torch.zeros((n, n), dtype=torch.float64, device="cuda")]
    Bl = [to.Tile.lowrank(_cu(torch, U), _cu(torch, S), _cu(torch, V), cap=16)]
    to.tpqrt(A, Bl, T)
    ra, rb, rt, info = lapack.dtpqrt(0, n, R0, S @ V)
    assert info == 0
    Ug, Sg, Vg = (x.cpu().numpy() for x in Bl[0].factors())
    _close(np.triu(A[0].cpu().numpy()), np.triu(ra), 1e-12)
    _close(Vg, rb, 1e-12)
    np.testing.assert_allclose(Sg, np.eye(r), atol=0)
    _close(Ug, U, 0)
    _close(np.triu(T[0].cpu().numpy()), np.triu(rt), 1e-12)


def _tpqrt_factors(rng, n, m):
    R0 = np.triu(rng.standard_normal((n, n))) + 4 * np.eye(n)
    _, vb, t, _ = lapack.dtpqrt(0, n, R0, rng.standard_normal((m, n)))
    return vb, t


@pytest.mark.parametrize("trans", [True, False])
def test_tpmqrt_all_dense_matches_lapack(torch_gpu, fa, trans):
    from frank_amd import tile_ops as to
    torch = torch_gpu
    n, m, nc = 128, 128, 96
    print(f"tpmqrt dense trans={trans}", flush=True)
    rng = _rng(11)
    vb, t = _tpqrt_factors(rng, n, m)
    A0, B0 = rng.standard_normal((n, nc)), rng.standard_normal((m, nc))
    A, B = [to.np_tile(A0)], [to.np_tile(B0)]
    to.tpmqrt([to.np_tile(vb)], [_cu(torch, t)], A, B, trans)
    ra, rb, info = lapack.dtpmqrt(0, vb, t, A0, B0, side=b"L", trans=b"T" if trans else b"N")
    assert info == 0
    _close(A[0].d.cpu().numpy(), ra, 1e-12)
    _close(B[0].d.cpu().numpy(), rb, 1e-12)


@pytest.mark.parametrize("kinds", ["LDD", "DLL", "LLL", "LDL", "LLD"])
def test_tpmqrt_lowrank_operands_match_dense_lapack(torch_gpu, fa, kinds):
    """Every low-rank specialisation of tpmqrt.cpp:61-170 (kinds = V, A, B)
    against dtpmqrt on the expanded dense operands, to the recompression
    accuracy of the rounded additions (eps = 1e-12)."""
    from frank_amd import tile_ops as to
    torch = torch_gpu
    n, m, nc, r, eps = 128, 128, 128, 8, 1e-12
    print(f"tpmqrt kinds={kinds}", flush=True)
    rng = _rng(17)
    # a low-rank reflector block: U (I) V with V the reflector parts of S V (tpqrt of a low-rank B)
    R0 = np.triu(rng.standard_normal((n, n))) + 4 * np.eye(n)
    Ub, Sb, Vb = _lr(rng, m, n, r)
    _, vpart, t, _ = lapack.dtpqrt(0, n, R0, Sb @ Vb)
    Vlr = (Ub, np.eye(r), vpart)
    Vd = Ub @ vpart
    if kinds[0] == "D":  # a dense reflector block of the same shape
        vb, t = _tpqrt_factors(rng, n, m)
        Vd, Vtile = vb, to.np_tile(vb)
    else:
        Vtile = to.Tile.lowrank(*(_cu(torch, x) for x in Vlr), cap=16)
    opnds = []
    dense = []
    for kd, rows in ((kinds[1], n), (kinds[2], m)):
        U, S, V = _lr(rng, rows, nc, r)
        dense.append(U @ S @ V)
        opnds.append(to.np_tile(U @ S @ V) if kd == "D" else
                     to.Tile.lowrank(_cu(torch, U), _cu(torch, S), _cu(torch, V), cap=64, eps=eps))
    to.tpmqrt([Vtile], [_cu(torch, t)], [opnds[0]], [opnds[1]], True)
    ra, rb, info = lapack.dtpmqrt(0, Vd, t, dense[0], dense[1], side=b"L", trans=b"T")
    assert info == 0
    ga, gb = (o.dense().cpu().numpy() for o in opnds)
    assert np.linalg.norm(ga - ra) <= 1e-9 * np.linalg.norm(ra)
    assert np.linalg.norm(gb - rb) <= 1e-9 * np.linalg.norm(rb)


@pytest.mark.parametrize("kinds", ["LLL", "DDL", "LDL", "DLL", "LLD", "LDD", "DLD", "DDD"])
@pytest.mark.parametrize("trans", [(False, False), (True, False), (False, True)])
def test_typed_gemm_tiles(torch_gpu, fa, kinds, trans):
    """gemm(A, B, C, alpha, beta, TransA, TransB) for every operand kind
    (gemm.cpp:126-272; LowRank x LowRank -> LowRank is :253-272) against the
    dense product; low-rank targets to the recompression accuracy."""
    from frank_amd import tile_ops as to
    torch = torch_gpu
    m, k, n, r, eps = 96, 80, 112, 6, 1e-12
    alpha, beta = -0.75, 0.5
    print(f"typed gemm kinds={kinds} trans={trans}", flush=True)
    rng = _rng(23)

    def make(kind, rows, cols, cap=32):
        U, S, V = _lr(rng, rows, cols, r)
        d = U @ S @ V
        if kind == "D":
            return to.np_tile(d), d
        return to.Tile.lowrank(_cu(torch, U), _cu(torch, S), _cu(torch, V), cap=cap, eps=eps), d

    ta, tb = trans
    A, Ad = make(kinds[0], k if ta else m, m if ta else k)
    B, Bd = make(kinds[1], n if tb else k, k if tb else n)
    Ct, Cd = make(kinds[2], m, n, cap=64)
    to.gemm([A], [B], [Ct], alpha=alpha, beta=beta, trans_a=ta, trans_b=tb)
    ref = alpha * ((Ad.T if ta else Ad) @ (Bd.T if tb else Bd)) + beta * Cd
    got = Ct.dense().cpu().numpy()
    tol = 1e-12 if kinds[2] == "D" else 1e-9
    assert np.linalg.norm(got - ref) <= tol * np.linalg.norm(ref)
