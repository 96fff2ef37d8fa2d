"""Adjoint of the BLR solve (no reference counterpart).

L = 1/2 ||A(ell)^-1 b||^2 with A(ell) = exp(-|x_i - x_j|^2 / (2 ell^2)) + 1e-4 I,
x in U[0,1]^3 (one mt19937 stream, Morton-ordered).  The gradients from
BlrSolve.backward (transposed block sweeps on the forward's factors plus the
matrix-free dA/d ell contraction) are checked against the dense FP64 LAPACK
gradient and against central finite differences (h = 1e-6) of the dense FP64
loss.  Tolerances: the BLR solve is accurate to about cond(A) * eps_BLR
(cond(A) <= 1e6 here, eps_BLR = 1e-13).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dense_gradients(pts, ell, b, shift=1e-4):
    d2 = sum((p[:, None] - p[None, :]) ** 2 for p in pts)
    G = np.exp(-d2 / (2.0 * ell * ell))
    A = G + shift * np.eye(len(pts[0]))
    u = np.linalg.solve(A, b)
    lam = np.linalg.solve(A.T, u)
    dA = G * d2 / ell ** 3
    return 0.5 * float((u * u).sum()), lam, -float((lam * (dA @ u)).sum())


@pytest.mark.parametrize("N,ell,nrhs", [(512, 0.1, 1), (1024, 0.2, 1), (1024, 0.3, 4)])
def test_adjoint_matches_dense_and_finite_differences(torch_gpu, fa, N, ell, nrhs):
    torch = torch_gpu
    from frank_amd.autograd import blr_solve
    from frank_amd.inputs import morton_sort, uniform_points
    print(f"adjoint N={N} ell={ell} nrhs={nrhs}", flush=True)
    pts, _ = morton_sort(uniform_points(N, 3), 4)
    b_np = np.random.default_rng(N + nrhs).standard_normal((N, nrhs))
    # 64-point tiles keep every rank within the 64-column slots
    nb, eps = N // 64, 1e-13
    b = torch.from_numpy(b_np).cuda().requires_grad_(True)
    ell_t = torch.tensor(ell, dtype=torch.float64, requires_grad=True)
    u = blr_solve(b, ell_t, pts, nb, eps=eps, kcap=64)
    loss = 0.5 * (u * u).sum()
    loss.backward()
    L_ref, lam_ref, dell_ref = _dense_gradients(list(pts), ell, b_np)
    assert abs(loss.item() - L_ref) <= 1e-4 * abs(L_ref)
    lam = b.grad.detach().cpu().numpy()
    assert np.linalg.norm(lam - lam_ref) <= 1e-4 * np.linalg.norm(lam_ref)
    assert abs(ell_t.grad.item() - dell_ref) <= 1e-4 * abs(dell_ref)
    h = 1e-6
    fd = (_dense_gradients(list(pts), ell + h, b_np)[0] - _dense_gradients(list(pts), ell - h, b_np)[0]) / (2 * h)
    assert abs(ell_t.grad.item() - fd) <= 1e-4 * abs(fd)


def test_dA_blocks_is_minus_lambda_u_transpose(torch_gpu, fa):
    from frank_amd.autograd import dA_blocks
    rng = np.random.default_rng(3)
    lam = rng.standard_normal((256, 2))
    u = rng.standard_normal((256, 2))
    offs = [0, 64, 128, 192, 256]
    G = dA_blocks(lam, u, offs, offs)
    np.testing.assert_allclose(G.dense(), -lam @ u.T, rtol=0, atol=1e-13)
    r = G.ranks()
    assert all(r[i, j] == 2 for i in range(4) for j in range(4) if i != j)
