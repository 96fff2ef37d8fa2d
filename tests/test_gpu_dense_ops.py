"""Dense operator surface on the GPU against the oracle (LAPACK semantics).

qr / rq / svd: reconstruction, orthogonality and the factors themselves
(Householder sign convention of LAPACK, so R and Q compare directly;
singular vectors compare through their products).  truncated_geqp3: same
rank as the oracle and Q R ~ A.  rsvd: the reference's Omega, so U S V
matches the oracle's to rounding.  ID / RID: the reference's contract
(test/dense_id_test.cpp:25-70, error < 1e-7 at rank 16), at n = 512 (the
library's tile limit; the reference test uses n = 2048).
"""
import numpy as np
import pytest

from oracle import frank_oracle as fo

pytestmark = pytest.mark.gpu


def _lap(m, n, off=0):
    x = fo.get_sorted_random_vector(m + n + off)
    return fo.laplacend([x], 0, m, m + off, n)


@pytest.mark.parametrize("m,n", [(64, 32), (200, 64), (512, 128), (37, 37)])
def test_qr_matches_lapack(torch_gpu, m, n):
    torch = torch_gpu
    from frank_amd import dense
    A = np.random.default_rng(m * n).standard_normal((m, n))
    Q, R = dense.qr(torch.from_numpy(A).cuda())
    Qr, Rr = fo.qr(A)
    np.testing.assert_allclose(R.cpu().numpy(), Rr, rtol=0, atol=1e-12 * np.abs(Rr).max())
    np.testing.assert_allclose(Q.cpu().numpy(), Qr, rtol=0, atol=1e-12)


@pytest.mark.parametrize("m,n", [(32, 64), (64, 200), (128, 512), (20, 20)])
def test_rq_matches_lapack(torch_gpu, m, n):
    torch = torch_gpu
    from frank_amd import dense
    A = np.random.default_rng(m + n).standard_normal((m, n))
    R, Q = dense.rq(torch.from_numpy(A).cuda())
    Rr, Qr = fo.rq(A)
    np.testing.assert_allclose(R.cpu().numpy(), Rr, rtol=0, atol=1e-12 * np.abs(Rr).max())
    np.testing.assert_allclose(Q.cpu().numpy(), Qr, rtol=0, atol=1e-12)


@pytest.mark.parametrize("m,n", [(64, 64), (300, 40), (40, 300), (512, 64)])
def test_svd_reconstructs_with_lapack_singular_values(torch_gpu, m, n):
    torch = torch_gpu
    from frank_amd import dense
    A = _lap(m, n)
    U, S, V = dense.svd(torch.from_numpy(A).cuda())
    U, S, V = U.cpu().numpy(), S.cpu().numpy(), V.cpu().numpy()
    s_ref = np.linalg.svd(A, compute_uv=False)
    np.testing.assert_allclose(np.diag(S), s_ref, rtol=0, atol=1e-13 * s_ref[0])
    np.testing.assert_allclose(S, np.diag(np.diag(S)), atol=0)
    np.testing.assert_allclose(U @ S @ V, A, rtol=0, atol=1e-13 * np.abs(A).max() * max(m, n))
    K = min(m, n)
    np.testing.assert_allclose(U.T @ U, np.eye(K), atol=1e-12)
    np.testing.assert_allclose(V @ V.T, np.eye(K), atol=1e-12)


@pytest.mark.parametrize("m,n,rank", [(1, 1, 1), (2, 2, 2), (3, 5, 3), (9, 9, 9), (16, 16, 16), (16, 7, 7),
                                       (7, 16, 7), (12, 12, 4), (13, 13, 5), (15, 15, 0),
                                       (32, 32, 4), (24, 30, 11), (40, 40, 7), (64, 64, 0), (100, 48, 20)])
def test_svd_small_cores(torch_gpu, m, n, rank):
    """K = min(m, n) <= 16: the one-wave Jacobi with the core in registers (odd K pads
    the round-robin with a zero column); K up to 64: the LDS Jacobi after QR
    preconditioning.  Rank-deficient inputs keep singular values that are zero,
    whose left vectors come from the orthonormal completion (K - rank of them)."""
    torch = torch_gpu
    from frank_amd import dense
    rng = np.random.default_rng(100 * m + n + rank)
    A = rng.standard_normal((m, rank)) @ rng.standard_normal((rank, n)) if rank else np.zeros((m, n))
    U, S, V = dense.svd(torch.from_numpy(A).cuda())
    U, S, V = U.cpu().numpy(), S.cpu().numpy(), V.cpu().numpy()
    s_ref = np.linalg.svd(A, compute_uv=False)
    scale = max(s_ref[0], 1.0)
    np.testing.assert_allclose(np.diag(S), s_ref, rtol=0, atol=1e-13 * scale)
    assert np.all(np.diff(np.diag(S)) <= 0)
    np.testing.assert_allclose(U @ S @ V, A, rtol=0, atol=1e-13 * scale * max(m, n))
    K = min(m, n)
    np.testing.assert_allclose(U.T @ U, np.eye(K), atol=1e-12)
    np.testing.assert_allclose(V @ V.T, np.eye(K), atol=1e-12)


@pytest.mark.parametrize("eps", [1e-6, 1e-10])
@pytest.mark.parametrize("m,n", [(128, 128), (256, 192)])
def test_truncated_geqp3_rank_and_reconstruction(torch_gpu, eps, m, n):
    torch = torch_gpu
    from frank_amd import dense
    A = _lap(m, n)
    Q, R = dense.truncated_geqp3(torch.from_numpy(A).cuda(), eps)
    Qr, Rr = fo.truncated_geqp3(A, eps)
    assert Q.shape[1] == Qr.shape[1]
    err = fo.l2_error(A, Q.cpu().numpy() @ R.cpu().numpy())
    assert abs(err - fo.l2_error(A, Qr @ Rr)) <= 10 * eps


@pytest.mark.parametrize("p", [8, 21])
def test_rsvd_matches_oracle(torch_gpu, p):
    torch = torch_gpu
    from frank_amd import dense
    A = _lap(256, 256)
    U, S, V = dense.rsvd(torch.from_numpy(A).cuda(), p)
    Ur, Sr, Vr = fo.rsvd(A, p)
    np.testing.assert_allclose(np.diag(S.cpu().numpy()), np.diag(Sr), rtol=0, atol=1e-12 * Sr[0, 0])
    np.testing.assert_allclose(U.cpu().numpy() @ S.cpu().numpy() @ V.cpu().numpy(), Ur @ Sr @ Vr, rtol=0,
                               atol=1e-12 * np.abs(A).max())


def test_one_sided_id_matches_oracle(torch_gpu):
    torch = torch_gpu
    from frank_amd import dense
    A = _lap(256, 300)
    V, cols = dense.one_sided_id(torch.from_numpy(A).cuda(), 16)
    Vr, colsr = fo.one_sided_id(A, 16)
    assert cols == [int(c) for c in colsr]
    # V = R11^-1 R12 inherits the conditioning of R11: compare to 1e-6 and by reconstruction
    np.testing.assert_allclose(V.cpu().numpy(), Vr, rtol=0, atol=1e-6 * np.abs(Vr).max())
    err = fo.l2_error(A, A[:, cols] @ V.cpu().numpy())
    assert err <= 10 * max(fo.l2_error(A, A[:, colsr] @ Vr), 1e-15)


def test_id_and_rid_contract(torch_gpu):
    """test/dense_id_test.cpp:25-70: ||A - U S V|| / ||A|| < 1e-7 at rank 16."""
    torch = torch_gpu
    from frank_amd import dense
    n = 512
    x = fo.get_sorted_random_vector(2 * n)
    A = fo.laplacend([x], 0, n, n, n)
    At = torch.from_numpy(A).cuda()
    U, S, V = dense.id(At, 16)
    assert fo.l2_error(A, U.cpu().numpy() @ S.cpu().numpy() @ V.cpu().numpy()) < 1e-7
    U, S, V = dense.rid(At, 16 + 5, 16)
    assert fo.l2_error(A, U.cpu().numpy() @ S.cpu().numpy() @ V.cpu().numpy()) < 1e-7
