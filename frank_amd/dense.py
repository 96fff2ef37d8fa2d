"""Dense operator surface on GPU tensors: qr, rq, svd, truncated_geqp3, rsvd,
one_sided_id, id, rid (include/FRANK/operations/LAPACK.h,
randomized_factorizations.h).  All arithmetic runs in the library's kernels;
torch only holds the device buffers.  Inputs are float64 CUDA tensors,
row-major (contiguous), at most 512 x 512.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Tuple

import numpy as np
import torch

from . import _lib as L
from .blr import Context, _dptr, _hptr, default_context, gemm
from .inputs import mt19937_uniform


def _arr(ts):
    return (C.c_void_p * len(ts))(*[_dptr(t) for t in ts])


def _i32(*v):
    """ctypes int32 array: passed as an argument object, it lives through the call
    (a bare address of a temporary numpy array would not)."""
    return (C.c_int32 * len(v))(*v)


def _i64(*v):
    return (C.c_int64 * len(v))(*v)


def _check(A: torch.Tensor):
    if A.dtype != torch.float64 or not A.is_cuda or A.dim() != 2:
        raise ValueError("expected a 2-D float64 CUDA tensor")
    return A.contiguous()


def qr(A: torch.Tensor, ctx: Optional[Context] = None) -> Tuple[torch.Tensor, torch.Tensor]:
    """qr(A, Q, R) (qr.cpp:65-90): thin Q (m x k), R (k x n), k = min(m, n)."""
    ctx = ctx or default_context()
    A = _check(A)
    m, n = A.shape
    k = min(m, n)
    Q = torch.empty((m, k), dtype=A.dtype, device=A.device)
    R = torch.empty((k, n), dtype=A.dtype, device=A.device)
    L.check(L.lib().frk_dqr_batched(ctx.h, 1, _i32(m), _i32(n), _arr([A]), _i64(n),
                                    _arr([Q]), _i64(k), _arr([R]), _i64(n)))
    return Q, R


def rq(A: torch.Tensor, ctx: Optional[Context] = None) -> Tuple[torch.Tensor, torch.Tensor]:
    """rq(A, R, Q) (rq.cpp:43-70): R (m x m upper), Q (m x n), m <= n."""
    ctx = ctx or default_context()
    A = _check(A)
    m, n = A.shape
    R = torch.empty((m, m), dtype=A.dtype, device=A.device)
    Q = torch.empty((m, n), dtype=A.dtype, device=A.device)
    L.check(L.lib().frk_drq_batched(ctx.h, 1, _i32(m), _i32(n), _arr([A]), _i64(n),
                                    _arr([R]), _i64(m), _arr([Q]), _i64(n)))
    return R, Q


def svd(A: torch.Tensor, ctx: Optional[Context] = None):
    """svd(Dense) (svd.cpp:23-44): U (m x K), S (K x K diagonal, descending), V (K x n)."""
    ctx = ctx or default_context()
    A = _check(A)
    m, n = A.shape
    K = min(m, n)
    U = torch.empty((m, K), dtype=A.dtype, device=A.device)
    S = torch.empty((K, K), dtype=A.dtype, device=A.device)
    V = torch.empty((K, n), dtype=A.dtype, device=A.device)
    L.check(L.lib().frk_dsvd_batched(ctx.h, 1, _i32(m), _i32(n), _arr([A]), _i64(n),
                                     _arr([U]), _i64(K), _arr([S]), _i64(K), _arr([V]),
                                     _i64(n)))
    return U, S, V


def truncated_geqp3(A: torch.Tensor, eps: float, cap: int = 64, ctx: Optional[Context] = None):
    """truncated_geqp3(A, eps) (geqp3.cpp:71-191): Q (m x r), R (r x n)."""
    ctx = ctx or default_context()
    A = _check(A)
    m, n = A.shape
    Q = torch.empty((m, cap), dtype=A.dtype, device=A.device)
    R = torch.zeros((cap, n), dtype=A.dtype, device=A.device)
    r = _i32(0)
    L.check(L.lib().frk_dtruncated_geqp3_batched(ctx.h, 1, _i32(m), _i32(n), _arr([A]),
                                                 _i64(n), eps, cap, _arr([Q]), _i64(cap),
                                                 _arr([R]), _i64(n), r))
    k = int(r[0])
    return Q[:, :k], R[:k]


def rsvd(A: torch.Tensor, p: int, ctx: Optional[Context] = None):
    """rsvd(A, p) (rsvd.cpp:17-29): U (m x p), S (p x p), V (p x n)."""
    ctx = ctx or default_context()
    A = _check(A)
    m, n = A.shape
    U = torch.empty((m, p), dtype=A.dtype, device=A.device)
    S = torch.empty((p, p), dtype=A.dtype, device=A.device)
    V = torch.empty((p, n), dtype=A.dtype, device=A.device)
    L.check(L.lib().frk_drsvd(ctx.h, m, n, _dptr(A), n, p, _dptr(U), p, _dptr(S), p, _dptr(V), n))
    return U, S, V


def one_sided_id(A: torch.Tensor, k: int, ctx: Optional[Context] = None) -> Tuple[torch.Tensor, List[int]]:
    """one_sided_id(A, k) (id.cpp:47-72): A ~ A[:, cols] V, V k x n."""
    ctx = ctx or default_context()
    A = _check(A)
    m, n = A.shape
    V = torch.zeros((k, n), dtype=A.dtype, device=A.device)
    cols = np.zeros(k, dtype=np.int32)
    L.check(L.lib().frk_did(ctx.h, m, n, _dptr(A), n, k, _dptr(V), n, _hptr(cols)))
    return V, [int(c) for c in cols]


def id(A: torch.Tensor, k: int, ctx: Optional[Context] = None):  # noqa: A001 (the reference's name)
    """id(A, k) (id.cpp:100-110): A ~ U A[rows, cols] V."""
    V, cols = one_sided_id(A, k, ctx)
    Ac = A[:, cols].contiguous()
    Ut, rows = one_sided_id(Ac.t().contiguous(), k, ctx)
    return Ut.t().contiguous(), Ac[rows, :].contiguous(), V


def rid(A: torch.Tensor, p: int, k: int, ctx: Optional[Context] = None):
    """rid(A, p, k) (rid.cpp:18-31): sketch with Omega = random_uniform(n, p), QR, ID of Q^T A."""
    A = _check(A)
    m, n = A.shape
    omega = torch.from_numpy(mt19937_uniform(0, n * p).reshape(n, p)).to(A.device)
    Y = torch.zeros((m, p), dtype=A.dtype, device=A.device)
    gemm(A, omega, Y, 1.0, 0.0, ctx=ctx)
    Q, _ = qr(Y, ctx)
    QtA = torch.zeros((Q.shape[1], n), dtype=A.dtype, device=A.device)
    gemm(Q, A, QtA, 1.0, 0.0, trans_a=True, ctx=ctx)
    Ub, S, V = id(QtA, k, ctx)
    U = torch.zeros((m, k), dtype=A.dtype, device=A.device)
    gemm(Q, Ub, U, 1.0, 0.0, ctx=ctx)
    return U, S, V
