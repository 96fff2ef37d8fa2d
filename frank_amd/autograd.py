"""Differentiable BLR solve: u = A(ell)^-1 b with a hand-written adjoint.

A(ell)_ij = exp(-|x_i - x_j|^2 / (2 ell^2)) + shift * delta_ij is compressed
to BLR form at accuracy eps and factorized on the GPU (the reference's
getrf / trsm path, getrf.cpp:45-65, trsm.cpp:40-125).  The forward pass keeps
the factors (a checkpoint of A's LU, not a trace of the solve).  The backward
pass for a cotangent g = dL/du:

    lambda   = A^-T g              (transposed block sweeps on the same factors)
    dL/db    = lambda
    dL/dell  = -lambda^T (dA/dell) u,   (dA/dell)_ij = A_ij^gauss |x_i - x_j|^2 / ell^3
    dL/dA    = -lambda u^T          (rank nrhs per block; see dA_blocks)

The contraction with dA/dell is matrix-free and exact (no compression).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib as L
from .blr import BlrMatrix, Context, _dptr, default_context


def kernel_apply(points: np.ndarray, x: torch.Tensor, kernel: str, ell: float = 0.0, shift: float = 0.0,
                 ctx: Optional[Context] = None) -> torch.Tensor:
    """y = K x for the exact kernel matrix of `points` (dim x N), matrix-free on the GPU."""
    ctx = ctx or default_context()
    kinds = {"laplace": L.KERNEL_LAPLACE, "gaussian": L.KERNEL_GAUSSIAN, "identity": L.KERNEL_IDENTITY,
             "gaussian_dell": L.KERNEL_GAUSSIAN_DELL}
    pts = np.ascontiguousarray(np.atleast_2d(np.asarray(points, dtype=np.float64)))
    dim, n = pts.shape
    x = x.contiguous()
    nrhs = 1 if x.dim() == 1 else x.shape[1]
    y = torch.empty_like(x)
    L.check(L.lib().frk_kernel_apply(ctx.h, kinds[kernel], dim, n, pts.ctypes.data, ell, shift, _dptr(x), _dptr(y),
                                     nrhs))
    return y


class BlrSolve(torch.autograd.Function):
    """u = A(ell)^-1 b for a gaussian kernel matrix in BLR form."""

    @staticmethod
    def forward(fctx, b: torch.Tensor, ell: torch.Tensor, points: np.ndarray, nb: int, eps: float, shift: float,
                kcap: int):
        ell_v = float(ell.detach().cpu())
        pts = np.ascontiguousarray(np.atleast_2d(np.asarray(points, dtype=np.float64)))
        A = BlrMatrix.from_kernel(list(pts), nb, eps=eps, kernel="gaussian", ell=ell_v, shift=shift, kcap=kcap)
        A.set_lra("rounded_addition")
        A.getrf()
        u = b.detach().clone().contiguous()
        A.solve(u)
        fctx.factors = A
        fctx.pts = pts
        fctx.ell = ell_v
        fctx.save_for_backward(u)
        return u

    @staticmethod
    def backward(fctx, g: torch.Tensor):
        (u,) = fctx.saved_tensors
        lam = g.detach().clone().contiguous()
        fctx.factors.solve(lam, trans=True)
        dA_u = kernel_apply(fctx.pts, u, "gaussian_dell", ell=fctx.ell)
        grad_ell = -(lam * dA_u).sum().reshape(())
        fctx.lam = lam
        return lam, grad_ell.to(dtype=torch.float64), None, None, None, None, None


def blr_solve(b: torch.Tensor, ell: torch.Tensor, points, nb: int, eps: float = 1e-10,
              shift: float = 1e-4, kcap: int = 0) -> torch.Tensor:
    """Differentiable u = A(ell)^-1 b (b: N x nrhs float64 on the GPU, ell: scalar tensor)."""
    return BlrSolve.apply(b, ell, points, nb, eps, shift, kcap)


class BlrSolveBatch(torch.autograd.Function):
    """A batch of independent systems u_s = A_s(ell_s)^-1 b_s (gaussian kernels,
    uneven N).  The systems are factorized and solved concurrently, one library
    context (its own streams) per worker thread; every system keeps its factors
    for the backward pass, which runs the transposed sweeps and the
    matrix-free dA/d ell contractions with one worker per context, the systems
    of one context in turn."""

    @staticmethod
    def forward(fctx, ells: torch.Tensor, systems: list, workers: int, *bs: torch.Tensor):
        import concurrent.futures as cf
        import queue

        ell_v = [float(e) for e in ells.detach().cpu()]
        ctxs = queue.Queue()
        for _ in range(max(1, min(workers, len(bs)))):
            ctxs.put(Context())

        def one(s):
            c = ctxs.get()
            try:
                sysd = systems[s]
                pts = np.ascontiguousarray(np.atleast_2d(np.asarray(sysd["points"], dtype=np.float64)))
                A = BlrMatrix.from_kernel(list(pts), sysd["nb"], eps=sysd["eps"], kernel="gaussian", ell=ell_v[s],
                                          shift=sysd["shift"], kcap=sysd["kcap"], ctx=c)
                A.set_lra("rounded_addition")
                A.getrf()
                u = bs[s].detach().clone().contiguous()
                A.solve(u)
                return A, pts, u
            finally:
                ctxs.put(c)

        with cf.ThreadPoolExecutor(max_workers=ctxs.qsize()) as ex:
            res = list(ex.map(one, range(len(bs))))
        fctx.factors = [r[0] for r in res]
        fctx.pts = [r[1] for r in res]
        fctx.ells = ell_v
        us = [r[2] for r in res]
        fctx.save_for_backward(*us)
        return tuple(us)

    @staticmethod
    def backward(fctx, *gs: torch.Tensor):
        import concurrent.futures as cf

        us = fctx.saved_tensors

        def one(s):
            A = fctx.factors[s]
            lam = gs[s].detach().clone().contiguous()
            A.solve(lam, trans=True)
            dA_u = kernel_apply(fctx.pts[s], us[s], "gaussian_dell", ell=fctx.ells[s], ctx=A.ctx)
            return lam, float(-(lam * dA_u).sum())

        # A context is used by one thread at a time (its host-side state and
        # streams are not shared safely): the systems factored on one context
        # run one after another on one worker, the contexts concurrently.
        groups = {}
        for s, A in enumerate(fctx.factors):
            groups.setdefault(id(A.ctx), []).append(s)

        def run(group):
            return [(s, one(s)) for s in group]

        res = [None] * len(us)
        with cf.ThreadPoolExecutor(max_workers=len(groups)) as ex:
            for out in ex.map(run, list(groups.values())):
                for s, r in out:
                    res[s] = r
        grad_ells = torch.tensor([r[1] for r in res], dtype=torch.float64)
        return (grad_ells, None, None) + tuple(r[0] for r in res)


def blr_solve_batched(bs: Sequence[torch.Tensor], ells: torch.Tensor, points: Sequence, nbs: Sequence[int],
                      eps: float = 1e-12, shift: float = 1e-4, kcap: int = 64, workers: int = 2) -> tuple:
    """Differentiable u_s = A_s(ell_s)^-1 b_s for a batch of gaussian-kernel
    systems of uneven sizes (b_s: N_s x nrhs float64 on the GPU; ells: tensor of
    the batch's length scales; points[s]: dim x N_s).  Returns the tuple of u_s;
    L = sum_s f(u_s) backpropagates to every b_s and every ell_s."""
    systems = [{"points": p, "nb": nb, "eps": eps, "shift": shift, "kcap": kcap} for p, nb in zip(points, nbs)]
    return BlrSolveBatch.apply(ells, systems, workers, *bs)


def dA_blocks(lam: np.ndarray, u: np.ndarray, row_offsets: Sequence[int], col_offsets: Sequence[int],
              ctx: Optional[Context] = None) -> BlrMatrix:
    """dL/dA = -lambda u^T as a BLR matrix: every off-diagonal block is LowRank
    (U = lambda_i, S = -I, V = u_j^T, rank = nrhs), diagonal blocks dense."""
    lam = np.atleast_2d(np.asarray(lam, dtype=np.float64).T).T
    u = np.atleast_2d(np.asarray(u, dtype=np.float64).T).T
    k = lam.shape[1]
    ro, co = list(row_offsets), list(col_offsets)
    tiles = []
    for i in range(len(ro) - 1):
        for j in range(len(co) - 1):
            li = lam[ro[i]:ro[i + 1]]
            uj = u[co[j]:co[j + 1]]
            if i == j:
                tiles.append(-li @ uj.T)
            else:
                tiles.append((li.copy(), -np.eye(k), np.ascontiguousarray(uj.T)))
    return BlrMatrix.from_tiles(ro, co, tiles, eps=0.0, rank=k, kcap=max(k, 1), ctx=ctx)
