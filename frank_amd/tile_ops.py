"""Tile-level operators on caller-owned GPU tiles (thin ctypes binding over
the C ABI's frk_dgemm_tiles_batched / frk_dgeqrt_batched / frk_dlarfb_batched /
frk_dtpqrt_batched / frk_dtpmqrt_batched).

A tile is a :class:`Tile`: Dense (one row-major float64 GPU tensor) or
LowRank in slots of capacity cap (U rows x cap, S cap x cap and V cap x cols,
row-major, plus rank and accuracy eps).  Operations that change
a low-rank tile update ``Tile.rank``.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence

import numpy as np

from . import _lib as L
from .blr import _LRA, Context, default_context


class _DTile(C.Structure):
    _fields_ = [("kind", C.c_int32), ("rows", C.c_int32), ("cols", C.c_int32), ("rank", C.c_int32),
                ("cap", C.c_int32), ("reserved", C.c_int32), ("eps", C.c_double), ("d", C.c_void_p),
                ("ld", C.c_int64), ("u", C.c_void_p), ("s", C.c_void_p), ("v", C.c_void_p)]


class Tile:
    """Dense(d) or LowRank(u_slot, s_slot, v_slot, rank, eps) on the GPU."""

    def __init__(self, d=None, u=None, s=None, v=None, rank: int = 0, eps: float = 0.0):
        self.d, self.u, self.s, self.v, self.rank, self.eps = d, u, s, v, int(rank), float(eps)
        if d is None and (u is None or s is None or v is None):
            raise ValueError("a tile is dense (d) or low rank (u, s, v)")

    @classmethod
    def lowrank(cls, U, S, V, cap: int, eps: float = 0.0):
        """Slots of capacity cap holding U (m x r), S (r x r), V (r x n) given as GPU tensors."""
        import torch
        m, r = U.shape
        n = V.shape[1]
        u = torch.zeros((m, cap), dtype=torch.float64, device=U.device)
        s = torch.zeros((cap, cap), dtype=torch.float64, device=U.device)
        v = torch.zeros((cap, n), dtype=torch.float64, device=U.device)
        u[:, :r] = U
        s[:r, :r] = S
        v[:r] = V
        return cls(u=u, s=s, v=v, rank=r, eps=eps)

    @property
    def is_dense(self) -> bool:
        return self.d is not None

    @property
    def shape(self):
        if self.is_dense:
            return tuple(self.d.shape)
        return (self.u.shape[0], self.v.shape[1])

    def factors(self):
        """(U, S, V) views of the current rank (low-rank tiles)."""
        r = self.rank
        return self.u[:, :r], self.s[:r, :r], self.v[:r]

    def dense(self):
        if self.is_dense:
            return self.d
        U, S, V = self.factors()
        return U @ S @ V

    def _desc(self) -> _DTile:
        t = _DTile()
        if self.is_dense:
            t.kind, (t.rows, t.cols) = L.TILE_DENSE, self.d.shape
            t.d, t.ld = self.d.data_ptr(), self.d.stride(0)
        else:
            t.kind = L.TILE_LOWRANK
            t.rows, t.cols = self.shape
            t.rank, t.cap, t.eps = self.rank, self.s.shape[0], self.eps
            t.u, t.s, t.v = self.u.data_ptr(), self.s.data_ptr(), self.v.data_ptr()
        return t


def _descs(tiles: Sequence[Tile]):
    import torch
    torch.cuda.current_stream().synchronize()
    return (_DTile * len(tiles))(*[t._desc() for t in tiles])


def _ranks_back(tiles: Sequence[Tile], arr):
    for t, d in zip(tiles, arr):
        if not t.is_dense:
            t.rank = int(d.rank)


def _ptrs(ts) -> C.Array:
    return (C.c_void_p * len(ts))(*[t.data_ptr() for t in ts])


def gemm(A: Sequence[Tile], B: Sequence[Tile], Cs: Sequence[Tile], alpha: float = 1.0, beta: float = 1.0,
         trans_a: bool = False, trans_b: bool = False, lra: str = "rounded_addition",
         ctx: Optional[Context] = None) -> List[Tile]:
    """gemm(A, B, C, alpha, beta, TransA, TransB) (gemm.cpp:126-272), batched."""
    ctx = ctx or default_context()
    da, db, dc = _descs(A), _descs(B), _descs(Cs)
    L.check(L.lib().frk_dgemm_tiles_batched(ctx.h, len(A), da, db, dc, alpha, beta, int(trans_a), int(trans_b),
                                            _LRA[lra]))
    _ranks_back(Cs, dc)
    return list(Cs)


def geqrt(A: Sequence, T: Sequence, ctx: Optional[Context] = None):
    """geqrt(A, T) (geqrt.cpp:34-43) on n x n GPU tensors, in place."""
    import torch
    ctx = ctx or default_context()
    n = A[0].shape[0]
    torch.cuda.current_stream().synchronize()
    L.check(L.lib().frk_dgeqrt_batched(ctx.h, len(A), n, _ptrs(A), A[0].stride(0), _ptrs(T), T[0].stride(0)))


def larfb(V: Sequence, T: Sequence, Cs: Sequence[Tile], trans: bool, ctx: Optional[Context] = None):
    """larfb(V, T, C, trans) (larfb.cpp:37-56); a low-rank C is updated through U."""
    ctx = ctx or default_context()
    n = V[0].shape[0]
    dc = _descs(Cs)
    L.check(L.lib().frk_dlarfb_batched(ctx.h, len(V), n, _ptrs(V), V[0].stride(0), _ptrs(T), T[0].stride(0), dc,
                                       int(trans)))


def tpqrt(A: Sequence, B: Sequence[Tile], T: Sequence, ctx: Optional[Context] = None):
    """tpqrt(A, B, T) (tpqrt.cpp:35-53) in place."""
    ctx = ctx or default_context()
    n = A[0].shape[0]
    db = _descs(B)
    L.check(L.lib().frk_dtpqrt_batched(ctx.h, len(A), n, _ptrs(A), A[0].stride(0), db, _ptrs(T), T[0].stride(0)))


def tpmqrt(V: Sequence[Tile], T: Sequence, A: Sequence[Tile], B: Sequence[Tile], trans: bool,
           lra: str = "rounded_addition", ctx: Optional[Context] = None):
    """tpmqrt(V, T, A, B, trans) (tpmqrt.cpp:45-170) in place."""
    ctx = ctx or default_context()
    n = T[0].shape[0]
    dv, da, db = _descs(V), _descs(A), _descs(B)
    L.check(L.lib().frk_dtpmqrt_batched(ctx.h, len(V), n, dv, _ptrs(T), T[0].stride(0), da, db, int(trans),
                                        _LRA[lra]))
    _ranks_back(A, da)
    _ranks_back(B, db)


def np_tile(x: np.ndarray) -> Tile:
    import torch
    return Tile(d=torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).cuda())
