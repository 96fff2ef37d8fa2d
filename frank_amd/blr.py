"""Python host mirror of FRANK's BLR (one-level Hierarchical) operator surface.

Tiles live in HBM inside the native engine; this module only holds handles.
Names follow the reference: ``getrf`` (src/operations/LAPACK/getrf.cpp:40-65),
``trsm``-based ``solve`` (src/operations/BLAS/trsm.cpp:40-125), ``gemm`` as
``matvec`` (gemm(Hierarchical, Dense, Dense)), ``l2_error``
(src/util/l2_error.cpp:35-41).  Device vectors are torch tensors on the GPU.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence, Tuple, Union

import numpy as np

from . import _lib as L

_ADMIS = {"position": L.ADMIS_POSITION, "geometry": L.ADMIS_GEOMETRY}
_PREC = {"f64": L.PREC_F64, "f32": L.PREC_F32}
_LRA = {"naive": L.LRA_NAIVE, "rounded_addition": L.LRA_ROUNDED, "fast_rounded_addition": L.LRA_FAST}


def _dptr(t, sync: bool = True) -> int:
    """Device pointer of a contiguous torch tensor on the GPU.  The library
    works on its own non-blocking streams, so the work torch has queued on the
    tensor's current stream is drained first (every entry point drains the
    library's stream before it returns, so results are ready for torch)."""
    if not t.is_cuda:
        raise ValueError("expected a GPU tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    if sync:
        import torch
        torch.cuda.current_stream(t.device).synchronize()
    return t.data_ptr()


def _hptr(a: np.ndarray) -> int:
    return a.ctypes.data


class Context:
    """Execution context (HIP stream + workspace); replaces FRANK::initialize()."""

    def __init__(self):
        h = C.c_void_p()
        L.check(L.lib().frk_ctx_create(C.byref(h)))
        self.h = h

    def synchronize(self):
        L.check(L.lib().frk_ctx_synchronize(self.h))

    @property
    def stream_ptr(self) -> int:
        p = C.c_void_p()
        L.check(L.lib().frk_ctx_stream(self.h, C.byref(p)))
        return p.value or 0

    def counters(self, reset: bool = False) -> dict:
        """Device event counters (drains the stream)."""
        out = (C.c_int64 * 16)()
        L.check(L.lib().frk_ctx_counters(self.h, out, 16, int(reset)))
        return {"svd_problems": out[0], "svd_sweeps": out[1], "svd_max_sweeps": out[2]}

    def time_class(self, cls: Optional[str]):
        """Bracket every launch of kernel class ``cls`` with HIP events (None: none)."""
        L.check(L.lib().frk_ctx_time_class(self.h, cls.encode() if cls else None))

    def op_stats(self) -> dict:
        """Op log per kernel class: launches, algorithmic flops/bytes, event-timed ms,
        roofline ms (sum over launches of max(flops / peak, bytes / HBM bandwidth))."""
        buf = C.create_string_buffer(4096)
        L.check(L.lib().frk_ctx_op_classes(self.h, buf, 4096))
        out = {}
        for cls in filter(None, buf.value.decode().split(",")):
            n, t = C.c_int64(), C.c_int64()
            f, b, ms = C.c_double(), C.c_double(), C.c_double()
            L.check(L.lib().frk_ctx_op_stats(self.h, cls.encode(), C.byref(n), C.byref(f), C.byref(b), C.byref(ms),
                                             C.byref(t)))
            roof = C.c_double()
            L.check(L.lib().frk_ctx_op_roofline(self.h, cls.encode(), C.byref(roof)))
            out[cls] = {"launches": n.value, "flops": f.value, "bytes": b.value, "ms": ms.value,
                        "timed_launches": t.value, "roofline_ms": roof.value}
        return out

    def set_rank_cap(self, truncate: bool):
        """Rank policy at the slot capacity for matrices built from now on:
        False raises on an eps-rank above kcap, True keeps the leading kcap terms."""
        L.check(L.lib().frk_ctx_set_rank_cap(self.h, 1 if truncate else 0))

    def reset_op_stats(self):
        L.check(L.lib().frk_ctx_reset_op_stats(self.h))

    def close(self):
        if self.h:
            L.lib().frk_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context()
    return _default_ctx


_KERNEL = {"laplace": L.KERNEL_LAPLACE, "gaussian": L.KERNEL_GAUSSIAN, "identity": L.KERNEL_IDENTITY}


class QrFactors:
    """T factors of a tiled Householder BLR-QR (the reference's Hierarchical T)."""

    def __init__(self, h, owner: "BlrMatrix"):
        self.h = h
        self.owner = owner

    def t(self, i: int, k: int):
        """(T(i, k), tau(i, k)) as host arrays."""
        _, _, rows, cols = self.owner.tile_info(k, k)
        T = np.empty((cols, cols))
        tau = np.empty(cols)
        L.check(L.lib().frk_qrf_get_t(self.h, i, k, _hptr(T), _hptr(tau)))
        return T, tau

    def close(self):
        if self.h:
            L.lib().frk_qrf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BlrMatrix:
    """One-level Hierarchical matrix of Dense / LowRank tiles resident on the GPU."""

    def __init__(self, handle, ctx: Context):
        self.h = handle
        self.ctx = ctx

    # -- construction -------------------------------------------------------
    @classmethod
    def from_kernel(cls, points: Sequence[np.ndarray], nb: int, *, eps: float = 0.0, rank: int = 0,
                    admis: float = 0.0, admis_type: str = "position", kcap: int = 0,
                    kernel: str = "laplace", ell: float = 0.0, shift: float = 0.0, precision: str = "f64",
                    ctx: Optional[Context] = None) -> "BlrMatrix":
        """Hierarchical(laplacend, params, N, N, nleaf, eps|rank, admis, nb, nb, admis_type)."""
        ctx = ctx or default_context()
        pts = np.ascontiguousarray(np.vstack([np.asarray(p, dtype=np.float64) for p in points]))
        dim, n = pts.shape
        h = C.c_void_p()
        L.check(L.lib().frk_blr_create_kernel(ctx.h, _KERNEL[kernel], dim, n, _hptr(pts), ell, shift, nb, eps, rank,
                                              admis, _ADMIS[admis_type], kcap, _PREC[precision], C.byref(h)))
        return cls(h, ctx)

    @classmethod
    def from_kernel_rect(cls, points: Sequence[np.ndarray], n_rows: int, n_cols: int, nb_r: int, nb_c: int, *,
                         eps: float = 0.0, rank: int = 0, admis: float = 0.0, admis_type: str = "position",
                         kcap: int = 0, kernel: str = "laplace", ell: float = 0.0, shift: float = 0.0,
                         precision: str = "f64", ctx: Optional[Context] = None) -> "BlrMatrix":
        """Hierarchical(kernel, params, n_rows, n_cols, nleaf, eps|rank, admis, nb_r, nb_c, admis_type)."""
        ctx = ctx or default_context()
        pts = np.ascontiguousarray(np.vstack([np.asarray(p, dtype=np.float64) for p in points]))
        dim, n = pts.shape
        h = C.c_void_p()
        L.check(L.lib().frk_blr_create_kernel_rect(ctx.h, _KERNEL[kernel], dim, n, _hptr(pts), n_rows, n_cols, nb_r,
                                                   nb_c, ell, shift, eps, rank, admis, _ADMIS[admis_type], kcap,
                                                   _PREC[precision], C.byref(h)))
        return cls(h, ctx)

    @classmethod
    def from_tiles(cls, row_offsets: Sequence[int], col_offsets: Sequence[int], tiles: list, *, eps: float = 0.0,
                   rank: int = 0, kcap: int = 64, precision: str = "f64",
                   ctx: Optional[Context] = None) -> "BlrMatrix":
        """Grid from host tiles: ndarray (Dense), (U, S, V) tuple (LowRank) or None (Empty); row-major list."""
        ctx = ctx or default_context()
        ro = np.asarray(row_offsets, dtype=np.int64)
        co = np.asarray(col_offsets, dtype=np.int64)
        nbr, nbc = ro.size - 1, co.size - 1
        kinds = np.empty(nbr * nbc, dtype=np.int32)
        for t, x in enumerate(tiles):
            kinds[t] = L.TILE_EMPTY if x is None else (L.TILE_DENSE if isinstance(x, np.ndarray) else L.TILE_LOWRANK)
        h = C.c_void_p()
        L.check(L.lib().frk_blr_create(ctx.h, nbr, nbc, _hptr(ro), _hptr(co), _hptr(kinds), kcap, eps, rank,
                                       _PREC[precision], C.byref(h)))
        self = cls(h, ctx)
        for t, x in enumerate(tiles):
            i, j = divmod(t, nbc)
            if x is None:
                continue
            if isinstance(x, np.ndarray):
                a = np.ascontiguousarray(x, dtype=np.float64)
                L.check(L.lib().frk_blr_set_dense(self.h, i, j, _hptr(a)))
            else:
                u, s, v = (np.ascontiguousarray(y, dtype=np.float64) for y in x)
                L.check(L.lib().frk_blr_set_lowrank(self.h, i, j, s.shape[0], _hptr(u), _hptr(s), _hptr(v)))
        return self

    def close(self):
        if self.h:
            L.lib().frk_blr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- inspection ---------------------------------------------------------
    @property
    def shape(self) -> Tuple[int, int]:
        a, b = C.c_int(), C.c_int()
        L.check(L.lib().frk_blr_shape(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def tile_info(self, i: int, j: int):
        kind, rank, rows, cols = C.c_int32(), C.c_int32(), C.c_int64(), C.c_int64()
        L.check(L.lib().frk_blr_tile_info(self.h, i, j, C.byref(kind), C.byref(rank), C.byref(rows),
                                          C.byref(cols)))
        return kind.value, rank.value, rows.value, cols.value

    def tile(self, i: int, j: int):
        """Host copy: ndarray for Dense, (U, S, V) for LowRank, None for Empty."""
        kind, rank, rows, cols = self.tile_info(i, j)
        if kind == L.TILE_DENSE:
            a = np.empty((rows, cols))
            L.check(L.lib().frk_blr_get_dense(self.h, i, j, _hptr(a)))
            return a
        if kind == L.TILE_LOWRANK:
            u = np.empty((rows, rank))
            s = np.empty((rank, rank))
            v = np.empty((rank, cols))
            L.check(L.lib().frk_blr_get_lowrank(self.h, i, j, _hptr(u), _hptr(s), _hptr(v)))
            return u, s, v
        return None

    def ranks(self) -> np.ndarray:
        nr, nc = self.shape
        out = np.full((nr, nc), -1, dtype=np.int64)
        for i in range(nr):
            for j in range(nc):
                kind, rank, _, _ = self.tile_info(i, j)
                if kind == L.TILE_LOWRANK:
                    out[i, j] = rank
        return out

    def dense(self) -> np.ndarray:
        nr, nc = self.shape
        rows = [self.tile_info(i, 0)[2] for i in range(nr)]
        cols = [self.tile_info(0, j)[3] for j in range(nc)]
        out = np.zeros((sum(rows), sum(cols)))
        r0 = 0
        for i in range(nr):
            c0 = 0
            for j in range(nc):
                t = self.tile(i, j)
                if t is not None:
                    blk = t if isinstance(t, np.ndarray) else (t[0] @ t[1]) @ t[2]
                    out[r0:r0 + rows[i], c0:c0 + cols[j]] = blk
                c0 += cols[j]
            r0 += rows[i]
        return out

    # -- structure and rank report --------------------------------------------
    def to_json(self) -> dict:
        """Structure and rank report in the layout of the reference's write_JSON
        (src/util/print.cpp:54-128): the root is the Hierarchical grid (type,
        dim, abs_pos [0, 0], level 0, children as rows of blocks); each block
        carries type, dim, abs_pos [i, j], level 1 and its singular values
        (a LowRank block those of its dense expansion, plus its rank).  Like the
        reference's, this is a host-side report: tiles are copied from the
        device and their singular values computed by host LAPACK.  An Empty
        block raises, as the reference has no to_json method for one."""
        return structure_json(self.shape, self.tile_info, self.tile)

    def write_json(self, filename: str) -> None:
        """write_JSON(A, filename) (src/util/print.cpp:124-128): the report of
        to_json as one compact JSON document with sorted keys."""
        import json
        with open(filename, "w") as f:
            json.dump(self.to_json(), f, sort_keys=True, separators=(",", ":"))

    def clone(self) -> "BlrMatrix":
        h = C.c_void_p()
        L.check(L.lib().frk_blr_clone(self.h, C.byref(h)))
        return BlrMatrix(h, self.ctx)

    def convert(self, precision: str) -> "BlrMatrix":
        """Same structure in another storage precision ('f64' or 'f32')."""
        h = C.c_void_p()
        L.check(L.lib().frk_blr_convert(self.h, _PREC[precision], C.byref(h)))
        return BlrMatrix(h, self.ctx)

    def copy_from(self, src: "BlrMatrix") -> "BlrMatrix":
        L.check(L.lib().frk_blr_copy(self.h, src.h))
        return self

    # -- hot path -------------------------------------------------------------
    def set_lra(self, lra: str):
        L.check(L.lib().frk_blr_set_lra(self.h, _LRA[lra]))

    def getrf(self, max_steps: Optional[int] = None) -> "BlrMatrix":
        """In-place tile LU: lower tiles -> L, upper tiles -> U (getrf.cpp:45-65).

        ``max_steps`` stops after that many elimination steps (a partial
        factorization for parity checks; the diagonal tile of the next step is
        factorized as well)."""
        if max_steps is None:
            L.check(L.lib().frk_blr_getrf(self.h))
        else:
            L.check(L.lib().frk_blr_getrf_steps(self.h, int(max_steps)))
        return self

    def qr(self) -> QrFactors:
        """tiled_householder_blr_qr(A, T) in place (qr.cpp:213-228); returns the T factors."""
        h = C.c_void_p()
        L.check(L.lib().frk_blr_qr(self.h, C.byref(h)))
        return QrFactors(h, self)

    def apply_q(self, F: QrFactors, C_: "BlrMatrix", trans: bool = False) -> "BlrMatrix":
        """left_multiply_tiled_reflector(self, T, C, trans): C <- Q C or Q^T C (qr.cpp:230-257)."""
        L.check(L.lib().frk_blr_apply_q(self.h, F.h, C_.h, int(trans)))
        return C_

    def solve(self, b, trans: bool = False):
        """In place on a GPU tensor b (N x nrhs, float64): L then U sweeps (U^T, L^T if trans)."""
        nrhs = 1 if b.dim() == 1 else b.shape[1]
        L.check(L.lib().frk_blr_solve(self.h, _dptr(b), nrhs, 1 if trans else 0))
        return b

    def matvec(self, x, y):
        """y += A x on GPU tensors (N x nrhs, float64)."""
        nrhs = 1 if x.dim() == 1 else x.shape[1]
        L.check(L.lib().frk_blr_matvec(self.h, _dptr(x), _dptr(y), nrhs))
        return y

    def refine(self, b, x, tol: float = 1e-10, max_iter: int = 50):
        """x <- refined solution of A x = b (A: the exact kernel matrix, FP64). Returns (iters, berr)."""
        nrhs = 1 if b.dim() == 1 else b.shape[1]
        it, be = C.c_int(), C.c_double()
        L.check(L.lib().frk_blr_refine(self.h, _dptr(b), _dptr(x), nrhs, tol, max_iter, C.byref(it), C.byref(be)))
        return it.value, be.value

    def kernel_residual(self, x, b, r):
        """r = b - A x with the exact (uncompressed) kernel matrix, FP64, on the GPU."""
        nrhs = 1 if x.dim() == 1 else x.shape[1]
        L.check(L.lib().frk_blr_kernel_residual(self.h, _dptr(x), _dptr(b), _dptr(r), nrhs))
        return r


# ---------------------------------------------------------------------------
# Tile-level operators (device tensors)
# ---------------------------------------------------------------------------
def getrf_nopiv_batched(tiles: list, ctx: Optional[Context] = None) -> np.ndarray:
    """getrf_omm(Dense) without pivot search on a list of square GPU tensors (in place)."""
    ctx = ctx or default_context()
    n = tiles[0].shape[0]
    ptrs = (C.c_void_p * len(tiles))(*[_dptr(t) for t in tiles])
    info = np.zeros(len(tiles), dtype=np.int32)
    rc = L.lib().frk_dgetrf_nopiv_batched(ctx.h, ptrs, len(tiles), n, tiles[0].stride(0), _hptr(info))
    if rc not in (0, 1):
        L.check(rc)
    return info


def trsm(A, B, uplo: str, side: str = "left", trans: bool = False, ctx: Optional[Context] = None):
    """trsm_omm(Dense, Dense): Lower => unit diagonal, Upper => non-unit (trsm.cpp:92-104)."""
    ctx = ctx or default_context()
    m, n = B.shape
    L.check(L.lib().frk_dtrsm(ctx.h, 0 if side == "left" else 1, 1 if uplo == "lower" else 0, 1 if trans else 0,
                              m, n, _dptr(A), A.stride(0), _dptr(B), B.stride(0)))
    return B


def gemm(A, B, C_, alpha: float = 1.0, beta: float = 1.0, trans_a: bool = False, trans_b: bool = False,
         ctx: Optional[Context] = None):
    """gemm_omm(Dense, Dense, Dense): C = alpha op(A) op(B) + beta C."""
    ctx = ctx or default_context()
    m, n = C_.shape
    k = A.shape[0] if trans_a else A.shape[1]
    L.check(L.lib().frk_dgemm(ctx.h, int(trans_a), int(trans_b), m, n, k, alpha, _dptr(A), A.stride(0),
                              _dptr(B), B.stride(0), beta, _dptr(C_), C_.stride(0)))
    return C_


def structure_json(shape: Tuple[int, int], tile_info, tile) -> dict:
    """The write_JSON document of a one-level grid (src/util/print.cpp:54-122),
    from accessors: shape = (tile rows, tile cols); tile_info(i, j) -> (kind,
    rank, rows, cols); tile(i, j) -> ndarray (Dense) or (U, S, V) (LowRank).
    Singular values are all min(m, n) of them, descending (get_singular_values,
    src/operations/LAPACK/svd.cpp:46-61)."""
    nr, nc = shape
    rows = [tile_info(i, 0)[2] for i in range(nr)]
    cols = [tile_info(0, j)[3] for j in range(nc)]
    children = []
    for i in range(nr):
        row = []
        for j in range(nc):
            kind, rank, m, n = tile_info(i, j)
            if kind == L.TILE_EMPTY:
                raise TypeError(f"to_json: no method for the Empty block ({i}, {j})")
            t = tile(i, j)
            node = {"type": "LowRank" if kind == L.TILE_LOWRANK else "Dense", "dim": [int(m), int(n)],
                    "abs_pos": [i, j], "level": 1}
            dense = (t[0] @ t[1]) @ t[2] if kind == L.TILE_LOWRANK else t
            node["svalues"] = [float(x) for x in np.linalg.svd(np.asarray(dense, dtype=np.float64),
                                                               compute_uv=False)]
            if kind == L.TILE_LOWRANK:
                node["rank"] = int(rank)
            row.append(node)
        children.append(row)
    return {"type": "Hierarchical", "dim": [int(sum(rows)), int(sum(cols))], "abs_pos": [0, 0], "level": 0,
            "children": children}


def lr_add_batched(As: list, Bs: list, *, eps: float = 0.0, lra: str = "fast_rounded_addition",
                   cap: int = 0, ctx: Optional[Context] = None) -> List[int]:
    """LowRank += LowRank for pairs of (U, S, V, rank) GPU slots (A in place)."""
    import torch
    ctx = ctx or default_context()
    nb = len(As)
    m = np.array([a[0].shape[0] for a in As], dtype=np.int32)
    n = np.array([a[2].shape[1] for a in As], dtype=np.int32)
    ranks = np.array([a[3] for a in As], dtype=np.int32)
    kb = np.array([b[1].shape[0] for b in Bs], dtype=np.int32)
    cap = cap or As[0][0].shape[1]

    torch.cuda.current_stream().synchronize()

    def arr(xs):
        return (C.c_void_p * nb)(*[_dptr(x, sync=False) for x in xs])

    fn = L.lib().frk_slr_add_batched if As[0][0].dtype == torch.float32 else L.lib().frk_dlr_add_batched
    L.check(fn(ctx.h, _LRA[lra], nb, _hptr(m), _hptr(n), eps, cap,
               arr([a[0] for a in As]), arr([a[1] for a in As]), arr([a[2] for a in As]),
               _hptr(ranks), arr([b[0] for b in Bs]), arr([b[1] for b in Bs]),
               arr([b[2] for b in Bs]), _hptr(kb)))
    return [int(r) for r in ranks]


def compress(A, *, eps: float = 0.0, rank: int = 0, cap: int = 64, ctx: Optional[Context] = None):
    """LowRank(Dense, eps) / LowRank(Dense, rank) on a GPU tensor (overwritten). Returns (U, S, V) views."""
    import torch
    ctx = ctx or default_context()
    m, n = A.shape
    U = torch.zeros((m, cap), dtype=torch.float64, device=A.device)
    S = torch.zeros((cap, cap), dtype=torch.float64, device=A.device)
    V = torch.zeros((cap, n), dtype=torch.float64, device=A.device)
    r = np.zeros(1, dtype=np.int32)
    L.check(L.lib().frk_dcompress(ctx.h, m, n, _dptr(A), A.stride(0), eps, rank, cap, _dptr(U), _dptr(S),
                                  _dptr(V), _hptr(r)))
    k = int(r[0])
    return U[:, :k], S[:k, :k], V[:k, :]
