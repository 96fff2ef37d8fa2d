"""CPU restatement (numpy + scipy LAPACK, FP64) of FRANK's BLR-LU hot path.

TEST INFRASTRUCTURE ONLY.  This module is the parity checker: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it, and only to check or time the reference algorithm.  The product path
(``frank_amd``) never imports it.

Every function follows one reference routine; the citation (path:line into the
FRANK sources) is given in its docstring.  The data model mirrors the
reference's:

* Dense      -> ``numpy.ndarray`` (2-D, float64, logical row-major semantics)
* LowRank    -> :class:`LowRank` (U m x k, S k x k, V k x n, rank, eps)
* Empty      -> ``None``
* Hierarchical (one level, BLR) -> :class:`Hier` (row-major grid of blocks)

How each function is pinned (tests/test_oracle_*.py):

* RNG: raw ``std::mt19937`` words and ``uniform_real_distribution<double>``
  values written by ``tests/golden/mt19937_golden.cpp`` (libstdc++), plus the
  C++ standard's 10000th-output check.
* getrf / trsm / gemm / qr / rq / svd: closed forms (L U == A, triangular
  back-substitution, dense expansion of U S V), scipy routines, and the
  tolerance contracts of the reference's own unit tests.
* find_svd_truncation_rank: brute-force evaluation of the H2 rule.
* truncated_geqp3: brute-force pivoted Gram-Schmidt with exact trailing column
  norms on small tiles (same pivots, same stop step), zero-matrix rule.
* rounded/fast/naive addition: dense SVD of the exact sum truncated by the
  same rule, orthogonality of the factors, rank preservation.
* BLR getrf / solve: dense LAPACK factor/solve of the same matrix, the
  reference's ``l2_error <= eps`` contract.
* to_json / get_singular_values: a document written out by hand for a grid
  whose blocks have known singular values (tests/test_structure_json.py).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple, Union

import numpy as np
import scipy.linalg as sla
from scipy.linalg import lapack

DBL_EPSILON = float(np.finfo(np.float64).eps)
# LAPACK dlamch('e'): relative machine precision for round-to-nearest (2^-53).
DLAMCH_E = 2.0 ** -53


# ---------------------------------------------------------------------------
# RNG: std::mt19937 + libstdc++ uniform_real_distribution<double>
# ---------------------------------------------------------------------------
def mt19937_words(seed: int, n: int) -> np.ndarray:
    """Raw 32-bit outputs of ``std::mt19937`` seeded with ``seed``.

    numpy's legacy MT19937 seeding is ``init_genrand(seed)``, the same as
    ``std::mt19937::seed(seed)``.
    """
    bg = np.random.MT19937()
    bg._legacy_seeding(seed)
    return bg.random_raw(n).astype(np.uint64)


def mt19937_uniform(seed: int, n: int) -> np.ndarray:
    """``uniform_real_distribution<double>(0, 1)`` over ``std::mt19937(seed)``.

    libstdc++ ``generate_canonical<double, 53>`` consumes two 32-bit words per
    value: ``(w0 + w1 * 2^32) / 2^64``, clamped below 1.
    """
    w = mt19937_words(seed, 2 * n).astype(np.float64)
    u = (w[0::2] + w[1::2] * 4294967296.0) / 18446744073709551616.0
    u[u >= 1.0] = np.nextafter(1.0, 0.0)
    return u


def get_sorted_random_vector(n: int) -> np.ndarray:
    """src/util/experiment_setup.cpp:12-23 (seed 0, U[0,1), sorted)."""
    return np.sort(mt19937_uniform(0, n))


def random_uniform(rows: int, cols: int) -> np.ndarray:
    """src/functions.cpp:52-66: row-major fill, generator re-seeded to 0."""
    return mt19937_uniform(0, rows * cols).reshape(rows, cols)


# ---------------------------------------------------------------------------
# Kernel generators
# ---------------------------------------------------------------------------
def identity(points, r0: int, nr: int, c0: int, nc: int) -> np.ndarray:
    """src/functions.cpp:25-34: A_ij = 1 where the global row equals the global column."""
    return (np.arange(r0, r0 + nr)[:, None] == np.arange(c0, c0 + nc)[None, :]).astype(np.float64)


def laplacend(points: Sequence[np.ndarray], r0: int, nr: int, c0: int, nc: int) -> np.ndarray:
    """src/functions.cpp:92-109: A_ij = 1 / (sqrt(sum_k (x_k[i]-x_k[j])^2) + 1e-3)."""
    rij = np.zeros((nr, nc))
    for x in points:
        d = x[r0:r0 + nr, None] - x[None, c0:c0 + nc]
        rij = rij + d * d
    return 1.0 / (np.sqrt(rij) + 1e-3)


def laplacend_shift(points, r0, nr, c0, nc, shift):
    """Laplace kernel plus ``shift`` on the global diagonal (shifted 1-D workload)."""
    a = laplacend(points, r0, nr, c0, nc)
    if shift != 0.0:
        i = np.arange(nr)[:, None] + r0
        j = np.arange(nc)[None, :] + c0
        a = a + np.where(i == j, shift, 0.0)
    return a


def gaussian(points, r0, nr, c0, nc, ell, shift):
    """Gaussian kernel exp(-|x_i-x_j|^2 / (2 ell^2)) + shift * delta_ij."""
    rij = np.zeros((nr, nc))
    for x in points:
        d = x[r0:r0 + nr, None] - x[None, c0:c0 + nc]
        rij = rij + d * d
    a = np.exp(-rij / (2.0 * ell * ell))
    if shift != 0.0:
        i = np.arange(nr)[:, None] + r0
        j = np.arange(nc)[None, :] + c0
        a = a + np.where(i == j, shift, 0.0)
    return a


# ---------------------------------------------------------------------------
# Misc
# ---------------------------------------------------------------------------
def seq_sum(a: np.ndarray) -> float:
    """Sequential left-to-right sum (the reference's plain accumulation loops)."""
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    if a.size == 0:
        return 0.0
    return float(np.add.accumulate(a)[-1])


def norm(a) -> float:
    """src/operations/misc/norm.cpp:24-44: SQUARED Frobenius norm (H1)."""
    if isinstance(a, LowRank):
        a = a.dense()
    return seq_sum(a * a)


def find_svd_truncation_rank(s_diag: np.ndarray, eps: float, cap: Optional[int] = None) -> int:
    """src/operations/misc/misc.cpp:133-146 (H2).

    threshold = eps * sqrt(sum sigma^2); r = smallest r >= 1 with
    sqrt((min_dim - r) * sigma_r^2) <= threshold (0-based sigma), else min_dim.
    The reference's inner loop adds sigma_r^2 (min_dim - r) times.

    cap (no reference counterpart: the reference's ranks are unbounded) keeps
    at most the cap leading terms, the rule of the GPU path's capped-rank
    mode (frk_ctx_set_rank_cap(ctx, 1)).
    """
    s = np.asarray(s_diag, dtype=np.float64)
    threshold = eps * math.sqrt(seq_sum(s * s))
    min_dim = s.size
    rank = 0
    while True:
        rank += 1
        err = 0.0
        for _ in range(rank, min_dim):
            err += s[rank] * s[rank]
        if not (rank < min_dim and math.sqrt(err) > threshold):
            return rank if cap is None else min(rank, cap)


def resize(a: np.ndarray, rows: int, cols: int) -> np.ndarray:
    """src/operations/misc/resize.cpp:26-34: top-left copy."""
    return np.array(a[:rows, :cols], dtype=np.float64, copy=True)


# ---------------------------------------------------------------------------
# Dense LAPACK-level ops
# ---------------------------------------------------------------------------
class PivotingError(RuntimeError):
    """dgetrf chose a row swap: the reference's discarded-ipiv LU is invalid (H3)."""


def getrf_dense(a: np.ndarray, check_pivots: bool = True) -> Tuple[np.ndarray, np.ndarray]:
    """src/operations/LAPACK/getrf.cpp:67-84: dgetrf, ipiv discarded, split to unit-L / U."""
    lu, piv, info = lapack.dgetrf(np.array(a, dtype=np.float64, copy=True))
    if check_pivots and not np.array_equal(piv, np.arange(piv.size)):
        raise PivotingError("dgetrf pivoted; no-pivot LU differs from the reference")
    L = np.tril(lu, -1) + np.eye(a.shape[0], a.shape[1])
    U = np.triu(lu)
    return L, U


def getrf_nopiv(a: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """Right-looking LU without pivoting (what getrf.cpp:67-84 yields when dgetrf does not swap)."""
    a = np.array(a, dtype=np.float64, copy=True)
    n = a.shape[0]
    for k in range(n - 1):
        a[k + 1:, k] /= a[k, k]
        a[k + 1:, k + 1:] -= np.outer(a[k + 1:, k], a[k, k + 1:])
    return np.tril(a, -1) + np.eye(n), np.triu(a)


def trsm_dense(a: np.ndarray, b: np.ndarray, uplo: str, side: str) -> np.ndarray:
    """src/operations/BLAS/trsm.cpp:92-104 (H4): Lower => unit diagonal, Upper => non-unit, NoTrans."""
    lower = uplo == "lower"
    unit = lower
    if side == "left":
        return sla.solve_triangular(a, b, lower=lower, unit_diagonal=unit)
    # right: X A = B  <=>  A^T X^T = B^T
    return sla.solve_triangular(a, b.T, lower=lower, unit_diagonal=unit, trans="T").T


def trsm_dense_trans(a, b, uplo, side):
    """Transposed-operator solve (op(A) = A^T) used by the adjoint sweeps (no reference counterpart)."""
    lower = uplo == "lower"
    unit = lower
    if side == "left":
        return sla.solve_triangular(a, b, lower=lower, unit_diagonal=unit, trans="T")
    return sla.solve_triangular(a, b.T, lower=lower, unit_diagonal=unit, trans="N").T


def qr(a: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """src/operations/LAPACK/qr.cpp:65-90: dgeqrf + dorgqr, Q m x k, R k x n, k = min(m, n)."""
    q, r = sla.qr(np.asarray(a, dtype=np.float64), mode="economic")
    return q, r


def rq(a: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """src/operations/LAPACK/rq.cpp:43-70: dgerqf + dorgrq, R m x k, Q k x n, k = min(m, n)."""
    r, q = sla.rq(np.asarray(a, dtype=np.float64), mode="economic")
    return r, q


def svd(a: np.ndarray) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """src/operations/LAPACK/svd.cpp:23-44: dgesvd 'S','S'; S returned as a diagonal matrix."""
    u, s, vt = sla.svd(np.asarray(a, dtype=np.float64), full_matrices=False,
                       lapack_driver="gesvd")
    return u, np.diag(s), vt


def dlarfg(n: int, alpha: float, x: np.ndarray) -> Tuple[float, np.ndarray, float]:
    """LAPACK dlarfg via scipy: returns (beta, v_tail, tau)."""
    if n <= 1:
        return alpha, x, 0.0
    beta, v, tau = lapack.dlarfg(n, alpha, np.array(x, dtype=np.float64, copy=True))
    return float(beta), v, float(tau)


def truncated_geqp3(a_in: np.ndarray, eps: float, cap: Optional[int] = None) -> Tuple[np.ndarray, np.ndarray]:
    """src/operations/LAPACK/geqp3.cpp:71-191 (H5).

    Column-pivoted Householder QR with LAWN-176 norm downdating, stopped when
    the largest partial column norm is <= eps * ||A||_F.  Returns the
    truncated Q (m x r) and R (r x n) with columns back in original order.
    cap (no reference counterpart) also stops after cap pivots, the rule of
    the GPU path's capped-rank mode.
    """
    a = np.array(a_in, dtype=np.float64, copy=True)
    m, n = a.shape
    tol = DLAMCH_E
    tol3z = math.sqrt(tol)
    min_dim = min(m, n)
    tau = np.zeros(min_dim)
    ipiv = np.arange(n)
    cnorm = np.array([np.linalg.norm(a[:, j]) for j in range(n)])
    partial = cnorm.copy()
    r = 0
    threshold = eps * math.sqrt(norm(a))
    max_cnorm = float(np.max(cnorm))
    if max_cnorm <= tol:
        q = np.zeros((m, 1))
        q[0, 0] = 1.0
        return q, np.zeros((1, n))
    while r < min_dim and (cap is None or r < cap) and max_cnorm > threshold:
        k = r + int(np.argmax(cnorm[r:]))
        a[:, [r, k]] = a[:, [k, r]]
        cnorm[[r, k]] = cnorm[[k, r]]
        partial[[r, k]] = partial[[k, r]]
        ipiv[[r, k]] = ipiv[[k, r]]
        if r < m - 1:
            beta, v, t = dlarfg(m - r, a[r, r], a[r + 1:, r])
            a[r, r] = beta
            a[r + 1:, r] = v
            tau[r] = t
        else:
            tau[r] = 0.0
        if r < min_dim - 1:
            vv = np.concatenate(([1.0], a[r + 1:, r]))
            w = a[r:, r + 1:].T @ vv
            a[r:, r + 1:] -= tau[r] * np.outer(vv, w)
        for j in range(r + 1, n):
            if cnorm[j] != 0.0:
                temp = abs(a[r, j] / cnorm[j])
                temp = max(0.0, (1.0 - temp) * (1.0 + temp))
                temp2 = temp * (cnorm[j] / partial[j]) * (cnorm[j] / partial[j])
                if temp2 > tol3z:
                    cnorm[j] = cnorm[j] * math.sqrt(temp)
                else:
                    if r < m - 1:
                        cnorm[j] = np.linalg.norm(a[r + 1:, j])
                        partial[j] = cnorm[j]
                    else:
                        cnorm[j] = 0.0
                        partial[j] = 0.0
        r += 1
        max_cnorm = float(np.max(cnorm[r:])) if r < n else 0.0
    # Q from the r reflectors (dorgqr)
    qf = np.tril(a[:, :r], -1)
    if r > 0:
        q, _, _ = lapack.dorgqr(qf, tau[:r])
        q = q[:, :r]
    else:
        q = np.zeros((m, 0))
    R = np.triu(a[:r, :])
    RP = np.zeros_like(R)
    RP[:, ipiv] = R
    return q, RP


def truncated_geqp3_bruteforce(a_in: np.ndarray, eps: float) -> Tuple[int, List[int]]:
    """Independent check of the stop rule: exact trailing column norms each step.

    Modified Gram-Schmidt with exact column norms; returns (rank, pivot order).
    The reference's downdated norms approximate exactly these values.
    """
    a = np.array(a_in, dtype=np.float64, copy=True)
    m, n = a.shape
    threshold = eps * math.sqrt(float(np.sum(a * a)))
    if float(np.max(np.linalg.norm(a, axis=0))) <= DLAMCH_E:
        return 1, [0]
    piv = list(range(n))
    r = 0
    while r < min(m, n):
        norms = np.linalg.norm(a[:, r:], axis=0)
        if float(np.max(norms)) <= threshold:
            break
        k = r + int(np.argmax(norms))
        a[:, [r, k]] = a[:, [k, r]]
        piv[r], piv[k] = piv[k], piv[r]
        q = a[:, r] / np.linalg.norm(a[:, r])
        a[:, r + 1:] -= np.outer(q, q @ a[:, r + 1:])
        a[:, r] = 0.0
        r += 1
    return r, piv[:r]


def geqp3(a: np.ndarray) -> Tuple[np.ndarray, List[int]]:
    """src/operations/LAPACK/geqp3.cpp:38-61: dgeqp3 (jpvt zero-initialised), R n x n upper, 0-based order."""
    a = np.array(a, dtype=np.float64, copy=True)
    m, n = a.shape
    r, order = sla.qr(a, mode="r", pivoting=True)  # dgeqp3 with all columns free (jpvt = 0)
    R = np.zeros((n, n))
    R[:min(m, n), :] = np.triu(r[:min(m, n), :])
    return R, [int(j) for j in order]


# ---------------------------------------------------------------------------
# LowRank
# ---------------------------------------------------------------------------
@dataclass
class LowRank:
    """include/FRANK/classes/low_rank.h:29-163: A ~= U S V."""
    U: np.ndarray
    S: np.ndarray
    V: np.ndarray
    eps: float = 0.0
    cap: Optional[int] = None  # rank cap of the capped-rank mode (no reference counterpart)

    @property
    def rank(self) -> int:
        return self.S.shape[0]

    @property
    def shape(self) -> Tuple[int, int]:
        return (self.U.shape[0], self.V.shape[1])

    def dense(self) -> np.ndarray:
        """src/classes/dense.cpp:64-92 fill_dense_from(LowRank): U S V."""
        return (self.U @ self.S) @ self.V

    def copy(self) -> "LowRank":
        return LowRank(self.U.copy(), self.S.copy(), self.V.copy(), self.eps, self.cap)

    @staticmethod
    def from_rank(a: np.ndarray, rank: int) -> "LowRank":
        """src/classes/low_rank.cpp:40-48: rsvd with oversampling min(rank+5, m, n), then resize."""
        m, n = a.shape
        u, s, v = rsvd(a, min(rank + 5, m, n))
        return LowRank(resize(u, m, rank), resize(s, rank, rank), resize(v, rank, n), 0.0)

    @staticmethod
    def from_eps(a: np.ndarray, eps: float, cap: Optional[int] = None) -> "LowRank":
        """src/classes/low_rank.cpp:50-59: truncated_geqp3 then RQ of R (S upper triangular)."""
        u, r = truncated_geqp3(a, eps, cap)
        s, v = rq(r)
        return LowRank(u, s, v, eps, cap)

    def transpose(self) -> "LowRank":
        """src/operations/misc/transpose.cpp: (U S V)^T = V^T S^T U^T."""
        return LowRank(self.V.T.copy(), self.S.T.copy(), self.U.T.copy(), self.eps, self.cap)


def rsvd(a: np.ndarray, p: int) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """src/operations/randomized/rsvd.cpp:17-29 (H6: Omega from mt19937 seed 0, U[0,1))."""
    m, n = a.shape
    rn = random_uniform(n, p)
    y = a @ rn
    q, _ = qr(y)
    qta = q.T @ a
    ub, s, v = svd(qta)
    return q @ ub, s, v


def rounded_addition(A: LowRank, B: LowRank) -> LowRank:
    """src/operations/arithmetic/addition.cpp:93-127 (QR + RQ + SVD, H2 rank if A.eps != 0)."""
    uc = np.hstack([A.U @ A.S, B.U @ B.S])
    qu, ru = qr(uc)
    vct = np.vstack([A.V, B.V])
    rvt, qvt = rq(vct)
    m = ru @ rvt
    rru, rrs, rrv = svd(m)
    rank = find_svd_truncation_rank(np.diag(rrs), A.eps, A.cap) if A.eps != 0 else A.rank
    S = resize(rrs, rank, rank)
    U = qu @ resize(rru, rru.shape[0], rank)
    V = resize(rrv, rank, rrv.shape[1]) @ qvt
    return LowRank(U, S, V, A.eps, A.cap)


def fast_rounded_addition(A: LowRank, B: LowRank) -> LowRank:
    """src/operations/arithmetic/addition.cpp:133-189 (A.U, A.V assumed orthonormal; keeps A.rank)."""
    if A.eps != 0.0:
        return rounded_addition(A, B)
    zu = A.U.T @ B.U
    yu = B.U - A.U @ zu
    qu, ru = qr(yu)
    uc = np.hstack([A.U, qu])
    zvt = B.V @ A.V.T
    yvt = B.V - zvt @ A.V
    rvt, qvt = rq(yvt)
    vct = np.vstack([A.V, qvt])
    ka, kb = A.S.shape[0], B.S.shape[0]
    M = np.zeros((ka + kb, ka + kb))
    zusb = zu @ B.S
    rusb = ru @ B.S
    M[:ka, :ka] = A.S + zusb @ zvt
    M[:ka, ka:] = zusb @ rvt
    M[ka:, :ka] = rusb @ zvt
    M[ka:, ka:] = rusb @ rvt
    um, sm, vmt = svd(M)
    r = A.rank
    return LowRank(uc @ resize(um, um.shape[0], r), resize(sm, r, r),
                   resize(vmt, r, vmt.shape[1]) @ vct, A.eps, A.cap)


def naive_addition(A: LowRank, B: LowRank) -> LowRank:
    """src/operations/arithmetic/addition.cpp:68-89: concatenate, dense recompress past min(m, n)."""
    m, n = A.shape
    if A.rank + B.rank >= min(m, n):
        return LowRank.from_rank(A.dense() + B.dense(), A.rank)
    ka, kb = A.rank, B.rank
    S = np.zeros((ka + kb, ka + kb))
    S[:ka, :ka] = A.S
    S[ka:, ka:] = B.S
    return LowRank(np.hstack([A.U, B.U]), S, np.vstack([A.V, B.V]), A.eps, A.cap)


LRA_DEFAULT = "fast_rounded_addition"


def lr_add(A: LowRank, B: LowRank, lra: str = LRA_DEFAULT) -> LowRank:
    """src/operations/arithmetic/addition.cpp:191-203: FRANK_LRA dispatch (H10)."""
    if lra == "naive":
        return naive_addition(A, B)
    if lra == "rounded_addition":
        return rounded_addition(A, B)
    return fast_rounded_addition(A, B)


def add_into(C, X, lra: str = LRA_DEFAULT):
    """operator+= for Dense/LowRank targets (addition.cpp:39-51, :191-203). Returns the new C."""
    if isinstance(C, LowRank):
        if isinstance(X, LowRank):
            return lr_add(C, X, lra)
        raise TypeError("operator+=(LowRank, Dense) undefined!")
    if isinstance(X, LowRank):
        return C + (X.U @ X.S) @ X.V
    return C + X


def _op(a, t):
    return a.T if t else a


def gemm(A, B, C, alpha: float = 1.0, beta: float = 1.0, ta: bool = False, tb: bool = False,
         lra: str = LRA_DEFAULT):
    """src/operations/BLAS/gemm.cpp:126-272: C = alpha op(A) op(B) + beta C for Dense/LowRank.

    Returns the updated C (numpy arrays are updated in place as well).
    """
    if isinstance(C, np.ndarray):
        if isinstance(A, np.ndarray) and isinstance(B, np.ndarray):  # D D D :126-146
            C[...] = alpha * (_op(A, ta) @ _op(B, tb)) + beta * C
            return C
        if isinstance(A, LowRank) and isinstance(B, np.ndarray):  # LR D D :148-162
            inner = alpha * (_op(A.U if ta else A.V, ta) @ _op(B, tb))
            as_b = _op(A.S, ta) @ inner
            C[...] = _op(A.V if ta else A.U, ta) @ as_b + beta * C
            return C
        if isinstance(A, np.ndarray) and isinstance(B, LowRank):  # D LR D :164-178
            a_b = alpha * (_op(A, ta) @ _op(B.V if tb else B.U, tb))
            a_bs = a_b @ _op(B.S, tb)
            C[...] = a_bs @ _op(B.U if tb else B.V, tb) + beta * C
            return C
        # LR LR D :180-200
        inner = alpha * (_op(A.U if ta else A.V, ta) @ _op(B.V if tb else B.U, tb))
        x = _op(A.S, ta) @ inner
        y = x @ _op(B.S, tb)
        z = _op(A.V if ta else A.U, ta) @ y
        C[...] = z @ _op(B.U if tb else B.V, tb) + beta * C
        return C
    # LowRank target
    if isinstance(A, np.ndarray) and isinstance(B, np.ndarray):  # D D LR :202-217
        C = LowRank(C.U, C.S * beta, C.V, C.eps, C.cap)
        prod = alpha * (_op(A, ta) @ _op(B, tb))
        X = LowRank.from_eps(prod, C.eps, C.cap) if C.eps != 0.0 else LowRank.from_rank(prod, C.rank)
        return lr_add(C, X, lra)
    if isinstance(A, LowRank) and isinstance(B, np.ndarray):  # LR D LR :219-234
        X = LowRank(A.V.T if ta else A.U, A.S.T if ta else A.S,
                    alpha * (_op(A.U if ta else A.V, ta) @ _op(B, tb)))
    elif isinstance(A, np.ndarray) and isinstance(B, LowRank):  # D LR LR :236-251
        X = LowRank(alpha * (_op(A, ta) @ _op(B.V if tb else B.U, tb)),
                    B.S.T if tb else B.S, B.U.T if tb else B.V)
    else:  # LR LR LR :253-272
        left = _op(A.S, ta) @ _op(A.U if ta else A.V, ta)
        right = _op(B.V if tb else B.U, tb) @ _op(B.S, tb)
        X = LowRank(A.V.T if ta else A.U, alpha * (left @ right), B.U.T if tb else B.V)
    C = LowRank(C.U, C.S * beta, C.V, C.eps, C.cap)
    return lr_add(C, X, lra)


def trsm(A: np.ndarray, B, uplo: str, side: str):
    """src/operations/BLAS/trsm.cpp:92-115: Dense target, or LowRank (Left -> U, Right -> V)."""
    if isinstance(B, LowRank):
        if side == "left":
            return LowRank(trsm_dense(A, B.U, uplo, side), B.S, B.V, B.eps, B.cap)
        return LowRank(B.U, B.S, trsm_dense(A, B.V, uplo, side), B.eps, B.cap)
    B[...] = trsm_dense(A, B, uplo, side)
    return B


# ---------------------------------------------------------------------------
# ID / RID (operator surface; tested at the reference's sizes)
# ---------------------------------------------------------------------------
def interleave_id(a: np.ndarray, P: List[int]) -> np.ndarray:
    """src/operations/LAPACK/id.cpp:29-39."""
    k = len(P) - a.shape[1]
    out = np.zeros((a.shape[0], len(P)))
    for j in range(len(P)):
        if j < k:
            out[:, P[j]] = (np.arange(a.shape[0]) == j).astype(np.float64)
        else:
            out[:, P[j]] = a[:, j - k]
    return out


def one_sided_id(a: np.ndarray, k: int) -> Tuple[np.ndarray, List[int]]:
    """src/operations/LAPACK/id.cpp:47-72: geqp3, R11^-1 R12 via trsm, interleave."""
    R, cols = geqp3(a)
    m, n = a.shape
    if k < min(m, n) or n > m:
        r11 = R[:k, :k]
        r12 = R[:k, k:]
        x = sla.solve_triangular(r11, r12, lower=False)
        basis = interleave_id(x, cols)
    else:
        basis = interleave_id(np.eye(k), cols)
    return basis, cols[:k]


def id_(a: np.ndarray, k: int) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """src/operations/LAPACK/id.cpp:100-110: two-sided ID (U, A[rows, cols], V)."""
    V, cols = one_sided_id(np.array(a, copy=True), k)
    ac = a[:, cols]
    ut, rows = one_sided_id(np.ascontiguousarray(ac.T), k)
    return ut.T, ac[rows, :], V


def rid(a: np.ndarray, p: int, k: int):
    """src/operations/randomized/rid.cpp:18-31: sketch, QR, ID of Q^T A."""
    m, n = a.shape
    rn = random_uniform(n, p)
    q, _ = qr(a @ rn)
    qta = q.T @ a
    ub, s, v = id_(qta, k)
    return q @ ub, s, v


# ---------------------------------------------------------------------------
# BLR (one-level Hierarchical) structure
# ---------------------------------------------------------------------------
def index_split(start: int, n: int, n_splits: int) -> List[Tuple[int, int]]:
    """src/classes/initialization_helpers/index_range.cpp:15-24 (ceil split, last smaller)."""
    out = []
    for i in range(n_splits):
        child_n = (n + n_splits - 1) // n_splits
        child_start = start + child_n * i
        if i == n_splits - 1:
            child_n = n - child_n * (n_splits - 1)
        out.append((child_start, child_n))
    return out


def is_admissible(rows, cols, bi, bj, admis: float, admis_type: str, points) -> bool:
    """src/classes/initialization_helpers/matrix_initializer.cpp:51-91."""
    ok = rows[1] > 1 and cols[1] > 1
    if admis_type == "position":
        return ok and abs(bi - bj) > int(admis)
    offset = 0.5
    max_len_r = max_len_c = 0.0
    dist = 0.0
    for x in points:
        rc = x[rows[0]:rows[0] + rows[1]]
        cc = x[cols[0]:cols[0] + cols[1]]
        mn_r, mx_r = rc.min() - offset, rc.max() + offset
        mn_c, mx_c = cc.min() - offset, cc.max() + offset
        ctr_r = mn_r + (mx_r - mn_r) / 2.0
        ctr_c = mn_c + (mx_c - mn_c) / 2.0
        max_len_r = max(max_len_r, mx_r - mn_r)
        max_len_c = max(max_len_c, mx_c - mn_c)
        d = abs(ctr_r - ctr_c)
        dist += d * d
    diam = max(max_len_r, max_len_c)
    return ok and (admis * admis * diam * diam) < dist


@dataclass
class Hier:
    """One-level Hierarchical (BLR) grid, row-major blocks (hierarchical.cpp:170-180)."""
    nb_r: int
    nb_c: int
    blocks: list
    row_ranges: List[Tuple[int, int]] = field(default_factory=list)
    col_ranges: List[Tuple[int, int]] = field(default_factory=list)

    def __getitem__(self, ij):
        i, j = ij
        return self.blocks[i * self.nb_c + j]

    def __setitem__(self, ij, v):
        i, j = ij
        self.blocks[i * self.nb_c + j] = v

    def copy(self) -> "Hier":
        def cp(b):
            if b is None:
                return None
            return b.copy()
        return Hier(self.nb_r, self.nb_c, [cp(b) for b in self.blocks],
                    list(self.row_ranges), list(self.col_ranges))

    def dense(self) -> np.ndarray:
        n_r = sum(r[1] for r in self.row_ranges)
        n_c = sum(c[1] for c in self.col_ranges)
        out = np.zeros((n_r, n_c))
        for i, (r0, rn) in enumerate(self.row_ranges):
            for j, (c0, cn) in enumerate(self.col_ranges):
                b = self[i, j]
                if b is None:
                    continue
                out[r0:r0 + rn, c0:c0 + cn] = b.dense() if isinstance(b, LowRank) else b
        return out

    def ranks(self) -> np.ndarray:
        r = np.full((self.nb_r, self.nb_c), -1, dtype=np.int64)
        for i in range(self.nb_r):
            for j in range(self.nb_c):
                b = self[i, j]
                if isinstance(b, LowRank):
                    r[i, j] = b.rank
        return r


def get_singular_values(a: np.ndarray) -> List[float]:
    """src/operations/LAPACK/svd.cpp:46-61: dgesvd 'N','N', all min(m, n) values, descending."""
    return [float(x) for x in sla.svd(np.asarray(a, dtype=np.float64), compute_uv=False, lapack_driver="gesvd")]


def to_json(H: Hier) -> dict:
    """src/util/print.cpp:54-122 (to_json over a one-level Hierarchical): every block
    carries type, dim, abs_pos and level; Hierarchical adds children (rows of
    blocks, abs_pos i_abs * dim0 + i), Dense adds svalues, LowRank adds svalues
    of its dense expansion and rank.  An Empty block has no to_json method in
    the reference (print.cpp:114-120 reports an error), so it raises here."""
    n_r = sum(r[1] for r in H.row_ranges)
    n_c = sum(c[1] for c in H.col_ranges)
    out = {"type": "Hierarchical", "dim": [n_r, n_c], "abs_pos": [0, 0], "level": 0, "children": []}
    for i, (_, rn) in enumerate(H.row_ranges):
        row = []
        for j, (_, cn) in enumerate(H.col_ranges):
            b = H[i, j]
            if b is None:
                raise TypeError("to_json: no method for an Empty block")
            node = {"type": "LowRank" if isinstance(b, LowRank) else "Dense", "dim": [rn, cn],
                    "abs_pos": [i, j], "level": 1}
            if isinstance(b, LowRank):
                node["svalues"] = get_singular_values(b.dense())
                node["rank"] = b.rank
            else:
                node["svalues"] = get_singular_values(b)
            row.append(node)
        out["children"].append(row)
    return out


def build_blr(kernel, n_rows: int, n_cols: int, nb_r: int, nb_c: int, *, eps: float = 0.0,
              rank: int = 0, admis: float = 0.0, admis_type: str = "position",
              points=None, cap: Optional[int] = None) -> Hier:
    """src/classes/hierarchical.cpp:49-109 for a one-level tree (every child a leaf).

    ``kernel(r0, nr, c0, nc)`` fills a block.  Fixed accuracy when ``eps`` != 0
    (LowRank(D, eps)), else fixed rank (LowRank(D, rank)).  ``cap`` bounds every
    eps-rank (the capped-rank mode; no reference counterpart).
    """
    rr = index_split(0, n_rows, nb_r)
    cr = index_split(0, n_cols, nb_c)
    blocks = []
    for i, r in enumerate(rr):
        for j, c in enumerate(cr):
            d = kernel(r[0], r[1], c[0], c[1])
            if is_admissible(r, c, i, j, admis, admis_type, points or []):
                blocks.append(LowRank.from_eps(d, eps, cap) if eps != 0.0 else LowRank.from_rank(d, rank))
            else:
                blocks.append(d)
    return Hier(nb_r, nb_c, blocks, rr, cr)


def blr_getrf(A: Hier, lra: str = LRA_DEFAULT, check_pivots: bool = True,
              op_log: Optional[list] = None, max_steps: Optional[int] = None) -> Tuple[Hier, Hier]:
    """src/operations/LAPACK/getrf.cpp:45-65: right-looking tile LU without pivoting.

    Consumes ``A`` (blocks moved into L), returns (L, U).  ``max_steps`` stops
    after that many elimination steps (a timing sample; the result is then a
    partial factorization).
    """
    nb = A.nb_r
    L = Hier(nb, A.nb_c, [None] * (nb * A.nb_c), list(A.row_ranges), list(A.col_ranges))
    for i in range(nb if max_steps is None else min(nb, max_steps)):
        Lii, Uii = getrf_dense(A[i, i], check_pivots)
        L[i, i] = Lii
        A[i, i] = Uii
        for ic in range(i + 1, nb):
            L[ic, i] = trsm(Uii, A[ic, i], "upper", "right")
            A[ic, i] = None
        for j in range(i + 1, A.nb_c):
            A[i, j] = trsm(Lii, A[i, j], "lower", "left")
        for ic in range(i + 1, nb):
            for k in range(i + 1, A.nb_c):
                A[ic, k] = gemm(L[ic, i], A[i, k], A[ic, k], -1.0, 1.0, lra=lra)
                if op_log is not None:
                    op_log.append((i, ic, k, A[ic, k].rank if isinstance(A[ic, k], LowRank) else -1))
    return L, A


def split_rows(b: np.ndarray, ranges) -> List[np.ndarray]:
    return [b[r0:r0 + rn] for (r0, rn) in ranges]


def blr_trsm_lower(L: Hier, b: np.ndarray) -> np.ndarray:
    """src/operations/BLAS/trsm.cpp:74-83, :117-125: block forward substitution (unit lower)."""
    b = np.array(b, dtype=np.float64, copy=True)
    if b.ndim == 1:
        b = b[:, None]
    bs = split_rows(b, L.row_ranges)
    for i in range(L.nb_r):
        for k in range(i):
            gemm(L[i, k], bs[k], bs[i], -1.0, 1.0)
        trsm(L[i, i], bs[i], "lower", "left")
    return b


def blr_trsm_upper(U: Hier, b: np.ndarray) -> np.ndarray:
    """src/operations/BLAS/trsm.cpp:47-54: block backward substitution (non-unit upper)."""
    b = np.array(b, dtype=np.float64, copy=True)
    if b.ndim == 1:
        b = b[:, None]
    bs = split_rows(b, U.row_ranges)
    for i in range(U.nb_r - 1, -1, -1):
        for j in range(U.nb_r - 1, i, -1):
            gemm(U[i, j], bs[j], bs[i], -1.0, 1.0)
        trsm(U[i, i], bs[i], "upper", "left")
    return b


def blr_trsm_upper_trans(U: Hier, b: np.ndarray) -> np.ndarray:
    """Solve U^T y = b (forward sweep with transposed tiles); adjoint use, no reference counterpart."""
    b = np.array(b, dtype=np.float64, copy=True)
    if b.ndim == 1:
        b = b[:, None]
    bs = split_rows(b, U.row_ranges)
    for i in range(U.nb_r):
        for k in range(i):
            gemm(U[k, i], bs[k], bs[i], -1.0, 1.0, ta=True)
        bs[i][...] = trsm_dense_trans(U[i, i], bs[i], "upper", "left")
    return b


def blr_trsm_lower_trans(L: Hier, b: np.ndarray) -> np.ndarray:
    """Solve L^T y = b (backward sweep, unit diagonal); adjoint use, no reference counterpart."""
    b = np.array(b, dtype=np.float64, copy=True)
    if b.ndim == 1:
        b = b[:, None]
    bs = split_rows(b, L.row_ranges)
    for i in range(L.nb_r - 1, -1, -1):
        for k in range(L.nb_r - 1, i, -1):
            gemm(L[k, i], bs[k], bs[i], -1.0, 1.0, ta=True)
        bs[i][...] = trsm_dense_trans(L[i, i], bs[i], "lower", "left")
    return b


def blr_matvec(A: Hier, x: np.ndarray) -> np.ndarray:
    """gemm(Hierarchical, Dense, Dense): y = A x block by block (gemm.cpp H D D)."""
    x = np.asarray(x, dtype=np.float64)
    if x.ndim == 1:
        x = x[:, None]
    n_r = sum(r[1] for r in A.row_ranges)
    y = np.zeros((n_r, x.shape[1]))
    for i, (r0, rn) in enumerate(A.row_ranges):
        for j, (c0, cn) in enumerate(A.col_ranges):
            blk = A[i, j]
            if blk is None:
                continue
            if isinstance(blk, LowRank):
                y[r0:r0 + rn] += blk.U @ (blk.S @ (blk.V @ x[c0:c0 + cn]))
            else:
                y[r0:r0 + rn] += blk @ x[c0:c0 + cn]
    return y


def l2_error(a, b) -> float:
    """src/util/l2_error.cpp:35-41: sqrt(||A-B||^2 / (||A||^2 + DBL_EPSILON))."""
    a = a.dense() if isinstance(a, (LowRank, Hier)) else np.asarray(a, dtype=np.float64)
    b = b.dense() if isinstance(b, (LowRank, Hier)) else np.asarray(b, dtype=np.float64)
    d = a - b
    return math.sqrt(norm(d) / (norm(a) + DBL_EPSILON))


def lu_product(L: Hier, U: Hier) -> np.ndarray:
    """Dense L @ U for the factorization-error metric ||A - LU|| / ||A||."""
    return L.dense() @ U.dense()


# ---------------------------------------------------------------------------
# Tiled Householder BLR-QR (src/operations/LAPACK/qr.cpp:213-257 and the
# geqrt / larfb / tpqrt / tpmqrt / trmm tile operations)
# ---------------------------------------------------------------------------
def trmm(A: np.ndarray, B, side: str, uplo: str, trans: str = "n", diag: str = "n", alpha: float = 1.0):
    """src/operations/BLAS/trmm.cpp:53-89: B = alpha op(tri(A)) B (left) or B op(tri(A)) (right).

    LowRank B: left acts on B.U, right on B.V.  Returns the updated B.
    """
    t = np.tril(A) if uplo == "lower" else np.triu(A)
    if diag == "u":
        t = t - np.diag(np.diag(t)) + np.eye(t.shape[0])
    op = t.T if trans == "t" else t
    if isinstance(B, LowRank):
        if side == "left":
            return LowRank(alpha * (op @ B.U), B.S, B.V, B.eps, B.cap)
        return LowRank(B.U, B.S, alpha * (B.V @ op), B.eps, B.cap)
    B[...] = alpha * (op @ B) if side == "left" else alpha * (B @ op)
    return B


def geqrt(A: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """src/operations/LAPACK/geqrt.cpp:34-43: dgeqrt3 (one block): A -> R above / V below, T (n x n)."""
    n = A.shape[1]
    a, t, info = lapack.dgeqrt(n, np.array(A, dtype=np.float64, copy=True))
    return a, t


def larfb(V: np.ndarray, T: np.ndarray, C, trans: bool):
    """src/operations/LAPACK/larfb.cpp:37-56: C <- (I - V T V^T)^{T|N} C ('L', 'F', 'C'); LowRank -> on C.U."""
    if isinstance(C, LowRank):
        return LowRank(larfb(V, T, C.U.copy(), trans), C.S, C.V, C.eps, C.cap)
    out, info = lapack.dgemqrt(V, T, np.array(C, dtype=np.float64, copy=True), side="L",
                               trans="T" if trans else "N")
    C[...] = out
    return C


def tpqrt(A: np.ndarray, B, T_out=None):
    """src/operations/LAPACK/tpqrt.cpp:35-53: dtpqrt2 (l = 0) of [A; B].

    A: R (upper triangle used); B dense m x n or LowRank (V <- S V, S <- I, QR of [R; V]).
    Returns (A, B, T).
    """
    if isinstance(B, LowRank):
        bv = B.S @ B.V
        A2, bv2, T = tpqrt(A, bv)
        return A2, LowRank(B.U, np.eye(B.rank), bv2, B.eps, B.cap), T
    n = A.shape[1]
    a, b, t, info = lapack.dtpqrt(0, n, np.array(A, dtype=np.float64, copy=True),
                                  np.array(B, dtype=np.float64, copy=True))
    return a, b, t


def _transpose_any(X):
    return X.transpose() if isinstance(X, LowRank) else np.ascontiguousarray(X.T)


def tpmqrt(V, T: np.ndarray, A, B, trans: bool, lra: str = "rounded_addition"):
    """src/operations/LAPACK/tpmqrt.cpp:45-170: apply the pentagonal reflector block to [A; B].

    All-dense: dtprfb ('L', T/N, 'F', 'C', l = 0).  Any LowRank operand:
    C = A + V^T B; C = T^{T|N} C; A -= C; B -= V C through the typed gemm /
    trmm / += of the reference.  Returns (A, B).
    """
    if all(isinstance(x, np.ndarray) for x in (V, A, B)):
        a, b, info = lapack.dtpmqrt(0, V, T, np.array(A, dtype=np.float64, copy=True),
                                    np.array(B, dtype=np.float64, copy=True), side="L",
                                    trans="T" if trans else "N")
        return a, b
    # C takes A's type when V is LowRank, else it is Dense(A) unless the
    # specialisation keeps a LowRank C (tpmqrt.cpp:61-170)
    if isinstance(A, LowRank) and (isinstance(V, LowRank) or isinstance(B, LowRank)):
        C = A.copy()
    else:
        C = A.dense() if isinstance(A, LowRank) else A.copy()
    Vt = _transpose_any(V)
    C = gemm(Vt, B, C, 1.0, 1.0, lra=lra)                       # C = A + V^T B
    C = trmm(T, C, "left", "upper", "t" if trans else "n", "n", 1.0)  # C = T^{T|N} C
    I = np.eye(T.shape[0])
    A = gemm(I, C, A, -1.0, 1.0, lra=lra)                       # A -= C
    B = gemm(V, C, B, -1.0, 1.0, lra=lra)                       # B -= V C
    return A, B


def tiled_householder_blr_qr(A: Hier, lra: str = "rounded_addition") -> Tuple[Hier, dict]:
    """src/operations/LAPACK/qr.cpp:213-228 (in place on A; returns A and the T tiles)."""
    Tt = {}
    for k in range(A.nb_c):
        A[k, k], Tt[k, k] = geqrt(A[k, k])
        for j in range(k + 1, A.nb_c):
            A[k, j] = larfb(A[k, k], Tt[k, k], A[k, j], True)
        for i in range(k + 1, A.nb_r):
            A[k, k], A[i, k], Tt[i, k] = tpqrt(A[k, k], A[i, k])
            for j in range(k + 1, A.nb_c):
                A[k, j], A[i, j] = tpmqrt(A[i, k], Tt[i, k], A[k, j], A[i, j], True, lra)
    return A, Tt


def left_multiply_tiled_reflector(Y: Hier, Tt: dict, C: Hier, trans: bool, lra: str = "rounded_addition") -> Hier:
    """src/operations/LAPACK/qr.cpp:230-257: C <- Q C (trans=False) or Q^T C from the stored reflectors."""
    if trans:
        for k in range(Y.nb_c):
            for j in range(k, Y.nb_c):
                C[k, j] = larfb(Y[k, k], Tt[k, k], C[k, j], True)
            for i in range(k + 1, Y.nb_r):
                for j in range(k, Y.nb_c):
                    C[k, j], C[i, j] = tpmqrt(Y[i, k], Tt[i, k], C[k, j], C[i, j], True, lra)
    else:
        for k in range(Y.nb_c - 1, -1, -1):
            for i in range(Y.nb_r - 1, k, -1):
                for j in range(k, Y.nb_c):
                    C[k, j], C[i, j] = tpmqrt(Y[i, k], Tt[i, k], C[k, j], C[i, j], False, lra)
            for j in range(k, Y.nb_c):
                C[k, j] = larfb(Y[k, k], Tt[k, k], C[k, j], False)
    return C


def blr_r_factor(A: Hier) -> np.ndarray:
    """Dense R from the upper tiles of the factored A (upper triangle of diagonal tiles)."""
    R = np.zeros((sum(r[1] for r in A.row_ranges[:A.nb_c]), sum(c[1] for c in A.col_ranges)))
    for i, (r0, rn) in enumerate(A.row_ranges[:A.nb_c]):
        for j, (c0, cn) in enumerate(A.col_ranges):
            if j < i:
                continue
            b = A[i, j]
            blk = b.dense() if isinstance(b, LowRank) else b
            R[r0:r0 + rn, c0:c0 + cn] = np.triu(blk) if i == j else blk
    return R
