"""Synthetic workload inputs with the reference's seeding.

``sorted_uniform_points`` reproduces get_sorted_random_vector
(src/util/experiment_setup.cpp:12-23): std::mt19937 seeded with 0,
uniform_real_distribution<double>(0, 1) (libstdc++ draws two 32-bit words
per double), sorted ascending.  numpy's legacy MT19937 seeding is the same
init_genrand(seed) as std::mt19937::seed(seed).
"""
from __future__ import annotations

import numpy as np


def mt19937_uniform(seed: int, n: int) -> np.ndarray:
    bg = np.random.MT19937()
    bg._legacy_seeding(seed)
    w = bg.random_raw(2 * n).astype(np.float64)
    u = (w[0::2] + w[1::2] * 4294967296.0) / 18446744073709551616.0
    u[u >= 1.0] = np.nextafter(1.0, 0.0)
    return u


def sorted_uniform_points(n: int) -> np.ndarray:
    return np.sort(mt19937_uniform(0, n))


def uniform_points(n: int, dim: int, seed: int = 0) -> np.ndarray:
    """dim x n points from ONE std::mt19937(seed) stream: coordinate d takes
    draws [d*n, (d+1)*n) (two get_sorted_random_vector calls would repeat the
    same stream, src/util/experiment_setup.cpp:12-23)."""
    return mt19937_uniform(seed, dim * n).reshape(dim, n)


def morton_sort(points: np.ndarray, level: int):
    """Points (dim x n) reordered along the Morton curve of the given level
    (src/operations/misc/morton_index.cpp:75-98: normalise into a box padded
    by 0.5, box index from the interleaved bits of floor(x * 2^level)).  The
    sort is stable (the reference's std::sort leaves ties in unspecified
    order).  Returns (sorted points, permutation)."""
    x = np.asarray(points, dtype=np.float64)
    dim, n = x.shape
    pad = 5e-1
    lo = x.min(axis=1) - pad
    size = (x.max(axis=1) - x.min(axis=1)).max() + 2 * pad
    xn = (x - lo[:, None]) / size
    ix = np.floor(xn * (1 << level)).astype(np.int64)
    key = np.zeros(n, dtype=np.int64)
    for lev in range(level):
        for d in range(dim):
            key += ((ix[d] >> lev) & 1) << (dim * lev + d)
    perm = np.argsort(key, kind="stable")
    return np.ascontiguousarray(x[:, perm]), perm


def normal_rhs(n: int, nrhs: int, seed: int = 1) -> np.ndarray:
    """Right-hand sides ~ N(0, 1), row-major n x nrhs."""
    return np.random.default_rng(seed).standard_normal((n, nrhs))
