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


def normal_rhs(n: int, nrhs: int, seed: int = 1) -> np.ndarray:
    """Right-hand sides ~ N(0, 1), row-major n x nrhs."""
    return np.random.default_rng(seed).standard_normal((n, nrhs))
