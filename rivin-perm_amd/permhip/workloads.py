"""Seeded synthetic inputs of the documented workloads (numpy default_rng).

  headline   32 x 32 integers U[-10, 10], seed 20261203 (8 CRT primes)
  config1    20 x 20 integers U[-10, 10], seed 1 (5 CRT primes)
  schur31    S_31 (no matrix: omega^(jk) per prime)
  binary36   36 x 36 Bernoulli(0.5), seed 3, redrawn until every row and
             column sum is at least 1
  batch4096  4096 matrices, n ~ U[12, 28], entries U[-100, 100], seed 4
  serve      Poisson arrivals (200/s over 60 s, seed 5), half integer
             requests (n ~ U[10, 34], entries U[-50, 50]) and half 0/1
             requests (n ~ U[16, 38], density ~ U[0.3, 0.7]); the 50/50 mix
             is this build's choice
"""

from __future__ import annotations

from typing import List, Tuple

import numpy as np


def headline() -> np.ndarray:
    return np.random.default_rng(20261203).integers(-10, 11, (32, 32)).astype(np.int64)


def config1() -> np.ndarray:
    return np.random.default_rng(1).integers(-10, 11, (20, 20)).astype(np.int64)


def binary_matrix(n: int, density: float, rng) -> np.ndarray:
    """0/1 matrix with i.i.d. Bernoulli(density) entries, redrawn until no row
    or column is empty."""
    while True:
        B = (rng.random((n, n)) < density).astype(np.uint8)
        if B.sum(axis=0).min() >= 1 and B.sum(axis=1).min() >= 1:
            return B


def binary36() -> np.ndarray:
    rng = np.random.default_rng(3)
    while True:
        B = rng.integers(0, 2, (36, 36)).astype(np.uint8)
        if B.sum(axis=0).min() >= 1 and B.sum(axis=1).min() >= 1:
            return B


def batch4096(count: int = 4096) -> List[np.ndarray]:
    rng = np.random.default_rng(4)
    mats = []
    for _ in range(count):
        n = int(rng.integers(12, 29))
        mats.append(rng.integers(-100, 101, (n, n)).astype(np.int64))
    return mats


def serve_stream(rate: float = 200.0, seconds: float = 60.0, seed: int = 5) -> List[Tuple[float, str, np.ndarray]]:
    """[(arrival time s, 'integer' | 'binary', matrix)] in arrival order."""
    rng = np.random.default_rng(seed)
    out = []
    t = 0.0
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= seconds:
            break
        if rng.random() < 0.5:
            n = int(rng.integers(10, 35))
            out.append((t, "integer", rng.integers(-50, 51, (n, n)).astype(np.int64)))
        else:
            n = int(rng.integers(16, 39))
            density = float(rng.uniform(0.3, 0.7))
            out.append((t, "binary", binary_matrix(n, density, rng)))
    return out
