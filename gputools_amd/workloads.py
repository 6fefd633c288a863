"""Synthetic point sets of the benchmark configurations (seeded, numpy).

  uniform        U[0,1)                           (n=1000 x 64 in the small case)
  normal         N(0,1)                           (n=8192 x 256, all methods)
  shifted        |N(0,1)| + 0.1                   (canberra: no 0/0 or x/0 term)
  bernoulli      0/1 with P(1) = 0.1              (binary)
  mixture        64 centres ~N(0, I), point i -> centre i mod 64, plus
                 1e-3 N(0, I) per point           (n=65536 x 512, cancellation)
  batched sets   sizes exp(U(ln 32, ln 2048)), d from {16, 32, 64, 128}
"""
from __future__ import annotations

import numpy as np


def uniform(n: int, k: int, seed: int = 1) -> np.ndarray:
    return np.random.default_rng(seed).random((n, k), dtype=np.float32)


def normal(n: int, k: int, seed: int = 2) -> np.ndarray:
    return np.random.default_rng(seed).standard_normal((n, k), dtype=np.float32)


def shifted(n: int, k: int, seed: int = 2) -> np.ndarray:
    return np.abs(normal(n, k, seed)) + np.float32(0.1)


def bernoulli(n: int, k: int, seed: int = 2, prob: float = 0.1) -> np.ndarray:
    return (np.random.default_rng(seed).random((n, k), dtype=np.float32) < prob).astype(np.float32)


def mixture(n: int, k: int, seed: int = 3, centers: int = 64, spread: float = 1e-3) -> np.ndarray:
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((centers, k), dtype=np.float32)
    x = rng.standard_normal((n, k), dtype=np.float32)
    x *= np.float32(spread)
    x += c[np.arange(n) % centers]
    return x


def method_input(method: str, n: int, k: int, seed: int = 2) -> np.ndarray:
    """The all-methods configuration's input for one method."""
    if method == "canberra":
        return shifted(n, k, seed)
    if method == "binary":
        return bernoulli(n, k, seed)
    return normal(n, k, seed)


def batched_sets(nsets: int = 4096, seed: int = 5, lo: int = 32, hi: int = 2048,
                 dims=(16, 32, 64, 128)):
    """Sizes and feature counts of the batched configuration."""
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), nsets))).astype(np.int64)
    ds = rng.choice(np.asarray(dims, dtype=np.int64), nsets)
    return sizes, ds
