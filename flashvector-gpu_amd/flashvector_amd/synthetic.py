"""Seeded synthetic workloads for the query-scan path.

Gaussian mixture of isotropic clusters with integer-valued coordinates: centers
~ U[0, 100]^d, point = clip(round(center[a] + N(0, sigma^2)), 0, 255) with
a ~ U{0..n_clusters-1}. Integer coordinates in [0, 255] at d <= 256 make every
FP32 partial sum of squared differences an exact integer below 2^24, and
FP16/BF16 store every coordinate exactly.

Seeds are fixed per workload so that the GPU run, the CPU baseline and the
tests see identical inputs.
"""
from __future__ import annotations

import numpy as np

# (centers/database seed, query seed, mask seed)
SEEDS = {
    "flat_fp32_100k": (1, 2, 3),
    "flat_mixed_1m": (11, 12, 13),
}


def mixture_centers(n_clusters: int, dim: int, seed: int, high: float = 100.0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, high, size=(n_clusters, dim)).astype(np.float64)


def mixture_points(centers: np.ndarray, n: int, seed: int, sigma: float = 8.0,
                   chunk: int = 1 << 18) -> np.ndarray:
    """n integer-valued float32 points drawn from the mixture (chunked to bound memory)."""
    rng = np.random.default_rng(seed)
    k, d = centers.shape
    out = np.empty((n, d), np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        a = rng.integers(0, k, size=e - s)
        pts = centers[a] + rng.normal(0.0, sigma, size=(e - s, d))
        np.clip(np.rint(pts), 0.0, 255.0, out=pts)
        out[s:e] = pts
    return out


def flat_workload(n: int, nq: int, dim: int = 128, n_clusters: int = 256,
                  seeds=(1, 2, 3)):
    """(database, queries) from one mixture: centers and rows from seeds[0],
    queries from seeds[1]."""
    centers = mixture_centers(n_clusters, dim, seeds[0])
    db = mixture_points(centers, n, seeds[0] + 1000)
    qs = mixture_points(centers, nq, seeds[1])
    return db, qs


def random_valid_mask(n: int, frac_valid: float, seed: int) -> np.ndarray:
    """Packed LSB-first validity words (u32) with Bernoulli(frac_valid) bits."""
    rng = np.random.default_rng(seed)
    bits = rng.random(n) < frac_valid
    return pack_bits(bits)


def pack_bits(bits: np.ndarray) -> np.ndarray:
    """bool[n] -> u32[ceil(n/32)], bit i at word i>>5, position i&31
    (crates/python/src/py_index.rs:73-81)."""
    bits = np.asarray(bits, dtype=bool)
    n = bits.shape[0]
    nw = (n + 31) // 32
    padded = np.zeros(nw * 32, dtype=bool)
    padded[:n] = bits
    by = np.packbits(padded.reshape(-1, 8), axis=1, bitorder="little").reshape(-1)
    return by.view(np.uint32).copy() if by.size % 4 == 0 else np.frombuffer(
        np.concatenate([by, np.zeros(4 - by.size % 4, np.uint8)]).tobytes(), np.uint32).copy()


def unpack_bits(words: np.ndarray, n: int) -> np.ndarray:
    by = np.ascontiguousarray(words, dtype=np.uint32).view(np.uint8)
    return np.unpackbits(by, bitorder="little")[:n].astype(bool)
