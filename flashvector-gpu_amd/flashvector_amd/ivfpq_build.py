"""Seeded IVF-PQ table construction on the host (numpy).

This is the step before the scan path (crates/engine/src/index.rs:405-458,
crates/engine/src/cluster.rs:28-287): a coarse k-means, residuals x - c(x),
per-subspace PQ codebooks of 256 centroids, argmin encoding (lowest code wins
ties, cluster.rs:272-275), and posting lists that hold vector ids in ascending
order per cluster. The reference trains with an unseeded RNG, so its tables are
not reproducible; here every step is seeded, and the scan treats the resulting
(centroids, codebooks, codes, lists) as inputs.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class IvfPqTables:
    centroids: np.ndarray   # [nlist, dim] f32
    codebooks: np.ndarray   # [m_pq, 256, sub_dim] f32
    codes: np.ndarray       # [n, m_pq] u8 (indexed by vector id)
    offsets: np.ndarray     # [nlist + 1] u32
    vec_ids: np.ndarray     # [n] u32, ascending within each list
    vectors: np.ndarray     # [n, dim] f32 rows for the exact rerank
    m_pq: int

    @property
    def nlist(self) -> int:
        return self.centroids.shape[0]

    @property
    def dim(self) -> int:
        return self.centroids.shape[1]


def _sq_dists(x: np.ndarray, c: np.ndarray) -> np.ndarray:
    x64, c64 = x.astype(np.float64), c.astype(np.float64)
    return (x64 * x64).sum(1)[:, None] + (c64 * c64).sum(1)[None, :] - 2.0 * x64 @ c64.T


def kmeans(x: np.ndarray, k: int, iters: int, seed: int) -> np.ndarray:
    """k-means++ seeding then Lloyd iterations (cluster.rs:28-189 shape), seeded."""
    rng = np.random.default_rng(seed)
    n = x.shape[0]
    if n <= k:
        reps = np.concatenate([x, x[rng.integers(0, n, size=k - n)]]) if n < k else x.copy()
        return reps.astype(np.float32)
    centers = np.empty((k, x.shape[1]), np.float64)
    centers[0] = x[rng.integers(0, n)]
    d2 = _sq_dists(x, centers[:1])[:, 0].clip(min=0.0)
    for i in range(1, k):
        p = d2 / d2.sum() if d2.sum() > 0 else np.full(n, 1.0 / n)
        centers[i] = x[rng.choice(n, p=p)]
        d2 = np.minimum(d2, _sq_dists(x, centers[i:i + 1])[:, 0].clip(min=0.0))
    for _ in range(iters):
        a = np.argmin(_sq_dists(x, centers), axis=1)
        sums = np.zeros_like(centers)
        np.add.at(sums, a, x.astype(np.float64))
        cnt = np.bincount(a, minlength=k)
        live = cnt > 0
        centers[live] = sums[live] / cnt[live, None]
    return centers.astype(np.float32)


def assign(x: np.ndarray, centroids: np.ndarray, chunk: int = 65536) -> np.ndarray:
    out = np.empty(x.shape[0], np.int64)
    for s in range(0, x.shape[0], chunk):
        out[s:s + chunk] = np.argmin(_sq_dists(x[s:s + chunk], centroids), axis=1)
    return out


def pq_encode(residuals: np.ndarray, codebooks: np.ndarray) -> np.ndarray:
    m_pq, _, sub = codebooks.shape
    n = residuals.shape[0]
    codes = np.empty((n, m_pq), np.uint8)
    for m in range(m_pq):
        codes[:, m] = assign(residuals[:, m * sub:(m + 1) * sub], codebooks[m]).astype(np.uint8)
    return codes


def build_ivfpq(vectors: np.ndarray, nlist: int, m_pq: int, seed: int, iters: int = 10,
                train_sample: int = 65536, integer_codebooks: bool = False,
                centroids: np.ndarray | None = None,
                assignment: np.ndarray | None = None) -> IvfPqTables:
    """Trains coarse and PQ quantizers on a seeded sample and encodes all rows.

    integer_codebooks rounds centroids and codebooks to integers, which makes
    every LUT entry and ADC sum exact in FP32 (used by order-independent tests).
    `centroids`/`assignment` may be supplied to skip the coarse k-means.
    """
    rng = np.random.default_rng(seed)
    x = np.ascontiguousarray(vectors, dtype=np.float32)
    n, dim = x.shape
    assert dim % m_pq == 0
    sub = dim // m_pq
    samp = x[rng.choice(n, size=min(n, train_sample), replace=False)]
    if centroids is None:
        centroids = kmeans(samp, nlist, iters, seed + 1)
    if integer_codebooks:
        centroids = np.rint(centroids).astype(np.float32)
    a = assign(x, centroids) if assignment is None else np.asarray(assignment, np.int64)
    resid = x - centroids[a]
    samp_res = resid[rng.choice(n, size=min(n, train_sample), replace=False)]
    codebooks = np.stack([kmeans(samp_res[:, m * sub:(m + 1) * sub], 256, iters, seed + 2 + m)
                          for m in range(m_pq)])
    if integer_codebooks:
        codebooks = np.rint(codebooks).astype(np.float32)
    codes = pq_encode(resid, codebooks)
    order = np.argsort(a, kind="stable")          # ascending vec_id within each list
    counts = np.bincount(a, minlength=nlist)
    offsets = np.zeros(nlist + 1, np.uint32)
    offsets[1:] = np.cumsum(counts)
    return IvfPqTables(centroids.astype(np.float32), codebooks.astype(np.float32), codes,
                       offsets, order.astype(np.uint32), x, m_pq)
