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
# (mean directions, rows, queries, per-query k) of the FP8 unit-vector workload
VMF_SEEDS = (31, 32, 33, 34)


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


# --------------------------------------------------------------------------
# von Mises-Fisher mixture on the unit sphere (torch; CPU or GPU)
# --------------------------------------------------------------------------

def vmf_mean_directions(n_clusters: int, dim: int, seed: int, device="cpu"):
    """Mean directions uniform on S^(dim-1) (normalised Gaussians), float64."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    mu = torch.randn((n_clusters, dim), generator=g, device=device, dtype=torch.float64)
    return mu / mu.norm(dim=1, keepdim=True)


def vmf_sample(mu, kappa: float, gen):
    """One vMF(mu_i, kappa) sample per row of mu (float64 [n, d]) by Wood's
    (1994) rejection sampler: W from the envelope with
    b = (-2k + sqrt(4k^2 + (d-1)^2)) / (d-1), x0 = (1-b)/(1+b),
    c = k*x0 + (d-1)*ln(1-x0^2); Z ~ Beta((d-1)/2, (d-1)/2) taken as
    (1 + G_1/|G|)/2 for G ~ N(0, I_d) (the first coordinate of a uniform point
    on the sphere); accept W = (1-(1+b)Z)/(1-(1-b)Z) when
    k*W + (d-1)*ln(1-x0*W) - c >= ln U. Then x = W*mu + sqrt(1-W^2)*V with V
    uniform on the great sphere orthogonal to mu."""
    import math

    import torch
    n, d = mu.shape
    dev = mu.device
    b = (-2.0 * kappa + math.sqrt(4.0 * kappa * kappa + (d - 1.0) ** 2)) / (d - 1.0)
    x0 = (1.0 - b) / (1.0 + b)
    c = kappa * x0 + (d - 1.0) * math.log(1.0 - x0 * x0)
    W = torch.empty(n, dtype=torch.float64, device=dev)
    todo = torch.arange(n, device=dev)
    while todo.numel():
        m = todo.numel()
        G = torch.randn((m, d), generator=gen, device=dev, dtype=torch.float64)
        z = 0.5 * (1.0 + G[:, 0] / G.norm(dim=1))
        u = torch.rand(m, generator=gen, device=dev, dtype=torch.float64)
        w = (1.0 - (1.0 + b) * z) / (1.0 - (1.0 - b) * z)
        ok = kappa * w + (d - 1.0) * torch.log(1.0 - x0 * w) - c >= torch.log(u)
        W[todo[ok]] = w[ok]
        todo = todo[~ok]
    V = torch.randn((n, d), generator=gen, device=dev, dtype=torch.float64)
    V = V - (V * mu).sum(1, keepdim=True) * mu
    V = V / V.norm(dim=1, keepdim=True)
    return W[:, None] * mu + torch.sqrt(1.0 - W * W)[:, None] * V


def vmf_points(mu, n: int, kappa: float, seed: int, chunk: int = 1 << 18):
    """n float32 unit rows of the mixture with means mu (cluster ~ uniform)."""
    import torch
    dev = mu.device
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    out = torch.empty((n, mu.shape[1]), dtype=torch.float32, device=dev)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        a = torch.randint(0, mu.shape[0], (e - s,), generator=g, device=dev)
        out[s:e] = vmf_sample(mu[a], kappa, g).to(torch.float32)
    return out


def log_uniform_k(nq: int, k_max: int, seed: int):
    """k_i = floor(exp(U[0, ln k_max])) clamped to [1, k_max] (int32 numpy)."""
    rng = np.random.default_rng(seed)
    k = np.floor(np.exp(rng.uniform(0.0, np.log(k_max), size=nq))).astype(np.int64)
    return np.clip(k, 1, k_max).astype(np.int32)


def vmf_workload(n: int, nq: int, dim: int = 768, n_clusters: int = 1024, kappa: float = 50.0,
                 seeds=VMF_SEEDS, device="cpu"):
    """(database, queries) as float32 torch tensors on `device`: unit rows of a
    vMF mixture (means from seeds[0], rows seeds[1], queries seeds[2])."""
    mu = vmf_mean_directions(n_clusters, dim, seeds[0], device)
    return vmf_points(mu, n, kappa, seeds[1]), vmf_points(mu, nq, kappa, seeds[2])
