"""Seeded IVF-PQ workload built on the GPU (torch), for the 10M-row batched
uneven scan: Gaussian-mixture rows around `nlist` list centers whose sizes
follow Zipf(s) (largest remainder to the exact total, randomly permuted over
lists), PQ codebooks of 256 centroids per subspace trained by seeded Lloyd
k-means on a residual sample, argmin encoding (lowest code wins ties), posting
lists in ascending vector id, and queries drawn from the mixture weighted by
list size with per-query nprobe and k.

The coarse centroids are the generating centers; every row is assigned to its
nearest centroid. Tables are returned as device tensors in the layout of
IvfPqGpuTables (kernels/include/types.h:85-97).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class IvfPqWorkload:
    centroids: "object"   # [nlist, d] f32
    codebooks: "object"   # [m_pq, 256*sub] f32
    codes: "object"       # [n, m_pq] u8 (by vector id)
    offsets: "object"     # [nlist+1] i32 (u32 bits)
    vec_ids: "object"     # [n] i32 (u32 bits), ascending per list
    vectors: "object"     # [n, d] f32
    queries: "object"     # [nq, d] f32
    top_k: "object"       # [nq] i32
    nprobe: "object"      # [nq] i32
    sizes: np.ndarray     # list sizes (host)
    m_pq: int


def zipf_sizes(total: int, nlist: int, s: float, seed: int) -> np.ndarray:
    w = np.arange(1, nlist + 1, dtype=np.float64) ** (-s)
    raw = w / w.sum() * total
    sizes = np.floor(raw).astype(np.int64)
    rem = total - int(sizes.sum())
    order = np.argsort(-(raw - sizes), kind="stable")
    sizes[order[:rem]] += 1
    rng = np.random.default_rng(seed)
    return sizes[rng.permutation(nlist)]


def build(n=10_000_000, dim=128, nlist=4096, zipf_s=1.1, m_pq=16, nq=8192, sigma=8.0,
          nprobe_choices=(8, 16, 32, 64), k_choices=(10, 100), seed=21, train=262144,
          iters=10, device="cuda") -> IvfPqWorkload:
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    centers = torch.rand((nlist, dim), generator=g, device=device) * 100.0
    sizes = zipf_sizes(n, nlist, zipf_s, seed + 1)
    labels = torch.repeat_interleave(torch.arange(nlist, device=device),
                                      torch.from_numpy(sizes).to(device))
    g.manual_seed(seed + 2)
    perm = torch.randperm(n, generator=g, device=device)
    labels = labels[perm]
    vectors = torch.empty((n, dim), dtype=torch.float32, device=device)
    for s in range(0, n, 1 << 20):
        e = min(n, s + (1 << 20))
        noise = torch.randn((e - s, dim), generator=g, device=device) * sigma
        vectors[s:e] = torch.clamp(torch.round(centers[labels[s:e]] + noise), 0.0, 255.0)
    from . import gpu_build as G
    # Exact nearest centers and deterministic k-means codebooks (fv_kmeans_*),
    # so the tables are the same on every run.
    assign = G.kmeans_assign(vectors, centers).long()
    sub = dim // m_pq
    g.manual_seed(seed + 3)
    samp = torch.randperm(n, generator=g, device=device)[:min(n, train)]
    resid_s = vectors[samp] - centers[assign[samp]]
    books = []
    for m in range(m_pq):
        books.append(G.kmeans_fit(resid_s[:, m * sub:(m + 1) * sub].contiguous(), 256, iters,
                                  seed=seed + 3 + m).centroids)
    codebooks = torch.stack(books)                      # [m_pq, 256, sub]
    codes = torch.empty((n, m_pq), dtype=torch.uint8, device=device)
    for s in range(0, n, 1 << 20):
        e = min(n, s + (1 << 20))
        r = vectors[s:e] - centers[assign[s:e]]
        for m in range(m_pq):
            codes[s:e, m] = G.kmeans_assign(r[:, m * sub:(m + 1) * sub].contiguous(),
                                            codebooks[m]).to(torch.uint8)
    order = torch.argsort(assign * n + torch.arange(n, device=device))   # ascending id per list
    counts = torch.bincount(assign, minlength=nlist)
    offsets = torch.zeros(nlist + 1, dtype=torch.int64, device=device)
    offsets[1:] = torch.cumsum(counts, 0)
    g.manual_seed(seed + 4)
    w = torch.from_numpy(sizes.astype(np.float64)).to(device)
    qlists = torch.multinomial(w / w.sum(), nq, replacement=True, generator=g)
    noise = torch.randn((nq, dim), generator=g, device=device) * sigma
    queries = torch.clamp(torch.round(centers[qlists] + noise), 0.0, 255.0).contiguous()
    g.manual_seed(seed + 5)
    npc = torch.tensor(nprobe_choices, device=device, dtype=torch.int32)
    kc = torch.tensor(k_choices, device=device, dtype=torch.int32)
    nprobe = npc[torch.randint(0, len(nprobe_choices), (nq,), generator=g, device=device)]
    top_k = kc[torch.randint(0, len(k_choices), (nq,), generator=g, device=device)]
    return IvfPqWorkload(centers.contiguous(), codebooks.reshape(m_pq, -1).contiguous(),
                         codes.contiguous(), offsets.to(torch.int32).contiguous(),
                         order.to(torch.int32).contiguous(), vectors, queries,
                         top_k.contiguous(), nprobe.contiguous(), counts.cpu().numpy(), m_pq)


def scanned_bytes(w: IvfPqWorkload, shortlist: int) -> float:
    """Algorithmic HBM bytes of one batched search: for every query the codes
    and ids of each probed list (M + 4 bytes per entry) and the FP32 rows of
    its shortlist (4d + 4), plus the coarse centroids and codebooks once.
    Probes are recomputed here from FP32 coarse distances."""
    import torch
    c, q = w.centroids, w.queries
    d = (c * c).sum(1)[None, :] - 2.0 * (q @ c.T)
    maxp = int(w.nprobe.max())
    probes = torch.topk(d, maxp, dim=1, largest=False).indices
    sizes = torch.from_numpy(w.sizes).to(q.device)
    sel = torch.arange(maxp, device=q.device)[None, :] < w.nprobe[:, None]
    entries = (sizes[probes] * sel).sum().item()
    m, dim = w.m_pq, c.shape[1]
    return float(entries * (m + 4) + q.shape[0] * shortlist * (4 * dim + 4) +
                 c.numel() * 4 + w.codebooks.numel() * 4)
