"""IVF-PQ tables built on the GPU (fv_build.h through ctypes).

The step before the scan: GpuVectorIndex::build's IVF-PQ stage
(crates/engine/src/index.rs:405-458) with KMeans::fit and ProductQuantizer
train/encode (crates/engine/src/cluster.rs:28-287), made deterministic by a
seed. Inputs and outputs are ROCm tensors; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

from . import lib as L


@dataclass
class KMeansResult:
    centroids: object   # [k, dim] f32
    assignments: object  # [n] int32 (uint32 values), the last assign step
    sizes: object        # [k] int32
    offsets: object      # [k + 1] int32, rows of each cluster in `members`
    members: object      # [n] int32, ascending within each cluster
    iterations: int


@dataclass
class GpuIvfPqTables:
    centroids: object    # [nlist, dim] f32
    codebooks: object    # [m_pq, 256 * sub_dim] f32
    codes: object        # [n, m_pq] u8 by row
    offsets: object      # [nlist + 1] int32
    vec_ids: object      # [n] int32 ascending per list
    iterations: list     # coarse, then one per sub-quantizer

    def index(self, vectors=None, list_order=True):
        """An IvfPqIndex over these tables (vectors: FP32 rows for the rerank)."""
        from .search import IvfPqIndex
        return IvfPqIndex(self.centroids, self.codebooks, self.codes, self.offsets,
                          self.vec_ids, vectors, list_order=list_order)


def _rows(x):
    torch = L._torch()
    if x.dtype != torch.float32 or not x.is_cuda or x.dim() != 2:
        raise L.FvError("rows must be a 2-D float32 ROCm tensor")
    return x.contiguous()


def kmeans_assign(x, centroids):
    """Nearest centroid of every row (first strict minimum)."""
    torch = L._torch()
    x, c = _rows(x), _rows(centroids)
    out = torch.empty(x.shape[0], dtype=torch.int32, device=x.device)
    L.check(L.load().fv_kmeans_assign(L.ptr(x), x.shape[0], x.shape[1], x.shape[1], L.ptr(c),
                                      c.shape[0], L.ptr(out), None, 0, L.stream_handle()),
            "fv_kmeans_assign")
    return out


def kmeans_fit(x, k, max_iters=25, seed=0) -> KMeansResult:
    torch = L._torch()
    x = _rows(x)
    n, dim = x.shape
    lib = L.load()
    ws_bytes = lib.fv_kmeans_workspace_size(n, dim, k)
    if ws_bytes == 0:
        raise L.FvError(f"kmeans_fit: unsupported shape n={n} dim={dim} k={k}")
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=x.device)
    dev = dict(dtype=torch.int32, device=x.device)
    cent = torch.empty((k, dim), dtype=torch.float32, device=x.device)
    assign, sizes = torch.empty(n, **dev), torch.empty(k, **dev)
    offsets, members = torch.empty(k + 1, **dev), torch.empty(n, **dev)
    it = ctypes.c_uint32(0)
    L.check(lib.fv_kmeans_fit(L.ptr(x), n, dim, dim, k, max_iters, seed, L.ptr(cent),
                              L.ptr(assign), L.ptr(sizes), L.ptr(offsets), L.ptr(members),
                              ctypes.byref(it), L.ptr(ws), ws_bytes, L.stream_handle()),
            "fv_kmeans_fit")
    return KMeansResult(cent, assign, sizes, offsets, members, it.value)


def build_ivfpq(vectors, nlist, m_pq, coarse_iters=25, pq_iters=25, seed=0) -> GpuIvfPqTables:
    """Coarse k-means, residual PQ (256 codewords per sub-space), codes and
    posting lists of `vectors` [n, dim] f32 (dim % m_pq == 0)."""
    torch = L._torch()
    x = _rows(vectors)
    n, dim = x.shape
    if m_pq <= 0 or dim % m_pq:
        raise L.FvError("dim must be a multiple of m_pq")
    lib = L.load()
    ws_bytes = lib.fv_ivf_pq_build_workspace_size(n, dim, nlist, m_pq)
    if ws_bytes == 0:
        raise L.FvError(f"build_ivfpq: unsupported shape n={n} dim={dim}")
    nl = max(1, min(nlist, n))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=x.device)
    cent = torch.empty((nl, dim), dtype=torch.float32, device=x.device)
    books = torch.empty((m_pq, 256 * (dim // m_pq)), dtype=torch.float32, device=x.device)
    codes = torch.empty((n, m_pq), dtype=torch.uint8, device=x.device)
    offsets = torch.empty(nl + 1, dtype=torch.int32, device=x.device)
    vec_ids = torch.empty(n, dtype=torch.int32, device=x.device)
    iters = (ctypes.c_uint32 * (1 + m_pq))()
    L.check(lib.fv_ivf_pq_build(L.ptr(x), n, dim, nlist, m_pq, coarse_iters, pq_iters, seed,
                                L.ptr(cent), L.ptr(books), L.ptr(codes), L.ptr(offsets),
                                L.ptr(vec_ids), iters, L.ptr(ws), ws_bytes, L.stream_handle()),
            "fv_ivf_pq_build")
    del ws
    return GpuIvfPqTables(cent, books, codes, offsets, vec_ids, list(iters))
