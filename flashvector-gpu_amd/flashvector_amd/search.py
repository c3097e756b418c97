"""Functional torch-facing wrappers over the C ABI.

Each function takes device tensors, calls the corresponding extern "C" entry
point on torch's current stream and returns device tensors. Results are int64
tensors holding GpuSearchResult pairs (low 32 bits id, high 32 bits distance
bits); `lib.split_results` turns them into numpy (ids, distances).
"""
from __future__ import annotations

import ctypes

from . import lib as L


def _metric(metric) -> int:
    return L.METRICS[metric] if isinstance(metric, str) else int(metric)


def flat_scan_batch(vectors, queries, top_k: int, metric="l2", valid_mask=None, vec_ids=None,
                    out=None):
    """cuda_flat_scan_batch (kernels/src/sparse_hybrid.cu:504-524): exact FP32,
    top_k clamped to 256."""
    lib = L.load()
    nq, dim = queries.shape
    width = max(1, min(int(top_k), 256))
    out = L.result_tensor(nq, width, queries.device) if out is None else out
    rc = lib.cuda_flat_scan_batch(L.ptr(vectors), L.ptr(vec_ids), vectors.shape[0], dim,
                                  L.ptr(queries), nq, top_k, _metric(metric), L.ptr(valid_mask),
                                  L.ptr(out), L.stream_handle())
    L.check(rc, "cuda_flat_scan_batch")
    return out


def flat_exact_search(vectors, queries, top_k: int, metric="l2", valid_mask=None, workspace=None):
    """fv_flat_exact_search: exact FP32 scan with k up to 1024 and a caller workspace."""
    torch = L._torch()
    lib = L.load()
    nq, dim = queries.shape
    n = vectors.shape[0]
    need = lib.fv_flat_exact_workspace_size(n, dim, nq, top_k, _metric(metric))
    if workspace is None:
        workspace = torch.empty(max(need, 1), dtype=torch.uint8, device=queries.device)
    out = L.result_tensor(nq, top_k, queries.device)
    rc = lib.fv_flat_exact_search(L.ptr(vectors), n, dim, L.ptr(queries), nq, top_k,
                                  _metric(metric), L.ptr(valid_mask), L.ptr(out),
                                  L.ptr(workspace), workspace.numel(), L.stream_handle())
    L.check(rc, "fv_flat_exact_search")
    return out


def merge_topk(shard_results, out_top_k: int):
    """cuda_merge_topk over an int64 [num_shards, num_queries, shard_top_k] tensor."""
    lib = L.load()
    s, nq, k = shard_results.shape
    width = max(1, min(int(out_top_k), 256))
    out = L.result_tensor(nq, width, shard_results.device)
    rc = lib.cuda_merge_topk(L.ptr(shard_results), s, nq, k, out_top_k, L.ptr(out),
                             L.stream_handle())
    L.check(rc, "cuda_merge_topk")
    return out


def compute_distances(queries, dataset, metric="l2"):
    """cuda_compute_distances_warp: dense [nq, n] exact FP32 distance matrix."""
    torch = L._torch()
    lib = L.load()
    out = torch.empty((queries.shape[0], dataset.shape[0]), dtype=torch.float32,
                      device=queries.device)
    rc = lib.cuda_compute_distances_warp(L.ptr(queries), L.ptr(dataset), queries.shape[0],
                                         dataset.shape[0], queries.shape[1], _metric(metric),
                                         L.ptr(out), L.stream_handle())
    L.check(rc, "cuda_compute_distances_warp")
    return out


def bitonic_sort_test(keys, vals):
    """cuda_bitonic_sort_test: one-wavefront register sort of up to 64 pairs."""
    torch = L._torch()
    lib = L.load()
    ko = torch.empty_like(keys)
    vo = torch.empty_like(vals)
    rc = lib.cuda_bitonic_sort_test(L.ptr(keys), L.ptr(vals), L.ptr(ko), L.ptr(vo),
                                    keys.numel(), L.stream_handle())
    L.check(rc, "cuda_bitonic_sort_test")
    return ko, vo


def device_count() -> int:
    lib = L.load()
    c = ctypes.c_int(0)
    L.check(lib.cuda_get_device_count(ctypes.byref(c)), "cuda_get_device_count")
    return c.value


FV_SEARCH_F64_BOUNDARY = 1
FV_SEARCH_FULL_RERANK = 2
# valid_count value: compact the mask with its popcount taken on the device.
VALID_COUNT_DEVICE = 0xFFFFFFFF


class FlatIndex:
    """Device-resident flat tables for the matrix-core scan (fv_flat_search).

    Holds the FP32 rows (for the exact rerank), their codes (FP16/BF16 padded
    to a multiple of 32 dims, or FP8 e4m3 with per-row scales padded to a
    multiple of 128), FP32 and FP64 norms, all encoded on the GPU.
    """

    def __init__(self, vectors, dtype="f16", keep_fp32=True, padded_dim=None):
        torch = L._torch()
        lib = L.load()
        self.vectors = vectors.contiguous()
        n, dim = self.vectors.shape
        self.dim, self.n = dim, n
        self.dtype = dtype
        fp8 = dtype == "fp8"
        step = 128 if fp8 else 32
        pd = padded_dim or (dim + step - 1) // step * step
        self.padded_dim = pd
        dev = self.vectors.device
        self.codes = torch.empty((n, pd), dtype=torch.uint8 if fp8 else torch.int16, device=dev)
        self.row_scales = torch.empty(n, dtype=torch.float32, device=dev) if fp8 else None
        self.norms = torch.empty(n, dtype=torch.float32, device=dev)
        self.norms64 = torch.empty(n, dtype=torch.float64, device=dev)
        L.check(lib.fv_encode_flat(L.ptr(self.vectors), n, dim, pd, L.DTYPES[dtype],
                                   L.ptr(self.codes), L.ptr(self.norms), L.ptr(self.norms64),
                                   L.ptr(self.row_scales), L.stream_handle()), "fv_encode_flat")
        self.keep_fp32 = keep_fp32
        # Per-row encoding error of 16-bit codes: tightens the bounded rerank.
        self.enc_err = None
        if not fp8:
            self.enc_err = torch.empty(n, dtype=torch.float32, device=dev)
            L.check(lib.fv_encode_error(L.ptr(self.vectors), L.ptr(self.codes), n, dim, pd,
                                        L.DTYPES[dtype], L.ptr(self.enc_err), L.stream_handle()),
                    "fv_encode_error")
        self.tables = L.FvFlatTables(L.ptr(self.codes), L.ptr(self.row_scales), L.ptr(self.norms),
                                     L.ptr(self.norms64),
                                     L.ptr(self.vectors) if keep_fp32 else None,
                                     n, dim, pd, L.DTYPES[dtype], L.ptr(self.enc_err))
        self._ws = None

    def params(self, top_k, metric="l2", valid_mask=None, valid_count=0, shortlist=0,
               shortlist_factor=0, top_k_per_query=None, f64_boundary=False, tuning=None,
               full_rerank=False, top_k_sum=0):
        """FvFlatSearchParams; `tuning` (an FvFlatTuning, e.g. from autotune)
        supplies the shortlist factor and plan overrides when the caller
        gives no shortlist of its own. `top_k_sum` (an upper bound on the
        sum of top_k_per_query) sizes per-query candidate buffers."""
        n_chunks = bound_div = 0
        if tuning is not None:
            n_chunks, bound_div = int(tuning.n_chunks), int(tuning.bound_div)
            if not shortlist and not shortlist_factor:
                shortlist_factor = int(tuning.shortlist_factor)
        return L.FvFlatSearchParams(L.ptr(valid_mask), L.ptr(top_k_per_query), int(valid_count),
                                    _metric(metric), int(top_k), int(shortlist_factor),
                                    int(shortlist),
                                    (FV_SEARCH_F64_BOUNDARY if f64_boundary else 0)
                                    | (FV_SEARCH_FULL_RERANK if full_rerank else 0),
                                    n_chunks, bound_div, int(top_k_sum))

    def autotune(self, queries, top_k, metric="l2", valid_mask=None, valid_count=0,
                 target_recall=0.999, trials=5, f64_boundary=False, shortlist_factor=0):
        """fv_flat_autotune on a query sample; returns an FvFlatTuning. A
        non-zero shortlist_factor is kept and only the plan is tuned."""
        lib = L.load()
        base = self.params(top_k, metric, valid_mask, valid_count,
                           shortlist_factor=shortlist_factor, f64_boundary=f64_boundary)
        best = L.FvFlatTuning()
        rc = lib.fv_flat_autotune(ctypes.byref(self.tables), L.ptr(queries), queries.shape[0],
                                  ctypes.byref(base), ctypes.c_float(target_recall), int(trials),
                                  ctypes.byref(best), L.stream_handle())
        L.check(rc, "fv_flat_autotune")
        return best

    def workspace_size(self, nq, params):
        return L.load().fv_flat_workspace_size(ctypes.byref(self.tables), nq, ctypes.byref(params))

    def search(self, queries, top_k, metric="l2", valid_mask=None, shortlist=0, out=None,
               valid_count=0, shortlist_factor=0, top_k_per_query=None, f64_boundary=False,
               tuning=None, full_rerank=False, top_k_sum=0):
        """fv_flat_search. `valid_count` (set bits of the mask, if known) lets
        sparse masks be compacted; `top_k_per_query` is an optional int32
        device tensor of per-query k (each <= top_k), `top_k_sum` an upper
        bound on its sum when the caller knows one (per-query candidate
        buffers; queries past an understated sum return no rows);
        `f64_boundary` sets FV_SEARCH_F64_BOUNDARY, `full_rerank`
        FV_SEARCH_FULL_RERANK."""
        torch = L._torch()
        lib = L.load()
        nq = queries.shape[0]
        prm = self.params(top_k, metric, valid_mask, valid_count, shortlist, shortlist_factor,
                          top_k_per_query, f64_boundary, tuning, full_rerank, top_k_sum)
        need = self.workspace_size(nq, prm)
        if need == 0:
            raise L.FvError("invalid flat search shape")
        if self._ws is None or self._ws.numel() < need:
            self._ws = torch.empty(need, dtype=torch.uint8, device=queries.device)
        out = L.result_tensor(nq, top_k, queries.device) if out is None else out
        rc = lib.fv_flat_search(ctypes.byref(self.tables), L.ptr(queries), nq, ctypes.byref(prm),
                                L.ptr(out), L.ptr(self._ws), self._ws.numel(), L.stream_handle())
        L.check(rc, "fv_flat_search")
        return out


def fp16_tc_search_batch(codes_fp16, norms, queries, top_k, dim, metric="l2", valid_mask=None,
                         vectors=None):
    """cuda_fp16_tc_search_batch over Fp16GpuTables built from caller tensors
    (codes: int16/uint16 [n, padded_dim], norms: float32 [n])."""
    lib = L.load()
    n, pd = codes_fp16.shape
    t = L.Fp16GpuTables(L.ptr(codes_fp16), L.ptr(norms), L.ptr(vectors), n, dim, pd)
    nq = queries.shape[0]
    width = max(1, min(int(top_k), 256))
    out = L.result_tensor(nq, width, queries.device)
    rc = lib.cuda_fp16_tc_search_batch(ctypes.byref(t), L.ptr(queries), nq, top_k,
                                       _metric(metric), L.ptr(valid_mask), L.ptr(out),
                                       L.stream_handle())
    L.check(rc, "cuda_fp16_tc_search_batch")
    return out


class IvfPqIndex:
    """Device-resident IVF-PQ tables (IvfPqGpuTables) plus a list-ordered copy
    of the codes for coalesced streaming (fv_ivf_pq_list_order_codes)."""

    def __init__(self, centroids, codebooks, codes, offsets, vec_ids, vectors=None,
                 list_order=True):
        torch = L._torch()
        lib = L.load()
        self.centroids = centroids.contiguous()
        self.codebooks = codebooks.contiguous()
        self.codes = codes.contiguous()
        self.offsets = offsets.contiguous()
        self.vec_ids = vec_ids.contiguous()
        self.vectors = None if vectors is None else vectors.contiguous()
        nlist, dim = self.centroids.shape
        n, m_pq = self.codes.shape
        self.tables = L.IvfPqGpuTables(
            L.ptr(self.centroids), L.ptr(self.codebooks), L.ptr(self.codes), L.ptr(self.offsets),
            L.ptr(self.vec_ids), L.ptr(self.vectors), n, dim, nlist, m_pq, dim // m_pq)
        self.codes_lo = None
        if list_order:
            self.codes_lo = torch.empty_like(self.codes)
            L.check(lib.fv_ivf_pq_list_order_codes(ctypes.byref(self.tables), L.ptr(self.codes_lo),
                                                   L.stream_handle()),
                    "fv_ivf_pq_list_order_codes")
        self._ws = None

    @classmethod
    def from_host(cls, t, device="cuda", keep_vectors=True, list_order=True):
        """From flashvector_amd.ivfpq_build.IvfPqTables (numpy)."""
        torch = L._torch()

        def d(a):
            return torch.from_numpy(a).to(device)
        return cls(d(t.centroids), d(t.codebooks.reshape(t.m_pq, -1)), d(t.codes),
                   d(t.offsets.view("int32")), d(t.vec_ids.view("int32")),
                   d(t.vectors) if keep_vectors else None, list_order=list_order)

    def search_batch(self, queries, top_k, nprobe, rerank_k, metric="l2", valid_mask=None):
        """cuda_ivf_pq_search_batch (kernels/src/ivf_pq_lookup.cu:450-473)."""
        lib = L.load()
        nq = queries.shape[0]
        width = max(1, min(int(top_k), 256))
        out = L.result_tensor(nq, width, queries.device)
        rc = lib.cuda_ivf_pq_search_batch(ctypes.byref(self.tables), L.ptr(queries), nq, top_k,
                                          nprobe, rerank_k, _metric(metric), L.ptr(valid_mask),
                                          L.ptr(out), L.stream_handle())
        L.check(rc, "cuda_ivf_pq_search_batch")
        return out

    def search(self, queries, max_top_k, max_nprobe, shortlist, top_k=None, nprobe=None,
               rerank=True, metric="l2", valid_mask=None, use_list_order=True, out=None,
               tuning=None):
        """fv_ivf_pq_search: optional per-query top_k / nprobe (uint32 device
        tensors); `tuning` (an FvIvfPqTuning) overrides the work-item plan."""
        torch = L._torch()
        lib = L.load()
        nq = queries.shape[0]
        need = lib.fv_ivf_pq_workspace_size(ctypes.byref(self.tables), nq, max_nprobe, shortlist,
                                            max_top_k)
        if need == 0:
            raise L.FvError("invalid IVF-PQ search shape")
        if self._ws is None or self._ws.numel() < need:
            self._ws = torch.empty(need, dtype=torch.uint8, device=queries.device)
        out = L.result_tensor(nq, max_top_k, queries.device) if out is None else out
        lo = self.codes_lo if use_list_order else None
        rc = lib.fv_ivf_pq_search_tuned(ctypes.byref(self.tables), L.ptr(lo), L.ptr(queries), nq,
                                        L.ptr(top_k), max_top_k, L.ptr(nprobe), max_nprobe,
                                        shortlist, int(rerank), _metric(metric),
                                        L.ptr(valid_mask), L.ptr(out), L.ptr(self._ws),
                                        self._ws.numel(),
                                        None if tuning is None else ctypes.byref(tuning),
                                        L.stream_handle())
        L.check(rc, "fv_ivf_pq_search_tuned")
        return out

    def autotune(self, queries, max_top_k, max_nprobe, shortlist, top_k=None, nprobe=None,
                 rerank=True, metric="l2", valid_mask=None, use_list_order=True, trials=3):
        """fv_ivf_pq_autotune for this batch shape; returns an FvIvfPqTuning."""
        lib = L.load()
        best = L.FvIvfPqTuning()
        lo = self.codes_lo if use_list_order else None
        rc = lib.fv_ivf_pq_autotune(ctypes.byref(self.tables), L.ptr(lo), L.ptr(queries),
                                    queries.shape[0], L.ptr(top_k), max_top_k, L.ptr(nprobe),
                                    max_nprobe, shortlist, int(rerank), _metric(metric),
                                    L.ptr(valid_mask), trials, ctypes.byref(best),
                                    L.stream_handle())
        L.check(rc, "fv_ivf_pq_autotune")
        return best
