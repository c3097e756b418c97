"""Python binding of the host search engine (include/flashvector/fv_host.h).

Counterpart of the reference's PyO3 class FlashVectorGPU
(crates/python/src/py_index.rs:361-436) and of its DLPack input path
(crates/python/src/dlpack.rs:15-121): numpy arrays and CPU tensors go through
the host path (pinned staging, one pooled stream); ROCm tensors are passed by
data pointer with no host copy and run on torch's current stream. Masks are
a bool array over logical ids or packed LSB-first u32 words
(py_index.rs:73-111); results are (ids uint32 [Q, k], distances float32
[Q, k]) like results_to_pair (py_index.rs:56-71).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import lib as L
from .synthetic import pack_bits

METRICS = {"l2": 0, "cosine": 1, "ip": 2}
DTYPES = {"f16": 1, "bf16": 2, "fp8": 3}
RESULT = np.dtype([("id", "<u4"), ("distance", "<f4")], align=True)


def _is_cuda(x) -> bool:
    return hasattr(x, "is_cuda") and bool(x.is_cuda)


def _host_f32(x) -> np.ndarray:
    if hasattr(x, "detach"):
        x = x.detach().cpu().numpy()
    return np.ascontiguousarray(x, dtype=np.float32)


def _host_u32(x) -> np.ndarray | None:
    if x is None:
        return None
    if hasattr(x, "detach"):
        x = x.detach().cpu().numpy()
    return np.ascontiguousarray(np.asarray(x).astype(np.uint32, copy=False))


def _mask_words(mask):
    """bool[n] over logical ids or packed u32 words -> u32 words (host or device)."""
    if mask is None:
        return None
    if _is_cuda(mask):
        torch = L._torch()
        if mask.dtype == torch.bool:
            raise L.FvError("device masks must be packed u32 words")
        return mask.contiguous()
    m = np.asarray(mask.detach().cpu().numpy() if hasattr(mask, "detach") else mask)
    if m.dtype == bool:
        return pack_bits(m)
    return np.ascontiguousarray(m.astype(np.uint32, copy=False))


class Engine:
    def __init__(self, dim, metric="l2", num_streams=4, max_batch=1024, max_top_k=256,
                 shortlist_factor=8, f64_boundary=False, ivf_nprobe=32, ivf_shortlist=1000):
        self._lib = L.load()
        cfg = L.FvEngineConfig(dim, METRICS[metric], num_streams, max_batch, max_top_k,
                               shortlist_factor, 1 if f64_boundary else 0, ivf_nprobe,
                               ivf_shortlist, 0)
        h = ctypes.c_void_p()
        L.check(self._lib.fvh_create(ctypes.byref(cfg), ctypes.byref(h)), "fvh_create")
        self._h = h
        self.dim = dim

    def close(self):
        if self._h:
            self._lib.fvh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load_flat(self, vectors, dtype="f16", ids=None):
        v = _host_f32(vectors)
        i = _host_u32(ids)
        L.check(self._lib.fvh_load_flat(self._h, v.ctypes.data, v.shape[0], DTYPES[dtype],
                                        None if i is None else i.ctypes.data), "fvh_load_flat")

    def load_ivfpq(self, centroids, codebooks, codes, offsets, vec_ids, vectors, ids=None):
        c, b, v = _host_f32(centroids), _host_f32(codebooks), _host_f32(vectors)
        cd = np.ascontiguousarray(np.asarray(codes), dtype=np.uint8)
        off, vid, i = _host_u32(offsets), _host_u32(vec_ids), _host_u32(ids)
        L.check(self._lib.fvh_load_ivfpq(self._h, c.ctypes.data, c.shape[0], b.ctypes.data,
                                         cd.shape[1], cd.ctypes.data, off.ctypes.data,
                                         vid.ctypes.data, v.ctypes.data, v.shape[0],
                                         None if i is None else i.ctypes.data), "fvh_load_ivfpq")

    def add(self, vectors, ids):
        v, i = _host_f32(vectors), _host_u32(ids)
        L.check(self._lib.fvh_add_delta(self._h, v.ctypes.data, i.ctypes.data, v.shape[0]),
                "fvh_add_delta")

    def add_vectors(self, vectors):
        """Appends rows under fresh logical ids (index.rs:647-694 add_vectors);
        returns the assigned ids (uint32 [n])."""
        v = _host_f32(vectors)
        out = np.zeros(v.shape[0], np.uint32)
        L.check(self._lib.fvh_add_vectors(self._h, v.ctypes.data, v.shape[0], out.ctypes.data),
                "fvh_add_vectors")
        return out

    @property
    def next_logical_id(self):
        n = ctypes.c_uint32(0)
        L.check(self._lib.fvh_next_logical_id(self._h, ctypes.byref(n)), "fvh_next_logical_id")
        return n.value

    def delete(self, ids):
        i = _host_u32(ids)
        L.check(self._lib.fvh_delete(self._h, i.ctypes.data, i.shape[0]), "fvh_delete")

    def search(self, queries, k, mask=None):
        """(ids uint32 [Q, k], distances float32 [Q, k]) as numpy arrays."""
        words = _mask_words(mask)
        if _is_cuda(queries):
            out = self.search_device(queries, k, words)
            return L.split_results(out)
        q = _host_f32(queries)
        nq = q.shape[0]
        res = np.zeros((nq, k), RESULT)
        if words is not None and _is_cuda(words):
            words = words.cpu().numpy()
        L.check(self._lib.fvh_search(self._h, q.ctypes.data, nq, k,
                                     None if words is None else words.ctypes.data,
                                     0 if words is None else words.shape[0], res.ctypes.data),
                "fvh_search")
        return res["id"].copy(), res["distance"].copy()

    def search_device(self, queries, k, mask_words=None, out=None):
        """Device queries (zero copy) -> int64 [Q, k] GpuSearchResult tensor on
        the device, on torch's current stream."""
        torch = L._torch()
        q = queries.detach().contiguous().float()
        if mask_words is not None and not _is_cuda(mask_words):
            mask_words = torch.from_numpy(np.asarray(mask_words, dtype=np.uint32).view(np.int32)).to(q.device)
        out = L.result_tensor(q.shape[0], k, q.device) if out is None else out
        L.check(self._lib.fvh_search_device(self._h, L.ptr(q), q.shape[0], k, L.ptr(mask_words),
                                            0 if mask_words is None else mask_words.numel(),
                                            L.ptr(out), L.stream_handle()), "fvh_search_device")
        return out

    def save_snapshot(self, path):
        """FLSHVEC1 file of this engine's tables, ids, tombstones and delta
        (GpuVectorIndex::save_snapshot, index.rs:845-1031)."""
        L.check(self._lib.fvh_save_snapshot(self._h, str(path).encode()), "fvh_save_snapshot")

    def load_snapshot(self, path, dtype=None):
        """Loads a FLSHVEC1 file (index.rs:1033-1258): its IVF-PQ tables when
        dtype is None, else a flat index of `dtype` encoded from its rows."""
        L.check(self._lib.fvh_load_snapshot(self._h, str(path).encode(),
                                            0 if dtype is None else DTYPES[dtype]),
                "fvh_load_snapshot")

    def stats(self):
        n = ctypes.c_uint64(0)
        ms = ctypes.c_double(0.0)
        L.check(self._lib.fvh_stats(self._h, ctypes.byref(n), ctypes.byref(ms)), "fvh_stats")
        return {"queries": n.value, "total_ms": ms.value}
