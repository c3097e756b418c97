"""`FlashVectorGPU`: the Python class of the reference's binding
(crates/python/src/py_index.rs:155-436), over the host engine of
include/flashvector/fv_host.h, for the paths this library serves.

Same constructor, argument names, defaults and return values as the
reference: `build`, `add`, `delete`, `save`, `load`, `search_ivf(query,
top_k=10, nprobe=8, rerank_k=32, mode="pq", filter_mask=None)` and
`search_fp16(query, top_k=10, filter_mask=None)` return
`(labels uint32 [Q, k], distances float32 [Q, k])` numpy arrays
(results_to_pair, py_index.rs:56-71); masks are a 1-D bool array over logical
ids or packed LSB-first uint32 words (mask_from_pyany, py_index.rs:83-111);
inputs are numpy arrays, CPU tensors or ROCm tensors (dlpack.rs:15-121).

`build` builds, on the GPU, the two kinds of tables these searches use: FP16
rows for the matrix-core flat scan and IVF-PQ tables (nlist lists, m_pq
sub-quantizers, 25 + 25 Lloyd iterations as index.rs:405-458). nprobe,
rerank_k and k are per call. With metric "cosine", rows and queries are
normalized as the reference's host does (index.rs:29-37, 1473-1483).

Differences: `search_fp16` reranks a shortlist of 8k rows exactly in FP32
(the reference returned FP16 scores); `search` (the reference's HNSW
traversal, out of scope here) is served by the same exact matrix-core scan,
so its results are at least as good, and `ef_search` is ignored;
`search_ivf` modes "sq8" and "rabitq" (codecs out of scope) raise
NotImplementedError.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import lib as L
from .engine import METRICS, RESULT, Engine, _is_cuda, _mask_words

_METRIC_NAMES = {"l2": "l2", "euclidean": "l2", "cosine": "cosine", "ip": "ip",
                 "inner_product": "ip", "dot": "ip"}
_IVF_MODES = {"pq": "pq", "ivf_pq": "pq", "ivf-pq": "pq", "sq8": "sq8", "sq": "sq8",
              "ivf_sq8": "sq8", "ivf-sq8": "sq8", "rabitq": "rabitq", "ivf_rabitq": "rabitq",
              "ivf-rabitq": "rabitq"}


def normalize_rows(x: np.ndarray) -> np.ndarray:
    """index.rs:29-37 normalize_slice_in_place: per row, norm = sqrt of the
    sequential float32 sum of x*x (at least 1e-8), then x * (1 / norm)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    if x.size == 0:
        return x.copy()
    ss = np.cumsum(x * x, axis=1, dtype=np.float32)[:, -1]
    norm = np.maximum(np.sqrt(ss), np.float32(1e-8)).astype(np.float32)
    inv = (np.float32(1.0) / norm).astype(np.float32)
    return (x * inv[:, None]).astype(np.float32)


def _host_rows(data, dim) -> np.ndarray:
    """TensorView::from_pyany + copy_view_to_host: a contiguous float32
    [n, dim] host array (ROCm tensors are copied to the host, as the
    reference does for build and add)."""
    if hasattr(data, "detach"):
        data = data.detach()
        if data.dtype.is_floating_point is False:
            raise TypeError("expected a float32 tensor")
        data = data.float().cpu().numpy()
    a = np.asarray(data)
    if a.ndim == 1:
        a = a.reshape(1, -1)
    if a.ndim != 2:
        raise ValueError("expected a 2-D [n, dim] array")
    if a.shape[1] != dim:
        raise ValueError(f"Dimension mismatch: expected {dim}, got {a.shape[1]}")
    return np.ascontiguousarray(a, dtype=np.float32)


def _ids(ids) -> np.ndarray:
    """ids_from_pyany (py_index.rs:113-136): uint32, int64 in range, or a list."""
    if hasattr(ids, "detach"):
        ids = ids.detach().cpu().numpy()
    a = np.asarray(ids)
    if a.dtype.kind in "iu":
        if a.size and (a.min() < 0 or a.max() > 0xFFFFFFFF):
            raise ValueError("id out of u32 range")
        return np.ascontiguousarray(a.astype(np.uint32).ravel())
    raise TypeError("ids must be integers")


class FlashVectorGPU:
    def __init__(self, dim, max_elements=100_000, m=32, ef_construction=128, nlist=256,
                 m_pq=16, metric="l2", *, flat_dtype="f16", num_streams=4, max_batch=1024,
                 max_top_k=256, seed=0):
        self.dim = int(dim)
        self.max_elements, self.m, self.ef_construction = max_elements, m, ef_construction
        self.nlist, self.m_pq = int(nlist), int(m_pq)
        # py_index.rs:177-182: unknown metric names fall back to L2.
        self.metric = _METRIC_NAMES.get(str(metric).lower(), "l2")
        self.flat_dtype = flat_dtype
        self.seed = seed
        # Engine defaults are the reference's search_ivf defaults (nprobe 8,
        # rerank_k 32); every call may override them.
        self._eng = Engine(self.dim, self.metric, num_streams=num_streams, max_batch=max_batch,
                           max_top_k=max_top_k, ivf_nprobe=8, ivf_shortlist=32)
        self._lib = L.load()
        self._built = False

    # ---- construction ---------------------------------------------------

    def _prep(self, rows: np.ndarray) -> np.ndarray:
        return normalize_rows(rows) if self.metric == "cosine" else rows

    def build(self, data):
        rows = _host_rows(data, self.dim)
        if rows.shape[0] == 0:
            raise ValueError("Dataset cannot be empty")
        rows = self._prep(rows)
        nlist, m_pq = self.nlist, self.m_pq
        if m_pq == 0 or self.dim % m_pq:
            raise ValueError(f"dim {self.dim} is not a multiple of m_pq {m_pq}")
        L.check(self._lib.fvh_build(self._eng._h, rows.ctypes.data, rows.shape[0],
                                    L.DTYPES[self.flat_dtype], nlist, m_pq, 25, 25, self.seed),
                "fvh_build")
        self._built = True

    def add(self, data):
        """Appends rows to the delta buffer; returns their assigned ids (uint32).
        Before a build, builds the index and returns 0..n-1 (index.rs:658-661)."""
        rows = _host_rows(data, self.dim)
        if rows.shape[0] == 0:
            return np.zeros(0, np.uint32)
        if not self._built:
            self.build(rows)
            return np.arange(rows.shape[0], dtype=np.uint32)
        return self._eng.add_vectors(self._prep(rows))

    def delete(self, ids) -> int:
        """Tombstones logical ids; returns how many were newly deleted."""
        before = self.info()["tombstones"]
        self._eng.delete(_ids(ids))
        return self.info()["tombstones"] - before

    def compact(self):
        """Rebuilds the index from its live rows under their logical ids and
        empties the delta buffer and tombstones (index.rs:716-752)."""
        L.check(self._lib.fvh_compact(self._eng._h), "fvh_compact")

    def set_mapped_rerank(self, enabled: bool):
        """The reference reranks from mapped host memory when enabled; rows stay
        in HBM here (288 GB per GPU), so this only records the setting."""
        self.mapped_rerank = bool(enabled)

    def save(self, path):
        self._eng.save_snapshot(path)

    @staticmethod
    def load(path, flat_dtype="f16", **kw):
        """Restores a FLSHVEC1 snapshot (index.rs:1033-1258): its IVF-PQ tables
        and FP16 rows encoded from its FP32 rows, delta buffer and tombstones."""
        from . import snapshot as SN
        info, _ = SN.read(str(path))
        metric = {0: "l2", 1: "cosine", 2: "ip"}[info["metric"]]
        obj = FlashVectorGPU(info["dim"], info["max_elements"], info["m"],
                             info["ef_construction"], info["nlist"], info["m_pq"], metric,
                             flat_dtype=flat_dtype, **kw)
        L.check(obj._lib.fvh_load_snapshot(obj._eng._h, str(path).encode(),
                                           L.DTYPES[flat_dtype] | 0x100),
                "fvh_load_snapshot")
        obj._built = True
        return obj

    # ---- state ----------------------------------------------------------

    def info(self):
        i = L.FvEngineInfo()
        L.check(self._lib.fvh_info(self._eng._h, ctypes.byref(i)), "fvh_info")
        return {f: getattr(i, f) for f, _ in L.FvEngineInfo._fields_}

    def num_vectors(self):
        return self.info()["num_vectors"]

    def delta_len(self):
        return self.info()["delta_len"]

    def tombstone_count(self):
        return self.info()["tombstones"]

    def next_logical_id(self):
        return self.info()["next_logical_id"]

    # ---- searches -------------------------------------------------------

    def _search(self, query, top_k, opts, filter_mask):
        if not self._built:
            raise RuntimeError("index is not built")
        top_k = int(top_k)
        words = _mask_words(filter_mask)
        if _is_cuda(query):
            torch = L._torch()
            q = query.detach().float()
            if q.dim() == 1:
                q = q.reshape(1, -1)
            if q.shape[1] != self.dim:
                raise ValueError(f"Dimension mismatch: expected {self.dim}, got {q.shape[1]}")
            if q.shape[0] == 0:
                return np.zeros((0, top_k), np.uint32), np.zeros((0, top_k), np.float32)
            if self.metric == "cosine":
                # Device-side row normalization (torch reduction order).
                norm = torch.clamp(torch.sqrt((q * q).sum(1, keepdim=True)), min=1e-8)
                q = q * (1.0 / norm)
            q = q.contiguous()
            if words is not None and not _is_cuda(words):
                words = torch.from_numpy(np.asarray(words, np.uint32).view(np.int32)).to(q.device)
            out = L.result_tensor(q.shape[0], top_k, q.device)
            L.check(self._lib.fvh_search_device_ex(
                self._eng._h, L.ptr(q), q.shape[0], top_k, ctypes.byref(opts), L.ptr(words),
                0 if words is None else words.numel(), L.ptr(out), L.stream_handle()),
                "fvh_search_device_ex")
            return L.split_results(out)
        q = self._prep(_host_rows(query, self.dim))
        nq = q.shape[0]
        if nq == 0:
            return np.zeros((0, top_k), np.uint32), np.zeros((0, top_k), np.float32)
        if words is not None and _is_cuda(words):
            words = words.cpu().numpy()
        res = np.zeros((nq, top_k), RESULT)
        L.check(self._lib.fvh_search_ex(self._eng._h, q.ctypes.data, nq, top_k,
                                        ctypes.byref(opts),
                                        None if words is None else words.ctypes.data,
                                        0 if words is None else words.shape[0], res.ctypes.data),
                "fvh_search_ex")
        return res["id"].copy(), res["distance"].copy()

    def search_ivf(self, query, top_k=10, nprobe=8, rerank_k=32, mode="pq", filter_mask=None):
        m = _IVF_MODES.get(str(mode).lower())
        if m is None:
            raise ValueError(f"Unsupported IVF mode {mode!r}; expected 'pq', 'sq8', or 'rabitq'")
        if m != "pq":
            raise NotImplementedError(f"IVF mode {m!r} is not served by this library")
        opts = L.FvSearchOptions(2, int(nprobe), int(rerank_k), 0)
        return self._search(query, top_k, opts, filter_mask)

    def search_fp16(self, query, top_k=10, filter_mask=None):
        return self._search(query, top_k, L.FvSearchOptions(1, 0, -1, 0), filter_mask)

    def search(self, query, top_k=10, ef_search=64, filter_mask=None):
        """The reference's HNSW search; served by the exact matrix-core flat scan."""
        del ef_search
        return self.search_fp16(query, top_k, filter_mask)


__all__ = ["FlashVectorGPU", "normalize_rows", "METRICS"]
