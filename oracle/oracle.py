"""ctypes binding of the CPU oracle (libfv_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline. The
product package never imports this module.

Every function takes and returns numpy arrays; result rows come back as two
arrays (ids uint32, distances float32) of shape (num_queries, row_width).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libfv_oracle.so")

RESULT_DTYPE = np.dtype([("id", "<u4"), ("distance", "<f4")], align=True)
METRICS = {"l2": 0, "cosine": 1, "ip": 2}


class IvfPqTablesC(ctypes.Structure):
    """Host-pointer mirror of IvfPqGpuTables (kernels/include/types.h:85-97)."""

    _fields_ = [
        ("d_centroids", ctypes.c_void_p),
        ("d_pq_codebooks", ctypes.c_void_p),
        ("d_pq_codes", ctypes.c_void_p),
        ("d_ivf_offsets", ctypes.c_void_p),
        ("d_ivf_vec_ids", ctypes.c_void_p),
        ("d_vectors", ctypes.c_void_p),
        ("num_vectors", ctypes.c_uint32),
        ("dim", ctypes.c_uint32),
        ("nlist", ctypes.c_uint32),
        ("m_pq", ctypes.c_uint32),
        ("sub_dim", ctypes.c_uint32),
    ]


def build(force: bool = False) -> str:
    """Compiles the oracle with its Makefile (gcc) and returns the .so path."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "-s"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P, U, I = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int
        L.fvo_distance.argtypes = [P, P, U, I]
        L.fvo_distance.restype = ctypes.c_float
        L.fvo_exact_knn.argtypes = [P, U, U, P, U, P]
        L.fvo_exact_knn.restype = U
        L.fvo_flat_scan_batch.argtypes = [P, P, U, U, P, U, U, I, P, P, I]
        L.fvo_flat_topk.argtypes = [P, U, U, P, U, U, I, P, P, I]
        L.fvo_coarse_probe.argtypes = [P, P, U, U, U, P, P]
        L.fvo_pq_lut.argtypes = [P, P, P, U, U, P]
        L.fvo_ivf_pq_search_batch.argtypes = [P, P, U, U, U, U, I, P, P, I]
        L.fvo_ivf_pq_search_uneven.argtypes = [P, P, U, P, P, U, I, I, P, P, U, I]
        L.fvo_merge_topk.argtypes = [P, U, U, U, U, P]
        L.fvo_f32_to_f16_bits.argtypes = [ctypes.c_float]
        L.fvo_f32_to_f16_bits.restype = ctypes.c_uint16
        L.fvo_f16_bits_to_f32.argtypes = [ctypes.c_uint16]
        L.fvo_f16_bits_to_f32.restype = ctypes.c_float
        L.fvo_encode_fp16_padded.argtypes = [P, U, U, P, P]
        L.fvo_encode_fp16_padded.restype = U
        L.fvo_distance_matrix.argtypes = [P, P, U, U, U, I, P]
        L.fvo_f32_to_e4m3.argtypes = [ctypes.c_float]
        L.fvo_f32_to_e4m3.restype = ctypes.c_uint8
        L.fvo_e4m3_to_f32.argtypes = [ctypes.c_uint8]
        L.fvo_e4m3_to_f32.restype = ctypes.c_float
        L.fvo_encode_fp8_rows.argtypes = [P, U, U, U, ctypes.c_float, P, P]
        L.fvo_flat_topk_f64.argtypes = [P, U, U, P, U, U, I, P, P, P, I]
        L.fvo_kmeans_assign.argtypes = [P, U, U, U, P, U, P]
        L.fvo_kmeans_fit.argtypes = [P, U, U, U, U, U, ctypes.c_uint64, P, P, P, P]
        L.fvo_ivf_pq_build.argtypes = [P, U, U, U, U, U, U, ctypes.c_uint64, P, P, P, P, P, P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _u32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.uint32)


def _split(res):
    return res["id"].copy(), res["distance"].copy()


def distance(q, x, metric="l2") -> float:
    q, x = _f32(q), _f32(x)
    return float(lib().fvo_distance(_p(q), _p(x), q.shape[-1], METRICS[metric]))


def distance_matrix(queries, dataset, metric="l2"):
    q, x = _f32(queries), _f32(dataset)
    out = np.empty((q.shape[0], x.shape[0]), np.float32)
    lib().fvo_distance_matrix(_p(q), _p(x), q.shape[0], x.shape[0], q.shape[1],
                              METRICS[metric], _p(out))
    return out


def exact_knn(dataset, query, top_k):
    x, q = _f32(dataset), _f32(query)
    out = np.zeros(top_k, RESULT_DTYPE)
    n = lib().fvo_exact_knn(_p(x), x.shape[0], x.shape[1], _p(q), top_k, _p(out))
    return _split(out[:n])


def flat_scan_batch(vectors, queries, top_k, metric="l2", valid_mask=None, vec_ids=None,
                    nthreads=0):
    """cuda_flat_scan_batch semantics; row width is min(top_k, 256)."""
    x, q = _f32(vectors), _f32(queries)
    width = min(top_k, 256)
    out = np.zeros((q.shape[0], max(width, 1)), RESULT_DTYPE)
    vids, m = _u32(vec_ids), _u32(valid_mask)
    rc = lib().fvo_flat_scan_batch(_p(x), _p(vids), x.shape[0], q.shape[1], _p(q), q.shape[0],
                                   top_k, METRICS[metric], _p(m), _p(out), nthreads)
    assert rc == 0
    return _split(out[:, :width])


def flat_topk(vectors, queries, top_k, metric="l2", valid_mask=None, nthreads=0):
    """Exact FP32 top-k by (distance, id) with no 256 clamp."""
    x, q = _f32(vectors), _f32(queries)
    out = np.zeros((q.shape[0], top_k), RESULT_DTYPE)
    m = _u32(valid_mask)
    rc = lib().fvo_flat_topk(_p(x), x.shape[0], x.shape[1], _p(q), q.shape[0], top_k,
                             METRICS[metric], _p(m), _p(out), nthreads)
    assert rc == 0
    return _split(out)


class IvfPqHost:
    """Keeps the host arrays alive behind an IvfPqTablesC."""

    def __init__(self, centroids, codebooks, codes, offsets, vec_ids, vectors, m_pq):
        self.centroids = _f32(centroids)
        self.codebooks = _f32(codebooks)
        self.codes = np.ascontiguousarray(codes, dtype=np.uint8)
        self.offsets = _u32(offsets)
        self.vec_ids = _u32(vec_ids)
        self.vectors = None if vectors is None else _f32(vectors)
        dim = self.centroids.shape[1]
        self.c = IvfPqTablesC(
            _p(self.centroids), _p(self.codebooks), _p(self.codes), _p(self.offsets),
            _p(self.vec_ids), _p(self.vectors), self.codes.shape[0], dim,
            self.centroids.shape[0], m_pq, dim // m_pq)


def coarse_probe(query, centroids, nprobe):
    q, c = _f32(query), _f32(centroids)
    probed = np.zeros(nprobe, np.uint32)
    dist = np.zeros(nprobe, np.float32)
    lib().fvo_coarse_probe(_p(q), _p(c), c.shape[1], c.shape[0], nprobe, _p(probed), _p(dist))
    return probed, dist


def pq_lut(query, centroid, codebooks, m_pq):
    q, c, b = _f32(query), _f32(centroid), _f32(codebooks)
    sub = q.shape[0] // m_pq
    lut = np.zeros((m_pq, 256), np.float32)
    lib().fvo_pq_lut(_p(q), _p(c), _p(b), m_pq, sub, _p(lut))
    return lut


def ivf_pq_search_batch(tables: IvfPqHost, queries, top_k, nprobe, rerank_k, metric="l2",
                        valid_mask=None, nthreads=0):
    q = _f32(queries)
    width = min(top_k, 256)
    out = np.zeros((q.shape[0], max(width, 1)), RESULT_DTYPE)
    m = _u32(valid_mask)
    rc = lib().fvo_ivf_pq_search_batch(ctypes.byref(tables.c), _p(q), q.shape[0], top_k,
                                       nprobe, rerank_k, METRICS[metric], _p(m), _p(out),
                                       nthreads)
    assert rc == 0
    return _split(out[:, :width])


def ivf_pq_search_uneven(tables: IvfPqHost, queries, k_per_query, nprobe_per_query,
                         shortlist, rerank=True, metric="l2", valid_mask=None,
                         out_stride=None, nthreads=0):
    q = _f32(queries)
    k = _u32(k_per_query)
    npb = _u32(nprobe_per_query)
    stride = int(out_stride or k.max())
    out = np.zeros((q.shape[0], stride), RESULT_DTYPE)
    m = _u32(valid_mask)
    rc = lib().fvo_ivf_pq_search_uneven(ctypes.byref(tables.c), _p(q), q.shape[0], _p(k),
                                        _p(npb), shortlist, int(rerank), METRICS[metric],
                                        _p(m), _p(out), stride, nthreads)
    assert rc == 0
    return _split(out)


def merge_topk(ids, dists, out_top_k):
    """ids/dists: (num_shards, num_queries, shard_top_k)."""
    s, nq, k = ids.shape
    inp = np.zeros((s, nq, k), RESULT_DTYPE)
    inp["id"], inp["distance"] = ids, dists
    width = min(out_top_k, 256)
    out = np.zeros((nq, max(width, 1)), RESULT_DTYPE)
    rc = lib().fvo_merge_topk(_p(inp), s, nq, k, out_top_k, _p(out))
    assert rc == 0
    return _split(out[:, :width])


def f32_to_f16_bits(x: float) -> int:
    return int(lib().fvo_f32_to_f16_bits(float(x)))


def f16_bits_to_f32(h: int) -> float:
    return float(lib().fvo_f16_bits_to_f32(int(h)))


def encode_fp16_padded(vectors):
    x = _f32(vectors)
    n, d = x.shape
    pd = (d + 15) // 16 * 16
    out = np.zeros((n, pd), np.uint16)
    norms = np.zeros(n, np.float32)
    lib().fvo_encode_fp16_padded(_p(x), n, d, _p(out), _p(norms))
    return out, norms


def f32_to_e4m3(x: float) -> int:
    return int(lib().fvo_f32_to_e4m3(float(x)))


def e4m3_to_f32(c: int) -> float:
    return float(lib().fvo_e4m3_to_f32(int(c)))


def e4m3_table() -> np.ndarray:
    """Decoded value of each of the 256 e4m3 codes (NaN for 0x7f/0xff)."""
    return np.array([e4m3_to_f32(c) for c in range(256)], np.float32)


def encode_fp8_rows(x, pdim=None, sign=1.0):
    """FP8 rows with per-row scales (fvo_encode_fp8_rows): (codes u8, scales f32)."""
    x = _f32(x)
    n, d = x.shape
    pdim = d if pdim is None else pdim
    codes = np.zeros((n, pdim), np.uint8)
    scales = np.zeros(n, np.float32)
    lib().fvo_encode_fp8_rows(_p(x), n, d, pdim, float(sign), _p(codes), _p(scales))
    return codes, scales


def fp8_scores(db, queries, metric="l2"):
    """Scores of the FP8 scan without rerank: c_row + (s_row * s_q) * acc,
    acc = the exact sum of code products (exact in FP64) rounded to FP32;
    c_row = 0.5|x|^2 (L2) else 0; cosine multiplies s_row by 1/|x|. Matches
    the device bit for bit whenever acc is exact in FP32."""
    t = e4m3_table().astype(np.float64)
    xc, xs = encode_fp8_rows(db)
    qc, qs = encode_fp8_rows(queries, sign=-1.0)
    acc = (t[qc] @ t[xc].T).astype(np.float32)
    x32 = _f32(db)
    norms = np.array([distance(np.zeros_like(r), r, "l2") for r in x32], np.float32)
    sx = xs.copy()
    if metric == "cosine":
        inv = np.where(norms > 0, (np.float32(1.0) / np.sqrt(norms)).astype(np.float32), np.float32(0))
        sx = (inv * xs).astype(np.float32)
    c = (np.float32(0.5) * norms) if metric == "l2" else np.zeros_like(norms)
    scale = (sx[None, :] * qs[:, None]).astype(np.float32)
    return (c[None, :] + acc * scale).astype(np.float32)


def flat_topk_f64(vectors, queries, top_k, metric="l2", valid_mask=None, nthreads=0):
    """Exact top-k by FP64 distance, ties to the lower id: (ids u32, dist f64)."""
    v, q = _f32(vectors), _f32(queries)
    n, d = v.shape
    nq = q.shape[0]
    ids = np.zeros((nq, top_k), np.uint32)
    dist = np.zeros((nq, top_k), np.float64)
    m = _u32(valid_mask)
    rc = lib().fvo_flat_topk_f64(_p(v), n, d, _p(q), nq, top_k, METRICS[metric], _p(m),
                                 _p(ids), _p(dist), nthreads)
    assert rc == 0
    return ids, dist


def kmeans_assign(x, centroids, dim=None, ld=None):
    """Nearest centroid of every row (first strict minimum). x is [n, ld];
    the first `dim` columns are the row."""
    x, c = _f32(x), _f32(centroids)
    n, ld_ = x.shape
    dim = dim or c.shape[1]
    out = np.empty(n, np.uint32)
    lib().fvo_kmeans_assign(_p(x), n, dim, ld or ld_, _p(c), c.shape[0], _p(out))
    return out


def kmeans_fit(x, k, max_iters, seed, dim=None):
    """(centroids [k, dim], assignments [n], sizes [k], iterations run)."""
    x = _f32(x)
    n, ld = x.shape
    dim = dim or ld
    cent = np.empty((k, dim), np.float32)
    assign = np.empty(n, np.uint32)
    sizes = np.empty(k, np.uint32)
    it = np.zeros(1, np.uint32)
    lib().fvo_kmeans_fit(_p(x), n, dim, ld, k, max_iters, seed, _p(cent), _p(assign),
                         _p(sizes), _p(it))
    return cent, assign, sizes, int(it[0])


def ivf_pq_build(x, nlist, m_pq, coarse_iters, pq_iters, seed):
    """dict of centroids, codebooks [m, 256*sub], codes [n, m], offsets, vec_ids, iters."""
    x = _f32(x)
    n, dim = x.shape
    nl = max(1, min(nlist, n))
    sub = dim // m_pq
    out = dict(centroids=np.empty((nl, dim), np.float32),
               codebooks=np.empty((m_pq, 256 * sub), np.float32),
               codes=np.empty((n, m_pq), np.uint8), offsets=np.empty(nl + 1, np.uint32),
               vec_ids=np.empty(n, np.uint32), iters=np.zeros(1 + m_pq, np.uint32))
    rc = lib().fvo_ivf_pq_build(_p(x), n, dim, nlist, m_pq, coarse_iters, pq_iters, seed,
                                _p(out["centroids"]), _p(out["codebooks"]), _p(out["codes"]),
                                _p(out["offsets"]), _p(out["vec_ids"]), _p(out["iters"]))
    if rc != 0:
        raise ValueError(f"fvo_ivf_pq_build: {rc}")
    return out
