"""ctypes binding of libflashvector_hip.so (the C ABI in include/flashvector/).

Device buffers are torch tensors on a ROCm device; pointers go through
`tensor.data_ptr()` and the stream is torch's current HIP stream, so calls
order correctly with surrounding torch work. There is no CPU fallback: if the
library is missing or no GPU is visible, the calls raise.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libflashvector_hip.so")
CSRC = os.path.join(os.path.dirname(_HERE), "csrc")

METRICS = {"l2": 0, "cosine": 1, "ip": 2}


class FvError(RuntimeError):
    pass


class IvfPqGpuTables(ctypes.Structure):
    """kernels/include/types.h:85-97 (72 bytes on LP64)."""

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


class Fp16GpuTables(ctypes.Structure):
    """kernels/include/types.h:131-138 (40 bytes on LP64)."""

    _fields_ = [
        ("d_vectors_fp16", ctypes.c_void_p),
        ("d_vec_norms_sq", ctypes.c_void_p),
        ("d_vectors", ctypes.c_void_p),
        ("num_vectors", ctypes.c_uint32),
        ("dim", ctypes.c_uint32),
        ("padded_dim", ctypes.c_uint32),
    ]


class FvFlatTables(ctypes.Structure):
    """include/flashvector/fv_ext.h FvFlatTables."""

    _fields_ = [
        ("d_codes", ctypes.c_void_p),
        ("d_row_scales", ctypes.c_void_p),
        ("d_norms_sq", ctypes.c_void_p),
        ("d_norms_sq_f64", ctypes.c_void_p),
        ("d_vectors", ctypes.c_void_p),
        ("num_vectors", ctypes.c_uint32),
        ("dim", ctypes.c_uint32),
        ("padded_dim", ctypes.c_uint32),
        ("dtype", ctypes.c_uint32),
        ("d_enc_err", ctypes.c_void_p),
    ]


class FvFlatSearchParams(ctypes.Structure):
    """include/flashvector/fv_ext.h FvFlatSearchParams."""

    _fields_ = [
        ("d_valid_mask", ctypes.c_void_p),
        ("d_top_k", ctypes.c_void_p),
        ("valid_count", ctypes.c_uint32),
        ("metric", ctypes.c_uint32),
        ("top_k", ctypes.c_uint32),
        ("shortlist_factor", ctypes.c_uint32),
        ("shortlist", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
        ("n_chunks", ctypes.c_uint32),
        ("bound_div", ctypes.c_uint32),
        ("top_k_sum", ctypes.c_uint32),
    ]


class FvEngineConfig(ctypes.Structure):
    """include/flashvector/fv_host.h FvEngineConfig."""

    _fields_ = [(n, ctypes.c_uint32) for n in (
        "dim", "metric", "num_streams", "max_batch", "max_top_k", "shortlist_factor", "flags",
        "ivf_nprobe", "ivf_shortlist", "reserved")]


class FvSearchOptions(ctypes.Structure):
    """include/flashvector/fv_host.h FvSearchOptions."""

    _fields_ = [("algo", ctypes.c_uint32), ("nprobe", ctypes.c_uint32),
                ("rerank_k", ctypes.c_int32), ("shortlist_factor", ctypes.c_uint32)]


class FvEngineInfo(ctypes.Structure):
    """include/flashvector/fv_host.h FvEngineInfo."""

    _fields_ = [(n, ctypes.c_uint32) for n in (
        "tables", "num_vectors", "delta_len", "tombstones", "next_logical_id", "nlist", "m_pq",
        "flat_dtype")]


class FvIvfPqTuning(ctypes.Structure):
    """include/flashvector/fv_ext.h FvIvfPqTuning."""

    _fields_ = [
        ("target_items", ctypes.c_uint32),
        ("chunk_min", ctypes.c_uint32),
        ("ms", ctypes.c_float),
        ("reserved", ctypes.c_uint32),
    ]


class FvFlatTuning(ctypes.Structure):
    """include/flashvector/fv_ext.h FvFlatTuning."""

    _fields_ = [
        ("shortlist_factor", ctypes.c_uint32),
        ("n_chunks", ctypes.c_uint32),
        ("bound_div", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("ms", ctypes.c_float),
        ("recall", ctypes.c_float),
    ]


class FvSnapshotInfo(ctypes.Structure):
    """include/flashvector/fv_host.h (the SnapshotManifest scalars of
    crates/engine/src/persistence.rs:35-48)."""

    _fields_ = [("version", ctypes.c_uint32), ("dim", ctypes.c_uint32),
                ("max_elements", ctypes.c_uint32), ("m", ctypes.c_uint32),
                ("ef_construction", ctypes.c_uint32), ("nlist", ctypes.c_uint32),
                ("m_pq", ctypes.c_uint32), ("metric", ctypes.c_uint32),
                ("num_vectors", ctypes.c_uint64), ("entry_point", ctypes.c_uint32),
                ("next_logical_id", ctypes.c_uint32), ("cagra_degree", ctypes.c_uint64),
                ("rabitq_padded_dim", ctypes.c_uint64), ("fp16_padded_dim", ctypes.c_uint64),
                ("sparse_num_docs", ctypes.c_uint64), ("sparse_nnz", ctypes.c_uint64),
                ("num_cagra_seeds", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class FvSnapshotSection(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("dtype", ctypes.c_char_p),
                ("len", ctypes.c_uint64), ("data", ctypes.c_void_p)]


DTYPES = {"f32": 0, "f16": 1, "bf16": 2, "fp8": 3}

P, U, I, SZ = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_size_t

# name -> (restype, argtypes)
SIGNATURES = {
    "cuda_init_device": (I, [I]),
    "cuda_set_device": (I, [I]),
    "cuda_get_device_count": (I, [P]),
    "cuda_get_device_memory": (I, [P, P]),
    "cuda_device_synchronize": (I, []),
    "cuda_enable_peer_access": (I, [I, I]),
    "cuda_memcpy_peer_async": (I, [P, I, P, I, SZ, P]),
    "cuda_malloc_device": (I, [P, SZ]),
    "cuda_free_device": (I, [P]),
    "cuda_malloc_host": (I, [P, SZ]),
    "cuda_free_host": (I, [P]),
    "cuda_memcpy_h2d_async": (I, [P, P, SZ, P]),
    "cuda_memcpy_d2h_async": (I, [P, P, SZ, P]),
    "cuda_memset_device_async": (I, [P, I, SZ, P]),
    "cuda_host_get_device_pointer": (I, [P, P]),
    "cuda_create_stream": (I, [P]),
    "cuda_destroy_stream": (I, [P]),
    "cuda_sync_stream": (I, [P]),
    "cuda_flat_scan_batch": (I, [P, P, U, U, P, U, U, I, P, P, P]),
    "cuda_merge_topk": (I, [P, U, U, U, U, P, P]),
    "fv_flat_scan_batch_noalloc": (I, [P, P, U, U, P, U, U, I, P, P, P]),
    "cuda_compute_distances_warp": (I, [P, P, U, U, U, I, P, P]),
    "cuda_bitonic_sort_test": (I, [P, P, P, P, U, P]),
    "fv_flat_exact_workspace_size": (SZ, [U, U, U, U, I]),
    "fv_flat_exact_search": (I, [P, U, U, P, U, U, I, P, P, P, SZ, P]),
    "cuda_fp16_tc_search_batch": (I, [P, P, U, U, I, P, P, P]),
    "fv_encode_flat": (I, [P, U, U, U, U, P, P, P, P, P]),
    "fv_encode_error": (I, [P, P, U, U, U, U, P, P]),
    "fv_flat_workspace_size": (SZ, [P, U, P]),
    "fv_flat_search": (I, [P, P, U, P, P, P, SZ, P]),
    "cuda_ivf_pq_search_batch": (I, [P, P, U, U, U, U, I, P, P, P]),
    "fv_ivf_pq_list_order_codes": (I, [P, P, P]),
    "fv_ivf_pq_workspace_size": (SZ, [P, U, U, U, U]),
    "fv_ivf_pq_search": (I, [P, P, P, U, P, U, P, U, U, I, I, P, P, P, SZ, P]),
    "fv_ivf_pq_search_tuned": (I, [P, P, P, U, P, U, P, U, U, I, I, P, P, P, SZ, P, P]),
    "fv_ivf_pq_autotune": (I, [P, P, P, U, P, U, P, U, U, I, I, P, U, P, P]),
    "fv_flat_autotune": (I, [P, P, U, P, ctypes.c_float, U, P, P]),
    "fvh_create": (I, [P, P]),
    "fvh_destroy": (None, [P]),
    "fvh_load_flat": (I, [P, P, U, U, P]),
    "fvh_load_ivfpq": (I, [P, P, U, P, U, P, P, P, P, U, P]),
    "fvh_add_delta": (I, [P, P, P, U]),
    "fvh_add_vectors": (I, [P, P, U, P]),
    "fvh_delete": (I, [P, P, U]),
    "fvh_next_logical_id": (I, [P, P]),
    "fvh_build": (I, [P, P, U, U, U, U, U, U, ctypes.c_uint64]),
    "fvh_compact": (I, [P]),
    "fvh_search": (I, [P, P, U, U, P, U, P]),
    "fvh_search_ex": (I, [P, P, U, U, P, P, U, P]),
    "fvh_search_device": (I, [P, P, U, U, P, U, P, P]),
    "fvh_search_device_ex": (I, [P, P, U, U, P, P, U, P, P]),
    "fvh_info": (I, [P, P]),
    "fvh_stats": (I, [P, P, P]),
    "fvh_snapshot_write": (I, [ctypes.c_char_p, P, P, U]),
    "fvh_snapshot_open": (I, [ctypes.c_char_p, P, P]),
    "fvh_snapshot_section": (I, [P, ctypes.c_char_p, P, P, P, P]),
    "fvh_snapshot_close": (None, [P]),
    "fvh_save_snapshot": (I, [P, ctypes.c_char_p]),
    "fvh_load_snapshot": (I, [P, ctypes.c_char_p, U]),
    "fvh_batcher_create": (I, [P, U, U, P]),
    "fvh_batcher_create_with_executor": (I, [P, P, U, U, U, P]),
    "fvh_batcher_submit": (I, [P, P, U, P, P]),
    "fvh_batcher_stats": (I, [P, P, P, P]),
    "fvh_batcher_destroy": (None, [P]),
    "fv_topk_ip_backward": (I, [P, U, P, U, U, P, U, P, P, ctypes.c_float, ctypes.c_float, P, P,
                                P, P, P]),
    "fv_kmeans_assign": (I, [P, U, U, U, P, U, P, P, U, P]),
    "fv_kmeans_workspace_size": (SZ, [U, U, U]),
    "fv_kmeans_fit": (I, [P, U, U, U, U, U, ctypes.c_uint64, P, P, P, P, P, P, P, SZ, P]),
    "fv_ivf_pq_build_workspace_size": (SZ, [U, U, U, U]),
    "fv_ivf_pq_build": (I, [P, U, U, U, U, U, U, ctypes.c_uint64, P, P, P, P, P, P, P, SZ, P]),
    "fv_kernel_timer_enable": (I, [I]),
    "fv_kernel_timer_read": (I, [P, P, I]),
}

_lib = None


def load():
    """Loads the library (once). Raises FvError if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FvError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          f"or `make -C {CSRC}`")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc: int, what: str):
    if rc != 0:
        raise FvError(f"{what} failed with code {rc}")


# --------------------------------------------------------------------------
# torch helpers
# --------------------------------------------------------------------------

def _torch():
    import torch
    if not torch.cuda.is_available():
        raise FvError("no ROCm GPU visible: the HIP path has no CPU fallback")
    return torch


def stream_handle(stream=None) -> int:
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def ptr(t) -> int | None:
    if t is None:
        return None
    if not t.is_contiguous():
        raise FvError("device buffers must be contiguous")
    return t.data_ptr()


def result_tensor(nq: int, width: int, device):
    """uint32 id / float32 distance pairs viewed as an int64 [nq, width] tensor."""
    torch = _torch()
    return torch.empty((nq, width), dtype=torch.int64, device=device)


def split_results(res):
    """int64 [nq, k] GpuSearchResult tensor -> (ids uint32 numpy, dists float32 numpy)."""
    import numpy as np
    a = res.cpu().numpy().view(np.uint32).reshape(res.shape[0], res.shape[1], 2)
    return a[..., 0].copy(), a[..., 1].copy().view(np.float32)


class KernelTimer:
    """Times the dominant scan kernel of each search call with HIP events."""

    def __enter__(self):
        lib = load()
        self._read(reset=True)
        check(lib.fv_kernel_timer_enable(1), "fv_kernel_timer_enable")
        return self

    def __exit__(self, *exc):
        load().fv_kernel_timer_enable(0)
        return False

    @staticmethod
    def _read(reset=False):
        ms = ctypes.c_double(0.0)
        n = ctypes.c_uint32(0)
        check(load().fv_kernel_timer_read(ctypes.byref(ms), ctypes.byref(n), int(reset)),
              "fv_kernel_timer_read")
        return ms.value, n.value

    def read(self, reset=False):
        """(total_ms, launches) recorded so far."""
        return self._read(reset)
