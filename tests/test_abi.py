"""C ABI checks that need no GPU: the library loads, exports every entry point
the headers declare, the POD layouts match the reference's, and NULL required
pointers are rejected with -1 before any device work."""
import ctypes
import os
import re

import pytest

from flashvector_amd import lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", "flashvector", h)
           for h in ("cuda_bridge.h", "fv_ext.h", "fv_host.h", "fv_build.h")]


def declared_symbols():
    names = []
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:int|size_t|void|uint32_t)\s+(fvh_\w+|fv_\w+|cuda_\w+)\s*\(",
                            text, flags=re.M)
    return names


@pytest.fixture(scope="module")
def so():
    if not os.path.exists(L.LIB_PATH):
        pytest.fail("libflashvector_hip.so is not built (run __graft_entry__.build())")
    return L.load()


def test_every_declared_symbol_is_exported(so):
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(so, n), n


def test_binding_covers_header():
    assert set(declared_symbols()) == set(L.SIGNATURES)


def test_struct_layouts():
    assert ctypes.sizeof(L.IvfPqGpuTables) == 72
    assert L.IvfPqGpuTables.num_vectors.offset == 48
    assert ctypes.sizeof(L.Fp16GpuTables) == 40
    assert L.Fp16GpuTables.num_vectors.offset == 24


def test_null_pointers_rejected(so):
    assert so.cuda_flat_scan_batch(None, None, 10, 8, None, 1, 5, 0, None, None, None) == -1
    assert so.cuda_merge_topk(None, 1, 1, 1, 1, None, None) == -1
    assert so.cuda_compute_distances_warp(None, None, 1, 1, 1, 0, None, None) == -1
    assert so.cuda_bitonic_sort_test(None, None, None, None, 4, None) == -1
    assert so.cuda_get_device_count(None) == -1
    assert so.cuda_malloc_device(None, 16) == -1
    assert so.fv_flat_exact_search(None, 1, 1, None, 1, 1, 0, None, None, None, 0, None) == -1


def test_workspace_sizes_are_host_only(so):
    assert so.fv_flat_exact_workspace_size(1000, 128, 10, 10, 0) > 0
    assert so.fv_flat_exact_workspace_size(1000, 128, 10, 2000, 0) == 0   # k above the limit
    assert so.fv_kmeans_workspace_size(10000, 128, 64) > 10000 * 16
    assert so.fv_kmeans_workspace_size(10, 128, 64) == 0                   # k above n
    assert so.fv_ivf_pq_build_workspace_size(10000, 128, 64, 16) > 10000 * 128 * 4
    assert so.fv_ivf_pq_build_workspace_size(10000, 130, 64, 16) == 0      # dim % m_pq


def test_index_build_rejects_null(so):
    assert so.fv_kmeans_assign(None, 1, 1, 1, None, 1, None, None, 0, None) == -1
    assert so.fv_kmeans_fit(None, 1, 1, 1, 1, 1, 0, None, None, None, None, None, None, None,
                            0, None) == -1
    assert so.fv_ivf_pq_build(None, 1, 1, 1, 1, 1, 1, 0, None, None, None, None, None, None,
                              None, 0, None) == -1


def test_host_engine_rejects_null_and_bad_config(so):
    assert so.fvh_create(None, None) == -1
    assert so.fvh_search(None, None, 1, 1, None, 0, None) == -1
    assert so.fvh_search_device(None, None, 1, 1, None, 0, None, None) == -1
    assert so.fvh_load_flat(None, None, 1, 1, None) == -1
    assert so.fvh_delete(None, None, 1) == -1
    assert so.fvh_stats(None, None, None) == -1
    assert ctypes.sizeof(L.FvEngineConfig) == 40
