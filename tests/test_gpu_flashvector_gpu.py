"""The FlashVectorGPU class (the reference's Python binding,
crates/python/src/py_index.rs:155-436) on the GPU.

The first tests restate the reference's own binding tests
(python/tests/test_bindings.py:22-169) as far as they concern the paths served
here: numpy / CPU tensor / ROCm tensor inputs, shapes and dtypes, the empty
build error, cosine distances in range, k = 100 through `search_ivf` sorted
and at least 90% valid, mask membership through `search_fp16` and
`search_ivf`, add / delete / save / load / compact. The others check parity:
`search_ivf` with per-call nprobe and rerank_k equals the oracle's IVF-PQ
search over the tables the index built (read back from its snapshot), and
`search_fp16` equals the oracle's exact top-k on integer data."""
import numpy as np
import pytest

import oracle as O
from flashvector_amd import lib as L
from flashvector_amd import snapshot as SN
from flashvector_amd import synthetic as S
from flashvector_amd.flashvector_gpu import FlashVectorGPU, normalize_rows

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a ROCm device"
    L.load()
    O.build()
    return t


def assert_same(got, exp):
    np.testing.assert_array_equal(got[0], exp[0])
    np.testing.assert_array_equal(got[1].view(np.uint32), exp[1].view(np.uint32))


def test_numpy_and_tensor_inputs(torch):
    rng = np.random.default_rng(0)
    data = rng.standard_normal((1000, 128)).astype(np.float32)
    query = rng.standard_normal((5, 128)).astype(np.float32)
    index = FlashVectorGPU(dim=128, max_elements=1000, m=16, ef_construction=64)
    index.build(data)
    labels, dists = index.search(query, top_k=10, ef_search=32)
    assert labels.shape == (5, 10) and dists.shape == (5, 10)
    assert labels.dtype == np.uint32 and dists.dtype == np.float32
    t_index = FlashVectorGPU(dim=128, max_elements=1000, m=16)
    t_index.build(torch.from_numpy(data))
    t_labels, _ = t_index.search(torch.from_numpy(query[:2]), top_k=5, ef_search=32)
    assert t_labels.shape == (2, 5)
    c_index = FlashVectorGPU(dim=128, max_elements=1000, m=16)
    c_index.build(torch.from_numpy(data).cuda())
    c_labels, c_dists = c_index.search(torch.from_numpy(query[:2]).cuda(), top_k=5)
    assert c_labels.shape == (2, 5) and c_dists.shape == (2, 5)
    np.testing.assert_array_equal(c_labels, t_labels)


def test_empty_build_raises(torch):
    index = FlashVectorGPU(dim=128, max_elements=100, m=16)
    with pytest.raises(ValueError):
        index.build(np.zeros((0, 128), np.float32))


def test_cosine_distances_in_range(torch):
    rng = np.random.default_rng(1)
    data = rng.standard_normal((200, 64)).astype(np.float32)
    query = rng.standard_normal((1, 64)).astype(np.float32)
    index = FlashVectorGPU(dim=64, max_elements=200, m=16, m_pq=8, metric="cosine")
    index.build(data)
    for labels, dists in (index.search(query, top_k=5), index.search_ivf(query, top_k=5)):
        assert labels.shape == (1, 5)
        assert np.all(dists >= -0.01) and np.all(dists <= 2.01)
    # The host normalizes rows and queries as the reference does; the flat
    # result is then the exact cosine top-k of the normalized rows.
    got = index.search_fp16(query, top_k=5)
    assert_same(got, O.flat_topk(normalize_rows(data), normalize_rows(query), 5, "cosine"))


def test_large_top_k_ivf_sorted(torch):
    rng = np.random.default_rng(2)
    data = rng.standard_normal((1000, 64)).astype(np.float32)
    query = rng.standard_normal((1, 64)).astype(np.float32)
    index = FlashVectorGPU(dim=64, max_elements=1000, m=16, nlist=16, m_pq=8)
    index.build(data)
    labels, dists = index.search_ivf(query, top_k=100, nprobe=16)
    assert labels.shape == (1, 100)
    assert np.sum(labels[0] != 0xFFFFFFFF) >= 90
    assert np.all(dists[0][:-1] <= dists[0][1:])
    labels, _ = index.search(query, top_k=100, ef_search=256)
    assert np.sum(labels[0] != 0xFFFFFFFF) >= 90


def test_masks_delta_delete_snapshot_compact(torch, tmp_path):
    rng = np.random.default_rng(7)
    dim, n = 64, 400
    data = rng.standard_normal((n, dim)).astype(np.float32)
    index = FlashVectorGPU(dim=dim, max_elements=n, m=16, nlist=8, m_pq=8)
    index.build(data)
    q = data[:2].copy()
    mask = np.zeros(n, dtype=bool)
    mask[10:200] = True
    words = S.pack_bits(mask)
    for labels, _ in [
        index.search_fp16(q, top_k=5, filter_mask=mask),
        index.search_fp16(q, top_k=5, filter_mask=words),
        index.search_ivf(q, top_k=5, nprobe=8, rerank_k=64, mode="pq", filter_mask=mask),
        index.search_ivf(q, top_k=5, nprobe=8, rerank_k=64, mode="ivf-pq", filter_mask=words),
        index.search(q, top_k=5, ef_search=48, filter_mask=mask),
    ]:
        assert labels.shape == (2, 5)
        valid = labels[labels != 0xFFFFFFFF]
        assert len(valid) > 0
        assert np.all((valid >= 10) & (valid < 200))
    for mode in ("sq8", "rabitq"):
        with pytest.raises(NotImplementedError):
            index.search_ivf(q, top_k=5, mode=mode)
    with pytest.raises(ValueError):
        index.search_ivf(q, top_k=5, mode="hnsw")

    extra = np.full((1, dim), 25.0, dtype=np.float32)
    assigned = index.add(extra)
    assert assigned.tolist() == [n]
    labels, _ = index.search(extra, top_k=3, ef_search=32)
    assert labels[0, 0] == n
    labels, _ = index.search_ivf(extra, top_k=3, nprobe=8)
    assert labels[0, 0] == n
    assert index.delete(np.array([0], dtype=np.uint32)) == 1
    assert index.delete(np.array([0, 100000], dtype=np.int64)) == 0
    labels_del, _ = index.search(data[:1], top_k=5, ef_search=32)
    assert 0 not in labels_del[0]

    snap = tmp_path / "bindings_snap.fv"
    index.save(str(snap))
    reloaded = FlashVectorGPU.load(str(snap))
    assert reloaded.tombstone_count() == 1 and reloaded.delta_len() == 1
    for a, b in [(index.search_fp16(data[:20], top_k=5), reloaded.search_fp16(data[:20], top_k=5)),
                 (index.search_ivf(data[:20], top_k=5, nprobe=4),
                  reloaded.search_ivf(data[:20], top_k=5, nprobe=4))]:
        assert_same(a, b)
    reloaded.compact()
    assert reloaded.tombstone_count() == 0
    assert reloaded.delta_len() == 0
    assert reloaded.num_vectors() == n          # n - 1 deleted + 1 added
    assert reloaded.next_logical_id() == n + 1
    labels, _ = reloaded.search(extra, top_k=3)
    assert labels[0, 0] == n                    # logical ids survive compaction
    labels, _ = reloaded.search(data[:1], top_k=5)
    assert 0 not in labels[0]


def test_search_ivf_per_call_params_vs_oracle(torch, tmp_path):
    """search_ivf with per-call nprobe and rerank_k (0 = k, 1 = no rerank,
    deeper than the engine default) equals the oracle over the index's own
    tables, and search_fp16 equals the exact top-k (integer data)."""
    db, qs = S.flat_workload(20000, 300, 64, 48, seeds=(301, 302, 303))
    index = FlashVectorGPU(dim=64, max_elements=20000, nlist=64, m_pq=16, seed=5)
    index.build(db)
    snap = str(tmp_path / "t.fv")
    index.save(snap)
    _, sec = SN.read(snap, ["ivf_centroids", "pq_codebooks", "pq_codes", "ivf_offsets",
                            "ivf_vec_ids", "vectors"])
    h = O.IvfPqHost(sec["ivf_centroids"].reshape(-1, 64), sec["pq_codebooks"],
                    sec["pq_codes"].reshape(20000, 16), sec["ivf_offsets"], sec["ivf_vec_ids"],
                    sec["vectors"].reshape(20000, 64), 16)
    d_q = torch.from_numpy(qs).cuda()
    for k, nprobe, rerank_k in [(10, 8, 32), (10, 3, 0), (20, 16, 1), (50, 40, 2000), (100, 8, 32)]:
        got = index.search_ivf(qs, top_k=k, nprobe=nprobe, rerank_k=rerank_k)
        assert_same(got, O.ivf_pq_search_batch(h, qs, k, nprobe, rerank_k))
        dev = index.search_ivf(d_q, top_k=k, nprobe=nprobe, rerank_k=rerank_k)
        assert_same(dev, got)
    mask = S.random_valid_mask(20000, 0.1, 304)
    got = index.search_ivf(qs, top_k=10, nprobe=12, rerank_k=200, filter_mask=mask)
    assert_same(got, O.ivf_pq_search_batch(h, qs, 10, 12, 200, valid_mask=mask))
    got = index.search_fp16(qs, top_k=25, filter_mask=mask)
    assert_same(got, O.flat_topk(db, qs, 25, "l2", valid_mask=mask))
