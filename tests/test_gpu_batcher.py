"""Micro-batcher over the host engine (fvh_batcher_create: SearchBatcher::for_index,
crates/server/src/batcher.rs:89-124): concurrent single-query submissions
with mixed k return exactly what one batched engine search of the same
queries returns, on flat and IVF-PQ engines (with deletions and a delta
buffer)."""
import numpy as np
import pytest

from flashvector_amd import engine as E
from flashvector_amd import lib as L
from flashvector_amd.batcher import SearchBatcher
from flashvector_amd.ivfpq_build import build_ivfpq

from test_batcher import run_threads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a ROCm device"
    L.load()
    return t


def check_batched(engine, qs, ks, max_wait_us, max_batch):
    want = {k: engine.search(qs, k) for k in set(ks)}
    got = {}
    with SearchBatcher.for_engine(engine, max_wait_us, max_batch) as b:
        run_threads(len(qs), lambda i: got.__setitem__(i, b.submit(qs[i], ks[i])))
        st = b.stats()
    for i, k in enumerate(ks):
        np.testing.assert_array_equal(got[i][0], want[k][0][i])
        np.testing.assert_array_equal(got[i][1].view(np.uint32), want[k][1][i].view(np.uint32))
        assert got[i][2] > 0.0
    assert st["total_requests"] == len(qs)
    assert st["total_batches"] >= len(set(ks))
    assert st["total_batches"] * max_batch >= len(qs)
    return st


def test_flat_engine(torch):
    rng = np.random.default_rng(61)
    n, dim = 20000, 64
    x = rng.integers(0, 256, size=(n, dim)).astype(np.float32)
    e = E.Engine(dim, "l2", max_top_k=64, shortlist_factor=8)
    e.load_flat(x, "f16", (7 + 3 * np.arange(n)).astype(np.uint32))
    e.add(rng.integers(0, 256, size=(200, dim)).astype(np.float32),
          (10 ** 6 + np.arange(200)).astype(np.uint32))
    e.delete(np.array([7, 10, 13, 10 ** 6 + 4], np.uint32))
    qs = np.concatenate([x[:8], rng.integers(0, 256, size=(40, dim)).astype(np.float32)])
    ks = [(5, 10, 32)[i % 3] for i in range(len(qs))]
    st = check_batched(e, qs, ks, max_wait_us=20000, max_batch=16)
    assert st["total_batches"] < len(qs)


def test_ivfpq_engine(torch):
    rng = np.random.default_rng(62)
    x = rng.normal(size=(6000, 32)).astype(np.float32)
    t = build_ivfpq(x, 24, 8, seed=63, iters=3)
    e = E.Engine(32, "l2", max_top_k=32, ivf_nprobe=6, ivf_shortlist=100)
    e.load_ivfpq(t.centroids, t.codebooks.reshape(8, -1), t.codes, t.offsets, t.vec_ids, x)
    e.delete(np.arange(0, 6000, 41, dtype=np.uint32))
    qs = rng.normal(size=(33, 32)).astype(np.float32)
    ks = [(10, 3)[i % 2] for i in range(len(qs))]
    check_batched(e, qs, ks, max_wait_us=5000, max_batch=64)
    check_batched(e, qs[:9], [10] * 9, max_wait_us=0, max_batch=1)
