"""IVF-PQ with more lists than one LDS sort holds: the coarse probe takes the
tiled exact scan (up to 1024 probes), a full sort in LDS (up to 16384 lists)
or a radix selection in the workspace (beyond), through the reference's entry
point (cuda_ivf_pq_search_batch, no nlist limit in ivf_pq_lookup.cu:339-362)
and the uneven-batch entry, bit-exact against the oracle. Zipf(1.3) sizes
over this many lists leave many lists empty."""
import numpy as np
import pytest

import oracle as O
from flashvector_amd import ivfpq_workload as W
from flashvector_amd import lib as L

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


def tables(torch, nlist, seed):
    w = W.build(n=120_000, dim=32, nlist=nlist, zipf_s=1.3, m_pq=8, nq=64, train=30000, iters=3,
                seed=seed, nprobe_choices=(4, 1100, 1500), k_choices=(10, 50))
    torch.cuda.synchronize()

    def h(t):
        return t.cpu().numpy()
    host = O.IvfPqHost(h(w.centroids), h(w.codebooks).reshape(8, 256, 4), h(w.codes),
                       h(w.offsets).view(np.uint32), h(w.vec_ids).view(np.uint32),
                       h(w.vectors), 8)
    from flashvector_amd import search as F
    idx = F.IvfPqIndex(w.centroids, w.codebooks, w.codes, w.offsets, w.vec_ids, w.vectors)
    return w, host, idx


@pytest.mark.parametrize("nlist", [16384, 20000])
def test_large_nlist_legacy_and_uneven(torch, nlist):
    w, host, idx = tables(torch, nlist, 501 + nlist)
    assert (w.sizes == 0).sum() > 0
    q = w.queries
    qn = q.cpu().numpy()
    # 8 probes: tiled coarse scan; 1100 / 1500: LDS sort (16384 lists) or
    # workspace selection (20000 lists).
    for k, nprobe, rr in [(10, 8, 100), (10, 1100, 300), (20, 1500, 1)]:
        got = L.split_results(idx.search_batch(q, k, nprobe, rr))
        assert_same(got, O.ivf_pq_search_batch(host, qn, k, nprobe, rr))
    ks = w.top_k.cpu().numpy().astype(np.uint32)
    nps = w.nprobe.cpu().numpy().astype(np.uint32)
    out = idx.search(q, 50, 1500, 400, top_k=w.top_k, nprobe=w.nprobe)
    assert_same(L.split_results(out), O.ivf_pq_search_uneven(host, qn, ks, nps, 400, out_stride=50))
