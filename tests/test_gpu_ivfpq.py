"""GPU parity of the IVF-PQ ADC scan against the oracle and the golden
fixture. Bar: bit-exact ids and distances given identical tables."""
import os

import numpy as np
import pytest

import oracle as O
from flashvector_amd import lib as L
from flashvector_amd import search as F
from flashvector_amd import synthetic as S
from flashvector_amd.ivfpq_build import IvfPqTables, build_ivfpq

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a ROCm device"
    L.load()
    O.build()
    return t


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def assert_same(got, exp):
    np.testing.assert_array_equal(got[0], exp[0])
    np.testing.assert_array_equal(got[1].view(np.uint32), exp[1].view(np.uint32))


def host_tables(t):
    return O.IvfPqHost(t.centroids, t.codebooks, t.codes, t.offsets, t.vec_ids, t.vectors, t.m_pq)


@pytest.fixture(scope="module")
def case(torch):
    db, qs = S.flat_workload(30000, 150, 64, 48, seeds=(41, 42, 43))
    t = build_ivfpq(db, 64, 16, seed=44, iters=4)
    return db, qs, t, F.IvfPqIndex.from_host(t), host_tables(t)


def test_golden_fixture(torch):
    g = np.load(os.path.join(GOLDEN, "ivfpq_small.npz"))
    db = g["db_u8"].astype(np.float32)
    qs = g["queries_u8"].astype(np.float32)
    idx = F.IvfPqIndex(dev(torch, g["centroids"]), dev(torch, g["codebooks"].reshape(8, -1)),
                       dev(torch, g["codes"]), dev(torch, g["offsets"].view(np.int32)),
                       dev(torch, g["vec_ids"].view(np.int32)), dev(torch, db))
    ci = 0
    while f"case{ci}_params" in g:
        k, npb, rr, masked = (int(v) for v in g[f"case{ci}_params"])
        m = dev(torch, g["mask"]) if masked else None
        got = L.split_results(idx.search_batch(dev(torch, qs), k, npb, rr, "l2", m))
        assert_same(got, (g[f"case{ci}_ids"], g[f"case{ci}_dist"]))
        ci += 1
    assert ci >= 5


@pytest.mark.parametrize("k,nprobe,rerank_k,metric", [
    (10, 8, 100, "l2"), (10, 8, 1, "l2"), (5, 3, 0, "l2"), (30, 16, 2000, "l2"),
    (10, 64, 9000, "l2"),          # deep shortlist: candidate buffer in global memory
    (20, 12, 300, "ip"), (20, 12, 300, "cosine"), (256, 20, 500, "l2"), (300, 4, 50, "l2"),
])
def test_legacy_entry_vs_oracle(torch, case, k, nprobe, rerank_k, metric):
    db, qs, t, idx, h = case
    got = L.split_results(idx.search_batch(dev(torch, qs), k, nprobe, rerank_k, metric))
    assert_same(got, O.ivf_pq_search_batch(h, qs, k, nprobe, rerank_k, metric))


def test_mask(torch, case):
    db, qs, t, idx, h = case
    mask = S.random_valid_mask(db.shape[0], 0.3, 5)
    got = L.split_results(idx.search_batch(dev(torch, qs), 10, 8, 200, "l2", dev(torch, mask)))
    assert_same(got, O.ivf_pq_search_batch(h, qs, 10, 8, 200, valid_mask=mask))
    bits = S.unpack_bits(mask, db.shape[0])
    assert bits[got[0][got[0] != 0xFFFFFFFF]].all()


@pytest.mark.parametrize("list_order", [True, False])
def test_uneven_batch_vs_oracle(torch, case, list_order):
    db, qs, t, idx, h = case
    rng = np.random.default_rng(7)
    ks = rng.choice([10, 100], size=qs.shape[0]).astype(np.uint32)
    nps = rng.choice([8, 16, 32, 64], size=qs.shape[0]).astype(np.uint32)
    out = idx.search(dev(torch, qs), 100, 64, 1000, top_k=dev(torch, ks.view(np.int32)),
                     nprobe=dev(torch, nps.view(np.int32)), use_list_order=list_order)
    got = L.split_results(out)
    assert_same(got, O.ivf_pq_search_uneven(h, qs, ks, nps, 1000, out_stride=100))


def test_integer_codebooks_order_independent(torch):
    db, qs = S.flat_workload(8000, 40, 32, 24, seeds=(51, 52, 53))
    t = build_ivfpq(db, 16, 8, seed=54, iters=3, integer_codebooks=True)
    idx = F.IvfPqIndex.from_host(t)
    got = L.split_results(idx.search_batch(dev(torch, qs), 15, 5, 1))
    assert_same(got, O.ivf_pq_search_batch(host_tables(t), qs, 15, 5, 1))


@pytest.mark.parametrize("dim,m_pq", [
    (128, 16),   # sub 8, the 10M-row workload's shape
    (64, 4),     # sub 16
    (96, 3),     # sub 32
    (36, 12),    # sub 3: scalar LUT build, LUT in dynamic LDS
    (48, 16),    # sub 3 with the static 16-sub-quantizer LUT
])
def test_lut_shapes_vs_oracle(torch, dim, m_pq):
    """Every LUT build shape (vector loads of 1, 2, 4 or 8 float4 per
    codeword, or the scalar loop) and both LUT placements, bit-exact, with
    shortlists shallow enough that the candidate buffer is cut many times."""
    db, qs = S.flat_workload(6000, 40, dim, 24, seeds=(61, 62, 63))
    t = build_ivfpq(db, 16, m_pq, seed=64, iters=3)
    idx = F.IvfPqIndex.from_host(t)
    h = host_tables(t)
    for k, nprobe, rr in [(10, 6, 40), (7, 16, 1), (25, 3, 600)]:
        got = L.split_results(idx.search_batch(dev(torch, qs), k, nprobe, rr))
        assert_same(got, O.ivf_pq_search_batch(h, qs, k, nprobe, rr))


def test_exhaustive_probe_matches_oracle_and_flat(torch, case):
    db, qs, t, idx, h = case
    got = L.split_results(idx.search_batch(dev(torch, qs[:20]), 10, t.nlist, 10240))
    assert_same(got, O.ivf_pq_search_batch(h, qs[:20], 10, t.nlist, 10240))
    flat = O.flat_scan_batch(db, qs[:20], 10)
    assert np.mean(got[0] == flat[0]) >= 0.99


def test_uneven_batch_graph_replay(torch, case):
    """The whole uneven batch captured as one HIP graph: every replay, with
    eager searches and synchronizations in between, equals the eager result
    and the oracle."""
    db, qs, t, idx, h = case
    rng = np.random.default_rng(91)
    ks = rng.integers(1, 21, size=qs.shape[0]).astype(np.int32)
    nps = rng.integers(1, 17, size=qs.shape[0]).astype(np.int32)
    d_q, d_k, d_np = dev(torch, qs), dev(torch, ks), dev(torch, nps)
    out = L.result_tensor(qs.shape[0], 20, d_q.device)

    def step():
        idx.search(d_q, 20, 16, 60, top_k=d_k, nprobe=d_np, out=out)

    step()
    torch.cuda.synchronize()
    eager = out.clone()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step()
    torch.cuda.current_stream().wait_stream(side)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        step()
    other = L.result_tensor(qs.shape[0], 5, d_q.device)
    for r in range(3):
        out.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, eager), f"replay {r}"
        idx.search(d_q, 5, 4, 20, out=other)
        torch.cuda.synchronize()
    e_ids, e_d = O.ivf_pq_search_uneven(h, qs, ks.astype(np.uint32), nps.astype(np.uint32),
                                        shortlist=60, out_stride=20)
    assert_same(L.split_results(eager), (e_ids, e_d))


def with_empty_lists(t, before):
    """The same tables with an empty posting list inserted before each list
    index in `before` (its centroid a copy of that list's, so coarse
    distances tie and the (distance, list id) order decides) and one more,
    far away, at the end."""
    cents, offs = [], [0]
    for li in range(t.nlist):
        if li in before:
            cents.append(t.centroids[li])
            offs.append(offs[-1])
        cents.append(t.centroids[li])
        offs.append(int(t.offsets[li + 1]))
    cents.append(t.centroids[-1] + 1000.0)
    offs.append(offs[-1])
    return IvfPqTables(np.asarray(cents, np.float32), t.codebooks, t.codes,
                       np.asarray(offs, np.uint32), t.vec_ids, t.vectors, t.m_pq)


def test_clamps_and_empty_lists(torch, case):
    """Parameter clamps of the legacy entry (ivf_pq_lookup.cu:351-362: nprobe
    0 -> 1, nprobe above nlist -> nlist, shortlist above 10240 -> 10240) as
    equalities between calls, and empty posting lists (tied with a populated
    twin, in the middle and at the end) on both entry points."""
    db, qs, t, idx, h = case
    q40 = qs[:40]
    q = dev(torch, q40)

    def run(ix, k, npb, rr):
        return L.split_results(ix.search_batch(q, k, npb, rr))

    assert_same(run(idx, 10, 0, 100), run(idx, 10, 1, 100))
    assert_same(run(idx, 10, 1000, 100), run(idx, 10, t.nlist, 100))
    assert_same(run(idx, 10, 8, 20000), run(idx, 10, 8, 10240))
    assert_same(run(idx, 10, 0, 100), O.ivf_pq_search_batch(h, q40, 10, 0, 100))
    e = with_empty_lists(t, {0, 10, 30})
    ei, eh = F.IvfPqIndex.from_host(e), host_tables(e)
    for k, npb, rr in [(10, 8, 200), (10, e.nlist, 500), (5, 3, 1)]:
        assert_same(run(ei, k, npb, rr), O.ivf_pq_search_batch(eh, q40, k, npb, rr))
    assert_same(run(ei, 10, e.nlist, 500), run(idx, 10, t.nlist, 500))
    rng = np.random.default_rng(93)
    ks = rng.integers(1, 30, size=40).astype(np.uint32)
    nps = rng.integers(1, 12, size=40).astype(np.uint32)
    out = ei.search(q, 32, 12, 300, top_k=dev(torch, ks.view(np.int32)),
                    nprobe=dev(torch, nps.view(np.int32)), use_list_order=True)
    assert_same(L.split_results(out), O.ivf_pq_search_uneven(eh, q40, ks, nps, 300, out_stride=32))


# Work-item plans: queries split into chunks of their probed entries, scanned
# by several workgroups and merged, must give exactly the unsplit results.
@pytest.mark.parametrize("target_items,chunk_min", [(0, 16384), (1024, 1), (65536, 1),
                                                    (65536, 300)])
@pytest.mark.parametrize("list_order,masked", [(True, False), (False, False), (True, True)])
def test_split_plans_vs_oracle(torch, case, target_items, chunk_min, list_order, masked):
    db, qs, t, idx, h = case
    rng = np.random.default_rng(47)
    ks = rng.choice([1, 5, 10, 64], size=qs.shape[0]).astype(np.uint32)
    nps = rng.choice([1, 4, 16, 64], size=qs.shape[0]).astype(np.uint32)
    mask = S.random_valid_mask(db.shape[0], 0.3, 48) if masked else None
    tu = L.FvIvfPqTuning(target_items, chunk_min, 0.0, 0)
    out = idx.search(dev(torch, qs), 64, 64, 200, top_k=dev(torch, ks.view(np.int32)),
                     nprobe=dev(torch, nps.view(np.int32)), use_list_order=list_order,
                     valid_mask=None if mask is None else dev(torch, mask), tuning=tu)
    torch.cuda.synchronize()
    exp = O.ivf_pq_search_uneven(h, qs, ks, nps, 200, out_stride=64, valid_mask=mask)
    assert_same(L.split_results(out), exp)


def test_autotune_plan_vs_oracle(torch, case):
    db, qs, t, idx, h = case
    best = idx.autotune(dev(torch, qs), 10, 32, 100, trials=1)
    assert best.ms > 0.0
    out = idx.search(dev(torch, qs), 10, 32, 100, tuning=best)
    torch.cuda.synchronize()
    ks = np.full(qs.shape[0], 10, np.uint32)
    nps = np.full(qs.shape[0], 32, np.uint32)
    exp = O.ivf_pq_search_uneven(h, qs, ks, nps, 100, out_stride=10)
    assert_same(L.split_results(out), exp)


@pytest.mark.parametrize("target_items,chunk_min", [(0, 16384), (4096, 1)])
def test_split_plans_deep_shortlist_vs_oracle(torch, case, target_items, chunk_min):
    """Split queries with a candidate buffer in global memory (shortlist 9000)
    and k up to 256: the merged item shortlists give the oracle's results."""
    db, qs, t, idx, h = case
    rng = np.random.default_rng(49)
    nq = 40
    ks = rng.choice([10, 256], size=nq).astype(np.uint32)
    nps = rng.choice([16, 64], size=nq).astype(np.uint32)
    tu = L.FvIvfPqTuning(target_items, chunk_min, 0.0, 0)
    out = idx.search(dev(torch, qs[:nq]), 256, 64, 9000, top_k=dev(torch, ks.view(np.int32)),
                     nprobe=dev(torch, nps.view(np.int32)), tuning=tu)
    torch.cuda.synchronize()
    exp = O.ivf_pq_search_uneven(h, qs[:nq], ks, nps, 9000, out_stride=256)
    assert_same(L.split_results(out), exp)
