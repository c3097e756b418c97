"""GPU parity of the exact FP32 flat scan, the merge and the KAT kernels
against the CPU oracle and the golden fixtures. Bar: bit-exact ids and
distances."""
import os

import numpy as np
import pytest

import oracle as O
from flashvector_amd import lib as L
from flashvector_amd import search as F
from flashvector_amd import synthetic as S

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a ROCm device"
    L.load()
    O.build()
    return t


def dev(torch, a, dtype=None):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def run_flat(torch, db, qs, k, metric="l2", mask=None, vec_ids=None):
    out = F.flat_scan_batch(dev(torch, db), dev(torch, qs), k, metric,
                            None if mask is None else dev(torch, mask),
                            None if vec_ids is None else dev(torch, vec_ids))
    torch.cuda.synchronize()
    return L.split_results(out)


def assert_same(got, exp):
    np.testing.assert_array_equal(got[0], exp[0])
    np.testing.assert_array_equal(got[1].view(np.uint32), exp[1].view(np.uint32))


def test_golden_flat_small(torch):
    g = np.load(os.path.join(GOLDEN, "flat_small.npz"))
    db, qs = g["db_u8"].astype(np.float32), g["queries_u8"].astype(np.float32)
    for metric in ("l2", "ip", "cosine"):
        assert_same(run_flat(torch, db, qs, 10, metric), (g[f"{metric}_ids"], g[f"{metric}_dist"]))
    assert_same(run_flat(torch, db, qs, 10, "l2", mask=g["mask"]),
                (g["l2_masked_ids"], g["l2_masked_dist"]))
    assert_same(run_flat(torch, db, qs, 300, "l2"), (g["l2_k300_ids"], g["l2_k300_dist"]))


@pytest.mark.parametrize("n,nq,dim,k,metric", [
    (5000, 100, 128, 10, "l2"),
    (777, 65, 100, 33, "l2"),
    (4096, 7, 17, 1, "ip"),
    (20000, 130, 64, 100, "cosine"),
    (300, 3, 128, 256, "l2"),
    (50, 5, 40, 64, "l2"),        # fewer rows than k: padded rows
])
def test_flat_random_vs_oracle(torch, n, nq, dim, k, metric):
    db, qs = S.flat_workload(n, nq, dim, 16, seeds=(n, nq, 0))
    assert_same(run_flat(torch, db, qs, k, metric), O.flat_scan_batch(db, qs, k, metric))


def test_full_size_100k_q1000_bit_exact(torch):
    """The CPU-runnable workload at full size (100,000 x 128 integer mixture
    rows, 1,000 queries, k=10, seeds of flat_fp32_100k): every query's ids and
    distances equal the oracle's; with a 50% mask and k=100 too."""
    db, qs = S.flat_workload(100_000, 1_000, 128, 256, seeds=S.SEEDS["flat_fp32_100k"])
    assert_same(run_flat(torch, db, qs, 10), O.flat_scan_batch(db, qs, 10))
    mask = S.random_valid_mask(100_000, 0.5, S.SEEDS["flat_fp32_100k"][2])
    assert_same(run_flat(torch, db, qs[:200], 100, mask=mask),
                O.flat_scan_batch(db, qs[:200], 100, valid_mask=mask))


def test_flat_non_integer_data_bit_exact(torch):
    rng = np.random.default_rng(4)
    db = rng.normal(size=(6000, 96)).astype(np.float32)
    qs = rng.normal(size=(40, 96)).astype(np.float32)
    for metric in ("l2", "ip", "cosine"):
        assert_same(run_flat(torch, db, qs, 20, metric), O.flat_scan_batch(db, qs, 20, metric))


def test_flat_ties_mask_remap(torch):
    db, qs = S.flat_workload(3000, 20, 128, 8, seeds=(5, 6, 7))
    db[100:400] = db[50]                     # many exact ties
    qs[0] = db[50]
    mask = S.random_valid_mask(3000, 0.4, 8)
    remap = (np.arange(3000, dtype=np.uint32) * 3 + 11)
    assert_same(run_flat(torch, db, qs, 50, "l2", mask=mask, vec_ids=remap),
                O.flat_scan_batch(db, qs, 50, "l2", valid_mask=mask, vec_ids=remap))
    zero = np.zeros_like(mask)
    ids, d = run_flat(torch, db, qs, 5, "l2", mask=zero)
    assert (ids == 0xFFFFFFFF).all() and (d == np.float32(1e30)).all()


def test_flat_empty_database_fills_ff(torch):
    db = np.zeros((0, 16), np.float32)
    qs = np.ones((3, 16), np.float32)
    q = dev(torch, qs)
    out = F.flat_scan_batch(torch.zeros((1, 16), device="cuda"), q, 4)
    # num_vectors = 0 through the raw entry point
    lib = L.load()
    rc = lib.cuda_flat_scan_batch(L.ptr(torch.zeros((1, 16), device="cuda")), None, 0, 16,
                                  L.ptr(q), 3, 4, 0, None, L.ptr(out), L.stream_handle())
    assert rc == 0
    torch.cuda.synchronize()
    assert (out.cpu().numpy().view(np.uint32) == 0xFFFFFFFF).all()
    del db


def test_flat_exact_search_large_k(torch):
    db, qs = S.flat_workload(8000, 33, 128, 16, seeds=(9, 10, 11))
    out = F.flat_exact_search(dev(torch, db), dev(torch, qs), 1000)
    torch.cuda.synchronize()
    assert_same(L.split_results(out), O.flat_topk(db, qs, 1000))


def test_distance_kat(torch):
    g = np.load(os.path.join(GOLDEN, "kat_distance.npz"))
    got = F.compute_distances(dev(torch, g["queries"]), dev(torch, g["dataset"])).cpu().numpy()
    assert (np.abs(got - g["exact_l2"]) / np.maximum(g["exact_l2"], 1.0) < 1e-4).all()
    np.testing.assert_array_equal(got, O.distance_matrix(g["queries"], g["dataset"]))
    rng = np.random.default_rng(1)
    q, x = rng.normal(size=(5, 33)).astype(np.float32), rng.normal(size=(9, 33)).astype(np.float32)
    for metric in ("ip", "cosine"):
        got = F.compute_distances(dev(torch, q), dev(torch, x), metric).cpu().numpy()
        np.testing.assert_array_equal(got, O.distance_matrix(q, x, metric))


def test_bitonic_kat(torch):
    g = np.load(os.path.join(GOLDEN, "kat_bitonic.npz"))
    ko, vo = F.bitonic_sort_test(dev(torch, g["keys"]), dev(torch, g["vals"]))
    np.testing.assert_array_equal(ko.cpu().numpy(), g["keys_sorted"])
    np.testing.assert_array_equal(vo.cpu().numpy(), g["vals_sorted"])
    rng = np.random.default_rng(2)
    keys = rng.integers(0, 20, 64).astype(np.float32)       # many equal keys
    vals = rng.permutation(64).astype(np.uint32)
    ko, vo = F.bitonic_sort_test(dev(torch, keys), dev(torch, vals))
    order = np.lexsort((vals, keys))
    np.testing.assert_array_equal(ko.cpu().numpy(), keys[order])
    np.testing.assert_array_equal(vo.cpu().numpy(), vals[order])


def test_merge_topk_vs_oracle(torch):
    db, qs = S.flat_workload(2400, 25, 64, 8, seeds=(12, 13, 14))
    parts = np.array_split(np.arange(2400), 4)
    res = []
    for p in parts:
        res.append(F.flat_scan_batch(dev(torch, db[p]), dev(torch, qs), 16,
                                     vec_ids=dev(torch, p.astype(np.uint32))))
    stacked = torch.stack(res)
    got = L.split_results(F.merge_topk(stacked, 16))
    ids = np.stack([L.split_results(r)[0] for r in res])
    ds = np.stack([L.split_results(r)[1] for r in res])
    assert_same(got, O.merge_topk(ids, ds, 16))
    assert_same(got, O.flat_scan_batch(db, qs, 16))
    # duplicate ids across shards keep their best distance
    dup = torch.stack([res[0], res[0]])
    assert_same(L.split_results(F.merge_topk(dup, 16)), L.split_results(res[0]))
