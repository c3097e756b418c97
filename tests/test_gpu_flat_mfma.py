"""GPU parity of the matrix-core flat scan with exact FP32 rerank.

On integer mixture data every FP16/BF16 score is exact, so the shortlist
holds the true top-S and the reranked output must equal the oracle's exact
FP32 top-k bit for bit. On real-valued data the bar is recall against the
oracle plus exact FP32 distances for every returned id."""
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


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def run(torch, db, qs, k, metric="l2", dtype="f16", mask=None, shortlist=0, keep_fp32=True):
    idx = F.FlatIndex(dev(torch, db), dtype, keep_fp32=keep_fp32)
    out = idx.search(dev(torch, qs), k, metric, None if mask is None else dev(torch, mask),
                     shortlist)
    torch.cuda.synchronize()
    return L.split_results(out)


def assert_same(got, exp):
    np.testing.assert_array_equal(got[0], exp[0])
    np.testing.assert_array_equal(got[1].view(np.uint32), exp[1].view(np.uint32))


@pytest.mark.parametrize("dtype", ["f16", "bf16"])
@pytest.mark.parametrize("n,nq,dim,k,metric", [
    (20000, 300, 128, 10, "l2"),
    (20000, 300, 128, 10, "ip"),
    (5000, 17, 100, 25, "l2"),      # padded dims, partial query block
    (3000, 40, 16, 5, "l2"),        # one K step
    (9000, 64, 256, 100, "l2"),     # eight K steps, deep shortlist
    (150, 9, 64, 40, "l2"),         # fewer rows than the shortlist
])
def test_integer_data_exact(torch, dtype, n, nq, dim, k, metric):
    db, qs = S.flat_workload(n, nq, dim, 32, seeds=(n + dim, nq, 1))
    assert_same(run(torch, db, qs, k, metric, dtype), O.flat_topk(db, qs, k, metric))


@pytest.mark.parametrize("dtype", ["f16", "bf16"])
def test_integer_data_masked(torch, dtype):
    db, qs = S.flat_workload(60000, 500, 128, 256, seeds=(11, 12, 13))
    mask = S.random_valid_mask(60000, 0.05, 13)
    got = run(torch, db, qs, 10, "l2", dtype, mask=mask)
    assert_same(got, O.flat_topk(db, qs, 10, "l2", valid_mask=mask))
    bits = S.unpack_bits(mask, 60000)
    assert bits[got[0][got[0] != 0xFFFFFFFF]].all()


def test_golden_fixture_fp16(torch):
    g = np.load(os.path.join(GOLDEN, "flat_small.npz"))
    db, qs = g["db_u8"].astype(np.float32), g["queries_u8"].astype(np.float32)
    assert_same(run(torch, db, qs, 10, "l2"), (g["l2_ids"], g["l2_dist"]))
    assert_same(run(torch, db, qs, 10, "ip"), (g["ip_ids"], g["ip_dist"]))
    assert_same(run(torch, db, qs, 10, "l2", mask=g["mask"]),
                (g["l2_masked_ids"], g["l2_masked_dist"]))


def test_real_valued_recall_and_exact_distances(torch):
    rng = np.random.default_rng(5)
    db = rng.normal(size=(30000, 96)).astype(np.float32)
    qs = rng.normal(size=(200, 96)).astype(np.float32)
    for metric in ("l2", "ip", "cosine"):
        ids, d = run(torch, db, qs, 10, metric, "f16")
        e_ids, e_d = O.flat_topk(db, qs, 10, metric)
        hits = sum(len(set(ids[i]) & set(e_ids[i])) for i in range(200))
        assert hits / (200 * 10) >= 0.999, metric
        exact = O.distance_matrix(qs, db, metric)
        np.testing.assert_array_equal(d, np.take_along_axis(exact, ids.astype(np.int64), 1))


def test_no_rerank_returns_approximate_scores(torch):
    db, qs = S.flat_workload(8000, 50, 128, 32, seeds=(3, 4, 5))
    ids, d = run(torch, db, qs, 10, "l2", keep_fp32=False)
    e_ids, e_d = O.flat_topk(db, qs, 10, "l2")
    np.testing.assert_array_equal(ids, e_ids)          # integer data: scores exact
    np.testing.assert_allclose(d, e_d, rtol=1e-6)


def test_fp16_tc_entry_point_reference_encoding(torch):
    """cuda_fp16_tc_search_batch fed with the reference's own FP16 encoding
    (crates/engine/src/quant.rs:306-327, via the oracle)."""
    db, qs = S.flat_workload(10000, 70, 120, 32, seeds=(21, 22, 23))
    codes, norms = O.encode_fp16_padded(db)
    mask = S.random_valid_mask(10000, 0.5, 24)
    out = F.fp16_tc_search_batch(dev(torch, codes.view(np.int16)), dev(torch, norms),
                                 dev(torch, qs), 12, 120, "l2", dev(torch, mask), dev(torch, db))
    torch.cuda.synchronize()
    assert_same(L.split_results(out), O.flat_topk(db, qs, 12, "l2", valid_mask=mask))


def test_mfma_layout_probe_matches_cpu(torch):
    """The kernel's fragment layouts reproduce a plain product: scores of the
    no-rerank path equal 0.5|x|^2 - q.x exactly on small integers."""
    rng = np.random.default_rng(9)
    db = rng.integers(0, 8, size=(64, 32)).astype(np.float32)
    qs = rng.integers(0, 8, size=(16, 32)).astype(np.float32)
    ids, d = run(torch, db, qs, 64, "ip", keep_fp32=False)
    ref = -(qs.astype(np.float64) @ db.astype(np.float64).T)
    np.testing.assert_array_equal(np.take_along_axis(ref, ids.astype(np.int64), 1), d)


def test_sparse_mask_compaction(torch):
    """A mask keeping 5% of rows with its popcount supplied: only valid rows
    are scored (compacted on the device); results equal the oracle."""
    db, qs = S.flat_workload(80000, 300, 128, 64, seeds=(61, 62, 63))
    mask = S.random_valid_mask(80000, 0.05, 64)
    vc = int(S.unpack_bits(mask, 80000).sum())
    idx = F.FlatIndex(dev(torch, db), "f16")
    # Compacted with the host count, dense masked scan, compacted with the
    # count taken on the device.
    for count in (vc, 0, F.VALID_COUNT_DEVICE):
        out = idx.search(dev(torch, qs), 10, "l2", dev(torch, mask), valid_count=count)
        torch.cuda.synchronize()
        assert_same(L.split_results(out), O.flat_topk(db, qs, 10, "l2", valid_mask=mask))
    dense = S.random_valid_mask(80000, 0.95, 66)          # device count, most rows valid
    out = idx.search(dev(torch, qs), 10, "l2", dev(torch, dense), valid_count=F.VALID_COUNT_DEVICE)
    torch.cuda.synchronize()
    assert_same(L.split_results(out), O.flat_topk(db, qs, 10, "l2", valid_mask=dense))
    odd = S.random_valid_mask(80001, 0.2, 65)[: (80000 + 31) // 32]   # tail word bits past n
    out = idx.search(dev(torch, qs), 10, "l2", dev(torch, odd), valid_count=1)
    torch.cuda.synchronize()
    assert_same(L.split_results(out), O.flat_topk(db, qs, 10, "l2", valid_mask=odd))


def test_per_query_k_and_factor(torch):
    db, qs = S.flat_workload(30000, 260, 128, 64, seeds=(71, 72, 73))
    rng = np.random.default_rng(74)
    ks = rng.integers(1, 65, size=260).astype(np.int32)
    idx = F.FlatIndex(dev(torch, db), "bf16")
    out = idx.search(dev(torch, qs), 64, "ip", shortlist_factor=4,
                     top_k_per_query=dev(torch, ks))
    torch.cuda.synchronize()
    ids, d = L.split_results(out)
    e_ids, e_d = O.flat_topk(db, qs, 64, "ip")
    for i in range(260):
        k = ks[i]
        np.testing.assert_array_equal(ids[i, :k], e_ids[i, :k])
        np.testing.assert_array_equal(d[i, :k].view(np.uint32), e_d[i, :k].view(np.uint32))
        assert (ids[i, k:] == 0xFFFFFFFF).all()


def test_deep_shortlist_workgroup_prunes(torch):
    """Shortlists above 960 rows switch the scan to workgroup-wide prunes."""
    db, qs = S.flat_workload(40000, 70, 64, 64, seeds=(81, 82, 83))
    got = run(torch, db, qs, 200, "l2", "f16", shortlist=2000)
    assert_same(got, O.flat_topk(db, qs, 200, "l2"))


def test_fp16_tc_entry_clamps_and_no_ops(torch):
    """cuda_fp16_tc_search_batch: top_k above 256 is clamped to 256 (the
    output row width); zero rows, zero queries or k = 0 return 0 and leave the
    output untouched (ivf_quant_lookup.cu:1047-1049); an all-zero mask
    yields empty slots."""
    import ctypes
    db, qs = S.flat_workload(3000, 20, 64, 16, seeds=(91, 92, 93))
    codes, norms = O.encode_fp16_padded(db)
    d_codes, d_norms = dev(torch, codes.view(np.int16)), dev(torch, norms)
    d_q, d_db = dev(torch, qs), dev(torch, db)
    wide = F.fp16_tc_search_batch(d_codes, d_norms, d_q, 300, 64, "l2", None, d_db)
    at256 = F.fp16_tc_search_batch(d_codes, d_norms, d_q, 256, 64, "l2", None, d_db)
    torch.cuda.synchronize()
    assert tuple(wide.shape) == (20, 256)
    assert_same(L.split_results(wide), L.split_results(at256))
    assert_same(L.split_results(wide), O.flat_topk(db, qs, 256, "l2"))
    lib = L.load()
    out = L.result_tensor(20, 8, d_q.device)
    out.fill_(0x1234)
    for n, nq, k in [(0, 20, 8), (3000, 0, 8), (3000, 20, 0)]:
        t = L.Fp16GpuTables(L.ptr(d_codes), L.ptr(d_norms), L.ptr(d_db), n, 64, codes.shape[1])
        assert lib.cuda_fp16_tc_search_batch(ctypes.byref(t), L.ptr(d_q), nq, k, 0, None,
                                             L.ptr(out), L.stream_handle()) == 0
    torch.cuda.synchronize()
    assert (out.cpu().numpy() == 0x1234).all()
    zero = np.zeros((3000 + 31) // 32, np.uint32)
    ids, d = L.split_results(F.fp16_tc_search_batch(d_codes, d_norms, d_q, 8, 64, "l2",
                                                    dev(torch, zero), d_db))
    assert (ids == 0xFFFFFFFF).all() and (d == np.float32(1e30)).all()


@pytest.mark.parametrize("dim,metric", [(1024, "l2"), (1536, "l2"), (1000, "ip"), (1536, "cosine"),
                                        (2048, "l2")])
def test_fp16_tc_entry_wide_rows(torch, dim, metric):
    """cuda_fp16_tc_search_batch on rows wider than 768 dims (the reference's
    kernel takes any padded width): 16 queries per wave up to 1536 padded
    dims, the exact FP32 scan beyond. Integer rows: every score is exact."""
    rng = np.random.default_rng(dim)
    n, nq = 3000, 40
    db = rng.integers(-6, 7, size=(n, dim)).astype(np.float32)
    qs = rng.integers(-6, 7, size=(nq, dim)).astype(np.float32)
    codes, norms = O.encode_fp16_padded(db)
    mask = S.random_valid_mask(n, 0.5, dim + 1)
    for m in (None, mask):
        out = F.fp16_tc_search_batch(dev(torch, codes.view(np.int16)), dev(torch, norms),
                                     dev(torch, qs), 10, dim, metric,
                                     None if m is None else dev(torch, m), dev(torch, db))
        torch.cuda.synchronize()
        assert_same(L.split_results(out), O.flat_topk(db, qs, 10, metric, valid_mask=m))


@pytest.mark.parametrize("dtype,dim", [("f16", 1024), ("bf16", 1536), ("fp8", 1024),
                                       ("fp8", 3072)])
def test_wide_rows_flat_search(torch, dtype, dim):
    """fv_flat_search on wide rows (16 or 32 queries per wave), k up to 100."""
    rng = np.random.default_rng(dim + 7)
    n, nq = 4000, 50
    hi = 14 if dtype == "fp8" else 6
    db = rng.integers(-hi, hi + 1, size=(n, dim)).astype(np.float32)
    db[:, 0] = hi
    qs = rng.integers(-hi, hi + 1, size=(nq, dim)).astype(np.float32)
    qs[:, 1] = hi
    idx = F.FlatIndex(dev(torch, db), dtype)
    for k in (1, 100):
        out = idx.search(dev(torch, qs), k, "ip")
        torch.cuda.synchronize()
        assert_same(L.split_results(out), O.flat_topk(db, qs, k, "ip"))


@pytest.mark.parametrize("dtype,k", [("f16", 1), ("f16", 10), ("f16", 64), ("bf16", 10),
                                     ("fp8", 10)])
def test_heavy_ties_exact_prune(torch, dtype, k):
    """Rows copied from 12 distinct integer vectors: thousands of rows share
    each distance, so the scan's buffers overflow with keys tied at the bound
    and are cut by exact (distance, row) selection. Ties go to the lower row.
    Values are exact in every storage type (FP8: row maximum 14, so each
    code is an integer multiple of 32)."""
    rng = np.random.default_rng(71)
    hi = 14 if dtype == "fp8" else 5
    base = rng.integers(-hi, hi + 1, size=(12, 128)).astype(np.float32)
    base[:, 0] = hi
    db = base[rng.integers(0, 12, size=60000)]
    qs = rng.integers(-hi, hi + 1, size=(48, 128)).astype(np.float32)
    qs[:, 1] = hi
    idx = F.FlatIndex(dev(torch, db), dtype)
    for metric in ("l2", "ip"):
        out = idx.search(dev(torch, qs), k, metric)
        torch.cuda.synchronize()
        assert_same(L.split_results(out), O.flat_topk(db, qs, k, metric))


@pytest.mark.parametrize("dtype", ["f16", "bf16"])
@pytest.mark.parametrize("metric", ["l2", "ip"])
@pytest.mark.parametrize("data", ["gauss", "near_dups"])
def test_bounded_rerank_equals_full_rerank(torch, dtype, metric, data):
    """The bounded rerank (only rows whose exact-distance interval reaches the
    k-th upper bound are reranked) returns exactly what reranking the whole
    shortlist returns: ids and FP32 distances, with and without a mask, on
    Gaussian rows and on near-duplicates whose spread is below the codes'
    rounding error."""
    rng = np.random.default_rng(81)
    n, nq, d = 20000, 64, 96
    if data == "gauss":
        db = rng.normal(0.0, 3.0, size=(n, d)).astype(np.float32)
        qs = rng.normal(0.0, 3.0, size=(nq, d)).astype(np.float32)
    else:
        base = rng.normal(0.0, 3.0, size=(40, d))
        db = (base[rng.integers(0, 40, size=n)] + rng.normal(0.0, 1e-3, size=(n, d))).astype(np.float32)
        qs = (base[rng.integers(0, 40, size=nq)] + rng.normal(0.0, 0.5, size=(nq, d))).astype(np.float32)
    idx = F.FlatIndex(dev(torch, db), dtype)
    mask = S.random_valid_mask(n, 0.4, 82)
    for m in (None, mask):
        for k in (1, 10, 32):
            dm = None if m is None else dev(torch, m)
            got = idx.search(dev(torch, qs), k, metric, dm, shortlist_factor=8)
            ref = idx.search(dev(torch, qs), k, metric, dm, shortlist_factor=8, full_rerank=True)
            torch.cuda.synchronize()
            assert_same(L.split_results(got), L.split_results(ref))


@pytest.mark.parametrize("dtype", ["f16", "bf16"])
@pytest.mark.parametrize("metric", ["l2", "ip"])
@pytest.mark.parametrize("data", ["gauss", "near_dups"])
def test_bounded_select_path_equals_full_rerank(torch, dtype, metric, data):
    """Shortlists deeper than the fused tail's (k=64, 8k = 512 rows, the
    select / rerank / finalize kernels): the bounded rerank set chosen in the
    select kernel returns exactly what reranking the whole shortlist returns,
    ids and FP32 distances, with and without a mask, with one k for all
    queries and per-query k. With the FP64 boundary the ids and their order
    are the same; a distance may be the FP64 value rounded to FP32 in one call
    and the FP32 rerank in the other (the near-tie band depends on the
    reranked rows), so distances agree to the FP32 error bound there."""
    rng = np.random.default_rng(91)
    n, nq, d = 30000, 96, 128
    if data == "gauss":
        db = rng.normal(0.0, 3.0, size=(n, d)).astype(np.float32)
        qs = rng.normal(0.0, 3.0, size=(nq, d)).astype(np.float32)
    else:
        base = rng.normal(0.0, 3.0, size=(40, d))
        db = (base[rng.integers(0, 40, size=n)] + rng.normal(0.0, 1e-3, size=(n, d))).astype(np.float32)
        qs = (base[rng.integers(0, 40, size=nq)] + rng.normal(0.0, 0.5, size=(nq, d))).astype(np.float32)
    idx = F.FlatIndex(dev(torch, db), dtype)
    mask = S.random_valid_mask(n, 0.4, 92)
    per_q = dev(torch, rng.integers(1, 65, size=nq).astype(np.int32))
    for m in (None, mask):
        dm = None if m is None else dev(torch, m)
        for f64 in (False, True):
            for kq in (None, per_q):
                got = idx.search(dev(torch, qs), 64, metric, dm, shortlist_factor=8,
                                 top_k_per_query=kq, f64_boundary=f64)
                ref = idx.search(dev(torch, qs), 64, metric, dm, shortlist_factor=8,
                                 top_k_per_query=kq, f64_boundary=f64, full_rerank=True)
                torch.cuda.synchronize()
                g, r = L.split_results(got), L.split_results(ref)
                if not f64:
                    assert_same(g, r)
                else:
                    np.testing.assert_array_equal(g[0], r[0])
                    scale = np.abs(r[1]).max() + 1.0
                    np.testing.assert_allclose(g[1], r[1], rtol=0, atol=1e-5 * scale)


def test_bounded_rerank_legacy_entry_without_encoding_error(torch):
    """cuda_fp16_tc_search_batch has no per-row encoding error (the bound
    falls back to 2^-10 |x|): same results as the full rerank."""
    rng = np.random.default_rng(83)
    db = rng.normal(0.0, 2.0, size=(15000, 64)).astype(np.float32)
    qs = rng.normal(0.0, 2.0, size=(50, 64)).astype(np.float32)
    codes, norms = O.encode_fp16_padded(db)
    idx = F.FlatIndex(dev(torch, db), "f16")
    for metric in ("l2", "ip"):
        got = F.fp16_tc_search_batch(dev(torch, codes.view(np.int16)), dev(torch, norms),
                                     dev(torch, qs), 10, 64, metric, None, dev(torch, db))
        ref = idx.search(dev(torch, qs), 10, metric, shortlist_factor=8, full_rerank=True)
        torch.cuda.synchronize()
        assert_same(L.split_results(got), L.split_results(ref))
