"""GPU parity of the FP8 e4m3 matrix-core scan, of the 16-bit scan at large
dims (32 queries per wave) and of the FP64 boundary re-scoring.

FP8 rows with integer entries in [-14, 14] and one |x| = 14 per row have
scale 1/32 and encode exactly, so their MFMA scores are exact and the
reranked output must equal the oracle's exact FP32 top-k bit for bit. On
real-valued unit vectors the bar is recall against the FP64 oracle (FP8
keeps 3 mantissa bits) plus exact FP32 distances for every returned row."""
import numpy as np
import pytest

import oracle as O
from flashvector_amd import lib as L
from flashvector_amd import search as F
from flashvector_amd import synthetic as S

pytestmark = pytest.mark.gpu


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


def int14(n, dim, seed, sign=1.0):
    """Integer rows in [-14, 14] with a +-14 in every row (exact FP8 at scale 1/32)."""
    rng = np.random.default_rng(seed)
    x = rng.integers(-14, 15, size=(n, dim)).astype(np.float32)
    x[np.arange(n), rng.integers(0, dim, size=n)] = 14.0 * sign
    return x


def search(torch, db, qs, k, metric, dtype="fp8", keep_fp32=True, **kw):
    idx = F.FlatIndex(dev(torch, db), dtype, keep_fp32=keep_fp32)
    out = idx.search(dev(torch, qs), k, metric, **kw)
    torch.cuda.synchronize()
    return L.split_results(out)


def test_fp8_encoder_matches_oracle(torch):
    rng = np.random.default_rng(1)
    x = (rng.normal(size=(300, 200)) * rng.uniform(0.01, 50, size=(300, 1))).astype(np.float32)
    x[7] = 0.0
    idx = F.FlatIndex(dev(torch, x), "fp8")
    torch.cuda.synchronize()
    codes, scales = O.encode_fp8_rows(x, pdim=idx.padded_dim)
    np.testing.assert_array_equal(idx.codes.cpu().numpy(), codes)
    np.testing.assert_array_equal(idx.row_scales.cpu().numpy().view(np.uint32), scales.view(np.uint32))


@pytest.mark.parametrize("n,nq,dim,k,metric", [
    (6000, 70, 64, 10, "l2"),       # one K step of 128 (padded)
    (6000, 70, 200, 10, "ip"),      # two K steps
    (5000, 40, 384, 25, "l2"),      # four K steps
    (5000, 300, 768, 10, "ip"),     # six K steps, several query blocks
    (5000, 33, 768, 100, "l2"),     # deep shortlist
])
def test_fp8_integer_data_exact(torch, n, nq, dim, k, metric):
    db = int14(n, dim, n + dim)
    qs = int14(nq, dim, nq + 7)
    assert_same(search(torch, db, qs, k, metric), O.flat_topk(db, qs, k, metric))


def test_fp8_scores_without_rerank(torch):
    """keep_fp32=False returns the FP8 scores: equal to the oracle's
    c_row + (s_row * s_q) * acc on exactly representable data."""
    db = int14(500, 256, 3)
    qs = int14(20, 256, 4)
    ids, d = search(torch, db, qs, 500, "ip", keep_fp32=False)
    s = O.fp8_scores(db, qs, "ip")
    for i in range(20):
        order = np.lexsort((np.arange(500), s[i]))
        np.testing.assert_array_equal(ids[i], order)
        np.testing.assert_array_equal(d[i], s[i][order])


def test_fp8_mixed_k_masked(torch):
    db = int14(20000, 768, 11)
    qs = int14(260, 768, 12)
    ks = S.log_uniform_k(260, 128, 13)
    mask = S.random_valid_mask(20000, 0.5, 14)
    idx = F.FlatIndex(dev(torch, db), "fp8")
    out = idx.search(dev(torch, qs), 128, "ip", dev(torch, mask), shortlist_factor=4,
                     top_k_per_query=dev(torch, ks))
    torch.cuda.synchronize()
    ids, d = L.split_results(out)
    e_ids, e_d = O.flat_topk(db, qs, 128, "ip", valid_mask=mask)
    for i in range(260):
        k = ks[i]
        np.testing.assert_array_equal(ids[i, :k], e_ids[i, :k])
        np.testing.assert_array_equal(d[i, :k].view(np.uint32), e_d[i, :k].view(np.uint32))
        assert (ids[i, k:] == 0xFFFFFFFF).all()


def test_fp8_vmf_recall_and_exact_distances(torch):
    """Unit vMF rows at d=768: every returned distance is the exact FP32
    distance of its row; recall@10 against the FP64 oracle with an 8x
    shortlist is high (FP8 noise only reorders near-ties)."""
    db, qs = S.vmf_workload(30000, 100, 768, 64, 50.0, device="cuda")
    dbn, qsn = db.cpu().numpy(), qs.cpu().numpy()
    idx = F.FlatIndex(db, "fp8")
    out = idx.search(qs, 10, "ip", shortlist_factor=8)
    torch.cuda.synchronize()
    ids, d = L.split_results(out)
    e_ids, _ = O.flat_topk_f64(dbn, qsn, 10, "ip")
    hits = sum(len(set(ids[i]) & set(e_ids[i])) for i in range(100))
    assert hits / 1000 >= 0.97
    exact = O.distance_matrix(qsn, dbn, "ip")
    np.testing.assert_array_equal(d, np.take_along_axis(exact, ids.astype(np.int64), 1))


@pytest.mark.parametrize("dtype", ["f16", "bf16"])
@pytest.mark.parametrize("dim", [384, 500, 768])
def test_16bit_large_dims_exact(torch, dtype, dim):
    """Widths of 12, 16 and 24 K steps (32 queries per wave)."""
    rng = np.random.default_rng(dim)
    db = rng.integers(0, 32, size=(4000, dim)).astype(np.float32)
    qs = rng.integers(0, 32, size=(150, dim)).astype(np.float32)
    assert_same(search(torch, db, qs, 20, "l2", dtype), O.flat_topk(db, qs, 20, "l2"))


def _tie_case(metric, dim=64):
    """Rows 0 and 1 tie in FP32 but not in exact arithmetic (row 1 is nearer),
    among 2000 far-away rows."""
    rng = np.random.default_rng(7)
    db = (rng.normal(size=(2002, dim)) * 1e-3).astype(np.float32)
    q = np.zeros((1, dim), np.float32)
    if metric == "ip":
        q[0, :2] = 1.0
        db[0] = 0.0
        db[0, 0] = 1.0                      # dot 1
        db[1] = 0.0
        db[1, 0], db[1, 1] = 1.0, 2.0 ** -30  # dot 1 + 2^-30, rounds to 1 in FP32
    else:
        q[0, 0] = 1.0
        db[0] = 0.0
        db[0, 1] = 2.0 ** -13                 # |q-x|^2 = 1 + 2^-26, rounds to 1
        db[1] = 0.0                           # exactly 1
        db[2:, 0] += 10.0
    return db, q


@pytest.mark.parametrize("dtype", ["f16", "fp8"])
@pytest.mark.parametrize("metric", ["ip", "l2"])
def test_f64_boundary_breaks_fp32_ties(torch, dtype, metric):
    db, q = _tie_case(metric)
    ids, d = search(torch, db, q, 1, metric, dtype, shortlist=64)
    assert ids[0, 0] == 0                     # FP32 tie -> lower id
    ids, d = search(torch, db, q, 2, metric, dtype, shortlist=64, f64_boundary=True)
    np.testing.assert_array_equal(ids[0], [1, 0])
    e_ids, e_d = O.flat_topk_f64(db, q, 2, metric)
    np.testing.assert_array_equal(ids[0], e_ids[0])
    np.testing.assert_array_equal(d[0], e_d[0].astype(np.float32))


@pytest.mark.parametrize("metric", ["ip", "l2", "cosine"])
def test_f64_boundary_equals_fp64_oracle(torch, metric):
    """Shortlist >= N: every row is reranked, so with FP64 boundary
    re-scoring the output is the FP64 top-k exactly; without it, the exact
    FP32 top-k."""
    db, qs = S.vmf_workload(3000, 50, 128, 16, 50.0, device="cuda")
    dbn, qsn = db.cpu().numpy(), qs.cpu().numpy()
    idx = F.FlatIndex(db, "bf16")
    out = idx.search(qs, 40, metric, shortlist=3072, f64_boundary=True)
    torch.cuda.synchronize()
    ids, d = L.split_results(out)
    e_ids, e_d = O.flat_topk_f64(dbn, qsn, 40, metric)
    np.testing.assert_array_equal(ids, e_ids)
    out = idx.search(qs, 40, metric, shortlist=3072)
    torch.cuda.synchronize()
    assert_same(L.split_results(out), O.flat_topk(dbn, qsn, 40, metric))


@pytest.mark.parametrize("f64_boundary", [False, True])
def test_fp8_deep_mixed_k_exact(torch, f64_boundary):
    """k log-uniform on [1, 1024] with a 4k shortlist (up to 4096 rows: the
    first pass takes 8192 subset minima per query, the scan prunes by
    workgroup). Integer data: equal to the oracle for every query."""
    db = int14(60000, 256, 21)
    qs = int14(100, 256, 22)
    ks = S.log_uniform_k(100, 1024, 23)
    idx = F.FlatIndex(dev(torch, db), "fp8")
    out = idx.search(dev(torch, qs), 1024, "ip", shortlist_factor=4,
                     top_k_per_query=dev(torch, ks), f64_boundary=f64_boundary)
    torch.cuda.synchronize()
    ids, d = L.split_results(out)
    e_ids, e_d = O.flat_topk(db, qs, 1024, "ip")
    for i in range(100):
        k = ks[i]
        np.testing.assert_array_equal(ids[i, :k], e_ids[i, :k])
        np.testing.assert_array_equal(d[i, :k].view(np.uint32), e_d[i, :k].view(np.uint32))
        assert (ids[i, k:] == 0xFFFFFFFF).all()


@pytest.mark.parametrize("dtype,metric,f64", [("fp8", "ip", True), ("fp8", "ip", False),
                                               ("bf16", "l2", False)])
def test_per_query_buffers_equal_uniform(torch, dtype, metric, f64):
    """Deep shortlists with per-query k: with the k sum given, candidate
    buffers are sized per query (far smaller workspace) and the results are
    those of the deepest-shortlist buffers, bit for bit; on integer data
    both equal the oracle."""
    db = int14(60000, 256, 41)
    qs = int14(300, 256, 42)
    ks = S.log_uniform_k(300, 1024, 43)
    idx = F.FlatIndex(dev(torch, db), dtype)
    d_q, d_k = dev(torch, qs), dev(torch, ks)
    kw = dict(shortlist_factor=4, top_k_per_query=d_k, f64_boundary=f64)
    ws_u = idx.workspace_size(300, idx.params(1024, metric, **kw))
    ws_p = idx.workspace_size(300, idx.params(1024, metric, top_k_sum=int(ks.sum()), **kw))
    assert 0 < ws_p < ws_u / 2, (ws_p, ws_u)
    uni = L.split_results(idx.search(d_q, 1024, metric, **kw))
    per = L.split_results(idx.search(d_q, 1024, metric, top_k_sum=int(ks.sum()), **kw))
    torch.cuda.synchronize()
    assert_same(per, uni)
    e_ids, e_d = O.flat_topk(db, qs, 1024, metric)
    for i in range(300):
        k = ks[i]
        np.testing.assert_array_equal(per[0][i, :k], e_ids[i, :k])
        np.testing.assert_array_equal(per[1][i, :k].view(np.uint32), e_d[i, :k].view(np.uint32))


def test_per_query_buffers_fixed_shortlist(torch):
    """A fixed shortlist (2048 rows, at least k_i) with per-query k: the
    per-query buffers take the shortlist rule's bound on the k sum and
    return the uniform buffers' results."""
    db = int14(50000, 256, 61)
    qs = int14(128, 256, 62)
    ks = S.log_uniform_k(128, 1024, 63)
    idx = F.FlatIndex(dev(torch, db), "fp8")
    d_q, d_k = dev(torch, qs), dev(torch, ks)
    kw = dict(shortlist=2048, top_k_per_query=d_k)
    ws_u = idx.workspace_size(128, idx.params(1024, "ip", **kw))
    ws_p = idx.workspace_size(128, idx.params(1024, "ip", top_k_sum=int(ks.sum()), **kw))
    assert 0 < ws_p < ws_u, (ws_p, ws_u)
    uni = L.split_results(idx.search(d_q, 1024, "ip", **kw))
    per = L.split_results(idx.search(d_q, 1024, "ip", top_k_sum=int(ks.sum()), **kw))
    torch.cuda.synchronize()
    assert_same(per, uni)


def test_per_query_buffers_understated_sum(torch):
    """A k sum below the true one: the buffers laid out within the planned
    total are whole, the queries past it return no rows (every slot empty),
    and nothing is written out of bounds (the queries that fit equal the
    uniform-buffer results)."""
    db = int14(40000, 256, 51)
    qs = int14(200, 256, 52)
    ks = S.log_uniform_k(200, 1024, 53)
    idx = F.FlatIndex(dev(torch, db), "fp8")
    d_q, d_k = dev(torch, qs), dev(torch, ks)
    kw = dict(shortlist_factor=4, top_k_per_query=d_k)
    uni = L.split_results(idx.search(d_q, 1024, "ip", **kw))
    per = L.split_results(idx.search(d_q, 1024, "ip", top_k_sum=max(1, int(ks.sum()) // 8), **kw))
    torch.cuda.synchronize()
    empty = (per[0] == 0xFFFFFFFF).all(axis=1)
    assert empty.any() and not empty.all()
    first = int(np.argmax(empty))
    assert empty[first:].all(), "queries past the planned total form a suffix"
    np.testing.assert_array_equal(per[0][:first], uni[0][:first])
    np.testing.assert_array_equal(per[1][:first].view(np.uint32), uni[1][:first].view(np.uint32))


def test_graph_replay_matches_eager(torch):
    """A whole mixed-k FP8 batch captured as one HIP graph: every replay,
    with other eager searches and synchronizations in between, reproduces
    the eager result exactly."""
    db = int14(30000, 768, 31)
    qs = int14(200, 768, 32)
    ks = dev(torch, S.log_uniform_k(200, 256, 33))
    idx = F.FlatIndex(dev(torch, db), "fp8")
    other = F.FlatIndex(dev(torch, int14(5000, 768, 34)), "fp8")
    d_q = dev(torch, qs)
    out = L.result_tensor(200, 256, d_q.device)

    def step():
        idx.search(d_q, 256, "ip", shortlist_factor=4, top_k_per_query=ks, f64_boundary=True,
                   out=out)

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
    for r in range(3):
        out.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, eager), f"replay {r}"
        other.search(d_q, 16, "ip")
        torch.cuda.synchronize()
    e_ids, _ = O.flat_topk(db, qs, 256, "ip")
    ids, _ = L.split_results(eager)
    k = ks.cpu().numpy()
    for i in range(200):
        np.testing.assert_array_equal(ids[i, :k[i]], e_ids[i, :k[i]])


def test_autotune_meets_recall_and_keeps_results(torch, tmp_path, monkeypatch):
    """fv_flat_autotune picks a shortlist factor that meets the recall target
    on its sample; searching with the tuned plan on integer data still equals
    the oracle; the cache returns the same tuning without re-measuring."""
    from flashvector_amd import autotune as T
    monkeypatch.setenv("FLASHVECTOR_TUNE_CACHE", str(tmp_path / "tune.json"))
    db, qs = S.flat_workload(60000, 256, 128, 64, seeds=(101, 102, 103))
    idx = F.FlatIndex(dev(torch, db), "f16")
    d_q = dev(torch, qs)
    best = T.tuned(idx, d_q, 10, "l2")
    assert best.recall >= 0.999 and best.shortlist_factor in (1, 2, 4, 8, 16)
    assert best.ms > 0   # n_chunks / bound_div 0 = the automatic plan was kept
    out = idx.search(d_q, 10, "l2", tuning=best)
    torch.cuda.synchronize()
    assert_same(L.split_results(out), O.flat_topk(db, qs, 10, "l2"))
    again = T.tuned(idx, d_q, 10, "l2")
    assert (again.shortlist_factor, again.n_chunks, again.bound_div) == \
        (best.shortlist_factor, best.n_chunks, best.bound_div)
