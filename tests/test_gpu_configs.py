"""GPU parity of the benchmark workloads as configured, at row counts the
oracle finishes in seconds:

* 1M x 128 integer mixture (256 clusters), 5% validity mask, FP16 and BF16
  storage, k = 10 and 100, 8k shortlist: bit-exact against the oracle's exact
  FP32 top-k over the allowed rows, on a 1,000-query slice of the 10k batch.
* IVF-PQ over nlist 4096 Zipf(1.1) list sizes, d = 128, M = 16 x 256, uneven
  per-query nprobe in {8, 16, 32, 64} and k in {10, 100}, shortlist 1000
  with FP32 rerank, at 400k rows: bit-exact against the oracle's ADC for both
  code layouts, with queries that probe the largest (head) list.
* FP8 inner product over 1,024-cluster vMF rows (kappa 50) at d = 768, k
  log-uniform on [1, 1024], 4k shortlist, FP64 boundary, the whole batch one
  captured HIP graph, at 100k rows: recall@k against the FP64 oracle, every
  returned distance the exact FP32 or FP64-rounded value of its row, FP64
  order.
"""
import numpy as np
import pytest

import oracle as O
from flashvector_amd import ivfpq_workload as W
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


# --------------------------------------------------------------------------
# 1M x 128, 5% mask, FP16 / BF16, k 10 / 100
# --------------------------------------------------------------------------

@pytest.fixture(scope="module")
def mixed_1m(torch):
    seeds = S.SEEDS["flat_mixed_1m"]
    db, qs = S.flat_workload(1_000_000, 10_000, 128, 256, seeds=seeds)
    mask = S.random_valid_mask(db.shape[0], 0.05, seeds[2])
    valid = int(np.unpackbits(mask.view(np.uint8)).sum())
    q = qs[:1000]
    d_db = dev(torch, db)
    return db, q, mask, valid, d_db, dev(torch, q), dev(torch, mask)


@pytest.mark.parametrize("dtype,k", [("f16", 100), ("bf16", 10), ("bf16", 100)])
def test_mixed_1m_mask5_exact(torch, mixed_1m, dtype, k):
    db, q, mask, valid, d_db, d_q, d_m = mixed_1m
    idx = F.FlatIndex(d_db, dtype)
    out = idx.search(d_q, k, "l2", d_m, valid_count=valid, shortlist_factor=8)
    torch.cuda.synchronize()
    got = L.split_results(out)
    exp = O.flat_topk(db, q, k, "l2", valid_mask=mask)
    assert_same(got, exp)
    bits = S.unpack_bits(mask, db.shape[0])
    assert bits[got[0].ravel()].all()


# --------------------------------------------------------------------------
# IVF-PQ, nlist 4096 Zipf(1.1), uneven batch
# --------------------------------------------------------------------------

@pytest.fixture(scope="module")
def zipf_ivf(torch):
    w = W.build(n=400_000, dim=128, nlist=4096, zipf_s=1.1, m_pq=16, nq=512, train=65536,
                iters=6)
    torch.cuda.synchronize()

    def h(t):
        return t.cpu().numpy()
    host = O.IvfPqHost(h(w.centroids), h(w.codebooks).reshape(16, 256, 8), h(w.codes),
                       h(w.offsets).view(np.uint32), h(w.vec_ids).view(np.uint32),
                       h(w.vectors), 16)
    idx = F.IvfPqIndex(w.centroids, w.codebooks, w.codes, w.offsets, w.vec_ids, w.vectors)
    return w, host, idx


def test_zipf_workload_shape(torch, zipf_ivf):
    """The list sizes follow Zipf(1.1) over 4096 lists (largest remainder
    to the exact total) and some queries probe the head list."""
    w, host, idx = zipf_ivf
    sizes = w.sizes
    assert sizes.sum() == 400_000 and len(sizes) == 4096
    assert sizes.max() == W.zipf_sizes(400_000, 4096, 1.1, 22).max()
    head = int(np.argmax(sizes))
    qs = w.queries.cpu().numpy()
    nps = w.nprobe.cpu().numpy()
    n_head = 0
    for i in range(qs.shape[0]):
        probes, _ = O.coarse_probe(qs[i], host.centroids, int(nps[i]))
        n_head += head in set(int(p) for p in probes)
    assert n_head >= 8, n_head


@pytest.mark.parametrize("list_order", [True, False])
def test_zipf_uneven_batch_exact(torch, zipf_ivf, list_order):
    w, host, idx = zipf_ivf
    out = idx.search(w.queries, 100, 64, 1000, top_k=w.top_k, nprobe=w.nprobe,
                     use_list_order=list_order)
    torch.cuda.synchronize()
    got = L.split_results(out)
    ks = w.top_k.cpu().numpy().astype(np.uint32)
    nps = w.nprobe.cpu().numpy().astype(np.uint32)
    exp = O.ivf_pq_search_uneven(host, w.queries.cpu().numpy(), ks, nps, 1000, out_stride=100)
    assert_same(got, exp)


# --------------------------------------------------------------------------
# FP8 IP, 1024-cluster vMF, d 768, k log-uniform to 1024, FP64 boundary
# --------------------------------------------------------------------------

def test_vmf_fp8_mixed_k_graph_f64(torch):
    nq = 128
    db, qs = S.vmf_workload(100_000, nq, 768, 1024, 50.0, device="cuda")
    ks = S.log_uniform_k(nq, 1024, S.VMF_SEEDS[3])
    d_k = dev(torch, ks)
    idx = F.FlatIndex(db, "fp8")
    out = L.result_tensor(nq, 1024, qs.device)

    def step():
        idx.search(qs, 1024, "ip", shortlist_factor=4, top_k_per_query=d_k, f64_boundary=True,
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
    for r in range(2):
        out.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, eager), f"replay {r}"
    ids, d = L.split_results(eager)
    dbn, qsn = db.cpu().numpy(), qs.cpu().numpy()
    e_ids, _ = O.flat_topk_f64(dbn, qsn, 1024, "ip")
    hits = total = 0
    for i in range(nq):
        k = int(ks[i])
        got = ids[i, :k]
        assert (ids[i, k:] == 0xFFFFFFFF).all()
        assert (got != 0xFFFFFFFF).all()
        hits += len(set(got.tolist()) & set(e_ids[i, :k].tolist()))
        total += k
        rows = dbn[got.astype(np.int64)]
        f32 = O.distance_matrix(qsn[i:i + 1], rows, "ip")[0]
        f64 = -(rows.astype(np.float64) @ qsn[i].astype(np.float64))
        # Each distance is its row's exact FP32 value, or (re-scored at the
        # boundary) its FP64 value rounded to FP32.
        ok = (d[i, :k].view(np.uint32) == f32.view(np.uint32)) | \
             (d[i, :k] == f64.astype(np.float32))
        assert ok.all(), i
        # Returned in exact (FP64) order, ties by id: neighbours closer than
        # the FP32 error band were re-scored, the others are further apart
        # than any FP32 error.
        assert np.all((f64[:-1] < f64[1:]) | ((f64[:-1] == f64[1:]) & (got[:-1] < got[1:]))), i
    recall = hits / total
    print(f"fp8 vmf 100k d768 mixed k: recall@k {recall:.5f} over {total} results")
    assert recall >= 0.99
