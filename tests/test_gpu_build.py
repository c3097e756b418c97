"""GPU index construction (fv_build.h) against the oracle: bit-exact
assignments, centroids, sizes, iteration counts, codebooks, codes and
posting lists on the same inputs and seed."""
import numpy as np
import pytest

import oracle as O
from flashvector_amd import gpu_build as G
from flashvector_amd import lib as L

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


def host(t):
    return t.cpu().numpy()


@pytest.mark.parametrize("n,dim,k", [(1000, 8, 256), (777, 13, 7), (3000, 128, 300), (65, 3, 1)])
def test_assign_matches_oracle(torch, n, dim, k):
    rng = np.random.default_rng(n + k)
    x = rng.normal(size=(n, dim)).astype(np.float32)
    c = rng.normal(size=(k, dim)).astype(np.float32)
    if k > 3:
        c[k - 1] = c[1]          # duplicate: ties to the lower index
    got = host(G.kmeans_assign(dev(torch, x), dev(torch, c))).view(np.uint32)
    np.testing.assert_array_equal(got, O.kmeans_assign(x, c))


def check_fit(torch, x, k, iters, seed):
    g = G.kmeans_fit(dev(torch, x), k, iters, seed)
    c, a, s, it = O.kmeans_fit(x, k, iters, seed)
    np.testing.assert_array_equal(host(g.centroids).view(np.uint32), c.view(np.uint32))
    np.testing.assert_array_equal(host(g.assignments).view(np.uint32), a)
    np.testing.assert_array_equal(host(g.sizes).view(np.uint32), s)
    assert g.iterations == it
    off, mem = host(g.offsets).view(np.uint32), host(g.members).view(np.uint32)
    exp_off = np.concatenate([[0], np.cumsum(np.bincount(a, minlength=k))]).astype(np.uint32)
    np.testing.assert_array_equal(off, exp_off)
    np.testing.assert_array_equal(mem, np.argsort(a, kind="stable").astype(np.uint32))
    return g


def test_fit_matches_oracle_normal(torch):
    x = np.random.default_rng(21).normal(size=(5000, 16)).astype(np.float32)
    check_fit(torch, x, 32, 25, 7)


def test_fit_matches_oracle_reference_bench_config(torch):
    """The reference's k-means benchmark input (benches/bench_kmeans.rs:5-14):
    10k x 128 rows of 0.001 * i, k = 64, 5 iterations."""
    x = (np.arange(10000 * 128, dtype=np.float32) * np.float32(0.001)).reshape(10000, 128)
    check_fit(torch, x, 64, 5, 1)


def test_fit_matches_oracle_with_empty_clusters(torch):
    base = np.random.default_rng(22).integers(-9, 10, size=(10, 4)).astype(np.float32)
    x = np.repeat(base, 30, axis=0)
    g = check_fit(torch, x, 12, 5, 4)
    assert (host(g.sizes) == 0).any()


def test_fit_zero_iterations_returns_seeds(torch):
    x = np.random.default_rng(23).normal(size=(900, 5)).astype(np.float32)
    g = check_fit(torch, x, 40, 0, 9)
    assert g.iterations == 0


def test_build_matches_oracle_and_searches(torch):
    from flashvector_amd import search as F
    x = (np.random.default_rng(24).normal(size=(6000, 32)) * 10).astype(np.float32)
    t = G.build_ivfpq(dev(torch, x), 16, 4, coarse_iters=10, pq_iters=8, seed=5)
    o = O.ivf_pq_build(x, 16, 4, 10, 8, 5)
    np.testing.assert_array_equal(host(t.centroids).view(np.uint32), o["centroids"].view(np.uint32))
    np.testing.assert_array_equal(host(t.codebooks).view(np.uint32), o["codebooks"].view(np.uint32))
    np.testing.assert_array_equal(host(t.codes), o["codes"])
    np.testing.assert_array_equal(host(t.offsets).view(np.uint32), o["offsets"])
    np.testing.assert_array_equal(host(t.vec_ids).view(np.uint32), o["vec_ids"])
    assert t.iterations == o["iters"].tolist()
    # The built tables feed the scan directly.
    idx = t.index(dev(torch, x))
    q = x[:50] + 0.25
    got = L.split_results(idx.search_batch(dev(torch, q), 10, 4, 200))
    h = O.IvfPqHost(o["centroids"], o["codebooks"], o["codes"], o["offsets"], o["vec_ids"], x, 4)
    exp = O.ivf_pq_search_batch(h, q, 10, 4, 200)
    np.testing.assert_array_equal(got[0], exp[0])
    np.testing.assert_array_equal(got[1].view(np.uint32), exp[1].view(np.uint32))


def test_build_small_n(torch):
    x = np.random.default_rng(25).integers(-3, 4, size=(100, 8)).astype(np.float32)
    t = G.build_ivfpq(dev(torch, x), 8, 2, coarse_iters=4, pq_iters=4, seed=1)
    o = O.ivf_pq_build(x, 8, 2, 4, 4, 1)
    np.testing.assert_array_equal(host(t.codebooks).view(np.uint32), o["codebooks"].view(np.uint32))
    np.testing.assert_array_equal(host(t.codes), o["codes"])
    np.testing.assert_array_equal(host(t.vec_ids).view(np.uint32), o["vec_ids"])
