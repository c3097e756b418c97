"""Pins of the oracle's IVF-PQ table construction (k-means, PQ train/encode,
posting lists; crates/engine/src/cluster.rs:28-287, index.rs:405-458) by
brute force and closed forms on data where FP32 arithmetic is exact."""
import numpy as np

import oracle as O


def int_rows(n, dim, lo, hi, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(lo, hi, size=(n, dim)).astype(np.float32)


def test_assign_is_first_strict_minimum_brute_force():
    x = int_rows(700, 9, -20, 21, 1)
    c = int_rows(40, 9, -20, 21, 2)
    c[17] = c[5]                     # duplicate centroid: ties go to the lower index
    d = ((x[:, None, :].astype(np.float64) - c[None]) ** 2).sum(-1)   # exact integers
    np.testing.assert_array_equal(O.kmeans_assign(x, c), d.argmin(1))
    # Rows read through a stride: the first 4 columns of a wider matrix.
    xs = np.concatenate([x[:, :4], x[:, 4:]], axis=1)
    d4 = ((xs[:, None, :4].astype(np.float64) - c[None, :, :4]) ** 2).sum(-1)
    np.testing.assert_array_equal(O.kmeans_assign(xs, np.ascontiguousarray(c[:, :4]), dim=4),
                                  d4.argmin(1))


def test_assign_never_picks_infinite_distances():
    x = int_rows(50, 4, -5, 6, 3)
    c = np.full((3, 4), np.finfo(np.float32).max, np.float32)
    c[2] = 0.0
    assert (O.kmeans_assign(x, c) == 2).all()
    c[2] = np.finfo(np.float32).max
    assert (O.kmeans_assign(x, c) == 0).all()    # none finite: cluster 0 (cluster.rs:109-121)


def test_one_lloyd_step_matches_numpy():
    """max_iters=0 returns the seeds; one more iteration must be assign
    (argmin) then centroid = float32 sum * float32(1/count), all exact on
    small integers."""
    x = int_rows(2000, 6, -50, 51, 4)
    k = 24
    seeds, _, sizes0, it0 = O.kmeans_fit(x, k, 0, 77)
    assert it0 == 0 and (sizes0 == 0).all()
    c1, a1, s1, it1 = O.kmeans_fit(x, k, 1, 77)
    d = ((x[:, None, :].astype(np.float64) - seeds[None]) ** 2).sum(-1)
    a = d.argmin(1)
    np.testing.assert_array_equal(a1, a)
    np.testing.assert_array_equal(s1, np.bincount(a, minlength=k))
    assert (s1 > 0).all()
    for c in range(k):
        sm = x[a == c].astype(np.float64).sum(0).astype(np.float32)   # exact: |sum| < 2^24
        exp = sm * (np.float32(1.0) / np.float32(s1[c]))
        np.testing.assert_array_equal(c1[c], exp)
    assert it1 == 1


def test_separated_blobs_give_exact_means():
    """Blobs 1e4 apart with power-of-two sizes: k-means++ seeds one row per
    blob, Lloyd converges to the blob means, which are exact in FP32."""
    rng = np.random.default_rng(5)
    sizes = [64, 128, 256, 32, 512, 16]
    rows, labels = [], []
    for b, s in enumerate(sizes):
        center = rng.integers(-3, 4, size=8) * 10000
        rows.append(center + rng.integers(-4, 5, size=(s, 8)))
        labels += [b] * s
    x = np.concatenate(rows).astype(np.float32)
    labels = np.array(labels)
    cent, assign, sz, it = O.kmeans_fit(x, len(sizes), 25, 3)
    assert it < 25
    for c in range(len(sizes)):
        members = labels[assign == c]
        assert len(set(members.tolist())) == 1
        b = members[0]
        np.testing.assert_array_equal(cent[c], x[labels == b].astype(np.float64).mean(0).astype(np.float32))
        assert sz[c] == sizes[b]


def test_seeding_with_k_equal_n_takes_every_row():
    x = int_rows(300, 5, -1000, 1000, 6)
    assert len({tuple(r) for r in x.tolist()}) == 300
    seeds, _, _, _ = O.kmeans_fit(x, 300, 0, 11)
    assert sorted(map(tuple, seeds.tolist())) == sorted(map(tuple, x.tolist()))


def test_fit_is_deterministic_and_seed_dependent():
    x = np.random.default_rng(7).normal(size=(1500, 12)).astype(np.float32)
    a = O.kmeans_fit(x, 20, 10, 1)
    b = O.kmeans_fit(x, 20, 10, 1)
    c = O.kmeans_fit(x, 20, 10, 2)
    np.testing.assert_array_equal(a[0], b[0])
    assert not np.array_equal(a[0], c[0])


def test_empty_clusters_are_refilled_with_data_rows():
    """10 distinct rows repeated, k = 12: seeding runs out of positive D^2,
    so some clusters start empty and are refilled with drawn rows."""
    base = int_rows(10, 4, -9, 10, 8)
    x = np.repeat(base, 30, axis=0)
    cent, assign, sz, it = O.kmeans_fit(x, 12, 5, 4)
    assert sz.sum() == x.shape[0]
    rows = {tuple(r) for r in base.tolist()}
    for c in range(12):
        if sz[c] == 0:
            assert tuple(cent[c].tolist()) in rows


def test_pq_encode_maps_codewords_to_themselves():
    rng = np.random.default_rng(9)
    book = rng.normal(size=(256, 8)).astype(np.float32)
    np.testing.assert_array_equal(O.kmeans_assign(book, book), np.arange(256))


def test_build_tables_structure_and_search():
    x = np.random.default_rng(10).normal(size=(3000, 32)).astype(np.float32) * 10
    t = O.ivf_pq_build(x, 16, 4, 10, 8, 5)
    off, ids = t["offsets"], t["vec_ids"]
    assert off[0] == 0 and off[-1] == 3000 and (np.diff(off.astype(np.int64)) >= 0).all()
    assert sorted(ids.tolist()) == list(range(3000))
    list_of = np.empty(3000, np.int64)
    for c in range(16):
        seg = ids[off[c]:off[c + 1]]
        assert (np.diff(seg.astype(np.int64)) > 0).all()
        list_of[seg] = c
    # Codes: nearest codeword of each residual sub-vector against its list's centroid.
    res = x - t["centroids"][list_of]
    sub = 8
    for m in range(4):
        book = t["codebooks"][m].reshape(256, sub)
        np.testing.assert_array_equal(
            t["codes"][:, m], O.kmeans_assign(np.ascontiguousarray(res[:, m * sub:(m + 1) * sub]), book))
    assert t["iters"][0] >= 1
    # The tables drive the scan: exhaustive probing equals the flat scan.
    h = O.IvfPqHost(t["centroids"], t["codebooks"], t["codes"], off, ids, x, 4)
    q = x[:20] + 0.5
    got = O.ivf_pq_search_batch(h, q, 10, 16, 3000)
    flat = O.flat_scan_batch(x, q, 10)
    np.testing.assert_array_equal(got[0], flat[0])


def test_build_small_n_fills_unused_codewords():
    x = int_rows(100, 8, -3, 4, 11)
    t = O.ivf_pq_build(x, 8, 2, 4, 4, 1)
    books = t["codebooks"].reshape(2, 256, 4)
    assert (books[:, 100:] == np.finfo(np.float32).max).all()
    assert (t["codes"] < 100).all()
