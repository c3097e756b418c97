"""Pins the CPU oracle against independent facts (closed forms, FP64 brute
force on integer data, the reference's own KATs, invariants). None of these
retype the oracle's formulas."""
import numpy as np
import pytest

import oracle as O
from flashvector_amd import synthetic as S
from flashvector_amd.ivfpq_build import build_ivfpq


def setup_module(_):
    O.build()


def _int_data(n, nq, d=128, seed=5, clusters=16):
    return S.flat_workload(n, nq, d, clusters, seeds=(seed, seed + 1, seed + 2))


def _f64_l2(q, x):
    q64, x64 = q.astype(np.float64), x.astype(np.float64)
    return (q64 * q64).sum(1)[:, None] + (x64 * x64).sum(1)[None, :] - 2 * q64 @ x64.T


def _lex_topk(dist_row, k, ids=None):
    ids = np.arange(dist_row.shape[0]) if ids is None else ids
    order = np.lexsort((ids, dist_row))
    return order[:k]


# --------------------------------------------------------------------------
# distances
# --------------------------------------------------------------------------

def test_distance_kat_closed_form():
    """tests/cuda_sanity_test.rs:44-113: q[i]=0.01 i, x[i]=0.02 i, Q=4, N=8,
    d=128; rel err < 1e-4 against the exact sum sum_d (c - 0.01 d)^2 with
    c = 1.28 qi - 2.56 vj."""
    d, nq, n = 128, 4, 8
    q = (np.arange(nq * d, dtype=np.float32) * np.float32(0.01)).reshape(nq, d)
    x = (np.arange(n * d, dtype=np.float32) * np.float32(0.02)).reshape(n, d)
    got = O.distance_matrix(q, x, "l2")
    sd, sd2 = d * (d - 1) / 2, (d - 1) * d * (2 * d - 1) / 6
    for i in range(nq):
        for j in range(n):
            c = 1.28 * i - 2.56 * j
            exact = d * c * c - 2 * c * 0.01 * sd + 1e-4 * sd2
            assert abs(got[i, j] - exact) / max(exact, 1.0) < 1e-4


def test_l2_ip_exact_on_integer_data():
    db, qs = _int_data(500, 8)
    for metric, ref in (("l2", _f64_l2(qs, db)),
                        ("ip", -(qs.astype(np.float64) @ db.astype(np.float64).T))):
        got = O.distance_matrix(qs, db, metric)
        np.testing.assert_array_equal(got.astype(np.float64), ref)


def test_cosine_close_to_fp64():
    rng = np.random.default_rng(0)
    q = rng.normal(size=(4, 96)).astype(np.float32)
    x = rng.normal(size=(50, 96)).astype(np.float32)
    q64, x64 = q.astype(np.float64), x.astype(np.float64)
    ref = 1 - (q64 @ x64.T) / (np.linalg.norm(q64, axis=1)[:, None] *
                              np.linalg.norm(x64, axis=1)[None, :] + 1e-8)
    got = O.distance_matrix(q, x, "cosine")
    np.testing.assert_allclose(got, ref, atol=2e-6)


def test_exact_knn_matches_fp64_and_stable_ties():
    db, qs = _int_data(2000, 5)
    db[7] = db[3]         # duplicate rows: equal distances, the lower id first
    db[11] = db[3]
    qs[0] = db[3]
    ref = _f64_l2(qs, db)
    for i in range(qs.shape[0]):
        ids, dist = O.exact_knn(db, qs[i], 10)
        exp = _lex_topk(ref[i], 10)
        np.testing.assert_array_equal(ids, exp)
        np.testing.assert_array_equal(dist.astype(np.float64), ref[i][exp])
    ids, _ = O.exact_knn(db, qs[0], 3)
    np.testing.assert_array_equal(ids, [3, 7, 11])


# --------------------------------------------------------------------------
# flat scan contract
# --------------------------------------------------------------------------

@pytest.mark.parametrize("metric", ["l2", "ip"])
def test_flat_scan_matches_brute_force(metric):
    db, qs = _int_data(3000, 12)
    ref = _f64_l2(qs, db) if metric == "l2" else -(qs.astype(np.float64) @ db.astype(np.float64).T)
    ids, dist = O.flat_scan_batch(db, qs, 20, metric)
    for i in range(qs.shape[0]):
        exp = _lex_topk(ref[i], 20)
        np.testing.assert_array_equal(ids[i], exp)
        np.testing.assert_array_equal(dist[i].astype(np.float64), ref[i][exp])


def test_flat_scan_mask_and_remap():
    db, qs = _int_data(1000, 6)
    rng = np.random.default_rng(3)
    bits = rng.random(1000) < 0.3
    mask = S.pack_bits(bits)
    ids, _ = O.flat_scan_batch(db, qs, 15, "l2", valid_mask=mask)
    assert bits[ids[ids != 0xFFFFFFFF]].all()
    ref = _f64_l2(qs, db)
    ref[:, ~bits] = np.inf
    for i in range(6):
        np.testing.assert_array_equal(ids[i], _lex_topk(ref[i], 15))
    full = S.pack_bits(np.ones(1000, bool))
    a = O.flat_scan_batch(db, qs, 15, "l2", valid_mask=full)
    b = O.flat_scan_batch(db, qs, 15, "l2")
    np.testing.assert_array_equal(a[0], b[0])
    remap = (np.arange(1000, dtype=np.uint32) * 7 + 100)
    r_ids, _ = O.flat_scan_batch(db, qs, 15, "l2", vec_ids=remap)
    np.testing.assert_array_equal(r_ids, remap[b[0]])


def test_flat_scan_edge_cases():
    db, qs = _int_data(300, 3)
    ids, dist = O.flat_scan_batch(db, qs, 400)          # clamps to 256
    assert ids.shape == (3, 256)
    ids, dist = O.flat_scan_batch(db[:5], qs, 8)        # fewer rows than k -> padding
    assert (ids[:, 5:] == 0xFFFFFFFF).all() and (dist[:, 5:] == np.float32(1e30)).all()
    ids, dist = O.flat_scan_batch(db[:0], qs, 4)        # N == 0 -> 0xFF bytes
    assert (ids == 0xFFFFFFFF).all() and np.isnan(dist).all()


# --------------------------------------------------------------------------
# bitonic KAT (tests/cuda_sanity_test.rs:115-159) as a golden vector
# --------------------------------------------------------------------------

def test_bitonic_kat_golden():
    g = np.load(__file__.replace("test_oracle.py", "golden/kat_bitonic.npz"))
    order = np.argsort(g["keys"], kind="stable")
    np.testing.assert_array_equal(g["keys_sorted"], g["keys"][order])
    np.testing.assert_array_equal(g["vals_sorted"], g["vals"][order])


# --------------------------------------------------------------------------
# binary16 (crates/engine/src/quant.rs:239-304)
# --------------------------------------------------------------------------

def test_fp16_reference_roundtrip_values():
    for v in [0.0, 1.0, -1.5, 0.125, 42.25, -100.5]:      # quant.rs:362-368
        assert abs(O.f16_bits_to_f32(O.f32_to_f16_bits(v)) - v) < 1e-2


def test_fp16_decode_exhaustive_vs_numpy():
    h = np.arange(65536, dtype=np.uint16)
    ref = h.view(np.float16).astype(np.float32)
    got = np.array([O.f16_bits_to_f32(int(b)) for b in h], np.float32)
    finite = np.isfinite(ref)
    np.testing.assert_array_equal(got[finite].view(np.uint32), ref[finite].view(np.uint32))
    assert np.isnan(got[np.isnan(ref)]).all()
    # normal and zero halves re-encode to themselves
    for b in h[finite][::7]:
        f = O.f16_bits_to_f32(int(b))
        e = (int(b) >> 10) & 0x1F
        if e != 0 or (int(b) & 0x3FF) == 0:
            assert O.f32_to_f16_bits(f) == int(b)


def test_fp16_encode_rounding_rule_vs_numpy():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(0, 100, 20000), rng.uniform(-1, 1, 5000) * 6e-5]).astype(np.float32)
    bits = x.view(np.uint32)
    rne = x.astype(np.float16).view(np.uint16)
    got = np.array([O.f32_to_f16_bits(float(v)) for v in x], np.uint16)
    normal = np.abs(x) >= 2.0 ** -14
    tie = (bits & 0x1FFF) == 0x1000
    # off ties both round to nearest
    np.testing.assert_array_equal(got[normal & ~tie], rne[normal & ~tie])
    # subnormals truncate toward zero: |h| = floor(|x| * 2^24)
    sub = ~normal
    mag = np.floor(np.abs(x[sub]).astype(np.float64) * 2.0 ** 24).astype(np.uint16)
    np.testing.assert_array_equal(got[sub] & 0x7FFF, mag)
    # exact ties round half up on magnitude
    t = np.array([1.0 + 2.0 ** -11, 3.0 + 3 * 2.0 ** -11], np.float32)
    for v in t:
        assert O.f16_bits_to_f32(O.f32_to_f16_bits(float(v))) > v


# --------------------------------------------------------------------------
# IVF-PQ
# --------------------------------------------------------------------------

def _ivf_case(n=3000, d=32, nlist=16, m_pq=8, integer=False, seed=7):
    db, qs = _int_data(n, 10, d=d, seed=seed, clusters=24)
    t = build_ivfpq(db, nlist, m_pq, seed=seed, iters=4, integer_codebooks=integer)
    h = O.IvfPqHost(t.centroids, t.codebooks, t.codes, t.offsets, t.vec_ids, t.vectors, m_pq)
    return db, qs, t, h


def test_ivf_exhaustive_probe_equals_flat():
    db, qs, t, h = _ivf_case()
    ids, dist = O.ivf_pq_search_batch(h, qs, 10, nprobe=t.nlist, rerank_k=10240)
    fids, fdist = O.flat_scan_batch(db, qs, 10)
    np.testing.assert_array_equal(ids, fids)
    np.testing.assert_array_equal(dist, fdist)


def test_ivf_integer_codebooks_adc_equals_decoded_brute_force():
    db, qs, t, h = _ivf_case(integer=True)
    k = 20
    ids, dist = O.ivf_pq_search_batch(h, qs, k, nprobe=4, rerank_k=1)   # ADC only
    sub = t.dim // t.m_pq
    a = np.zeros(db.shape[0], np.int64)
    for c in range(t.nlist):
        a[t.vec_ids[t.offsets[c]:t.offsets[c + 1]]] = c
    recon = t.centroids[a].astype(np.float64)
    for m in range(t.m_pq):
        recon[:, m * sub:(m + 1) * sub] += t.codebooks[m][t.codes[:, m]]
    for i in range(qs.shape[0]):
        probed, _ = O.coarse_probe(qs[i], t.centroids, 4)
        # coarse probe = 4 smallest exact distances, ties by cluster id
        cd = _f64_l2(qs[i:i + 1], t.centroids)[0]
        np.testing.assert_array_equal(probed, _lex_topk(cd, 4))
        cand = np.isin(a, probed)
        d64 = ((qs[i].astype(np.float64) - recon) ** 2).sum(1)
        d64[~cand] = np.inf
        exp = _lex_topk(d64, k)
        exp = exp[np.isfinite(d64[exp])]
        np.testing.assert_array_equal(ids[i][:len(exp)], exp)
        np.testing.assert_array_equal(dist[i][:len(exp)].astype(np.float64), d64[exp])


def test_ivf_list_permutation_invariance_and_mask():
    db, qs, t, h = _ivf_case(seed=9)
    base = O.ivf_pq_search_batch(h, qs, 10, nprobe=5, rerank_k=64)
    rng = np.random.default_rng(2)
    vid = t.vec_ids.copy()
    for c in range(t.nlist):
        s, e = t.offsets[c], t.offsets[c + 1]
        vid[s:e] = rng.permutation(vid[s:e])
    h2 = O.IvfPqHost(t.centroids, t.codebooks, t.codes, t.offsets, vid, t.vectors, t.m_pq)
    perm = O.ivf_pq_search_batch(h2, qs, 10, nprobe=5, rerank_k=64)
    np.testing.assert_array_equal(base[0], perm[0])
    np.testing.assert_array_equal(base[1], perm[1])
    bits = rng.random(db.shape[0]) < 0.5
    ids, _ = O.ivf_pq_search_batch(h, qs, 10, nprobe=5, rerank_k=64, valid_mask=S.pack_bits(bits))
    assert bits[ids[ids != 0xFFFFFFFF]].all()


def test_ivf_parameter_clamps():
    db, qs, t, h = _ivf_case()
    a = O.ivf_pq_search_batch(h, qs, 10, nprobe=0, rerank_k=0)
    b = O.ivf_pq_search_batch(h, qs, 10, nprobe=1, rerank_k=10)
    np.testing.assert_array_equal(a[0], b[0])
    a = O.ivf_pq_search_batch(h, qs, 10, nprobe=10**6, rerank_k=5)   # S raised to k
    b = O.ivf_pq_search_batch(h, qs, 10, nprobe=t.nlist, rerank_k=10)
    np.testing.assert_array_equal(a[0], b[0])


def test_ivf_uneven_matches_per_query_calls():
    db, qs, t, h = _ivf_case(seed=11)
    ks = np.array([1, 5, 10, 3, 7, 10, 2, 9, 4, 6], np.uint32)
    nps = np.array([1, 2, 4, 8, 16, 3, 5, 7, 2, 1], np.uint32)
    ids, dist = O.ivf_pq_search_uneven(h, qs, ks, nps, shortlist=50, out_stride=10)
    for i in range(qs.shape[0]):
        e_ids, e_dist = O.ivf_pq_search_batch(h, qs[i:i + 1], int(ks[i]), int(nps[i]), 50)
        np.testing.assert_array_equal(ids[i, :ks[i]], e_ids[0])
        assert (ids[i, ks[i]:] == 0xFFFFFFFF).all()


# --------------------------------------------------------------------------
# merge
# --------------------------------------------------------------------------

def test_merge_of_shards_equals_full_scan():
    db, qs = _int_data(2400, 7)
    full_ids, full_d = O.flat_scan_batch(db, qs, 12)
    parts = np.array_split(np.arange(2400), 3)
    sid, sd = [], []
    for p in parts:
        i, d = O.flat_scan_batch(db[p], qs, 12, vec_ids=p.astype(np.uint32))
        sid.append(i)
        sd.append(d)
    m_ids, m_d = O.merge_topk(np.stack(sid), np.stack(sd), 12)
    np.testing.assert_array_equal(m_ids, full_ids)
    np.testing.assert_array_equal(m_d, full_d)


def test_merge_dedup_keeps_best():
    ids = np.array([[[5, 9, 0xFFFFFFFF]], [[9, 5, 2]]], np.uint32)
    d = np.array([[[1.0, 4.0, 1e30]], [[2.0, 3.0, 5.0]]], np.float32)
    m_ids, m_d = O.merge_topk(ids, d, 4)
    np.testing.assert_array_equal(m_ids[0], [5, 9, 2, 0xFFFFFFFF])
    np.testing.assert_array_equal(m_d[0], np.array([1.0, 2.0, 5.0, 1e30], np.float32))


# --------------------------------------------------------------------------
# FP8 e4m3 and the FP64 top-k
# --------------------------------------------------------------------------

def test_e4m3_decode_matches_torch_float8():
    import torch
    ref = torch.arange(256, dtype=torch.uint8).view(torch.float8_e4m3fn).float().numpy()
    got = O.e4m3_table()
    np.testing.assert_array_equal(np.isnan(got), np.isnan(ref))
    np.testing.assert_array_equal(got[~np.isnan(got)], ref[~np.isnan(ref)])


def test_e4m3_encode_matches_torch_and_brute_force():
    """Round to nearest even inside +-448 (torch's float8_e4m3fn cast), on
    random values, every code, and every midpoint between adjacent codes."""
    import torch
    tab = O.e4m3_table()
    pos = np.unique(tab[np.isfinite(tab) & (tab >= 0)])
    mid = ((pos[1:].astype(np.float64) + pos[:-1]) / 2).astype(np.float32)
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-448, 448, 20000), rng.normal(0, 1e-2, 5000),
                        rng.normal(0, 1e-3, 5000), pos, -pos, mid, -mid]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    got = np.array([O.f32_to_e4m3(v) for v in x], np.uint8)
    np.testing.assert_array_equal(got, ref)
    # Brute force: the nearest finite code, ties to the even mantissa.
    fin = np.nonzero(np.isfinite(tab))[0]
    for v in x[:3000]:
        d = np.abs(tab[fin].astype(np.float64) - float(v))
        best = fin[d == d.min()]
        if len(best) > 1:
            best = best[(best & 1) == 0]
        assert tab[O.f32_to_e4m3(v)] == tab[best[0]], v


def test_e4m3_saturates_and_keeps_nan():
    assert O.f32_to_e4m3(1e6) == 0x7E and O.f32_to_e4m3(-1e6) == 0xFE
    assert O.f32_to_e4m3(460.0) == 0x7E
    assert O.f32_to_e4m3(float("inf")) == 0x7E
    assert O.f32_to_e4m3(float("nan")) & 0x7F == 0x7F


def test_fp8_rows_scale_and_exact_integers():
    """Rows with integer entries in [-14, 14] and a 14 in every row encode
    exactly (scale 1/32): decode * scale reproduces the row."""
    rng = np.random.default_rng(4)
    x = rng.integers(-14, 15, size=(50, 40)).astype(np.float32)
    x[:, 0] = 14.0
    codes, scales = O.encode_fp8_rows(x, pdim=48)
    np.testing.assert_array_equal(scales, np.float32(14.0) / np.float32(448.0))
    tab = O.e4m3_table()
    np.testing.assert_array_equal(tab[codes[:, :40]] * scales[:, None], x)
    assert (codes[:, 40:] == 0).all()
    z, zs = O.encode_fp8_rows(np.zeros((2, 8), np.float32))
    assert (z == 0).all() and (zs == 1.0).all()


def test_fp8_scores_closed_form_on_integers():
    rng = np.random.default_rng(5)
    db = rng.integers(-14, 15, size=(60, 64)).astype(np.float32)
    qs = rng.integers(-14, 15, size=(7, 64)).astype(np.float32)
    db[:, 3] = 14.0
    qs[:, 5] = -14.0
    s = O.fp8_scores(db, qs, "ip")
    np.testing.assert_array_equal(s, -(qs.astype(np.float64) @ db.T.astype(np.float64)))
    s2 = O.fp8_scores(db, qs, "l2")
    ref = 0.5 * (db.astype(np.float64) ** 2).sum(1)[None, :] - qs.astype(np.float64) @ db.T
    np.testing.assert_array_equal(s2, ref)


def test_flat_topk_f64_matches_numpy():
    rng = np.random.default_rng(6)
    db = rng.normal(size=(700, 24)).astype(np.float32)
    qs = rng.normal(size=(9, 24)).astype(np.float32)
    db[10] = db[3]                              # exact tie: lower id first
    for metric in ("l2", "ip", "cosine"):
        ids, d = O.flat_topk_f64(db, qs, 15, metric)
        x, q = db.astype(np.float64), qs.astype(np.float64)
        if metric == "l2":
            full = ((q[:, None, :] - x[None, :, :]) ** 2).sum(-1)
        elif metric == "ip":
            full = -(q @ x.T)
        else:
            full = 1.0 - (q @ x.T) / (np.linalg.norm(q, axis=1)[:, None] * np.linalg.norm(x, axis=1)[None, :] + 1e-8)
        for i in range(9):
            order = np.lexsort((np.arange(700), full[i]))[:15]
            np.testing.assert_array_equal(ids[i], order)
            np.testing.assert_allclose(d[i], full[i][order], rtol=1e-12, atol=1e-12)
