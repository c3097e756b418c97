"""The reference's entry points (cuda_flat_scan_batch,
cuda_fp16_tc_search_batch, cuda_ivf_pq_search_batch) take no workspace: they
lease a library-owned buffer per (device, stream) that grows to the largest
call. Calls of growing and shrinking sizes, on several streams and from
several host threads at once, all equal the oracle."""
import threading

import numpy as np
import pytest

import oracle as O
from flashvector_amd import lib as L
from flashvector_amd import search as F
from flashvector_amd import synthetic as S
from flashvector_amd.ivfpq_build import build_ivfpq

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


@pytest.fixture(scope="module")
def case(torch):
    db, qs = S.flat_workload(40000, 600, 128, 64, seeds=(111, 112, 113))
    codes, norms = O.encode_fp16_padded(db)
    mask = S.random_valid_mask(40000, 0.05, 114)
    return db, qs, mask, dev(torch, codes.view(np.int16)), dev(torch, norms), dev(torch, db)


def test_sizes_grow_and_shrink(torch, case):
    db, qs, mask, d_codes, d_norms, d_db = case
    d_m = dev(torch, mask)
    for nq, k in [(40, 10), (600, 100), (7, 3), (300, 10)]:
        q = qs[:nq]
        got = L.split_results(F.fp16_tc_search_batch(d_codes, d_norms, dev(torch, q), k, 128,
                                                     "l2", d_m, d_db))
        assert_same(got, O.flat_topk(db, q, k, "l2", valid_mask=mask))
        got = L.split_results(F.flat_scan_batch(d_db, dev(torch, q), k, "l2", d_m))
        assert_same(got, O.flat_scan_batch(db, q, k, "l2", valid_mask=mask))


def test_device_counted_mask_matches_dense_scan(torch, case):
    """The legacy FP16 entry compacts the mask with its count taken on the
    device: same results as the workspace entry's dense masked scan."""
    db, qs, mask, d_codes, d_norms, d_db = case
    d_q, d_m = dev(torch, qs), dev(torch, mask)
    legacy = L.split_results(F.fp16_tc_search_batch(d_codes, d_norms, d_q, 10, 128, "l2", d_m,
                                                    d_db))
    idx = F.FlatIndex(d_db, "f16")
    dense = L.split_results(idx.search(d_q, 10, "l2", d_m, valid_count=0))
    assert_same(legacy, dense)
    for frac in (0.0, 0.5, 1.0):
        m = S.random_valid_mask(40000, frac, 115)
        got = L.split_results(F.fp16_tc_search_batch(d_codes, d_norms, d_q[:50], 10, 128, "l2",
                                                     dev(torch, m), d_db))
        assert_same(got, O.flat_topk(db, qs[:50], 10, "l2", valid_mask=m))


def test_streams_and_threads(torch, case):
    db, qs, mask, d_codes, d_norms, d_db = case
    d_m = dev(torch, mask)
    t = build_ivfpq(db, 32, 16, seed=116, iters=3)
    ivf = F.IvfPqIndex.from_host(t)
    h = O.IvfPqHost(t.centroids, t.codebooks, t.codes, t.offsets, t.vec_ids, t.vectors, t.m_pq)
    exp_fp16 = O.flat_topk(db, qs, 10, "l2", valid_mask=mask)
    exp_ivf = O.ivf_pq_search_batch(h, qs, 10, 8, 100)
    errors = []

    def worker(w):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                for r in range(4):
                    lo, hi = (w * 97 + r * 31) % 500, 600
                    d_q = dev(torch, qs[lo:hi])
                    a = F.fp16_tc_search_batch(d_codes, d_norms, d_q, 10, 128, "l2", d_m, d_db)
                    b = ivf.search_batch(d_q, 10, 8, 100)
                    s.synchronize()
                    assert_same(L.split_results(a), (exp_fp16[0][lo:hi], exp_fp16[1][lo:hi]))
                    assert_same(L.split_results(b), (exp_ivf[0][lo:hi], exp_ivf[1][lo:hi]))
        except Exception as ex:  # noqa: BLE001
            errors.append(ex)

    threads = [threading.Thread(target=worker, args=(w,)) for w in range(6)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors[0]
