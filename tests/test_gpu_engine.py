"""GPU checks of the host search engine (fvh_*, include/flashvector/fv_host.h)
against the oracle: logical-id remapping, tombstones, caller masks over
logical ids, the delta buffer merged with the main results, the host and
device (zero-copy) paths, IVF-PQ tables, and concurrent callers."""
import threading

import numpy as np
import pytest

import oracle as O
from flashvector_amd import engine as E
from flashvector_amd import lib as L
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


def expected(main, main_ids, main_alive, delta, delta_ids, delta_alive, qs, k, metric, allow):
    """Exact FP32 top-k over the live, allowed rows of main and delta, with
    logical ids, ordered by (distance, logical id)."""
    rows = np.concatenate([main, delta]) if len(delta) else main
    ids = np.concatenate([main_ids, delta_ids]) if len(delta) else main_ids
    live = np.concatenate([main_alive, delta_alive]) if len(delta) else main_alive
    live = live & allow(ids)
    sel = np.flatnonzero(live)
    r_ids, r_d = O.flat_topk(rows[sel], qs, k, metric)
    out_ids = np.where(r_ids == 0xFFFFFFFF, 0xFFFFFFFF, ids[sel][np.minimum(r_ids, len(sel) - 1)])
    # Reorder each row by (distance, logical id).
    for i in range(len(qs)):
        o = np.lexsort((out_ids[i], r_d[i]))
        out_ids[i], r_d[i] = out_ids[i][o], r_d[i][o]
    return out_ids.astype(np.uint32), r_d


def test_flat_engine_remap_tombstones_mask_delta(torch):
    rng = np.random.default_rng(5)
    n, nd, dim, k = 20000, 500, 64, 10
    main = rng.normal(size=(n, dim)).astype(np.float32)
    main_ids = (1000 + 3 * rng.permutation(n)).astype(np.uint32)
    delta = rng.normal(size=(nd, dim)).astype(np.float32)
    delta_ids = (200000 + np.arange(nd)).astype(np.uint32)
    qs = rng.normal(size=(300, dim)).astype(np.float32)
    eng = E.Engine(dim, "l2", max_batch=128, max_top_k=64, shortlist_factor=16)
    eng.load_flat(main, "f16", main_ids)
    eng.add(delta, delta_ids)
    dead = np.concatenate([rng.choice(main_ids, 300, replace=False),
                           rng.choice(delta_ids, 50, replace=False)])
    eng.delete(dead)
    allow_bits = rng.random(int(delta_ids.max()) + 1) < 0.5
    main_alive = ~np.isin(main_ids, dead)
    delta_alive = ~np.isin(delta_ids, dead)
    exp = expected(main, main_ids, main_alive, delta, delta_ids, delta_alive, qs, k, "l2",
                   lambda ids: allow_bits[ids])
    got = eng.search(qs, k, mask=allow_bits)                 # host path, 3 batches of 128
    np.testing.assert_array_equal(got[0], exp[0])
    np.testing.assert_array_equal(got[1].view(np.uint32), exp[1].view(np.uint32))
    dev = eng.search(torch.from_numpy(qs).cuda(), k, mask=S.pack_bits(allow_bits))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dev[0], got[0])
    np.testing.assert_array_equal(dev[1].view(np.uint32), got[1].view(np.uint32))
    # No mask: only tombstones apply.
    exp2 = expected(main, main_ids, main_alive, delta, delta_ids, delta_alive, qs, k, "l2",
                    lambda ids: np.ones(len(ids), bool))
    got2 = eng.search(qs, k)
    np.testing.assert_array_equal(got2[0], exp2[0])
    assert eng.stats()["queries"] == 900


def test_flat_engine_identity_ids_no_delta(torch):
    db, qs = S.flat_workload(30000, 200, 128, 64, seeds=(7, 8, 9))
    eng = E.Engine(128, "ip", max_top_k=32)
    eng.load_flat(db, "bf16")
    got = eng.search(qs, 32)
    exp = O.flat_topk(db, qs, 32, "ip")
    np.testing.assert_array_equal(got[0], exp[0])
    np.testing.assert_array_equal(got[1].view(np.uint32), exp[1].view(np.uint32))


def test_ivfpq_engine_matches_oracle(torch):
    db, qs = S.flat_workload(20000, 120, 64, 48, seeds=(11, 12, 13))
    t = build_ivfpq(db, 64, 16, seed=14, iters=4)
    rng = np.random.default_rng(15)
    lids = (50 + 2 * rng.permutation(20000)).astype(np.uint32)
    eng = E.Engine(64, "l2", max_top_k=20, ivf_nprobe=16, ivf_shortlist=100)
    eng.load_ivfpq(t.centroids, t.codebooks, t.codes, t.offsets, t.vec_ids, t.vectors, lids)
    dead = rng.choice(lids, 500, replace=False)
    eng.delete(dead)
    got = eng.search(qs, 20)
    slot_alive = ~np.isin(lids, dead)
    h = O.IvfPqHost(t.centroids, t.codebooks, t.codes, t.offsets, t.vec_ids, t.vectors, t.m_pq)
    ks = np.full(120, 20, np.uint32)
    nps = np.full(120, 16, np.uint32)
    e_ids, e_d = O.ivf_pq_search_uneven(h, qs, ks, nps, 100, valid_mask=S.pack_bits(slot_alive),
                                        out_stride=20)
    exp_ids = np.where(e_ids == 0xFFFFFFFF, 0xFFFFFFFF, lids[np.minimum(e_ids, 19999)])
    np.testing.assert_array_equal(got[0], exp_ids)
    np.testing.assert_array_equal(got[1].view(np.uint32), e_d.view(np.uint32))


def test_concurrent_callers_match_sequential(torch):
    db, qs = S.flat_workload(20000, 400, 64, 32, seeds=(21, 22, 23))
    eng = E.Engine(64, "l2", num_streams=3, max_batch=64, max_top_k=16)
    eng.load_flat(db, "f16")
    eng.add(db[:100] + 1.0, np.arange(100000, 100100, dtype=np.uint32))
    parts = np.array_split(qs, 8)
    seq = [eng.search(p, 16) for p in parts]
    res = [None] * 8
    errs = []

    def worker(i):
        try:
            for _ in range(2):
                res[i] = eng.search(parts[i], 16)
        except Exception as ex:  # noqa: BLE001 - reported below
            errs.append(ex)

    th = [threading.Thread(target=worker, args=(i,)) for i in range(8)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    assert not errs
    for a, b in zip(res, seq):
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1].view(np.uint32), b[1].view(np.uint32))


def test_delete_survives_later_adds_and_snapshot(torch, tmp_path):
    """index.rs:647-714: rows deleted from the delta buffer stay deleted when
    more rows are added later (alive bits follow the tombstones), the deleted
    id is written to the snapshot's tombstones, and ids are never reused."""
    rng = np.random.default_rng(17)
    n, dim, k = 5000, 32, 10
    main = rng.integers(0, 50, size=(n, dim)).astype(np.float32)
    qs = rng.integers(0, 50, size=(64, dim)).astype(np.float32)
    eng = E.Engine(dim, "l2", max_batch=64, max_top_k=32)
    eng.load_flat(main, "f16")
    assert eng.next_logical_id == n
    # First add: rows equal to the queries themselves, so each query's best
    # match is its own delta row.
    a_ids = eng.add_vectors(qs[:32])
    np.testing.assert_array_equal(a_ids, np.arange(n, n + 32))
    dead = np.array([a_ids[3], a_ids[7], 11], np.uint32)      # two delta rows, one main row
    eng.delete(dead)
    # Second add: more rows under fresh ids; the deleted rows must not return.
    b_ids = eng.add_vectors(qs[32:])
    np.testing.assert_array_equal(b_ids, np.arange(n + 32, n + 64))
    delta = qs.copy()
    delta_ids = np.concatenate([a_ids, b_ids]).astype(np.uint32)
    main_ids = np.arange(n, dtype=np.uint32)
    allow_all = lambda ids: np.ones(len(ids), bool)  # noqa: E731
    exp = expected(main, main_ids, ~np.isin(main_ids, dead), delta, delta_ids,
                   ~np.isin(delta_ids, dead), qs, k, "l2", allow_all)
    got = eng.search(qs, k)
    np.testing.assert_array_equal(got[0], exp[0])
    np.testing.assert_array_equal(got[1].view(np.uint32), exp[1].view(np.uint32))
    assert not np.isin(got[0], dead).any()
    # Used and deleted ids are rejected, and a failed add changes nothing.
    for bad in ([a_ids[3]], [b_ids[0]], [5], [n + 100, n + 100]):
        with pytest.raises(L.FvError):
            eng.add(qs[:len(bad)], np.array(bad, np.uint32))
    np.testing.assert_array_equal(eng.search(qs, k)[0], exp[0])
    # Snapshot round trip keeps the deletions and the id counter.
    path = tmp_path / "snap.fv"
    eng.save_snapshot(path)
    eng2 = E.Engine(dim, "l2", max_batch=64, max_top_k=32)
    eng2.load_snapshot(path, "f16")
    got2 = eng2.search(qs, k)
    np.testing.assert_array_equal(got2[0], exp[0])
    np.testing.assert_array_equal(got2[1].view(np.uint32), exp[1].view(np.uint32))
    assert eng2.next_logical_id == n + 64
    c_ids = eng2.add_vectors(qs[:2])
    np.testing.assert_array_equal(c_ids, [n + 64, n + 65])
