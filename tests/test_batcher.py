"""Micro-batcher (fvh_batcher_*) on the CPU with Python executors: the
reference's own tests of SearchBatcher (crates/server/src/batcher.rs:270-361)
restated, plus the batch cap, per-caller slices under mixed keys, error
propagation, and shutdown."""
import ctypes
import threading

import numpy as np
import pytest

from flashvector_amd import lib as L
from flashvector_amd.batcher import EXECUTOR_RAISED, SearchBatcher


def run_threads(n, target):
    start = threading.Barrier(n)
    errors = []

    def body(i):
        try:
            start.wait()
            target(i)
        except Exception as e:   # surfaced below, in the test's own thread
            errors.append(e)

    ts = [threading.Thread(target=body, args=(i,)) for i in range(n)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(60)
    assert not any(t.is_alive() for t in ts)
    if errors:
        raise errors[0]


def test_coalesces_concurrent_queries():
    """batcher.rs:270-325: 8 concurrent callers, tag-derived results, fewer
    than 8 executor calls, and as many calls as the batch counter says."""
    calls = []

    def execute(k, q):
        calls.append(q.shape[0])
        tags = q[:, 0].astype(np.uint32)
        ids = tags[:, None] * 100 + np.arange(k, dtype=np.uint32)[None, :]
        return ids, np.broadcast_to(np.arange(k, dtype=np.float32), ids.shape)

    with SearchBatcher.with_executor(execute, dim=4, max_wait_us=20000, max_batch=32) as b:
        def one(qid):
            ids, dists, _ = b.submit(np.full(4, qid, np.float32), 4)
            assert len(ids) == 4 and ids[0] == qid * 100
            np.testing.assert_array_equal(ids, qid * 100 + np.arange(4))
            np.testing.assert_array_equal(dists, np.arange(4, dtype=np.float32))

        run_threads(8, one)
        st = b.stats()
    assert st["total_requests"] == 8
    assert st["total_batches"] < 8
    assert len(calls) == st["total_batches"] and sum(calls) == 8
    assert st["avg_batch_size"] == pytest.approx(8 / st["total_batches"])


def test_separates_distinct_keys():
    """batcher.rs:327-361: two concurrent callers with different keys get
    results of their own length."""
    seen = []

    def execute(k, q):
        seen.append((k, q.shape[0]))
        return (np.full((q.shape[0], k), k, np.uint32), np.full((q.shape[0], k), 0.5, np.float32))

    with SearchBatcher.with_executor(execute, dim=4, max_wait_us=2000, max_batch=16) as b:
        out = {}

        def one(i):
            k = (2, 5)[i]
            out[k] = b.submit(np.full(4, i + 1, np.float32), k)

        run_threads(2, one)
    assert len(out[2][0]) == 2 and set(out[2][0]) == {2}
    assert len(out[5][0]) == 5 and set(out[5][0]) == {5}
    assert all(n == 1 for _, n in seen) and sorted(k for k, _ in seen) == [2, 5]


def test_mixed_keys_and_batch_cap():
    """Every caller gets the rows computed from its own query, whatever the
    grouping; no executor call exceeds max_batch or mixes keys."""
    calls = []

    def execute(k, q):
        calls.append((k, q.shape[0]))
        base = (q[:, 0] * 1000 + q[:, 1]).astype(np.uint32)
        ids = base[:, None] + np.arange(k, dtype=np.uint32)[None, :] * 7
        return ids, q[:, 2:3] + np.arange(k, dtype=np.float32)[None, :]

    ks = [1, 3, 8, 3, 1]
    with SearchBatcher.with_executor(execute, dim=3, max_wait_us=3000, max_batch=5) as b:
        def one(i):
            k = ks[i % len(ks)]
            q = np.array([i % 7, i, 0.25 * i], np.float32)
            ids, d, lat = b.submit(q, k)
            np.testing.assert_array_equal(ids, (i % 7) * 1000 + i + 7 * np.arange(k))
            np.testing.assert_array_equal(d, np.float32(0.25 * i) + np.arange(k, dtype=np.float32))
            assert lat >= 0.0

        run_threads(40, one)
        st = b.stats()
    assert st["total_requests"] == 40 == sum(n for _, n in calls)
    assert all(n <= 5 for _, n in calls)
    assert st["total_batches"] == len(calls)


def test_zero_wait_still_serves_everyone():
    with SearchBatcher.with_executor(
            lambda k, q: (np.tile(q[:, :1].astype(np.uint32), (1, k)), np.zeros((len(q), k))),
            dim=2, max_wait_us=0, max_batch=1) as b:
        got = []
        run_threads(12, lambda i: got.append(b.submit(np.array([i, 0], np.float32), 2)[0][0]))
        assert sorted(got) == list(range(12))
        assert b.stats()["total_batches"] == 12


def test_executor_errors_reach_every_caller_of_the_batch():
    def execute(k, q):
        if k == 9:
            raise ValueError("boom")
        return np.zeros((len(q), k), np.uint32), np.zeros((len(q), k), np.float32)

    with SearchBatcher.with_executor(execute, dim=2, max_wait_us=5000, max_batch=8) as b:
        codes = []

        def one(i):
            try:
                b.submit(np.zeros(2, np.float32), 9 if i < 3 else 4)
                codes.append(0)
            except L.FvError as e:
                codes.append(e.code)

        run_threads(6, one)
    assert sorted(codes) == [EXECUTOR_RAISED] * 3 + [0] * 3


def test_argument_checks_and_shutdown():
    lib = L.load()
    b = SearchBatcher.with_executor(lambda k, q: (np.zeros((len(q), k)), np.zeros((len(q), k))),
                                    dim=4, max_wait_us=100, max_batch=4)
    with pytest.raises(L.FvError):
        b.submit(np.zeros(3, np.float32), 2)            # wrong length
    res = (ctypes.c_uint64 * 4)()
    q = np.zeros(4, np.float32)
    assert lib.fvh_batcher_submit(b._h, q.ctypes.data, 0, res, None) == 1   # top_k 0
    assert lib.fvh_batcher_submit(b._h, None, 2, res, None) == -1
    assert lib.fvh_batcher_submit(None, q.ctypes.data, 2, res, None) == -1
    assert lib.fvh_batcher_stats(None, None, None, None) == -1
    assert b.stats() == {"total_requests": 0, "total_batches": 0, "avg_batch_size": 0.0}
    b.close()
    b.close()
    h = ctypes.c_void_p()
    assert lib.fvh_batcher_create(None, 350, 64, ctypes.byref(h)) == -1
    assert lib.fvh_batcher_create_with_executor(None, None, 4, 0, 1, ctypes.byref(h)) == -1
    lib.fvh_batcher_destroy(None)
