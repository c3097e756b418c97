"""Micro-batcher of the host engine (fvh_batcher_*, include/flashvector/fv_host.h).

Counterpart of the reference server's SearchBatcher
(crates/server/src/batcher.rs:80-262): threads submit one query each; the
library's worker thread collects queries for up to max_wait_us (or until
max_batch are held), splits them by top_k, runs one batched search per group
and returns each caller its own results. `for_engine` searches an Engine
(SearchBatcher::for_index); `with_executor` takes any Python callable
(SearchBatcher::with_executor), which the worker thread calls with
(top_k, queries float32 [nq, dim]) and which returns (ids [nq, top_k],
distances [nq, top_k]).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import lib as L
from .engine import RESULT

EXECUTOR = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_uint32,
                            ctypes.POINTER(ctypes.c_float), ctypes.c_uint32, ctypes.c_void_p)

# Return code of a batch whose Python executor raised.
EXECUTOR_RAISED = -100


class SearchBatcher:
    def __init__(self, handle, dim, keep):
        self._lib = L.load()
        self._h = handle
        self.dim = dim
        self._keep = keep  # engine or callback: must outlive the worker

    @classmethod
    def for_engine(cls, engine, max_wait_us=350, max_batch=64):
        """Batches fvh_search calls on `engine` (the reference server's
        defaults: 350 us, 64 queries)."""
        lib = L.load()
        h = ctypes.c_void_p()
        L.check(lib.fvh_batcher_create(engine._h, max_wait_us, max_batch, ctypes.byref(h)),
                "fvh_batcher_create")
        return cls(h, engine.dim, engine)

    @classmethod
    def with_executor(cls, fn, dim, max_wait_us, max_batch):
        lib = L.load()

        def call(_user, k, q_ptr, nq, out_ptr):
            try:
                q = np.ctypeslib.as_array(q_ptr, shape=(nq, dim)).copy()
                ids, dists = fn(k, q)
                buf = (ctypes.c_char * (nq * k * RESULT.itemsize)).from_address(out_ptr)
                res = np.frombuffer(buf, dtype=RESULT)
                res["id"] = np.asarray(ids, dtype=np.uint32).ravel()
                res["distance"] = np.asarray(dists, dtype=np.float32).ravel()
                return 0
            except L.FvError as e:
                return getattr(e, "code", EXECUTOR_RAISED)
            except Exception:
                return EXECUTOR_RAISED

        cb = EXECUTOR(call)
        h = ctypes.c_void_p()
        L.check(lib.fvh_batcher_create_with_executor(ctypes.cast(cb, ctypes.c_void_p), None, dim,
                                                     max_wait_us, max_batch, ctypes.byref(h)),
                "fvh_batcher_create_with_executor")
        return cls(h, dim, cb)

    def submit(self, query, top_k):
        """Blocks until the query's batch has run: (ids uint32 [top_k],
        distances float32 [top_k], batch executor time in microseconds)."""
        q = np.ascontiguousarray(query, dtype=np.float32).ravel()
        if q.size != self.dim:
            raise L.FvError(f"query has {q.size} values, the batcher takes {self.dim}")
        res = np.zeros(top_k, RESULT)
        lat = ctypes.c_double(0.0)
        rc = self._lib.fvh_batcher_submit(self._h, q.ctypes.data, top_k, res.ctypes.data,
                                          ctypes.byref(lat))
        if rc != 0:
            err = L.FvError(f"fvh_batcher_submit failed with code {rc}")
            err.code = rc
            raise err
        return res["id"].copy(), res["distance"].copy(), lat.value

    def stats(self):
        """BatcherStats (batcher.rs:64-69)."""
        nr, nb, avg = ctypes.c_uint64(0), ctypes.c_uint64(0), ctypes.c_double(0.0)
        L.check(self._lib.fvh_batcher_stats(self._h, ctypes.byref(nr), ctypes.byref(nb),
                                            ctypes.byref(avg)), "fvh_batcher_stats")
        return {"total_requests": nr.value, "total_batches": nb.value,
                "avg_batch_size": avg.value}

    def close(self):
        """Serves the queries already queued, then stops the worker."""
        if self._h:
            self._lib.fvh_batcher_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
