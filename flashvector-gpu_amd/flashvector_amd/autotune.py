"""Cached autotuning of the flat search plan per (d, N, k, dtype, metric).

fv_flat_autotune (C++, in the library) measures the choices on the device;
this module keys its results by the shape that determines them — dims, a
power-of-two bucket of the scanned rows, k, dtype, metric and whether the FP64
boundary is on — and keeps them in a JSON file (FLASHVECTOR_TUNE_CACHE, or
~/.cache/flashvector_amd/flat_tuning.json) so that a later process reuses them.
"""
from __future__ import annotations

import json
import math
import os

from . import lib as L


def _cache_path() -> str:
    return os.environ.get("FLASHVECTOR_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "flashvector_amd",
                                       "flat_tuning.json"))


def _bucket(v: int) -> int:
    return 1 << max(0, int(math.ceil(math.log2(max(v, 1)))))


def tuning_key(index, top_k, metric, rows, nq, f64_boundary=False, shortlist_factor=0) -> str:
    return (f"d{index.dim}_n{_bucket(rows)}_q{_bucket(nq)}_k{top_k}_{index.dtype}_{metric}"
            f"_b{int(bool(f64_boundary))}_s{shortlist_factor}")


def _load() -> dict:
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(table: dict) -> None:
    path = _cache_path()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


# Plans measured on MI355X for the benchmark shapes and committed with the
# package, so that a benchmark run times a fixed, reproducible plan.
COMMITTED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_plans.json")


def committed(index, top_k, metric, rows, nq, f64_boundary=False, shortlist_factor=0):
    """(the committed FvFlatTuning for this shape or None, its key)."""
    key = tuning_key(index, top_k, metric, rows, nq, f64_boundary, shortlist_factor)
    try:
        with open(COMMITTED) as f:
            entry = json.load(f).get(key)
    except (OSError, ValueError):
        return None, key
    return (to_tuning(entry) if entry else None), key


def committed_entry(key: str):
    """The committed plan entry (a dict) under `key`, or None."""
    try:
        with open(COMMITTED) as f:
            return json.load(f).get(key)
    except (OSError, ValueError):
        return None


def to_tuning(entry: dict):
    t = L.FvFlatTuning()
    t.shortlist_factor = entry["shortlist_factor"]
    t.n_chunks = entry["n_chunks"]
    t.bound_div = entry["bound_div"]
    t.ms = entry.get("ms", 0.0)
    t.recall = entry.get("recall", 0.0)
    return t


def tuned(index, sample_queries, top_k, metric="l2", valid_mask=None, valid_count=0,
          target_recall=0.999, f64_boundary=False, refresh=False, shortlist_factor=0):
    """The cached FvFlatTuning for this shape (rows, batch size, k, dtype,
    metric), measuring it on `sample_queries` — a batch of the size that will
    be searched, since the work shape depends on it — when absent (or when
    refresh is set). A non-zero shortlist_factor is kept fixed."""
    rows = valid_count if (valid_mask is not None and valid_count) else index.n
    key = tuning_key(index, top_k, metric, rows, sample_queries.shape[0], f64_boundary,
                     shortlist_factor)
    table = _load()
    if key in table and not refresh:
        return to_tuning(table[key])
    best = index.autotune(sample_queries, top_k, metric, valid_mask, valid_count, target_recall,
                          f64_boundary=f64_boundary, shortlist_factor=shortlist_factor)
    table[key] = {"shortlist_factor": best.shortlist_factor, "n_chunks": best.n_chunks,
                  "bound_div": best.bound_div, "ms": round(best.ms, 5),
                  "recall": round(best.recall, 6), "target_recall": target_recall}
    _store(table)
    return best
