"""FLSHVEC1 snapshot files (the reference's on-disk index format,
crates/engine/src/persistence.rs:3-200) through the library's C ABI
(fvh_snapshot_*). Host-side: numpy arrays in and out, no GPU needed."""
from __future__ import annotations

import ctypes

import numpy as np

from . import lib as L

_NP = {"f32": np.float32, "u32": np.uint32, "u16": np.uint16, "u8": np.uint8}
_NAMES = {np.dtype(v): k for k, v in _NP.items()}

INFO_FIELDS = [f for f, _ in L.FvSnapshotInfo._fields_ if f != "reserved"]


def _check(rc, what):
    if rc != 0:
        raise L.FvError(f"{what} failed with code {rc}")


def write(path: str, sections, **info) -> None:
    """Writes `sections` (list of (name, array); u8/u16/u32/f32 arrays, in
    file order) under a manifest built from `info` (FvSnapshotInfo fields;
    version defaults to 1)."""
    lib = L.load()
    inf = L.FvSnapshotInfo()
    inf.version = 1
    for k, v in info.items():
        setattr(inf, k, v)
    keep, arr = [], (L.FvSnapshotSection * max(1, len(sections)))()
    for i, (name, a) in enumerate(sections):
        a = np.ascontiguousarray(a)
        dt = _NAMES.get(a.dtype)
        if dt is None:
            raise L.FvError(f"section {name}: unsupported dtype {a.dtype}")
        keep.append(a)
        arr[i] = L.FvSnapshotSection(name.encode(), dt.encode(), a.size,
                                     a.ctypes.data if a.size else None)
    _check(lib.fvh_snapshot_write(path.encode(), ctypes.byref(inf), arr, len(sections)),
           "fvh_snapshot_write")


def read(path: str, names=None):
    """(info dict, {name: array copy}) of a snapshot; `names` limits the
    sections returned (absent names are skipped)."""
    lib = L.load()
    h = ctypes.c_void_p()
    inf = L.FvSnapshotInfo()
    _check(lib.fvh_snapshot_open(path.encode(), ctypes.byref(h), ctypes.byref(inf)),
           "fvh_snapshot_open")
    try:
        out = {}
        for name in names or []:
            data, bl, ln = ctypes.c_void_p(), ctypes.c_uint64(), ctypes.c_uint64()
            dt = ctypes.create_string_buffer(8)
            rc = lib.fvh_snapshot_section(h, name.encode(), ctypes.byref(data), ctypes.byref(bl),
                                          ctypes.byref(ln), dt)
            if rc == 1:
                continue
            _check(rc, f"fvh_snapshot_section({name})")
            np_t = _NP[dt.value.decode()]
            raw = ctypes.string_at(data, bl.value) if bl.value else b""
            out[name] = np.frombuffer(raw, dtype=np_t).copy()
        return {f: getattr(inf, f) for f in INFO_FIELDS}, out
    finally:
        lib.fvh_snapshot_close(h)


def open_error(path: str) -> int:
    """Return code of fvh_snapshot_open (0 = valid) without keeping the file."""
    lib = L.load()
    h = ctypes.c_void_p()
    rc = lib.fvh_snapshot_open(path.encode(), ctypes.byref(h), None)
    if rc == 0:
        lib.fvh_snapshot_close(h)
    return rc
