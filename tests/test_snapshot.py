"""FLSHVEC1 snapshot format (crates/engine/src/persistence.rs:3-200) through
the library's host-only C ABI: byte-identical to a restatement of the
reference writer, read back intact, and rejected exactly where
MmapSnapshot::open rejects."""
import json
import os

import numpy as np
import pytest

from flashvector_amd import lib as L
from flashvector_amd import snapshot as SN

INFO = dict(dim=8, max_elements=5, m=32, ef_construction=128, nlist=2, m_pq=4, metric=0,
            num_vectors=5, entry_point=3, cagra_degree=0, rabitq_padded_dim=0,
            fp16_padded_dim=16, sparse_num_docs=0, sparse_nnz=0, next_logical_id=9)


def reference_bytes(info, sections):
    """write_aligned_snapshot (persistence.rs:51-109): sections 64-aligned
    in the given order, manifest = serde_json compact of SnapshotManifest
    (declaration order; sections as a name-sorted map)."""
    cursor, metas = 0, {}
    for name, a in sections:
        cursor = (cursor + 63) // 64 * 64
        metas[name] = {"offset": cursor, "byte_len": a.nbytes,
                       "dtype": {np.dtype(np.float32): "f32", np.dtype(np.uint32): "u32",
                                 np.dtype(np.uint16): "u16", np.dtype(np.uint8): "u8"}[a.dtype],
                       "len": a.size}
        cursor += a.nbytes
    man = {"version": 1,
           "config": {k: info[k] for k in ("dim", "max_elements", "m", "ef_construction",
                                           "nlist", "m_pq", "metric")},
           **{k: info[k] for k in ("num_vectors", "entry_point", "cagra_degree")},
           "cagra_seeds": [],
           **{k: info[k] for k in ("rabitq_padded_dim", "fp16_padded_dim", "sparse_num_docs",
                                   "sparse_nnz", "next_logical_id")},
           "sections": {k: metas[k] for k in sorted(metas)}}
    js = json.dumps(man, separators=(",", ":")).encode()
    out = bytearray(b"FLSHVEC1" + len(js).to_bytes(8, "little") + js)
    out += b"\0" * ((-len(out)) % 64)
    written = 0
    for _, a in sections:
        pad = (-written) % 64
        out += b"\0" * pad
        written += pad
        out += a.tobytes()
        written += a.nbytes
    return bytes(out)


def sample_sections(seed=0):
    rng = np.random.default_rng(seed)
    return [
        ("vectors", rng.normal(size=40).astype(np.float32)),
        ("logical_ids", np.array([4, 0, 7, 2, 8], np.uint32)),
        ("main_valid_words", np.array([0b11011], np.uint32)),
        ("hnsw_adjacency", np.zeros(0, np.uint32)),
        ("pq_codes", rng.integers(0, 256, size=21, dtype=np.uint8)),
        ("vectors_fp16", rng.integers(0, 65535, size=7, dtype=np.uint16)),
        ("tombstones", np.array([2, 5], np.uint32)),
    ]


@pytest.fixture(scope="module")
def so():
    return L.load()


def test_writer_matches_reference_layout(so, tmp_path):
    secs = sample_sections()
    p = str(tmp_path / "a.flshvec")
    SN.write(p, secs, **INFO)
    assert open(p, "rb").read() == reference_bytes(INFO, secs)


def test_round_trip(so, tmp_path):
    secs = sample_sections(1)
    p = str(tmp_path / "b.flshvec")
    SN.write(p, secs, **INFO)
    info, got = SN.read(p, [n for n, _ in secs] + ["missing"])
    for k, v in INFO.items():
        assert info[k] == v, k
    assert info["version"] == 1 and info["num_cagra_seeds"] == 0
    assert "missing" not in got
    for n, a in secs:
        np.testing.assert_array_equal(got[n], a)


def test_no_sections_and_unaligned_header(so, tmp_path):
    for nlist in (1, 12, 123, 1234):   # manifest lengths that hit every header padding
        info = dict(INFO, nlist=nlist)
        p = str(tmp_path / f"c{nlist}.flshvec")
        SN.write(p, [], **info)
        assert open(p, "rb").read() == reference_bytes(info, [])
        assert SN.read(p)[0]["nlist"] == nlist


def test_reader_accepts_any_json_layout(so, tmp_path):
    """serde_json accepts whitespace and any key order; so must the reader."""
    secs = sample_sections(2)
    raw = reference_bytes(INFO, secs)
    hl = int.from_bytes(raw[8:16], "little")
    man = json.loads(raw[16:16 + hl])
    data_start = (16 + hl + 63) // 64 * 64
    payload = raw[data_start:]
    pretty = json.dumps(dict(reversed(list(man.items()))), indent=2).encode()
    head = b"FLSHVEC1" + len(pretty).to_bytes(8, "little") + pretty
    head += b"\0" * ((-len(head)) % 64)
    p = str(tmp_path / "d.flshvec")
    open(p, "wb").write(head + payload)
    info, got = SN.read(p, [n for n, _ in secs])
    assert info["entry_point"] == INFO["entry_point"]
    for n, a in secs:
        np.testing.assert_array_equal(got[n], a)


def test_rejections(so, tmp_path):
    secs = sample_sections(3)
    good = reference_bytes(INFO, secs)
    hl = int.from_bytes(good[8:16], "little")

    def rc_of(data):
        p = str(tmp_path / "e.flshvec")
        open(p, "wb").write(data)
        return SN.open_error(p)

    assert rc_of(good) == 0
    assert rc_of(b"FLSHVEC2" + good[8:]) == -3                 # magic
    assert rc_of(good[:12]) == -3                              # shorter than 16 bytes
    assert rc_of(good[:8] + (len(good)).to_bytes(8, "little") + good[16:]) == -3  # header past EOF
    bad_json = bytearray(good)
    bad_json[16] = ord("[")
    assert rc_of(bytes(bad_json)) == -3                        # not a JSON object
    assert rc_of(good[:-1]) == -3                              # last section truncated
    man = json.loads(good[16:16 + hl])
    del man["config"]["m_pq"]
    js = json.dumps(man, separators=(",", ":")).encode()
    assert rc_of(b"FLSHVEC1" + len(js).to_bytes(8, "little") + js + b"\0" * 64) == -3
    assert SN.open_error(str(tmp_path / "absent.flshvec")) == -2


def test_null_arguments(so):
    assert so.fvh_snapshot_write(None, None, None, 0) == -1
    assert so.fvh_snapshot_open(None, None, None) == -1
    assert so.fvh_save_snapshot(None, None) == -1
    assert so.fvh_load_snapshot(None, None, 0) == -1
