"""FLSHVEC1 snapshots of the host engine (fvh_save_snapshot /
fvh_load_snapshot): saved sections hold the engine's state in the
reference's layout, and an engine loaded from the file returns the same
results as the one that wrote it; a file laid out like the reference writer's
(index.rs:866-1028 section set) loads too."""
import numpy as np
import pytest

import oracle as O
from flashvector_amd import engine as E
from flashvector_amd import lib as L
from flashvector_amd import snapshot as SN
from flashvector_amd.ivfpq_build import build_ivfpq

pytestmark = pytest.mark.gpu

REF_SECTIONS = ["vectors", "logical_ids", "main_valid_words", "hnsw_adjacency", "hnsw_degrees",
                "cagra_adjacency", "ivf_centroids", "pq_codebooks", "pq_codes", "ivf_offsets",
                "ivf_vec_ids", "sq_mins", "sq_scales", "sq_codes", "rabitq_signs",
                "rabitq_centroids", "rabitq_bits", "rabitq_meta", "vectors_fp16", "vec_norms_sq"]


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a ROCm device"
    L.load()
    O.build()
    return t


def same(a, b):
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1].view(np.uint32), b[1].view(np.uint32))


def test_flat_engine_round_trip(torch, tmp_path):
    rng = np.random.default_rng(31)
    n, nd, dim, k = 6000, 300, 40, 10      # dim 40: FP16 rows padded to 48 in the file
    main = rng.normal(size=(n, dim)).astype(np.float32)
    ids = (50 + 2 * rng.permutation(n)).astype(np.uint32)
    delta = rng.normal(size=(nd, dim)).astype(np.float32)
    dids = (100000 + np.arange(nd)).astype(np.uint32)
    qs = rng.normal(size=(64, dim)).astype(np.float32)
    a = E.Engine(dim, "l2", max_top_k=32, shortlist_factor=16)
    a.load_flat(main, "f16", ids)
    a.add(delta, dids)
    dead = np.concatenate([rng.choice(ids, 100, replace=False), dids[:7]])
    a.delete(dead)
    ref = a.search(qs, k)
    p = str(tmp_path / "flat.flshvec")
    a.save_snapshot(p)
    info, sec = SN.read(p, REF_SECTIONS + ["tombstones", "delta_vectors", "delta_ids"])
    assert set(REF_SECTIONS) <= set(sec)
    assert info["dim"] == dim and info["num_vectors"] == n and info["fp16_padded_dim"] == 48
    assert info["next_logical_id"] == int(dids.max()) + 1
    np.testing.assert_array_equal(sec["vectors"].reshape(n, dim), main)
    np.testing.assert_array_equal(sec["logical_ids"], ids)
    np.testing.assert_array_equal(sec["tombstones"], np.unique(dead))
    np.testing.assert_array_equal(sec["delta_ids"], dids)
    np.testing.assert_array_equal(sec["delta_vectors"].reshape(nd, dim), delta)
    # FP16 rows and norms as the reference encodes them (quant.rs:306-327).
    f16, norms = O.encode_fp16_padded(main)
    np.testing.assert_array_equal(sec["vectors_fp16"].reshape(n, -1), f16)
    np.testing.assert_array_equal(sec["vec_norms_sq"].view(np.uint32), norms.view(np.uint32))
    alive = ~np.isin(ids, dead)
    words = sec["main_valid_words"]
    assert all(bool((words[s >> 5] >> (s & 31)) & 1) == alive[s] for s in range(n))
    b = E.Engine(dim, "l2", max_top_k=32, shortlist_factor=16)
    b.load_snapshot(p, "f16")
    same(b.search(qs, k), ref)


def test_ivfpq_engine_round_trip_and_flat_reload(torch, tmp_path):
    rng = np.random.default_rng(32)
    x = rng.normal(size=(8000, 32)).astype(np.float32)
    t = build_ivfpq(x, 32, 8, seed=33, iters=3)
    qs = rng.normal(size=(50, 32)).astype(np.float32)
    a = E.Engine(32, "l2", max_top_k=32, ivf_nprobe=8, ivf_shortlist=200)
    a.load_ivfpq(t.centroids, t.codebooks.reshape(8, -1), t.codes, t.offsets, t.vec_ids, x)
    a.delete(np.arange(0, 8000, 97, dtype=np.uint32))
    ref = a.search(qs, 10)
    p = str(tmp_path / "ivf.flshvec")
    a.save_snapshot(p)
    info, sec = SN.read(p, ["ivf_centroids", "pq_codes", "ivf_vec_ids", "tombstones"])
    assert info["nlist"] == 32 and info["m_pq"] == 8
    np.testing.assert_array_equal(sec["pq_codes"].reshape(8000, 8), t.codes)
    np.testing.assert_array_equal(sec["ivf_vec_ids"], t.vec_ids)
    b = E.Engine(32, "l2", max_top_k=32, ivf_nprobe=8, ivf_shortlist=200)
    b.load_snapshot(p)
    same(b.search(qs, 10), ref)
    # The same file as a flat index equals a flat engine over the same rows.
    c = E.Engine(32, "l2", max_top_k=32, shortlist_factor=16)
    c.load_snapshot(p, "bf16")
    d = E.Engine(32, "l2", max_top_k=32, shortlist_factor=16)
    d.load_flat(x, "bf16")
    d.delete(np.arange(0, 8000, 97, dtype=np.uint32))
    same(c.search(qs, 10), d.search(qs, 10))


def test_loads_reference_style_file(torch, tmp_path):
    """A file written section by section in the reference's order, with
    empty graph/SQ8/RaBitQ sections and no tombstones section: deletions come
    from the cleared bits of main_valid_words (index.rs:1195-1206)."""
    rng = np.random.default_rng(34)
    n, dim = 3000, 16
    x = rng.normal(size=(n, dim)).astype(np.float32)
    f16, norms = O.encode_fp16_padded(x)
    valid = np.full((n + 31) // 32, 0xFFFFFFFF, np.uint32)
    valid[-1] = (1 << (n % 32)) - 1 if n % 32 else 0xFFFFFFFF
    valid[0] &= ~np.uint32(0b1010)             # slots 1 and 3 deleted
    empty_u32, empty_f32 = np.zeros(0, np.uint32), np.zeros(0, np.float32)
    secs = [("vectors", x.ravel()), ("logical_ids", np.arange(n, dtype=np.uint32)),
            ("main_valid_words", valid), ("hnsw_adjacency", empty_u32),
            ("hnsw_degrees", empty_u32), ("cagra_adjacency", empty_u32),
            ("ivf_centroids", empty_f32), ("pq_codebooks", empty_f32),
            ("pq_codes", np.zeros(0, np.uint8)), ("ivf_offsets", empty_u32),
            ("ivf_vec_ids", empty_u32), ("sq_mins", empty_f32), ("sq_scales", empty_f32),
            ("sq_codes", np.zeros(0, np.uint8)), ("rabitq_signs", empty_f32),
            ("rabitq_centroids", empty_f32), ("rabitq_bits", empty_u32),
            ("rabitq_meta", empty_f32), ("vectors_fp16", f16.ravel()), ("vec_norms_sq", norms)]
    p = str(tmp_path / "ref.flshvec")
    SN.write(p, secs, dim=dim, max_elements=n, m=24, ef_construction=200, nlist=256, m_pq=16,
             metric=0, num_vectors=n, fp16_padded_dim=16, next_logical_id=n + 40)
    qs = np.concatenate([x[:4], rng.normal(size=(20, dim)).astype(np.float32)])
    b = E.Engine(dim, "l2", max_top_k=16, shortlist_factor=16)
    b.load_snapshot(p, "f16")
    got = b.search(qs, 5)
    d = E.Engine(dim, "l2", max_top_k=16, shortlist_factor=16)
    d.load_flat(x, "f16")
    d.delete(np.array([1, 3], np.uint32))
    same(got, d.search(qs, 5))
    assert got[0][0][0] == 0 and got[0][2][0] == 2          # live rows find themselves
    assert 1 not in got[0][1] and 3 not in got[0][3]
    # The manifest's id counter (ids handed out and compacted away by the
    # reference) is kept: new rows get ids above it, and a save carries it
    # and the graph fields the engine does not use.
    assert b.next_logical_id == n + 40
    np.testing.assert_array_equal(b.add_vectors(x[:2]), [n + 40, n + 41])
    p2 = str(tmp_path / "resaved.flshvec")
    b.save_snapshot(p2)
    info, _ = SN.read(p2)
    assert info["next_logical_id"] == n + 42
    assert info["m"] == 24 and info["ef_construction"] == 200 and info["max_elements"] == n


def test_fp16_row_codes_match_reference_encoder(torch):
    """fv_encode_flat's FP16 rows equal the reference's host conversion
    (quant.rs:237-273, restated by the oracle) on exact halfway values,
    subnormal results, overflow, infinities, NaN and random values."""
    rng = np.random.default_rng(35)
    ties = (rng.integers(0x38000000, 0x47000000, size=4000, dtype=np.uint32) & ~np.uint32(0x1fff)) | 0x1000
    sub = rng.integers(0x30000000, 0x38800000, size=4000, dtype=np.uint32)
    vals = np.concatenate([ties.view(np.float32), sub.view(np.float32),
                           rng.normal(size=7000).astype(np.float32) * 300,
                           np.array([65504, 65520, 65536, 1e9, np.inf, -np.inf, np.nan, 0.0, -0.0,
                                     6.1e-5, 5.9e-8, 3e-8], np.float32)])
    vals = np.concatenate([vals, -vals[:8000]])
    dim = 32
    vals = np.resize(vals, (len(vals) // dim) * dim).reshape(-1, dim)
    n = vals.shape[0]
    d_x = torch.from_numpy(vals).cuda()
    codes = torch.empty((n, dim), dtype=torch.int16, device="cuda")
    norms = torch.empty(n, dtype=torch.float32, device="cuda")
    L.check(L.load().fv_encode_flat(L.ptr(d_x), n, dim, dim, 1, L.ptr(codes), L.ptr(norms), None,
                                    None, L.stream_handle()), "fv_encode_flat")
    torch.cuda.synchronize()
    exp, _ = O.encode_fp16_padded(vals)
    np.testing.assert_array_equal(codes.cpu().numpy().view(np.uint16), exp[:, :dim])
