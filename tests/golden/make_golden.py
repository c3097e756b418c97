"""Writes the small golden fixtures under tests/golden/.

Inputs are seeded synthetic data; expected outputs come from the CPU oracle
(oracle/), except the reference KATs whose expected values are written out
literally below. Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "flashvector-gpu_amd")]

import oracle as O  # noqa: E402
from flashvector_amd import synthetic as S  # noqa: E402
from flashvector_amd.ivfpq_build import build_ivfpq  # noqa: E402


def kat_bitonic():
    # tests/cuda_sanity_test.rs:120-121 inputs; ascending output by hand.
    keys = np.array([45.2, 1.2, 99.0, 3.15, 0.5, 12.8, 88.3, 7.7], np.float32)
    vals = np.arange(8, dtype=np.uint32)
    keys_sorted = np.array([0.5, 1.2, 3.15, 7.7, 12.8, 45.2, 88.3, 99.0], np.float32)
    vals_sorted = np.array([4, 1, 3, 7, 5, 0, 6, 2], np.uint32)
    np.savez(os.path.join(HERE, "kat_bitonic.npz"), keys=keys, vals=vals,
             keys_sorted=keys_sorted, vals_sorted=vals_sorted)


def kat_distance():
    # tests/cuda_sanity_test.rs:50-62 inputs; closed-form exact L2.
    d, nq, n = 128, 4, 8
    q = (np.arange(nq * d, dtype=np.float32) * np.float32(0.01)).reshape(nq, d)
    x = (np.arange(n * d, dtype=np.float32) * np.float32(0.02)).reshape(n, d)
    sd, sd2 = d * (d - 1) / 2, (d - 1) * d * (2 * d - 1) / 6
    c = 1.28 * np.arange(nq)[:, None] - 2.56 * np.arange(n)[None, :]
    exact = d * c * c - 2 * c * 0.01 * sd + 1e-4 * sd2
    np.savez(os.path.join(HERE, "kat_distance.npz"), queries=q, dataset=x, exact_l2=exact)


def flat_small():
    db, qs = S.flat_workload(3000, 16, 128, 32, seeds=(21, 22, 23))
    mask = S.random_valid_mask(3000, 0.3, 24)
    out = {"db_u8": db.astype(np.uint8), "queries_u8": qs.astype(np.uint8), "mask": mask}
    for metric in ("l2", "ip", "cosine"):
        i, dd = O.flat_scan_batch(db, qs, 10, metric)
        out[f"{metric}_ids"], out[f"{metric}_dist"] = i, dd
    i, dd = O.flat_scan_batch(db, qs, 10, "l2", valid_mask=mask)
    out["l2_masked_ids"], out["l2_masked_dist"] = i, dd
    i, dd = O.flat_scan_batch(db, qs, 300, "l2")
    out["l2_k300_ids"], out["l2_k300_dist"] = i, dd
    np.savez_compressed(os.path.join(HERE, "flat_small.npz"), **out)


def ivfpq_small():
    db, qs = S.flat_workload(4000, 12, 32, 24, seeds=(31, 32, 33))
    t = build_ivfpq(db, 16, 8, seed=34, iters=4)
    h = O.IvfPqHost(t.centroids, t.codebooks, t.codes, t.offsets, t.vec_ids, t.vectors, 8)
    out = {"db_u8": db.astype(np.uint8), "queries_u8": qs.astype(np.uint8),
           "centroids": t.centroids, "codebooks": t.codebooks, "codes": t.codes,
           "offsets": t.offsets, "vec_ids": t.vec_ids}
    mask = S.random_valid_mask(4000, 0.5, 35)
    out["mask"] = mask
    cases = [(10, 4, 32, None), (10, 4, 1, None), (10, 8, 200, None), (20, 3, 0, None),
             (10, 6, 100, mask)]
    for ci, (k, npb, rr, m) in enumerate(cases):
        i, dd = O.ivf_pq_search_batch(h, qs, k, npb, rr, valid_mask=m)
        out[f"case{ci}_params"] = np.array([k, npb, rr, m is not None], np.uint32)
        out[f"case{ci}_ids"], out[f"case{ci}_dist"] = i, dd
    np.savez_compressed(os.path.join(HERE, "ivfpq_small.npz"), **out)


if __name__ == "__main__":
    O.build()
    kat_bitonic()
    kat_distance()
    flat_small()
    ivfpq_small()
    print("golden fixtures written to", HERE)
