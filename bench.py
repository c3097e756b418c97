"""Benchmark of the query-scan hot path on one MI355X.

Workload (default): exact flat L2 search over N=1,000,000 x d=128 integer
mixture rows (256 clusters), Q=10,000 queries, k=10, with a packed validity
bitmask that keeps 5% of the rows, evaluated inside the kernels. One step is
one batched search of all Q queries: matrix-core scan in FP16 (or BF16) with
FP32 accumulation, shortlist of 8k per query, exact FP32 rerank, top-k.

Prints one JSON line. `value` is queries per second with inputs resident in
HBM. Beside it: the plain exact FP32 GPU scan on the same inputs, recall@k and
max |distance error| against an FP64 oracle on a query sample, the live
roofline of the dominant kernel, and the CPU oracle timed on a query sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "flashvector-gpu_amd"), os.path.join(ROOT, "oracle")]

MFMA_PEAK_TFLOPS = {"f16": 2500.0, "bf16": 2500.0}
FP32_PEAK_TFLOPS = 157.3


def measured_traffic(key):
    """HBM bytes per launch of a kernel/workload from the committed rocprofv3
    counter summaries (profiles/traffic.json), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            return json.load(f).get(key, {}).get("hbm_bytes_per_launch")
    except OSError:
        return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def best_of(fn, nq, reps=5):
    """The reference harness's reading (scripts/run_benchmarks.py:60-70: one
    warm-up, then the best of 5 calls), taken after the main timed loop with
    HIP events around each call; also the median. `value` stays the mean over
    the timed steps."""
    import statistics

    import torch
    fn()
    ts = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        torch.cuda.synchronize()
        ts.append(a.elapsed_time(b))
    best, med = min(ts), statistics.median(ts)
    return {"reps": reps, "best_ms": round(best, 4), "median_ms": round(med, 4),
            "best_qps": round(nq / (best / 1e3), 1), "median_qps": round(nq / (med / 1e3), 1)}


def fp64_reference(db, qs, mask_bits, k):
    """Exact top-k in FP64 (ties by lower id) for a query sample."""
    import numpy as np
    x64 = db.astype(np.float64)
    xn = (x64 * x64).sum(1)
    ids = np.empty((qs.shape[0], k), np.int64)
    dist = np.empty((qs.shape[0], k), np.float64)
    for s in range(0, qs.shape[0], 16):
        q64 = qs[s:s + 16].astype(np.float64)
        d = (q64 * q64).sum(1)[:, None] + xn[None, :] - 2.0 * (q64 @ x64.T)
        d[:, ~mask_bits] = np.inf
        for i in range(d.shape[0]):
            kth = np.partition(d[i], k - 1)[k - 1]
            cand = np.flatnonzero(d[i] <= kth)
            order = cand[np.lexsort((cand, d[i][cand]))][:k]
            ids[s + i], dist[s + i] = order, d[i][order]
    return ids, dist


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--dtype", choices=["f16", "bf16"], default="f16")
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--n", type=int, default=None, help="rows (default per workload)")
    ap.add_argument("--nq", type=int, default=None, help="queries (default per workload)")
    ap.add_argument("--valid-frac", type=float, default=0.05)
    ap.add_argument("--fp32-steps", type=int, default=2)
    ap.add_argument("--recall-sample", type=int, default=100)
    ap.add_argument("--cpu-sample", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--alt-valid-frac", type=float, default=0.95,
                    help="second mask density reported beside the headline (0 = skip)")
    ap.add_argument("--workload", choices=["flat", "ivfpq", "fp8", "backward", "build", "exact"],
                    default="flat",
                    help="flat: the headline matrix-core scan; ivfpq: 10M-row IVF-PQ uneven batch; "
                         "build: IVF-PQ table construction on the GPU")
    ap.add_argument("--cpu-rows", type=int, default=50_000,
                    help="build workload: rows of the CPU oracle's timed sample")
    ap.add_argument("--autotune", action="store_true",
                    help="measure the scan plan on this batch instead of using the committed "
                         "plan (flashvector_amd/tuned_plans.json)")
    ap.add_argument("--auto-plan", action="store_true",
                    help="use the planner's automatic plan (no committed or tuned plan)")
    ap.add_argument("--ivf-plan", default="committed",
                    help="IVF-PQ work-item plan: committed | default | tune | TARGET_ITEMS:CHUNK_MIN")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run a 1 GiB device copy first (reference point for traffic counters)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the scan would shard database rows per GPU and merge top-k lists).",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "ivfpq":
        return bench_ivfpq(args)
    if args.workload == "fp8":
        return bench_fp8(args)
    if args.workload == "backward":
        return bench_backward(args)
    if args.workload == "build":
        return bench_build(args)
    if args.workload == "exact":
        return bench_exact(args)
    args.n = args.n or 1_000_000
    args.nq = args.nq or 10_000

    import numpy as np
    import torch

    from flashvector_amd import lib as L
    from flashvector_amd import search as F
    from flashvector_amd import synthetic as S

    assert torch.cuda.is_available(), "bench.py needs a ROCm GPU"
    torch.cuda.set_device(0)
    n, nq, dim, k = args.n, args.nq, 128, args.k
    log(f"generating workload n={n} nq={nq} d={dim}")
    db, qs = S.flat_workload(n, nq, dim, 256, seeds=S.SEEDS["flat_mixed_1m"])
    mask = S.random_valid_mask(n, args.valid_frac, S.SEEDS["flat_mixed_1m"][2])
    mask_bits = S.unpack_bits(mask, n)
    d_db = torch.from_numpy(db).cuda()
    d_q = torch.from_numpy(qs).cuda()
    d_m = torch.from_numpy(mask).cuda()
    valid_count = int(mask_bits.sum())
    if args.calibrate_copy:
        src = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        torch.empty_like(src).copy_(src)
        del src
    idx = F.FlatIndex(d_db, args.dtype)
    out = L.result_tensor(nq, k, d_q.device)
    torch.cuda.synchronize()
    from flashvector_amd import autotune as T
    tuning, plan_src = None, "automatic"
    if args.autotune:
        log("autotuning the scan plan on this batch (shortlist fixed at 8k)")
        tuning = T.tuned(idx, d_q, k, "l2", d_m, valid_count, refresh=True, shortlist_factor=8)
        plan_src = "autotuned in this run"
    elif not args.auto_plan:
        tuning, key = T.committed(idx, k, "l2", valid_count, nq, False, 8)
        plan_src = f"committed ({key})" if tuning is not None else "automatic"

    def search(m=d_m, vc=valid_count, tn=tuning):
        idx.search(d_q, k, "l2", m, out=out, valid_count=vc, shortlist_factor=8, tuning=tn)

    log("warmup")
    for _ in range(args.warmup):
        search()
    torch.cuda.synchronize()

    log(f"timing {args.steps} steps")
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with L.KernelTimer() as kt:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        start.record()
        for _ in range(args.steps):
            search()
        stop.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        scan_ms, launches = kt.read()
    step_ms = start.elapsed_time(stop) / args.steps
    qps = nq / (step_ms / 1e3)
    scan_avg_ms = scan_ms / max(launches, 1)
    per_call = best_of(search, nq)
    # Two flop conventions for the dominant kernel: the contraction it performs
    # (2*Q*valid*d: the mask is compacted, masked rows are never scored) and
    # 2*Q*N*d over every row, which credits skipped rows as done and can pass
    # the peak; only the first is a roofline fraction.
    flops = 2.0 * nq * valid_count * dim
    flops_all = 2.0 * nq * n * dim
    achieved = flops / (scan_avg_ms / 1e3) / 1e12
    peak = MFMA_PEAK_TFLOPS[args.dtype]

    ids, dists = L.split_results(out)
    log("recall vs FP64 oracle")
    rs = min(args.recall_sample, nq)
    ref_ids, ref_d = fp64_reference(db, qs[:rs], mask_bits, k)
    hits = sum(len(set(ids[i].tolist()) & set(ref_ids[i].tolist())) for i in range(rs))
    recall = hits / (rs * k)
    max_err = float(np.max(np.abs(dists[:rs].astype(np.float64) - ref_d)))

    fp32 = None
    if args.fp32_steps > 0:
        # The plain exact FP32 GPU scan, twice: over the same valid rows only
        # (compacted on the host, ids mapped back: like for like with the
        # compacted matrix-core scan) and as a dense masked scan of all rows.
        def time_exact(rows, m):
            ws = torch.empty(L.load().fv_flat_exact_workspace_size(rows.shape[0], dim, nq, k, 0),
                             dtype=torch.uint8, device="cuda")
            F.flat_exact_search(rows, d_q, k, "l2", m, workspace=ws)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(args.fp32_steps):
                o = F.flat_exact_search(rows, d_q, k, "l2", m, workspace=ws)
            e1.record()
            torch.cuda.synchronize()
            return e0.elapsed_time(e1) / args.fp32_steps, o

        log("plain FP32 exact scan of the valid rows, then of all rows with the mask")
        valid_idx = np.flatnonzero(mask_bits)
        d_valid = d_db[torch.from_numpy(valid_idx).cuda()].contiguous()
        c_ms, c_out = time_exact(d_valid, None)
        c_ids, _ = L.split_results(c_out)
        c_ids = np.where(c_ids == 0xFFFFFFFF, 0xFFFFFFFF, valid_idx[np.minimum(c_ids, len(valid_idx) - 1)])
        del d_valid
        m_ms, m_out = time_exact(d_db, d_m)
        m_ids, _ = L.split_results(m_out)
        fp32 = {"valid_rows": {"ms_per_step": round(c_ms, 3), "qps": round(nq / (c_ms / 1e3), 1),
                               "speedup_of_value": round(c_ms / step_ms, 2),
                               "ids_agree_fraction": round(float(np.mean(c_ids == ids)), 5)},
                "all_rows_masked": {"ms_per_step": round(m_ms, 3),
                                    "qps": round(nq / (m_ms / 1e3), 1),
                                    "speedup_of_value": round(m_ms / step_ms, 2),
                                    "ids_agree_fraction": round(float(np.mean(m_ids == ids)), 5)}}

    legacy = None
    if args.dtype == "f16":
        # The reference's entry point (cuda_fp16_tc_search_batch) on the same
        # tables: the mask is compacted with its count taken on the device and
        # the scratch is the stream's leased buffer.
        log("reference entry point cuda_fp16_tc_search_batch")

        def legacy_call():
            F.fp16_tc_search_batch(idx.codes, idx.norms, d_q, k, dim, "l2", d_m, d_db)

        legacy_call()
        torch.cuda.synchronize()
        l0, l1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        l0.record()
        for _ in range(args.steps):
            legacy_call()
        l1.record()
        torch.cuda.synchronize()
        l_ms = l0.elapsed_time(l1) / args.steps
        l_ids, _ = L.split_results(F.fp16_tc_search_batch(idx.codes, idx.norms, d_q, k, dim, "l2",
                                                          d_m, d_db))
        legacy = {"entry_point": "cuda_fp16_tc_search_batch", "ms_per_step": round(l_ms, 4),
                  "qps": round(nq / (l_ms / 1e3), 1), "ratio_to_value": round(step_ms / l_ms, 4),
                  "ids_equal_value": bool(np.array_equal(l_ids, ids))}

    alt = None
    if args.alt_valid_frac > 0:
        log(f"second reading: mask keeping {args.alt_valid_frac:.0%} of rows")
        amask = S.random_valid_mask(n, args.alt_valid_frac, S.SEEDS["flat_mixed_1m"][2] + 100)
        abits = S.unpack_bits(amask, n)
        d_am = torch.from_numpy(amask).cuda()
        avc = int(abits.sum())
        atun = None
        if args.autotune:
            atun = T.tuned(idx, d_q, k, "l2", d_am, avc, refresh=True, shortlist_factor=8)
        elif not args.auto_plan:
            atun, _ = T.committed(idx, k, "l2", avc, nq, False, 8)
        search(d_am, avc, atun)
        torch.cuda.synchronize()
        a0, a1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        asteps = max(1, args.steps // 2)
        with L.KernelTimer() as akt:
            a0.record()
            for _ in range(asteps):
                search(d_am, avc, atun)
            a1.record()
            torch.cuda.synchronize()
            a_scan_ms, a_launches = akt.read()
        ams = a0.elapsed_time(a1) / asteps
        a_scan_avg = a_scan_ms / max(a_launches, 1)
        a_ids, a_d = L.split_results(out)
        r_ids, r_d = fp64_reference(db, qs[:rs], abits, k)
        ahits = sum(len(set(a_ids[i].tolist()) & set(r_ids[i].tolist())) for i in range(rs))
        alt = {"valid_fraction": args.alt_valid_frac, "ms_per_step": round(ams, 4),
               "qps": round(nq / (ams / 1e3), 1), "recall_at_k": round(ahits / (rs * k), 5),
               "max_abs_distance_error": float(np.max(np.abs(a_d[:rs].astype(np.float64) - r_d))),
               "end_to_end_tflops": round(2.0 * nq * avc * dim / (ams / 1e3) / 1e12, 2),
               "scan_kernel_avg_ms": round(a_scan_avg, 4),
               "scan_frac_of_peak": round(2.0 * nq * avc * dim / (a_scan_avg / 1e3) / 1e12
                                          / MFMA_PEAK_TFLOPS[args.dtype], 4)}

    cpu = None
    if not args.no_cpu:
        import oracle as O
        cores = len(os.sched_getaffinity(0))
        cs = min(args.cpu_sample, nq)
        log(f"CPU oracle on {cs} queries, {cores} threads")
        O.build()
        t0 = time.perf_counter()
        c_ids, _ = O.flat_topk(db, qs[:cs], k, "l2", valid_mask=mask, nthreads=cores)
        cpu_s = time.perf_counter() - t0
        cpu = {"value": round(cs / cpu_s, 2), "unit": "queries/s", "cores": cores,
               "kind": "port",
               "sample": f"{cs} of the {nq} queries against all {n} rows, exact FP32 "
                         f"(oracle/fv_oracle.c, OpenMP over queries), {cpu_s:.2f} s"}
        cpu["ids_match_gpu"] = bool(np.array_equal(c_ids, ids[:cs]))

    res = {
        "metric": "QPS @ recall@10>=0.999, 1Mx128 FP16-MFMA vs plain FP32 scan; % of roofline",
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16" if args.dtype == "f16" else "bf16",
        "data": "synthetic (seeded integer Gaussian mixture, 256 clusters)",
        "config": {"workload": "flat_l2_mfma_rerank", "n": n, "dim": dim, "num_queries": nq,
                   "top_k": k, "shortlist": 8 * k, "valid_fraction": args.valid_frac,
                   "storage": args.dtype, "parallelism": "single_gpu",
                   "plan": {"source": plan_src,
                            "n_chunks": 0 if tuning is None else int(tuning.n_chunks),
                            "bound_div": 0 if tuning is None else int(tuning.bound_div)}},
        "recall_at_k": round(recall, 5),
        "valid_rows": valid_count,
        "recall_sample_queries": rs,
        "max_abs_distance_error": max_err,
        "plain_fp32_scan": fp32,
        "reference_entry_point": legacy,
        "second_mask_reading": alt,
        "roofline": {"bound": "mfma", "achieved": round(achieved, 2), "peak": peak,
                     "unit": "TFLOP/s", "frac": round(achieved / peak, 4),
                     "traffic": measured_traffic("flat_mfma_scan_kernel/flat_l2_mfma_rerank"),
                     "kernel": "flat_mfma_scan_kernel", "kernel_avg_ms": round(scan_avg_ms, 4),
                     "algorithmic_flop_per_launch": flops,
                     "convention": "2*Q*valid*d: the contraction the kernel performs over the "
                                   "compacted valid rows",
                     "all_rows_2QNd": {"flop_per_launch": flops_all,
                                       "effective_tflops": round(flops_all / (scan_avg_ms / 1e3)
                                                                 / 1e12, 1),
                                       "note": "credits the 95% masked rows as done; not a "
                                               "roofline fraction"},
                     "frac_at_95pct_valid": None if alt is None else alt["scan_frac_of_peak"]},
        "reference_best_of_5": per_call,
        "cpu_baseline": cpu,
        "wall_s_timed": round(wall, 3),
    }
    print(json.dumps(res), flush=True)


def bench_ivfpq(args):
    """IVF-PQ ADC scan, batched and uneven: N=10M x 128 Gaussian mixture over
    4096 lists with Zipf(1.1) sizes, PQ 16 x 256, Q=8192 with per-query nprobe
    in {8,16,32,64} and k in {10,100}, shortlist 1000 reranked exactly in FP32,
    one launch per batch over list-ordered codes."""
    import numpy as np
    import torch

    from flashvector_amd import ivfpq_workload as W
    from flashvector_amd import lib as L
    from flashvector_amd import search as F

    torch.cuda.set_device(0)
    log("building the 10M-row IVF-PQ workload on the GPU")
    t0 = time.perf_counter()
    w = W.build()
    torch.cuda.synchronize()
    log(f"built in {time.perf_counter() - t0:.1f} s; largest list {int(w.sizes.max())}")
    if args.calibrate_copy:
        src = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        torch.empty_like(src).copy_(src)
        del src
    idx = F.IvfPqIndex(w.centroids, w.codebooks, w.codes, w.offsets, w.vec_ids, w.vectors)
    nq, S = w.queries.shape[0], 1000
    out = L.result_tensor(nq, 100, w.queries.device)

    # Work-item plan: the committed measured plan, the library default, an
    # explicit "target_items:chunk_min", or one autotuned in this run.
    from flashvector_amd import autotune as AT
    plan = {"source": args.ivf_plan}
    tuning = None
    if args.ivf_plan == "tune":
        tuning = idx.autotune(w.queries, 100, 64, S, top_k=w.top_k, nprobe=w.nprobe, trials=3)
        plan["source"] = "fv_ivf_pq_autotune in this run"
    elif args.ivf_plan == "committed":
        entry = AT.committed_entry("ivfpq_uneven_10m_q8192_s1000")
        if entry is not None:
            tuning = L.FvIvfPqTuning(entry["target_items"], entry["chunk_min"], 0.0, 0)
            plan["source"] = "committed (ivfpq_uneven_10m_q8192_s1000)"
        else:
            plan["source"] = "library default (no committed plan)"
    elif args.ivf_plan != "default":
        a_, b_ = (int(v) for v in args.ivf_plan.split(":"))
        tuning = L.FvIvfPqTuning(a_, b_, 0.0, 0)
    if tuning is not None:
        plan.update(target_items=int(tuning.target_items), chunk_min=int(tuning.chunk_min))

    def step():
        idx.search(w.queries, 100, 64, S, top_k=w.top_k, nprobe=w.nprobe, out=out, tuning=tuning)

    for _ in range(max(1, args.warmup)):
        step()
    torch.cuda.synchronize()
    eager = out.clone()
    steps = max(1, args.steps // 4)
    with L.KernelTimer() as kt:
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        scan_ms, launches = kt.read()
    # The whole uneven batch as one captured HIP graph; replays are the step.
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step()
    torch.cuda.current_stream().wait_stream(side)
    with torch.cuda.graph(g):
        step()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(steps):
        g.replay()
    e1.record()
    torch.cuda.synchronize()
    step_ms = e0.elapsed_time(e1) / steps
    graph_matches_eager = bool(torch.equal(out, eager))
    scan_avg = scan_ms / max(launches, 1)
    per_call = best_of(g.replay, nq)
    nbytes = W.scanned_bytes(w, S)
    ids, dists = L.split_results(out)
    ks = w.top_k.cpu().numpy()

    log("recall@k against the exact FP32 scan on a query sample")
    rs = min(args.recall_sample, nq)
    ws = torch.empty(L.load().fv_flat_exact_workspace_size(w.vectors.shape[0], 128, rs, 100, 0),
                     dtype=torch.uint8, device="cuda")
    ex = F.flat_exact_search(w.vectors, w.queries[:rs].contiguous(), 100, workspace=ws)
    ex_ids, _ = L.split_results(ex)
    del ws
    hits = tot = 0
    for i in range(rs):
        k = int(ks[i])
        hits += len(set(ids[i, :k].tolist()) & set(ex_ids[i, :k].tolist()))
        tot += k
    cpu = None
    ids_match = None
    if not args.no_cpu:
        import oracle as O
        O.build()
        cs = min(32, nq)
        log(f"CPU oracle (ADC + rerank) on {cs} queries")
        h = O.IvfPqHost(w.centroids.cpu().numpy(), w.codebooks.cpu().numpy(),
                        w.codes.cpu().numpy(), w.offsets.cpu().numpy().view(np.uint32),
                        w.vec_ids.cpu().numpy().view(np.uint32), w.vectors.cpu().numpy(), w.m_pq)
        cores = len(os.sched_getaffinity(0))
        t0 = time.perf_counter()
        c_ids, c_d = O.ivf_pq_search_uneven(h, w.queries[:cs].cpu().numpy(), ks[:cs],
                                            w.nprobe[:cs].cpu().numpy(), S, out_stride=100,
                                            nthreads=cores)
        cpu_s = time.perf_counter() - t0
        ids_match = bool(np.array_equal(c_ids, ids[:cs]) and np.array_equal(c_d, dists[:cs]))
        cpu = {"value": round(cs / cpu_s, 2), "unit": "queries/s", "cores": cores, "kind": "port",
               "sample": f"{cs} of the {nq} queries, same tables (oracle/fv_oracle.c ADC + rerank, "
                         f"OpenMP over queries), {cpu_s:.2f} s"}
    achieved = nbytes / (scan_avg / 1e3) / 1e9
    ivf_traffic = measured_traffic("ivf_scan_kernel/ivfpq_uneven_10m")
    res = {
        "metric": "IVF-PQ batched uneven QPS (10M x 128, nlist 4096 Zipf, M=16, shortlist 1000)",
        "value": round(nq / (step_ms / 1e3), 1), "unit": "queries/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": round(step_ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded Gaussian mixture, Zipf list sizes)",
        "config": {"workload": "ivfpq_uneven_10m", "n": int(w.vectors.shape[0]), "dim": 128,
                   "nlist": 4096, "m_pq": 16, "num_queries": nq, "shortlist": S,
                   "nprobe": "per query in {8,16,32,64}", "top_k": "per query in {10,100}",
                   "plan": plan},
        "recall_at_k_vs_exact": round(hits / tot, 5), "recall_sample_queries": rs,
        "ids_bit_exact_vs_cpu_adc": ids_match,
        "hip_graph": True, "graph_replay_matches_eager": graph_matches_eager,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(achieved / 8000.0, 4),
                     "traffic": ivf_traffic,
                     "kernel": "ivf_scan_kernel", "kernel_avg_ms": round(scan_avg, 4),
                     "algorithmic_bytes_per_launch": nbytes,
                     "convention": "achieved = the probed lists' code and id bytes, query-major "
                                   "(every probed entry once per query), including lists that the "
                                   "LUT-minimum bound skips without reading them",
                     "measured_traffic_rate_gbps": (round(ivf_traffic / (scan_avg / 1e3) / 1e9, 1)
                                                    if ivf_traffic else None)},
        "reference_best_of_5": per_call,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


def bench_exact(args):
    """Exact FP32 flat scan through the reference's entry point
    (cuda_flat_scan_batch): N=100,000 x d=128 integer mixture rows, Q=1,000
    queries, k=10, L2, no mask. One step is one batched call over all Q
    queries (its scratch is the stream's leased library buffer).
    Every query's ids and distances are checked bit-exact against the CPU
    oracle, which is also the CPU baseline (all 1,000 queries)."""
    import numpy as np
    import torch

    from flashvector_amd import lib as L
    from flashvector_amd import search as F
    from flashvector_amd import synthetic as S

    torch.cuda.set_device(0)
    n, nq, dim, k = args.n or 100_000, args.nq or 1_000, 128, args.k
    log(f"generating workload n={n} nq={nq} d={dim}")
    db, qs = S.flat_workload(n, nq, dim, 256, seeds=S.SEEDS["flat_fp32_100k"])
    d_db = torch.from_numpy(db).cuda()
    d_q = torch.from_numpy(qs).cuda()
    out = L.result_tensor(nq, k, d_q.device)
    if args.calibrate_copy:
        src = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        torch.empty_like(src).copy_(src)
        del src

    def search():
        F.flat_scan_batch(d_db, d_q, k, "l2", out=out)

    log("warmup")
    for _ in range(args.warmup):
        search()
    torch.cuda.synchronize()
    log(f"timing {args.steps} steps")
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with L.KernelTimer() as kt:
        torch.cuda.synchronize()
        start.record()
        for _ in range(args.steps):
            search()
        stop.record()
        torch.cuda.synchronize()
        scan_ms, launches = kt.read()
    step_ms = start.elapsed_time(stop) / args.steps
    scan_avg = scan_ms / max(launches, 1)
    per_call = best_of(search, nq)
    # The flat-scan convention 2*Q*N*d; the kernel issues 3 unfused FP32 ops
    # per (query, row, dim) (sub, mul, add: the CPU path's rounding), whose
    # rate is reported beside it.
    flops = 2.0 * nq * n * dim
    ops = 3.0 * nq * n * dim
    achieved = flops / (scan_avg / 1e3) / 1e12
    ids, dists = L.split_results(out)

    import oracle as O
    O.build()
    cores = len(os.sched_getaffinity(0))
    log(f"CPU oracle on all {nq} queries, {cores} threads")
    t0 = time.perf_counter()
    ref_ids, ref_d = O.flat_scan_batch(db, qs, k, "l2")
    cpu_s = time.perf_counter() - t0
    exact = bool(np.array_equal(ids, ref_ids) and
                 np.array_equal(dists.view(np.uint32), ref_d.view(np.uint32)))
    res = {
        "metric": "QPS, exact FP32 flat L2 top-10 (100k x 128, Q=1000), bit-exact vs CPU oracle",
        "value": round(nq / (step_ms / 1e3), 1), "unit": "queries/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(step_ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded integer Gaussian mixture, 256 clusters)",
        "config": {"workload": "flat_l2_exact_fp32", "n": n, "dim": dim, "num_queries": nq,
                   "top_k": k, "entry_point": "cuda_flat_scan_batch",
                   "parallelism": "single_gpu"},
        "bit_exact_vs_oracle": exact, "queries_checked": nq,
        "roofline": {"bound": "fp32-valu", "achieved": round(achieved, 2),
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                     "traffic": measured_traffic("flat_exact_scan_kernel"),
                     "kernel": "flat_exact_scan_kernel", "kernel_avg_ms": round(scan_avg, 4),
                     "launches": launches, "algorithmic_flop_per_launch": flops,
                     "convention": "2*Q*N*d",
                     "issued_unfused_ops": {"per_launch": ops,
                                            "tops": round(ops / (scan_avg / 1e3) / 1e12, 2)}},
        "reference_best_of_5": per_call,
        "cpu_baseline": {"value": round(nq / cpu_s, 1), "unit": "queries/s", "cores": cores,
                         "kind": "port",
                         "sample": f"all {nq} queries against all {n} rows, exact FP32 "
                                   f"(oracle/fv_oracle.c, OpenMP over queries), {cpu_s:.2f} s"},
    }
    print(json.dumps(res), flush=True)
    if not exact:
        sys.exit(1)


def bench_build(args):
    """IVF-PQ table construction on the GPU: N=1M x 128 Gaussian-mixture rows
    (integers 0-255), nlist 256, M=16, 25 coarse and 25 PQ Lloyd iterations at
    most (the reference's SIFT1M build settings, README.md:188, index.rs:414,428),
    seed 1. One build is the step: k-means++ seeding, Lloyd iterations,
    residuals, 16 sub-space k-means, encoding and posting lists."""
    import numpy as np
    import torch

    from flashvector_amd import gpu_build as G
    from flashvector_amd import lib as L
    from flashvector_amd import synthetic as S

    torch.cuda.set_device(0)
    n, dim, nlist, m_pq, iters = args.n or 1_000_000, 128, 256, 16, 25
    log(f"generating {n} x {dim} rows")
    db, _ = S.flat_workload(n, 1, dim, 256, seeds=S.SEEDS["flat_mixed_1m"])
    x = torch.from_numpy(db).cuda()
    G.build_ivfpq(x[:20000].contiguous(), 64, m_pq, 2, 2, seed=1)   # loads the kernels
    torch.cuda.synchronize()
    log("timed build")
    with L.KernelTimer() as kt:
        t0 = time.perf_counter()
        t = G.build_ivfpq(x, nlist, m_pq, iters, iters, seed=1)
        torch.cuda.synchronize()
        build_s = time.perf_counter() - t0
        assign_ms, launches = kt.read()
    sizes = np.diff(t.offsets.cpu().numpy().astype(np.int64))
    assign_avg = assign_ms / max(launches, 1)
    flops = 2.0 * n * nlist * dim          # 2*N*nlist*d (the kernel issues 3 unfused ops each)
    achieved = flops / (assign_avg / 1e3) / 1e12
    cpu = None
    if not args.no_cpu:
        import oracle as O
        O.build()
        ns = min(args.cpu_rows, n)
        cores = len(os.sched_getaffinity(0))
        log(f"CPU oracle build on the first {ns} rows")
        t0 = time.perf_counter()
        o = O.ivf_pq_build(db[:ns], nlist, m_pq, iters, iters, 1)
        cpu_s = time.perf_counter() - t0
        cpu = {"value": round(ns / cpu_s, 1), "unit": "rows/s", "cores": cores, "kind": "port",
               "sample": f"first {ns} of the {n} rows, same nlist/M/iterations/seed "
                         f"(oracle/fv_oracle.c, OpenMP over rows in the assign steps): "
                         f"{cpu_s:.2f} s, iterations {o['iters'][0]} coarse"}
    res = {
        "metric": "IVF-PQ table build time (1M x 128, nlist 256, M=16, 25+25 iterations)",
        "value": round(build_s, 3), "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(build_s * 1e3, 1), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic (seeded Gaussian mixture)",
        "config": {"workload": "ivfpq_build_1m", "n": n, "dim": dim, "nlist": nlist, "m_pq": m_pq,
                   "coarse_iters_max": iters, "pq_iters_max": iters, "seed": 1},
        "rows_per_s": round(n / build_s, 1), "iterations": t.iterations,
        "list_sizes": {"min": int(sizes.min()), "max": int(sizes.max())},
        "roofline": {"bound": "fp32-valu", "achieved": round(achieved, 2),
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": None,
                     "kernel": "km_assign_kernel (coarse)", "kernel_avg_ms": round(assign_avg, 4),
                     "launches": launches, "algorithmic_flops_per_launch": flops,
                     "convention": "2*N*nlist*d",
                     "issued_unfused_tops": round(1.5 * flops / (assign_avg / 1e3) / 1e12, 2)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


def bench_fp8(args):
    """FP8 flat inner-product scan with mixed k: N=5,000,000 unit rows at
    d=768 from a von Mises-Fisher mixture (1024 clusters, kappa 50), stored as
    FP8 e4m3 with per-row FP32 scales and scored on FP8 MFMA with FP32
    accumulation; Q=4096 queries with k_i log-uniform on [1, 1024]; shortlist
    4 k_i reranked exactly in FP32 with FP64 re-scoring at the boundary. One
    step is one replay of the whole batch captured as a HIP graph."""
    import numpy as np
    import torch

    from flashvector_amd import lib as L
    from flashvector_amd import search as F
    from flashvector_amd import synthetic as S

    torch.cuda.set_device(0)
    n, nq, dim, kmax = args.n or 5_000_000, args.nq or 4096, 768, 1024
    sd = S.VMF_SEEDS
    log(f"generating vMF workload n={n} nq={nq} d={dim} on the GPU")
    t0 = time.perf_counter()
    mu = S.vmf_mean_directions(1024, dim, sd[0], "cuda")
    db = S.vmf_points(mu, n, 50.0, sd[1])
    qs = S.vmf_points(mu, nq, 50.0, sd[2])
    ks = S.log_uniform_k(nq, kmax, sd[3])
    d_k = torch.from_numpy(ks).cuda()
    if args.calibrate_copy:
        src = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        torch.empty_like(src).copy_(src)
        del src
    idx = F.FlatIndex(db, "fp8")
    torch.cuda.synchronize()
    log(f"generated and encoded in {time.perf_counter() - t0:.1f} s")
    out = L.result_tensor(nq, kmax, qs.device)

    # The k sum sizes the candidate buffers per query (each query's own
    # shortlist instead of the deepest one for all).
    k_sum = int(ks.sum())

    def step():
        idx.search(qs, kmax, "ip", shortlist_factor=4, top_k_per_query=d_k, f64_boundary=True,
                   out=out, top_k_sum=k_sum)

    for _ in range(max(1, args.warmup)):
        step()
    torch.cuda.synchronize()
    eager = out.clone()
    # Main-pass kernel time from eager calls (events around the kernel).
    ksteps = max(1, args.steps // 4)
    with L.KernelTimer() as kt:
        for _ in range(ksteps):
            step()
        torch.cuda.synchronize()
        scan_ms, launches = kt.read()
    scan_avg = scan_ms / max(launches, 1)
    log("capturing the batch as one HIP graph")
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step()
    torch.cuda.current_stream().wait_stream(side)
    with torch.cuda.graph(g):
        step()
    out.zero_()
    g.replay()
    torch.cuda.synchronize()
    graph_matches_eager = bool(torch.equal(out, eager))
    steps = max(1, args.steps // 2)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    e0.record()
    for _ in range(steps):
        g.replay()
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    step_ms = e0.elapsed_time(e1) / steps
    graph_matches_eager = graph_matches_eager and bool(torch.equal(out, eager))
    per_call = best_of(g.replay, nq)
    ids, dists = L.split_results(out)

    rs = min(args.recall_sample, nq)
    log(f"FP64 oracle on {rs} queries (GPU FP64 products)")
    q64 = qs[:rs].double()
    sc = torch.empty((rs, n), dtype=torch.float64, device="cuda")
    for c0 in range(0, n, 1 << 19):
        c1 = min(n, c0 + (1 << 19))
        sc[:, c0:c1] = q64 @ db[c0:c1].double().T
    top = torch.topk(sc, kmax, dim=1).indices.cpu().numpy()
    hits = tot = 0
    rec = []
    err = 0.0
    for i in range(rs):
        k = int(ks[i])
        h = len(set(ids[i, :k].tolist()) & set(top[i, :k].tolist()))
        hits, tot = hits + h, tot + k
        rec.append(h / k)
        got = torch.from_numpy(ids[i, :k].astype(np.int64)).cuda()
        ref = -sc[i, got].cpu().numpy()
        err = max(err, float(np.max(np.abs(dists[i, :k].astype(np.float64) - ref))))
    small = [r for r, k in zip(rec, ks[:rs]) if k <= 10]
    del sc

    cpu = None
    if not args.no_cpu:
        import oracle as O
        O.build()
        cores = len(os.sched_getaffinity(0))
        cs = min(cores, nq)
        log(f"CPU oracle (exact FP32 scan) on {cs} queries, {cores} threads")
        dbn = db.cpu().numpy()
        qn = qs[:cs].cpu().numpy()
        t0 = time.perf_counter()
        O.flat_topk(dbn, qn, int(ks[:cs].max()), "ip", nthreads=cores)
        cpu_s = time.perf_counter() - t0
        cpu = {"value": round(cs / cpu_s, 3), "unit": "queries/s", "cores": cores, "kind": "port",
               "sample": f"{cs} of the {nq} queries against all {n} rows, exact FP32 inner "
                         f"product (oracle/fv_oracle.c, OpenMP over queries), {cpu_s:.2f} s"}
        del dbn

    flops = 2.0 * nq * n * dim
    achieved = flops / (scan_avg / 1e3) / 1e12
    peak = 5000.0
    res = {
        "metric": "FP8 flat IP QPS, 5M x 768 vMF, mixed k in [1,1024], 4k shortlist + FP32 rerank "
                  "+ FP64 boundary, one HIP graph; recall vs FP64",
        "value": round(nq / (step_ms / 1e3), 1), "unit": "queries/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": round(step_ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp8_e4m3",
        "data": "synthetic (seeded vMF mixture, 1024 clusters, kappa 50, unit rows)",
        "config": {"workload": "flat_ip_fp8_mixed_k", "n": n, "dim": dim, "num_queries": nq,
                   "top_k": "per query, log-uniform on [1,1024]", "shortlist": "4 k_i",
                   "f64_boundary": True, "hip_graph": True},
        "recall_at_k": round(hits / tot, 5), "recall_at_k_mean_per_query": round(float(np.mean(rec)), 5),
        "recall_at_k_queries_k_le_10": round(float(np.mean(small)), 5) if small else None,
        "recall_sample_queries": rs, "max_abs_distance_error_vs_fp64": err,
        "graph_replay_matches_eager": graph_matches_eager,
        "workspace_gb": round(idx._ws.numel() / 1e9, 3),
        "roofline": {"bound": "mfma", "achieved": round(achieved, 2), "peak": peak,
                     "unit": "TFLOP/s", "frac": round(achieved / peak, 4),
                     "traffic": measured_traffic("flat_mfma_scan_kernel/flat_ip_fp8_mixed_k"),
                     "kernel": "flat_mfma_scan_kernel", "kernel_avg_ms": round(scan_avg, 4),
                     "algorithmic_flop_per_launch": flops},
        "reference_best_of_5": per_call,
        "cpu_baseline": cpu,
        "wall_s_timed": round(wall, 3),
    }
    print(json.dumps(res), flush=True)


def bench_backward(args):
    """Differentiable top-k retrieval: B=4096 fp32 queries, a database of
    N=1,000,000 BF16-valued unit rows at d=768 (vMF mixture), k=64 inner
    product, loss = mean CE(softmax(s/0.05)) against a seeded target position
    per query. One step = forward search (BF16 MFMA scan, FP32 rerank, FP64
    boundary) + loss + hand-written backward (dQ and scatter-add dX, scores
    recomputed)."""
    import numpy as np
    import torch

    from flashvector_amd import autograd as A
    from flashvector_amd import lib as L
    from flashvector_amd import search as F
    from flashvector_amd import synthetic as S

    torch.cuda.set_device(0)
    n, b, dim, k, tau = args.n or 1_000_000, args.nq or 4096, 768, 64, 0.05
    log(f"generating vMF workload n={n} b={b} d={dim}")
    mu = S.vmf_mean_directions(1024, dim, 41, "cuda")
    x = S.vmf_points(mu, n, 50.0, 42).bfloat16().float().requires_grad_(True)
    q = S.vmf_points(mu, b, 50.0, 41 + 1000).requires_grad_(True)
    rng = np.random.default_rng(43)
    t = torch.from_numpy(rng.integers(0, k, size=b).astype(np.int32)).cuda()
    idx = F.FlatIndex(x.detach(), "bf16")
    torch.cuda.synchronize()

    def step():
        q.grad = None
        x.grad = None
        loss = A.topk_softmax_ce(q, x, k, t, tau, index=idx)
        loss.backward()
        return loss

    for _ in range(max(1, args.warmup)):
        step()
    torch.cuda.synchronize()
    steps = max(1, args.steps // 4)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    fwd_ms = bwd_ms = 0.0
    for _ in range(steps):
        q.grad = None
        x.grad = None
        ev[0].record()
        loss = A.topk_softmax_ce(q, x, k, t, tau, index=idx)
        ev[1].record()
        loss.backward()
        ev[2].record()
        torch.cuda.synchronize()
        fwd_ms += ev[0].elapsed_time(ev[1])
        bwd_ms += ev[1].elapsed_time(ev[2])
    fwd_ms /= steps
    bwd_ms /= steps
    step_ms = fwd_ms + bwd_ms
    with L.KernelTimer() as kt:
        A.topk_inner_product(q, x, k, index=idx)
        torch.cuda.synchronize()
        scan_ms, launches = kt.read()
    # Backward kernel alone (dX zeroed outside the timed region).
    scores, ids = A.topk_inner_product(q, x, k, index=idx)
    dq = torch.empty_like(q)
    dx = torch.zeros_like(x)
    g1 = torch.ones(1, device="cuda")
    lib = L.load()
    import ctypes
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 5
    e0.record()
    for _ in range(reps):
        lib.fv_topk_ip_backward(L.ptr(q.detach()), b, L.ptr(x.detach()), n, dim, L.ptr(ids), k, None,
                                L.ptr(t), ctypes.c_float(1.0 / tau), ctypes.c_float(1.0 / b),
                                L.ptr(g1), L.ptr(dq), L.ptr(dx), None, L.stream_handle())
    e1.record()
    torch.cuda.synchronize()
    kern_ms = e0.elapsed_time(e1) / reps
    bwd_bytes = b * k * dim * (4 + 8) + 2 * b * dim * 4

    log("dense FP64 reference gradient on a 20,000-row subset")
    sub = 20000
    xs = x.detach()[:sub].clone().requires_grad_(True)
    qs = q.detach()[:64].clone().requires_grad_(True)
    ls = A.topk_softmax_ce(qs, xs, k, t[:64], tau)
    ls.backward()
    q64 = qs.detach().double().cpu().requires_grad_(True)
    x64 = xs.detach().double().cpu().requires_grad_(True)
    s64 = q64 @ x64.T
    top = torch.topk(s64.detach(), k, dim=1).indices
    z = s64.gather(1, top) / tau
    ref = (torch.logsumexp(z, 1) - z.gather(1, t[:64].long().cpu()[:, None])[:, 0]).mean()
    ref.backward()

    def rel(a, r):
        a, r = a.detach().cpu().double(), r.detach()
        return float((a - r).abs().max() / r.abs().max())

    flops = 2.0 * b * n * dim
    achieved = flops / (scan_ms / max(launches, 1) / 1e3) / 1e12
    res = {
        "metric": "differentiable top-k retrieval: forward+backward queries/s (B=4096, 1M x 768 BF16, k=64)",
        "value": round(b / (step_ms / 1e3), 1), "unit": "queries/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": round(step_ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16 scan, fp32 rerank and gradients",
        "data": "synthetic (seeded vMF mixture, 1024 clusters, kappa 50; BF16-valued rows)",
        "config": {"workload": "topk_backward_ce", "n": n, "dim": dim, "num_queries": b, "top_k": k,
                   "tau": tau, "shortlist": "8 k", "f64_boundary": True},
        "forward_ms": round(fwd_ms, 4), "backward_ms": round(bwd_ms, 4),
        "backward_kernel_ms": round(kern_ms, 4),
        "backward_kernel_gbps": round(bwd_bytes / (kern_ms / 1e3) / 1e9, 1),
        "grad_check_subset": {"rows": sub, "queries": 64,
                              "loss_rel_err": abs(ls.item() - ref.item()) / abs(ref.item()),
                              "dq_max_rel_err": rel(qs.grad, q64.grad),
                              "dx_max_rel_err": rel(xs.grad, x64.grad)},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 2), "peak": 2500.0,
                     "unit": "TFLOP/s", "frac": round(achieved / 2500.0, 4),
                     "traffic": measured_traffic("flat_mfma_scan_kernel/topk_backward_ce"),
                     "kernel": "flat_mfma_scan_kernel (forward)",
                     "kernel_avg_ms": round(scan_ms / max(launches, 1), 4),
                     "algorithmic_flop_per_launch": flops},
        "cpu_baseline": None,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
