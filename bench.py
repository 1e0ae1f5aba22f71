#!/usr/bin/env python3
"""Filtered top-100 benchmark on one MI355X.

One step = one fbf_search over the whole query batch of the workload (host
planning, scan, merge and -- for the approximate modes -- the exact FP32
rerank), with base vectors, query vectors and outputs resident in HBM.

Prints ONE JSON line (the driver's contract) with a ``roofline`` object for
the dominant kernel and a ``cpu_baseline`` object timed on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X published peaks.
PEAK_HBM_GBS = 8000.0
PEAK_TFLOPS = {"bf16": 2500.0, "fp8": 5000.0, "int8": 5000.0, "fp32": 157.3}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--config", type=int, default=1)
    p.add_argument("--mode", default="bf16", choices=["exact", "sq8", "bf16", "fp8", "sq8sym", "exact_bf16"])
    p.add_argument("--kprime", type=int, default=0)
    p.add_argument("--k", type=int, default=100)
    p.add_argument("--n", type=int, default=0, help="override base rows (smaller runs)")
    p.add_argument("--nq", type=int, default=0, help="override query count")
    p.add_argument("--chunk-rows", type=int, default=0)
    p.add_argument("--group-queries", type=int, default=0)
    p.add_argument("--persistent", action="store_true")
    p.add_argument("--prepass-div", type=int, default=0)
    p.add_argument("--prepass-margin", type=int, default=0)
    p.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline sample time")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--exact-bf16", action=argparse.BooleanOptionalAction, default=True,
                   help="in BF16 runs, also time the certified exact mode (results identical to exact FP32)")
    p.add_argument("--recall-sample", type=int, default=512,
                   help="type-stratified queries checked against the FP64 exact answer")
    p.add_argument("--autotune", action=argparse.BooleanOptionalAction, default=True,
                   help="pick group_queries / chunk_rows / kprime with the project's autotuner (untimed) "
                        "unless they are given (approximate modes)")
    p.add_argument("--graph", action=argparse.BooleanOptionalAction, default=None,
                   help="also time replays of the batch captured as one HIP graph (fbf_search_graph_*); "
                        "default: on for the approximate modes")
    p.add_argument("--dropin", action="store_true",
                   help="only time the reference-shaped bruteforce() drop-in on the router-routed slices "
                        "(selectivity below bf_thr / bf_thr3) beside the CPU bruteforce(), and exit")
    p.add_argument("--subinterval", action="store_true",
                   help="only time the type-2 sub-interval path (fbf_search_subinterval) on the type-2 queries "
                        "the router sends to the graph side, beside the oracle's restatement on the host, and exit")
    p.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                   help="only run a plain device copy of MIB MiB (HBM counter calibration) and exit")
    return p.parse_args()


def main() -> int:
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; multi-GPU sharding is a design "
                          "(query-sharded replicas or row shards + one top-k gather), not a run"}))
        return 2
    import torch
    import __graft_entry__ as ge

    if a.calibrate_copy:
        # A plain copy kernel of known size: under rocprofv3 --pmc FETCH_SIZE /
        # WRITE_SIZE its counters give the bytes-per-counter-unit scale.
        n = a.calibrate_copy * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": 4 * n, "launches": 3}), flush=True)
        return 0

    pkg = ge.load_pkg()
    fbf, dg = pkg.fbf, pkg.datagen
    cfg = dg.CONFIGS[a.config]
    n = a.n or cfg.n
    nq = a.nq or cfg.nq
    t0 = time.time()
    log(f"generating {cfg.name}: n={n} nq={nq}")
    vecs, labels, ts = dg.make_base(cfg, n=n)
    q, metas = dg.make_queries(cfg, nq=nq)
    log(f"data ready in {time.time() - t0:.1f}s")
    if a.dropin:
        return dropin(a, ge, fbf, cfg, vecs, labels, ts, q, metas)
    if a.subinterval:
        return subinterval_bench(a, ge, fbf, cfg, vecs, labels, ts, q, metas)

    mode = a.mode
    repr_flags = fbf.MODE_REPR[fbf.MODES[mode]]
    if mode != "exact":
        # FP32 rows for the rerank; BF16 tiles for the certified exact search
        # that gives the whole batch's exact answer for the recall check.
        repr_flags |= fbf.REPR_F32 | fbf.REPR_BF16
    dev = torch.device("cuda:0")
    t0 = time.time()
    ds = fbf.Dataset(vecs, labels, ts, repr_flags=repr_flags)
    log(f"dataset resident in {time.time() - t0:.1f}s")
    qd = torch.from_numpy(q).to(dev)
    out = (torch.empty((nq, a.k), dtype=torch.int32, device=dev),
           torch.empty((nq, a.k), dtype=torch.float32, device=dev))
    kw = dict(k=a.k, mode=mode, kprime=a.kprime, chunk_rows=a.chunk_rows, group_queries=a.group_queries,
              persistent=a.persistent, out=out, prepass_div=a.prepass_div, prepass_margin=a.prepass_margin)
    tuned = None
    if a.autotune and mode not in ("exact", "exact_bf16") and not (a.kprime or a.chunk_rows or a.group_queries):
        t0 = time.time()
        res = pkg.autotune.tune(ds, qd, metas, k=a.k, mode=mode, log=log)
        tuned = dict(res.opts, sample_recall=res.recall)
        kw.update(res.opts)
        log(f"autotune picked {res.opts} ({res.ms:.1f} ms, sample recall {res.recall:.4f}) in {time.time() - t0:.1f}s")

    # Uneven batches (the C4 workload): the timed step runs all batches as one
    # search -- one persistent launch per pass over a work queue of (query
    # group, row chunk) items; for comparison the same batches are also run as
    # one search each (untimed for the result line, logged).
    per_batch = None
    if cfg.uneven_batches:
        sizes = dg.batch_sizes(cfg, nq)
        off = np.concatenate([[0], np.cumsum(sizes)])
        kwb = dict(kw)
        kwb.pop("out")
        ds.search(qd, metas, **kwb)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(len(sizes)):
            ds.search(qd[off[i]:off[i + 1]], metas[off[i]:off[i + 1]], **kwb)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        per_batch = {"batches": len(sizes), "qps": nq / dt, "ms": 1e3 * dt}
        log(f"{len(sizes)} uneven batches (sizes {min(sizes)}..{max(sizes)}) one search each: "
            f"{nq / dt:.0f} QPS ({1e3 * dt:.1f} ms)")

    for i in range(a.warmup):
        ds.search(qd, metas, **kw)
        torch.cuda.synchronize()
        log(f"warmup step {i} done")
    ds.profile(True)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(a.steps):
        ds.search(qd, metas, **kw)
    torch.cuda.synchronize()
    t_end = time.perf_counter()
    ds.profile(False)
    elapsed = t_end - t_start
    stats = ds.last_stats()
    pairs_prefix, pairs_main = ds.last_pass_pairs()
    rescans = ds.last_rescans()
    fallbacks = ds.last_fallbacks()
    n_main, ms_main = ds.profile_collect("scan_main")
    n_pre, ms_pre = ds.profile_collect("scan_prefix")
    n_resc, ms_resc = ds.profile_collect("scan_rescan")
    n_scan, ms_scan = n_main + n_pre + n_resc, ms_main + ms_pre + ms_resc
    n_merge, ms_merge = ds.profile_collect("merge")
    n_rr, ms_rr = ds.profile_collect("rerank")
    n_plan, ms_plan = ds.profile_collect("plan")
    n_merge += n_rr
    ms_merge += ms_rr
    ds.profile_collect("all")
    ms_step = 1000.0 * elapsed / a.steps
    qps = nq * a.steps / elapsed
    log(f"{a.steps} steps in {elapsed:.3f}s: {qps:.0f} QPS; scan {ms_scan / max(1, a.steps):.1f} ms/step "
        f"({n_scan} launches; main {ms_main / max(1, a.steps):.1f}, prefix {ms_pre / max(1, a.steps):.1f}, "
        f"rescan {ms_resc / max(1, a.steps):.2f}), merge+rerank {ms_merge / max(1, a.steps):.1f} ms/step, "
        f"{stats['work_items']} work items, {stats['pairs_computed'] / max(1, stats['pairs']):.3f}x pairs computed, "
        f"{rescans} queries rescanned, host planning {ms_plan / max(1, a.steps):.1f} ms/step"
        + (f", {fallbacks} queries left to the FP32 scan by the certificate" if mode == "exact_bf16" else ""))

    ids = out[0].cpu().numpy().view(np.uint32)
    dists = out[1].cpu().numpy()

    # The same batch captured once as a HIP graph (planning and upload
    # untimed), then replayed: the device work alone, with no host planning.
    graph_qps = None
    if mode != "exact_bf16" and (a.graph or (a.graph is None and mode != "exact")):
        kwg = {key: kw[key] for key in ("k", "mode", "kprime", "chunk_rows", "group_queries", "persistent",
                                       "prepass_div", "prepass_margin")}
        g = pkg.fbf.SearchGraph(ds, qd, metas, **kwg)
        try:
            g.launch()
            torch.cuda.synchronize()
            same = bool(np.array_equal(g.ids.cpu().numpy().view(np.uint32), ids))
            t0 = time.perf_counter()
            for i in range(a.steps):
                g.launch()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
        finally:
            g.close()
        graph_qps = nq * a.steps / dt
        log(f"graph replays: {a.steps} in {dt:.3f}s: {graph_qps:.0f} QPS ({1e3 * dt / a.steps:.1f} ms/step), "
            f"ids equal to the eager search: {same}")

    # Any-batch graph (metas in device memory): every replay routes the
    # batch, plans it on the device and searches -- no host work at all.
    any_qps = None
    if mode in ("bf16", "fp8", "sq8", "sq8sym") and a.graph is not False:
        kwg = {key: kw[key] for key in ("k", "mode", "kprime", "chunk_rows", "group_queries", "persistent",
                                       "prepass_div", "prepass_margin")}
        md = torch.from_numpy(np.ascontiguousarray(metas)).cuda()
        g = pkg.fbf.SearchGraph(ds, qd, md, **kwg)
        try:
            g.launch()
            torch.cuda.synchronize()
            same = bool(np.array_equal(g.ids.cpu().numpy().view(np.uint32), ids))
            info = g.plan_info()
            t0 = time.perf_counter()
            for i in range(a.steps):
                g.launch()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
        finally:
            g.close()
            del md
        any_qps = nq * a.steps / dt if not info["overflow"] else None
        log(f"any-batch graph replays (device router + planner): {a.steps} in {dt:.3f}s: "
            + (f"{any_qps:.0f} QPS ({1e3 * dt / a.steps:.1f} ms/step)" if any_qps else "plan over the graph's capacity")
            + f", plan {info}, ids equal to the eager search: {same}")

    # The certified exact mode on the same dataset and batch (BF16 scan + FP32
    # rerank + certificate + FP32 fallback): exact results, timed like the
    # main line (1 warmup, then min(steps, 3) timed searches).
    exact_cert = None
    exact_ids = None
    if mode == "bf16" and a.exact_bf16:
        kwe = dict(k=a.k, mode="exact_bf16", out=out)
        ds.search(qd, metas, **kwe)
        torch.cuda.synchronize()
        exact_ids = out[0].cpu().numpy().view(np.uint32).copy()
        ne = min(a.steps, 3)
        t0 = time.perf_counter()
        for i in range(ne):
            ds.search(qd, metas, **kwe)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        exact_cert = {"qps": nq * ne / dt, "ms_per_step": 1e3 * dt / ne, "steps": ne,
                      "fallbacks": ds.last_fallbacks()}
        log(f"certified exact (exact_bf16): {ne} steps in {dt:.3f}s: {nq * ne / dt:.0f} QPS, "
            f"{exact_cert['fallbacks']} queries left to the FP32 scan")
        ds.search(qd, metas, **kw)  # the main line's results back in out for the recall check
        torch.cuda.synchronize()
        ids = out[0].cpu().numpy().view(np.uint32)
        dists = out[1].cpu().numpy()

    # Recall@k of the timed mode over the WHOLE batch against the exact
    # answer (the certified exact mode: identical to exact FP32), outside the
    # timed region.
    orc = ge.load_oracle()
    full_recall = None
    if mode not in ("exact", "exact_bf16"):
        if exact_ids is None:
            eo = (torch.empty_like(out[0]), torch.empty_like(out[1]))
            ds.search(qd, metas, k=a.k, mode="exact_bf16", out=eo)
            torch.cuda.synchronize()
            exact_ids = eo[0].cpu().numpy().view(np.uint32)
            del eo
        full_recall, per_q = orc.recall(ids, exact_ids)
        by_type = {str(t): float(orc.recall(ids[metas[:, 0] == t], exact_ids[metas[:, 0] == t])[0])
                   for t in range(4)}
        worst = int(np.argmin(per_q))
        full_recall = {"recall": float(full_recall), "queries": int(nq), "by_type": by_type,
                       "worst_query": worst, "worst_hits": int(per_q[worst]), "reference": "exact_bf16"}
        log(f"recall@{a.k} vs exact over all {nq} queries: {full_recall['recall']:.5f}; by type {by_type}")

    # Recall@k and distance errors against FP64 exact on a type-stratified
    # sample (CPU oracle over the sorted ranges).
    t0 = time.time()
    lay = orc.Layouts(labels.astype(np.int32), ts)
    vb = np.ascontiguousarray(vecs[lay.perm_both])
    vt = np.ascontiguousarray(vecs[lay.perm_time])
    rng = np.random.default_rng(0)
    sample = []
    for t in range(4):
        idx = np.nonzero(metas[:, 0] == t)[0]
        if idx.size:
            sample += rng.choice(idx, min(idx.size, max(1, a.recall_sample // 4)), replace=False).tolist()
    sample = np.array(sorted(sample))
    gi, gd = orc.exact_ranges_f64_prepared(vb, vt, lay, q[sample], metas[sample], a.k)
    rec, per_s = orc.recall(ids[sample], gi)
    fp64 = {"queries": int(sample.size), "recall": float(rec), "by_type": {}}
    kth_err = 0.0
    for t in range(4):
        sel = metas[sample, 0] == t
        if not sel.any():
            continue
        # k-th distance relative error: the k-th returned distance against
        # the FP64 k-th distance (queries with at least k candidates)
        dk, gk = dists[sample][sel, a.k - 1].astype(np.float64), gd[sel, a.k - 1]
        ok = np.isfinite(gk) & (gk > 0)
        e = float(np.max(np.abs(dk[ok] - gk[ok]) / gk[ok])) if ok.any() else 0.0
        r_t = float(orc.recall(ids[sample][sel], gi[sel])[0])
        wi = int(np.argmin(per_s[sel]))
        fp64["by_type"][str(t)] = {"recall": r_t, "kth_rel_err": e, "queries": int(sel.sum()),
                                   "worst_query": int(sample[sel][wi]), "worst_hits": int(per_s[sel][wi])}
        kth_err = max(kth_err, e)
    fp64["kth_rel_err"] = kth_err
    log(f"recall@{a.k} vs FP64 on {sample.size} stratified queries: {rec:.5f}, max k-th distance rel err "
        f"{kth_err:.2e}; {fp64['by_type']} ({time.time() - t0:.1f}s)")

    # Roofline of the dominant kernel, from its launches measured on the
    # search's own stream (HIP events).  MFMA modes: the main scan pass (one
    # launch per step; 2d = 200 flop per (query, candidate) pair of that
    # pass), with the figure over all scan launches beside it.  Exact FP32:
    # the VALU scan's launches (one per sorted layout), 3d = 300 flop/pair.
    pairs = stats["pairs"]
    if mode in ("exact",):
        flops = 3.0 * 100 * pairs / max(1, n_scan // a.steps)
        avg_launch_ms = ms_scan / max(1, n_scan)
        achieved = flops / (avg_launch_ms * 1e-3) / 1e12
        roof = {"bound": "valu", "achieved": achieved, "peak": PEAK_TFLOPS["fp32"], "unit": "TFLOP/s",
                "frac": achieved / PEAK_TFLOPS["fp32"], "traffic": committed_traffic(cfg.name, mode, a, kw),
                "kernel": "scan_f32_kernel", "launch_ms": avg_launch_ms}
    else:
        pk = PEAK_TFLOPS["bf16"] if mode in ("bf16", "sq8", "exact_bf16") else PEAK_TFLOPS["fp8"] if mode == "fp8" else \
            PEAK_TFLOPS["int8"]
        unit = "TOP/s" if mode == "sq8sym" else "TFLOP/s"
        main_ms = ms_main / max(1, n_main)
        achieved = 2.0 * 100 * pairs_main / (main_ms * 1e-3) / 1e12
        all_scans = 2.0 * 100 * pairs * a.steps / (ms_scan * 1e-3) / 1e12
        kernel = "scan_mfma_kernel (main pass)"
        if mode == "exact_bf16":
            # Its main-pass events also hold the FP32 fallback's scans: the
            # figure is over all scan launches of a step, timed per step.
            achieved, main_ms = all_scans, ms_scan / a.steps
            kernel = "scan_mfma_kernel + FP32 fallback scan_f32_kernel (all scan launches of a step)"
        roof = {"bound": "mfma", "achieved": achieved, "peak": pk, "unit": unit, "frac": achieved / pk,
                "traffic": committed_traffic(cfg.name, mode, a, kw), "kernel": kernel,
                "launch_ms": main_ms, "pairs_per_launch": pairs_main,
                "all_scan_launches": {"achieved": all_scans, "frac": all_scans / pk,
                                      "ms_per_step": ms_scan / a.steps}}

    cpu = None
    cpu_router = None
    cpu_ref = None
    if not a.no_cpu:
        cpu = cpu_baseline(orc, vb, vt, lay, q, metas, a.k, a.cpu_seconds)
        cpu_router = cpu_router_bf(orc, vecs, lay, q, metas, a.k, a.cpu_seconds / 2)
        cpu_ref = cpu_reference_router(orc, vecs, lay, q, metas, a.k, a.cpu_seconds / 2)
    del vb, vt

    line = {
        "metric": "filtered top-100 QPS at recall@100>=0.99 (BF16/SQ8 vs FP32), 1 MI355X",
        "value": qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"exact": "f32", "sq8": "u8", "bf16": "bf16", "fp8": "fp8e4m3", "sq8sym": "i8",
                  "exact_bf16": "bf16+f32"}[mode],
        "data": "synthetic (seeded Gaussian mixture, Zipf categories, uniform timestamps)",
        "config": {"workload": cfg.name, "n": n, "d": 100, "nq": nq, "k": a.k, "mode": mode,
                   "kprime": kw["kprime"], "group_queries": kw["group_queries"], "chunk_rows": kw["chunk_rows"],
                   "autotuned": tuned, "per_batch_searches": per_batch, "graph_replay_qps": graph_qps,
                   "any_batch_graph_replay_qps": any_qps,
                   "pairs_per_step": pairs, "work_items": stats["work_items"],
                   "certificate_fallbacks": fallbacks if mode == "exact_bf16" else None,
                   "certified_exact": exact_cert},
        "recall_at_k_vs_fp64": rec,
        "recall_sample": int(sample.size),
        "kth_rel_err_vs_fp64": kth_err,
        "fp64_check": fp64,
        "recall_full_batch_vs_exact": full_recall,
        "roofline": roof,
        "cpu_baseline": cpu,
        "cpu_baseline_router_bf": cpu_router,
        "cpu_baseline_reference_router_bf": cpu_ref,
    }
    print(json.dumps(line), flush=True)
    ds.close()
    return 0


def subinterval_bench(a, ge, fbf, cfg, vecs, labels, ts, q, metas) -> int:
    """The router's type-2 path (hybrid_graph.cpp:317-494) for the type-2
    queries it sends to the graph side (selectivity >= bf_thr = 0.045):
    SplitInterval over the ten time buckets, the SMALL buckets' int8 scans with
    bucket-encoded queries, the candidate pool and the global FP32 refine
    (fbf_search_subinterval; the FULL and MEDIUM buckets' graph searches are
    the caller's).  Times the whole set on the GPU (1 warmup, then a.steps
    timed searches, queries resident in HBM) and, as the CPU baseline, the
    oracle's restatement (orc_subinterval, all host threads) on a sample whose
    ids it compares."""
    import torch
    ds = fbf.Dataset(vecs, labels, ts, repr_flags=fbf.REPR_F32 | fbf.REPR_TBUCKETS)
    md = torch.from_numpy(np.ascontiguousarray(metas)).cuda()
    ranges, _ = ds.route(md, order=False)
    torch.cuda.synchronize()
    r = ranges.cpu().numpy().view(np.uint32)
    idx = np.nonzero((metas[:, 0] == 2) & (r[:, 3] == 0))[0]
    m = np.ascontiguousarray(metas[idx])
    qd = torch.from_numpy(np.ascontiguousarray(q[idx])).cuda()
    log(f"sub-interval: {idx.size} graph-routed type-2 queries")
    ids, _, types = ds.search_subinterval(qd, m, k=a.k)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        ids, _, types = ds.search_subinterval(qd, m, k=a.k)
    torch.cuda.synchronize()
    gpu_qps = idx.size * a.steps / (time.perf_counter() - t0)
    ids = ids.cpu().numpy().view(np.uint32)
    orc = ge.load_oracle()
    lay = orc.Layouts(labels.astype(np.int32), ts)
    tb = orc.TimeBuckets(vecs, lay)
    s = np.random.default_rng(3).choice(idx.size, min(512, idx.size), replace=False)
    t0 = time.perf_counter()
    oid, _, _ = orc.subinterval(vecs, lay, tb, q[idx[s]], m[s], a.k)
    cpu_qps = s.size / (time.perf_counter() - t0)
    same = int((ids[s] == oid).all(1).sum())
    line = {"subinterval": cfg.name, "queries": int(idx.size), "steps": a.steps, "gpu_qps": gpu_qps,
            "small_slices_per_query": float((types == 0).sum(1).mean()), "cpu_port_qps": cpu_qps,
            "cpu_sample": int(s.size), "ids_equal_on_sample": f"{same}/{s.size}", **host_cpu()}
    log(f"sub-interval: GPU {gpu_qps:.0f} QPS, CPU port {cpu_qps:.1f} QPS on {s.size} queries, ids equal {same}/{s.size}")
    print(json.dumps(line), flush=True)
    ds.close()
    return 0


def dropin(a, ge, fbf, cfg, vecs, labels, ts, q, metas) -> int:
    """The router's per-query brute-force calls (hybrid_graph.cpp:200-211):
    queries with selectivity below bf_thr = 0.045 (type 3: bf_thr3 = 0.08)
    call bruteforce(dim, min(140, e - s), ..., codes + s * 112, ...) on their
    range of the symmetric int8 codes.  Times, on the same slices and query
    codes: the GPU drop-in fbf_bruteforce called from 16 host threads, the
    batched fbf_bruteforce_batch (one call for all of them), and the CPU
    bruteforce() -- the reference's own kernel (oracle/_ref) when this host
    can run it, else the oracle's heap restatement -- from 16 host threads."""
    from concurrent.futures import ThreadPoolExecutor
    orc = ge.load_oracle()
    lay = orc.Layouts(labels.astype(np.int32), ts)
    alpha = orc.sq8sym_train(vecs)
    codes = orc.sq8sym_encode(vecs, alpha)
    cb = np.ascontiguousarray(codes[lay.perm_both]); ct = np.ascontiguousarray(codes[lay.perm_time])
    del codes
    rr = lay.ranges(metas)
    cnt = rr[:, 2] - rr[:, 1]
    sel = cnt / float(vecs.shape[0])
    routed = np.nonzero((sel < 0.045) | ((metas[:, 0] == 3) & (sel < 0.08)))[0]
    routed = routed[cnt[routed] > 0][: a.nq or 20000]
    qc = orc.sq8sym_encode(q[routed], alpha)
    log(f"drop-in: {routed.size} routed queries, slices of {int(cnt[routed].min())}-{int(cnt[routed].max())} rows "
        f"(mean {cnt[routed].mean():.0f})")
    fbf.register_codes(cb); fbf.register_codes(ct)
    R = orc.ref()

    def gpu_one(j):
        i = routed[j]
        arr = cb if rr[i, 0] == 0 else ct
        return fbf.bruteforce(int(min(140, cnt[i])), int(cnt[i]), arr, alpha, qc[j], offset_rows=int(rr[i, 1]))

    def cpu_one(j):
        i = routed[j]
        arr = cb if rr[i, 0] == 0 else ct
        s, e = int(rr[i, 1]), int(rr[i, 2])
        K = min(140, e - s)
        out = np.zeros(K, np.int32)
        if R is not None:
            R.ref_bruteforce(K, orc._p(out), e - s, orc._p(arr[s:e]), alpha, orc._p(qc[j]))
        else:
            orc.bf_heap_ref(arr[s:e], qc[j], K)
        return out

    res = {}
    with ThreadPoolExecutor(16) as ex:
        list(ex.map(gpu_one, range(min(64, routed.size))))
        t = time.perf_counter()
        got = list(ex.map(gpu_one, range(routed.size)))
        res["gpu_per_call_16_threads"] = routed.size / (time.perf_counter() - t)
        t = time.perf_counter()
        ncpu = min(routed.size, 4000)
        ref_out = list(ex.map(cpu_one, range(ncpu)))
        res["cpu_per_call_16_threads"] = ncpu / (time.perf_counter() - t)
    # the batched form over both layouts (two calls, one per code array)
    t = time.perf_counter()
    for L, arr in ((0, cb), (1, ct)):
        m = rr[routed, 0] == L
        if m.any():
            fbf.bruteforce_batch(arr, rr[routed[m], 1].astype(np.uint64), cnt[routed[m]].astype(np.uint64), qc[m], 140)
    res["gpu_batch"] = routed.size / (time.perf_counter() - t)
    same = sum(set(got[j].tolist()) == set(ref_out[j].tolist()) for j in range(ncpu))
    fbf.unregister_codes(cb); fbf.unregister_codes(ct)
    # The same calls from a C++ program (16 OpenMP threads, no Python in the
    # loop): the drop-in header's bruteforce() on the GPU, then the
    # reference's own bruteforce() from oracle/_ref on this host's CPU.
    cpp = dropin_cpp(ge, cb, ct, rr[routed], qc, alpha, R is not None)
    if cpp:
        res.update(cpp)
    line = {"dropin": cfg.name, "routed_queries": int(routed.size), "calls_per_s": res,
            "cpu_kind": "reference" if R is not None else "port", "id_sets_equal": f"{same}/{ncpu}",
            "mean_slice_rows": float(cnt[routed].mean()), **host_cpu()}
    print(json.dumps(line), flush=True)
    return 0


DROPIN_CPP = r"""
#define FBF_REPLACE_REFERENCE_BRUTEFORCE
#include "fbf_bruteforce.hpp"
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <dlfcn.h>
#include <omp.h>
#include <vector>

// argv: cb.bin ct.bin q.bin slices.bin (n x 3 u64: layout, s, e) alpha ref_so|- threads
static bool slurp(const char *p, std::vector<char> &v) {
    FILE *f = std::fopen(p, "rb"); if (!f) return false;
    std::fseek(f, 0, SEEK_END); long n = std::ftell(f); std::fseek(f, 0, SEEK_SET);
    v.resize((size_t)n); bool ok = std::fread(v.data(), 1, (size_t)n, f) == (size_t)n; std::fclose(f); return ok;
}
typedef void (*ref_bf_t)(uint64_t, int32_t *, uint64_t, const int8_t *, float, const int8_t *);
int main(int argc, char **argv) {
    if (argc != 8) return 2;
    std::vector<char> cb, ct, q, sl;
    if (!slurp(argv[1], cb) || !slurp(argv[2], ct) || !slurp(argv[3], q) || !slurp(argv[4], sl)) return 3;
    const float alpha = std::strtof(argv[5], nullptr);
    const int threads = std::atoi(argv[7]);
    const uint64_t *se = reinterpret_cast<const uint64_t *>(sl.data());
    const size_t nq = q.size() / 112;
    fbf_ref::check_or_abort(fbf_register_codes(cb.data(), cb.size()), "fbf_register_codes");
    fbf_ref::check_or_abort(fbf_register_codes(ct.data(), ct.size()), "fbf_register_codes");
    std::vector<int> gpu(nq * 140, -1), cpu(nq * 140, -1);
    omp_set_num_threads(threads);
    auto run = [&](std::vector<int> &out, ref_bf_t fn, size_t count) {
#pragma omp parallel for schedule(dynamic)
        for (size_t i = 0; i < count; ++i) {
            char *arr = se[3 * i] == 0 ? cb.data() : ct.data();
            const size_t s = se[3 * i + 1], e = se[3 * i + 2], K = std::min<size_t>(140, e - s);
            int8_t *qi = reinterpret_cast<int8_t *>(q.data() + i * 112);
            if (fn) fn(K, out.data() + i * 140, e - s, reinterpret_cast<const int8_t *>(arr + s * 112), alpha, qi);
            else bruteforce(100, K, out.data() + i * 140, e - s, arr + s * 112, alpha, 1, qi);
        }
    };
    run(gpu, nullptr, std::min<size_t>(nq, 64));  // warm-up: streams, engines
    auto t0 = std::chrono::steady_clock::now();
    run(gpu, nullptr, nq);
    double tg = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    double tc = -1.0;
    size_t same = 0;
    if (argv[6][0] != '-') {
        void *h = dlopen(argv[6], RTLD_NOW);
        ref_bf_t fn = h ? (ref_bf_t)dlsym(h, "ref_bruteforce") : nullptr;
        if (fn) {
            t0 = std::chrono::steady_clock::now();
            run(cpu, fn, nq);
            tc = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            for (size_t i = 0; i < nq; ++i) {
                std::vector<int> a(gpu.begin() + i * 140, gpu.begin() + i * 140 + 140);
                std::vector<int> b(cpu.begin() + i * 140, cpu.begin() + i * 140 + 140);
                std::sort(a.begin(), a.end()); std::sort(b.begin(), b.end());
                same += a == b;
            }
        }
    }
    std::printf("{\"cpp_gpu_per_call\": %.1f, \"cpp_cpu_per_call\": %.1f, \"cpp_id_sets_equal\": %zu, \"cpp_calls\": %zu, \"cpp_threads\": %d}\n",
                nq / tg, tc > 0 ? nq / tc : -1.0, same, nq, threads);
    return 0;
}
"""


def dropin_cpp(ge, cb, ct, ranges, qc, alpha, have_ref):
    """Build and run DROPIN_CPP (g++, the drop-in header, libfbf.so); None
    when no C++ compiler is available."""
    import shutil
    import subprocess
    import tempfile
    if shutil.which("g++") is None:
        return None
    lib_dir = os.path.join(ROOT, "sigmod-2024-contest_amd", "lib")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "dropin.cpp")
        with open(src, "w") as f:
            f.write(DROPIN_CPP)
        exe = os.path.join(d, "dropin")
        b = subprocess.run(["g++", "-std=c++17", "-O2", "-fopenmp", f"-I{os.path.join(ROOT, 'include')}", src,
                            f"-L{lib_dir}", "-lfbf", f"-Wl,-rpath,{lib_dir}", "-ldl", "-o", exe],
                           capture_output=True, text=True)
        if b.returncode != 0:
            log(f"drop-in C++ program did not build: {b.stderr[-400:]}")
            return None
        files = []
        for name, arr in (("cb", cb), ("ct", ct), ("q", qc), ("se", np.ascontiguousarray(ranges[:, :3], np.uint64))):
            path = os.path.join(d, name + ".bin")
            arr.tofile(path)
            files.append(path)
        ref_so = ge.load_oracle().REF_SO if have_ref else "-"
        r = subprocess.run([exe, *files, repr(float(alpha)), ref_so, "16"], capture_output=True, text=True,
                           timeout=600)
        if r.returncode != 0:
            log(f"drop-in C++ program failed: rc={r.returncode} {r.stderr[-400:]}")
            return None
        out = json.loads(r.stdout.strip().splitlines()[-1])
        log(f"drop-in from C++ (16 OpenMP threads): {out}")
        return out


def committed_traffic(workload, mode, a, kw):
    """HBM bytes per scan launch measured with rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, calibrated on a plain copy kernel) for this workload, mode and
    search options, from the committed profile summary; None otherwise.  The
    summary holds a `search_options:` / `traffic_bytes_per_launch:` pair for
    its main run and may list more option sets as `traffic_for: {...} = B`."""
    if a.k != 100 or a.prepass_div or a.prepass_margin or a.n or a.nq:
        return None
    name = f"scan_f32_exact_{workload}.md" if mode == "exact" else f"scan_mfma_{mode}_{workload}.md"
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    opts = {key: int(kw[key]) for key in ("group_queries", "chunk_rows", "kprime")}
    traffic, profiled = None, None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                traffic = float(line.split(":", 1)[1])
            if line.startswith("search_options:"):
                profiled = json.loads(line.split(":", 1)[1])
            if line.startswith("traffic_for:"):
                spec, val = line.split(":", 1)[1].rsplit("=", 1)
                if json.loads(spec) == opts:
                    return float(val)
    return traffic if profiled == opts else None


def host_cpu() -> dict:
    """The host's logical CPU count, the CPUs this process may use, and the
    CPU model (lscpu's "Model name", from /proc/cpuinfo)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"nproc": usable, "logical_cpus": os.cpu_count(), "cpu_model": model}


def cpu_baseline(orc, vb, vt, L, q, metas, k, target_s):
    """Exact canonical-FP32 top-k over the sorted contiguous ranges (the CPU
    port of the path), all host threads, on a type-stratified sample sized to
    about target_s seconds; QPS extrapolates over the sample."""
    rng = np.random.default_rng(1)
    order = np.concatenate([rng.permutation(np.nonzero(metas[:, 0] == t)[0]) for t in range(4)])
    # interleave types so every prefix is stratified
    per = [rng.permutation(np.nonzero(metas[:, 0] == t)[0]) for t in range(4)]
    m = min(len(p) for p in per)
    order = np.stack([p[:m] for p in per], 1).reshape(-1)
    done, spent, size = 0, 0.0, 8
    while done < order.size and spent < target_s:
        idx = order[done:done + size]
        t = time.perf_counter()
        orc.exact_ranges_prepared(vb, vt, L, q[idx], metas[idx], k)
        spent += time.perf_counter() - t
        done += idx.size
        size = min(size * 2, 4096)
    threads = orc.num_threads()
    log(f"cpu baseline: {done} queries in {spent:.1f}s on {threads} threads ({host_cpu()})")
    return dict({"value": done / spent, "unit": "queries/s", "cores": threads, "kind": "port",
                 "sample": f"{done} type-stratified queries of the same workload, exact canonical FP32 over "
                           f"sorted ranges (oracle/fbf_oracle.c orc_exact_ranges, OpenMP)"}, **host_cpu())


def cpu_router_bf(orc, vecs, L, q, metas, k, target_s):
    """The reference's own brute-force route on the CPU (hybrid_graph.cpp:
    203-228 restated in the oracle, orc_router_bf): symmetric int8 codes,
    top-140 scan of the filtered range, exact FP32 rerank to k, all host
    threads, on a type-stratified sample sized to about target_s seconds."""
    t0 = time.time()
    alpha = orc.sq8sym_train(vecs)
    codes = orc.sq8sym_encode(vecs, alpha)
    cb = np.ascontiguousarray(codes[L.perm_both])
    ct = np.ascontiguousarray(codes[L.perm_time])
    del codes
    log(f"cpu router path: codes and layouts in {time.time() - t0:.1f}s")
    rng = np.random.default_rng(2)
    per = [rng.permutation(np.nonzero(metas[:, 0] == t)[0]) for t in range(4)]
    m = min(len(p) for p in per)
    order = np.stack([p[:m] for p in per], 1).reshape(-1)
    done, spent, size = 0, 0.0, 4
    n = vecs.shape[0]
    while done < order.size and spent < target_s:
        idx = np.sort(order[done:done + size])
        qq = np.ascontiguousarray(q[idx])
        mm = np.ascontiguousarray(metas[idx])
        ids = np.zeros((idx.size, k), np.uint32)
        dd = np.zeros((idx.size, k), np.float32)
        t = time.perf_counter()
        orc.lib().orc_router_bf(orc._p(cb), orc._p(ct), orc._p(L.perm_both), orc._p(L.lab_b), orc._p(L.ts_b),
                                orc._p(L.perm_time), orc._p(L.ts_t), orc._p(vecs), n, 100, alpha,
                                orc._p(qq), orc._p(mm), idx.size, 140, k, orc._p(ids), orc._p(dd))
        spent += time.perf_counter() - t
        done += idx.size
        size = min(size * 2, 4096)
    threads = orc.num_threads()
    log(f"cpu router path (SQ8-sym top-140 + FP32 rerank): {done} queries in {spent:.1f}s on {threads} threads")
    return dict({"value": done / spent, "unit": "queries/s", "cores": threads, "kind": "port",
                 "sample": f"{done} type-stratified queries; the reference's brute-force route (symmetric int8 "
                           f"top-140 scan + exact FP32 rerank, oracle/fbf_oracle.c orc_router_bf, OpenMP)"},
                **host_cpu())


def cpu_reference_router(orc, vecs, L, q, metas, k, target_s):
    """The reference's own kernels (bruteforce(), L2SqrSQ8_sym, L2SqrRef and
    maxPQIFCS, compiled from its headers into oracle/_ref) driven as its
    router drives them (hybrid_graph.cpp:168-228, one OpenMP task per query),
    on a type-stratified sample; None when that library is absent or the host
    lacks the AVX-512 subsets it was built for."""
    R = orc.ref()
    if R is None:
        log("reference kernels not available on this host: no reference-kind CPU baseline")
        return None
    alpha = orc.sq8sym_train(vecs)
    codes = orc.sq8sym_encode(vecs, alpha)
    cb = np.ascontiguousarray(codes[L.perm_both])
    ct = np.ascontiguousarray(codes[L.perm_time])
    del codes
    rng = np.random.default_rng(3)
    per = [rng.permutation(np.nonzero(metas[:, 0] == t)[0]) for t in range(4)]
    m = min(len(p) for p in per)
    order = np.stack([p[:m] for p in per], 1).reshape(-1)
    done, spent, size = 0, 0.0, 4
    while done < order.size and spent < target_s:
        idx = np.sort(order[done:done + size])
        qq = np.ascontiguousarray(q[idx])
        rr = np.ascontiguousarray(L.ranges(metas[idx]).astype(np.int64))
        qc = orc.sq8sym_encode(qq, alpha)
        ids = np.zeros((idx.size, k), np.uint32)
        t = time.perf_counter()
        R.ref_router_bf(orc._p(cb), orc._p(ct), orc._p(L.perm_both), orc._p(L.perm_time), orc._p(vecs), 100,
                        alpha, orc._p(qq), orc._p(qc), orc._p(rr), idx.size, 140, k, orc._p(ids))
        spent += time.perf_counter() - t
        done += idx.size
        size = min(size * 2, 4096)
    threads = orc.num_threads()
    log(f"reference kernels, router route: {done} queries in {spent:.1f}s on {threads} threads")
    return dict({"value": done / spent, "unit": "queries/s", "cores": threads, "kind": "reference",
                 "sample": f"{done} type-stratified queries; the reference's own bruteforce() top-140 + "
                           f"L2SqrRef rerank (oracle/_ref, built from its sources), OpenMP over queries"},
                **host_cpu())


if __name__ == "__main__":
    sys.exit(main())
