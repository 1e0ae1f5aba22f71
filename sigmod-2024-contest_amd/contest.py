"""Contest-format driver: read the contest's base and query files, answer every
query with the GPU library, write the contest's k-NN output file, and
optionally score it against a ground-truth file.

The file formats are the reference harness's (base: uint32 n, then per row
label, timestamp and d floats, utils.h:257-276; queries: uint32 nq, then per
query type, v, l, r and d floats; output and ground truth: nq x k uint32 ids,
utils.h:550-566), and the score is GetKNNRecall's (utils.h:224-253).  The
reference's main (baseline.cpp) reads the same files, builds its index, runs
the router and saves the output; here every query goes through fbf_search
(exact FP32 by default: the brute-force answer, which is also what the
ground-truth generator writes).

    python3 contest.py --data dummy-data.bin --queries dummy-queries.bin --out output.bin \\
        [--mode exact|exact_bf16|bf16|sq8|fp8|sq8sym] [--k 100] [--gt gt.bin]
"""
from __future__ import annotations

import argparse
import importlib.util
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _pkg():
    """This package's fbf and datagen modules (the directory name is not a
    Python identifier, so it is registered as fbf_amd)."""
    if "fbf_amd" not in sys.modules:
        spec = importlib.util.spec_from_file_location("fbf_amd", os.path.join(HERE, "__init__.py"),
                                                      submodule_search_locations=[HERE])
        mod = importlib.util.module_from_spec(spec)
        sys.modules["fbf_amd"] = mod
        spec.loader.exec_module(mod)
    return sys.modules["fbf_amd"]


def run(data: str, queries: str, out: str, mode: str = "exact", k: int = 100, gt: str | None = None,
        log=print) -> dict:
    pkg = _pkg()
    fbf, dg = pkg.fbf, pkg.datagen
    t0 = time.time()
    vecs, labels, ts = dg.read_base_bin(data)
    q, metas = dg.read_query_bin(queries)
    log(f"read {vecs.shape[0]} rows and {q.shape[0]} queries in {time.time() - t0:.2f}s")
    flags = fbf.MODE_REPR[fbf.MODES[mode]] | (fbf.REPR_F32 if mode != "exact" else 0)
    t0 = time.time()
    ds = fbf.Dataset(vecs, labels, ts, repr_flags=flags)
    try:
        log(f"dataset resident in {time.time() - t0:.2f}s")
        t0 = time.time()
        ids, dists = ds.search_host(q, metas, k=k, mode=mode)
        dt = time.time() - t0
    finally:
        ds.close()
    dg.write_knn_bin(out, ids)
    res = {"queries": int(q.shape[0]), "seconds": dt, "qps": q.shape[0] / dt if dt > 0 else float("inf")}
    log(f"searched in {dt:.3f}s ({res['qps']:.0f} queries/s); wrote {out}")
    if gt:
        truth = dg.read_gt_bin(gt, k=k)
        hit = sum(np.intersect1d(a, b).size for a, b in zip(ids, truth))
        res["recall"] = hit / float(truth.size)
        log(f"recall@{k} vs {gt}: {res['recall']:.4f}")
    return res


def main(argv=None) -> int:
    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    p.add_argument("--data", required=True)
    p.add_argument("--queries", required=True)
    p.add_argument("--out", required=True)
    p.add_argument("--mode", default="exact", choices=["exact", "exact_bf16", "bf16", "sq8", "fp8", "sq8sym"])
    p.add_argument("--k", type=int, default=100)
    p.add_argument("--gt", default=None)
    a = p.parse_args(argv)
    run(a.data, a.queries, a.out, a.mode, a.k, a.gt)
    return 0


if __name__ == "__main__":
    sys.exit(main())
