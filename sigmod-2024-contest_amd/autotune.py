"""Autotuner for the approximate search modes.

Picks, for one dataset, one query batch and one mode, the search options --
queries per work item (group_queries: the query-tile width of the MFMA
scans), rows per work item (chunk_rows: the candidate-tile depth) and the
rerank depth k' -- that give the highest throughput while recall@k against the
exact FP32 search stays at or above a target.

Recall is measured on a type-stratified sample of the batch against the
library's own exact FP32 mode (bit-exact with the canonical CPU search), so
the tuner needs a dataset built with FBF_REPR_F32 besides the mode's own
representation.  Each candidate is timed on the whole batch (one warm-up
search, then the mean of `repeats` searches), since the item plan and the
threshold seeding depend on the whole batch.  Candidates are tried in order
of increasing k' within each tile shape, and a tile shape stops at the first
k' that meets the target: larger k' only costs time.  Among the candidates
that meet the target, the first one tried (smaller tile width, then fewer
rows per item) whose time is within `tie` of the fastest wins, so that timing
noise does not flip the choice between near-equal shapes.
"""
from __future__ import annotations

import itertools
import time
from dataclasses import dataclass, field

import numpy as np

from . import fbf


@dataclass
class TuneResult:
    opts: dict                       # best options: group_queries, chunk_rows, kprime
    ms: float                        # mean search time of the best options
    recall: float                    # its recall on the sample
    table: list = field(default_factory=list)  # (opts, ms, recall) of every candidate tried


def stratified_sample(metas: np.ndarray, n: int, seed: int = 0) -> np.ndarray:
    """Up to n query indices, the same share from every filter type."""
    rng = np.random.default_rng(seed)
    out = []
    types = [t for t in range(4) if np.any(metas[:, 0] == t)]
    for t in types:
        idx = np.nonzero(metas[:, 0] == t)[0]
        out.append(rng.choice(idx, min(idx.size, max(1, n // len(types))), replace=False))
    return np.sort(np.concatenate(out)) if out else np.zeros(0, np.int64)


def recall_at_k(ids: np.ndarray, ref: np.ndarray) -> float:
    """Mean |ids ∩ ref| / |ref| over queries (padding ids excluded), the
    reference's GetKNNRecall (utils.h:224-253) restricted to valid entries."""
    hit = tot = 0
    for a, b in zip(ids, ref):
        b = b[b != fbf.PAD_ID]
        if b.size == 0:
            continue
        hit += np.intersect1d(a, b).size
        tot += b.size
    return hit / tot if tot else 1.0


def tune(ds, queries, metas, k: int = 100, mode: str = "bf16", recall_target: float = 0.995,
         sample: int = 256, groups=(128, 256), chunks=(32768, 65536, 131072),
         kprime_factors=(1.0, 1.1, 1.2, 1.5, 2.0, 4.0, 6.0, 9.0), repeats: int = 1, tie: float = 0.03,
         log=None) -> TuneResult:
    """queries: torch device float32 [nq, d]; metas: numpy float32 [nq, 4]."""
    import torch
    metas = np.ascontiguousarray(metas, np.float32)
    idx = stratified_sample(metas, sample)
    ref_ids, _ = ds.search(queries[torch.from_numpy(idx).to(queries.device)], metas[idx], k=k, mode="exact")
    torch.cuda.synchronize()
    ref = ref_ids.cpu().numpy().view(np.uint32)

    def run(opts):
        kw = dict(k=k, mode=mode, **opts)
        ds.search(queries, metas, **kw)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(repeats):
            ids, _ = ds.search(queries, metas, **kw)
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t) / repeats
        got = ids.cpu().numpy().view(np.uint32)[idx]
        return ms, recall_at_k(got, ref)

    table = []
    ok = []
    for g, c in itertools.product(groups, chunks):
        for f in kprime_factors:
            kp = min(896, max(k, int(round(f * k))))
            opts = {"group_queries": g, "chunk_rows": c, "kprime": kp}
            ms, rec = run(opts)
            table.append((opts, ms, rec))
            if log:
                log(f"autotune {mode}: {opts} {ms:.1f} ms recall {rec:.4f}")
            if rec >= recall_target:
                ok.append((opts, ms, rec))
                break
    if ok:
        fastest = min(r[1] for r in ok)
        best = next(r for r in ok if r[1] <= fastest * (1.0 + tie))
    else:  # nothing met the target: the most accurate candidate
        best = max(table, key=lambda r: (r[2], -r[1]))
    return TuneResult(opts=best[0], ms=best[1], recall=best[2], table=table)
