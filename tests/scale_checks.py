"""Checks shared by the full-scale GPU tests (N = 10 M): the invariants every
search mode must hold on every query of a batch, and a type-stratified
sample for the oracle's exact answers.  Test infrastructure (not collected:
no test functions here)."""
from __future__ import annotations

import numpy as np

PAD = 0xFFFFFFFF


def stratified_sample(metas: np.ndarray, counts: np.ndarray, per_type: int, seed: int = 0,
                      smallest: int = 32, largest: int = 8) -> np.ndarray:
    """per_type random queries of each filter type, plus the queries with the
    fewest candidates (fewer than k, none at all) and the largest filtered
    ranges of types 1 and 3."""
    rng = np.random.default_rng(seed)
    pick = []
    for t in range(4):
        idx = np.nonzero(metas[:, 0] == t)[0]
        if idx.size:
            pick += rng.choice(idx, min(per_type, idx.size), replace=False).tolist()
    order = np.argsort(counts, kind="stable")
    pick += order[:smallest].tolist()
    for t in (1, 3):
        idx = np.nonzero(metas[:, 0] == t)[0]
        pick += idx[np.argsort(counts[idx])[-largest:]].tolist()
    return np.array(sorted(set(pick)), np.int64)


def check_invariants(orc, vecs, lab, ts, q, metas, counts, ids, d, k, kprime=None, exact_ids=None,
                     exact_d=None):
    """Every returned (id, distance) passes the query's filter and carries the
    row's canonical FP32 distance; rows are ascending by (distance, id); a
    query is padded exactly when fewer than k rows pass its filter; a query
    admitting at most k' rows equals the exact answer.  Returns the number of
    queries the last rule covered."""
    nq = ids.shape[0]
    valid = ids != PAD
    nvalid = valid.sum(1)
    want = np.minimum(counts, k)
    bad = np.nonzero(nvalid != want)[0]
    assert bad.size == 0, f"padding: queries {bad[:8]} have {nvalid[bad[:8]]} ids, want {want[bad[:8]]}"
    # the valid ids form a prefix
    assert np.all(valid[:, :-1] | ~valid[:, 1:]), "pad ids before valid ones"
    # filters
    safe = np.where(valid, ids, 0).astype(np.int64)
    rl = lab[safe]
    rt = ts[safe]
    t = metas[:, 0:1].astype(np.int32)
    v = metas[:, 1:2].astype(np.int32)
    lo = metas[:, 2:3]
    hi = metas[:, 3:4]
    ok_lab = (t == 0) | (t == 2) | (rl == v)
    ok_ts = (t == 0) | (t == 1) | ((rt >= lo) & (rt <= hi))
    badf = valid & ~(ok_lab & ok_ts)
    assert not badf.any(), f"rows outside the filter at queries {np.nonzero(badf.any(1))[0][:8]}"
    # distances: the canonical FP32 distance of each returned row
    canon = orc.l2_canon_rows(vecs, q, ids)
    badd = valid & (canon != d)
    assert not badd.any(), f"distance not canonical at queries {np.nonzero(badd.any(1))[0][:8]}"
    assert np.all(np.isinf(d[~valid]))
    # order: strictly ascending (distance, id) keys over the valid prefix
    keys = orc.make_keys(d.reshape(-1), ids.reshape(-1)).reshape(nq, -1)
    both = valid[:, 1:] & valid[:, :-1]
    assert np.all((keys[:, 1:] > keys[:, :-1]) | ~both), "rows not ascending by (distance, id)"
    covered = 0
    if kprime is not None and exact_ids is not None:
        small = counts <= kprime
        covered = int(small.sum())
        if covered:
            assert np.array_equal(ids[small], exact_ids[small]), "a query with <= k' rows differs from exact"
            assert np.array_equal(d[small], exact_d[small])
    return covered


def recall_vs(orc, ids, ref_ids) -> float:
    r, _ = orc.recall(ids, ref_ids)
    return float(r)
