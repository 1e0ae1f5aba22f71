"""The sorted layouts built on the device (stable radix sorts of packed
(label, timestamp) and timestamp keys, row id as the last key) against the
oracle's own layouts (SortDataset order, hybrid_graph.cpp:505-524): every
filter's row count must match, and so must the exact search, on data made of
key ties -- few labels (negative ones too), few timestamps, +0.0 and -0.0 --
with filter bounds placed exactly on timestamp values.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_layouts_with_key_ties(pkg, orc):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    fbf = pkg.fbf
    rng = np.random.default_rng(11)
    n, nq, k = 60_000, 200, 20
    vecs = rng.standard_normal((n, 100)).astype(np.float32)
    labels = rng.choice(np.array([-3, -1, 0, 2, 5, 7], np.float32), n)
    tvals = np.array([-0.0, 0.0, 0.25, 0.5, 0.5000001, 0.75, 1.0], np.float32)
    ts = rng.choice(tvals, n)
    q = rng.standard_normal((nq, 100)).astype(np.float32)
    metas = np.zeros((nq, 4), np.float32)
    metas[:, 0] = rng.integers(0, 4, nq)
    metas[:, 1] = rng.choice(np.array([-3, -1, 0, 2, 5, 7, 4], np.float32), nq)
    lo = rng.choice(tvals, nq)
    hi = rng.choice(tvals, nq)
    metas[:, 2] = np.minimum(lo, hi)
    metas[:, 3] = np.maximum(lo, hi)
    _say(f"case: N={n} tied keys, {nq} filters")
    ds = fbf.Dataset(vecs, labels, ts, repr_flags=fbf.REPR_F32)
    try:
        L = orc.Layouts(labels.astype(np.int32), ts)
        for m in metas:
            _, s, e = L.query_range(*m)
            c, _ = ds.filter_count(m)
            assert c == e - s, (m, c, e - s)
        qd = torch.from_numpy(q).cuda()
        ids, d = ds.search(qd, metas, k=k, mode="exact")
        torch.cuda.synchronize()
        ids = ids.cpu().numpy().view(np.uint32)
        d = d.cpu().numpy()
    finally:
        ds.close()
    rid, rd = orc.exact_predicate(vecs, labels.astype(np.int32), ts, q, metas, k=k, mode="canon")
    assert np.array_equal(ids, rid)
    assert np.array_equal(d, rd.astype(np.float32))
