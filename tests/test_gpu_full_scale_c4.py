"""Full-scale parity at the uneven-batch workload (10,000,000 rows, 10,000
Zipf(1.3) categories, 100,000 queries with log-uniform timestamp widths in
[0.001, 0.8], arriving in 193 batches of 1-4096 queries):

* exact FP32 over the whole batch: bit-identical to the oracle on a
  stratified sample that includes the rarest categories' queries and the
  queries with fewer than k candidates; the invariants on all queries;
* exact_bf16 equals exact FP32 on all queries;
* BF16 + FP32 rerank with the bench's tuned options: all batches fused into
  one search return exactly what one search per batch returns, the fused
  search captured as a HIP graph replays to the same result, and recall@100
  against the exact results over the whole batch is >= 0.99.
"""
import sys

import numpy as np
import pytest

import scale_checks as sc

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]
K = 100
TUNED = dict(group_queries=256, chunk_rows=32768, kprime=110)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def c4(pkg, orc):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    fbf, dg = pkg.fbf, pkg.datagen
    cfg = dg.CONFIGS[4]
    _say("fixture: generating the uneven-batch N=10M workload")
    vecs, labels, ts = dg.make_base(cfg)
    q, metas = dg.make_queries(cfg)
    lab = labels.astype(np.int32)
    ds = fbf.Dataset(vecs, labels, ts, repr_flags=fbf.REPR_F32 | fbf.REPR_BF16)
    L = orc.Layouts(lab, ts)
    rng = L.ranges(metas)
    counts = (rng[:, 2] - rng[:, 1]).astype(np.int64)
    sample = sc.stratified_sample(metas, counts, per_type=48, smallest=48, largest=8)
    sizes = dg.batch_sizes(cfg, q.shape[0])
    qd = torch.from_numpy(q).cuda()
    yield dict(vecs=vecs, lab=lab, ts=ts, q=q, metas=metas, ds=ds, L=L, counts=counts, sample=sample,
               sizes=sizes, qd=qd)
    ds.close()


def _search(st, qd=None, metas=None, **kw):
    import torch
    qd = st["qd"] if qd is None else qd
    metas = st["metas"] if metas is None else metas
    ids, d = st["ds"].search(qd, metas, k=K, **kw)
    torch.cuda.synchronize()
    return ids.cpu().numpy().view(np.uint32), d.cpu().numpy()


@pytest.fixture(scope="module")
def exact(c4):
    _say("case: C4 exact FP32 search, whole batch")
    return _search(c4, mode="exact")


def test_c4_exact_vs_oracle_and_invariants(orc, c4, exact):
    ids, d = exact
    s = c4["sample"]
    L = c4["L"]
    v = c4["vecs"]
    rare = int((c4["counts"][s] < K).sum())
    _say(f"case: C4 exact vs oracle on {s.size} queries ({rare} with fewer than {K} rows)")
    # the oracle over the predicate's ranges, rows gathered per layout
    vb = np.ascontiguousarray(v[L.perm_both]); vt = np.ascontiguousarray(v[L.perm_time])
    oid, od = orc.exact_ranges_prepared(vb, vt, L, c4["q"][s], c4["metas"][s], K)
    del vb, vt
    assert np.array_equal(ids[s], oid)
    assert np.array_equal(d[s], od)
    sc.check_invariants(orc, v, c4["lab"], c4["ts"], c4["q"], c4["metas"], c4["counts"], ids, d, K)


def test_c4_exact_bf16_equals_exact(c4, exact):
    _say("case: C4 exact_bf16 vs exact FP32, whole batch")
    bi, bd = _search(c4, mode="exact_bf16")
    _say(f"  {c4['ds'].last_fallbacks()} queries left to the FP32 scan")
    assert np.array_equal(bi, exact[0]) and np.array_equal(bd, exact[1])


def test_c4_fused_equals_per_batch_and_graph(pkg, orc, c4, exact):
    import torch
    sizes = c4["sizes"]
    off = np.concatenate([[0], np.cumsum(sizes)])
    _say(f"case: C4 bf16 {TUNED}: {len(sizes)} batches fused vs one search each")
    fi, fd = _search(c4, mode="bf16", **TUNED)
    pi = np.zeros_like(fi); pd = np.zeros_like(fd)
    for i in range(len(sizes)):
        a, b = int(off[i]), int(off[i + 1])
        pi[a:b], pd[a:b] = _search(c4, qd=c4["qd"][a:b], metas=c4["metas"][a:b], mode="bf16", **TUNED)
    assert np.array_equal(fi, pi) and np.array_equal(fd, pd)
    sc.check_invariants(orc, c4["vecs"], c4["lab"], c4["ts"], c4["q"], c4["metas"], c4["counts"], fi, fd, K,
                        kprime=TUNED["kprime"], exact_ids=exact[0], exact_d=exact[1])
    rec = sc.recall_vs(orc, fi, exact[0])
    _say(f"  recall@100 vs exact over 100k: {rec:.5f}")
    assert rec >= 0.99
    _say("case: C4 fused batch captured as one HIP graph, replayed twice")
    g = pkg.fbf.SearchGraph(c4["ds"], c4["qd"], c4["metas"], k=K, mode="bf16", **TUNED)
    try:
        for r in range(2):
            gi, gd = g.launch()
            torch.cuda.synchronize()
            assert np.array_equal(gi.cpu().numpy().view(np.uint32), fi), r
            assert np.array_equal(gd.cpu().numpy(), fd), r
    finally:
        g.close()
