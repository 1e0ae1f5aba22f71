"""Any-batch captured searches (fbf_search_graph_create with metas in device
memory): the graph routes, plans and searches on the device at every replay,
so a replay after the metas and the query vectors were refilled in place
must give exactly what an eager search of the new batch gives (approximate
modes keep the same keys whatever the plan).  Batches differ in size of
ranges and filter mix; other eager work runs between replays; unknown filter
types give padded rows and leave the other rows unchanged.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 200_000
NQ = 300


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def data(pkg):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    dg, fbf = pkg.datagen, pkg.fbf
    cfg = dg.CONFIGS[0]
    vecs, labels, ts = dg.make_base(cfg, n=N)
    qa, ma = dg.make_queries(cfg, nq=NQ)
    ma[:60, 0] = 0  # unfiltered: whole layout, seeded prefix pass
    rng = np.random.default_rng(11)
    qb = (qa[rng.permutation(NQ)] + 0.3 * rng.standard_normal(qa.shape)).astype(np.float32)
    mb = ma[rng.permutation(NQ)].copy()
    mb[:, 2:] = np.sort(rng.random((NQ, 2), dtype=np.float32), axis=1)  # other time windows
    mb[mb[:, 0] == 0, 0] = 2  # batch B has no unfiltered query
    ds = fbf.Dataset(vecs, labels, ts, repr_flags=fbf.REPR_BF16 | fbf.REPR_FP8 | fbf.REPR_SQ8 | fbf.REPR_SQ8SYM)
    yield (qa, ma), (qb, mb), ds
    ds.close()


def _host(t):
    return t.cpu().numpy()


def _eager(ds, q, m, mode, kw):
    import torch
    i, d = ds.search(torch.from_numpy(np.ascontiguousarray(q)).cuda(), m, k=100, mode=mode, **kw)
    torch.cuda.synchronize()
    return _host(i), _host(d)


@pytest.mark.parametrize("mode,kw", [("bf16", {}), ("bf16", {"prepass_margin": 10}), ("bf16", {"prepass_div": 1}),
                                     ("bf16", {"group_queries": 64, "chunk_rows": 8192}), ("fp8", {}),
                                     ("sq8", {}), ("sq8sym", {})])
def test_any_batch_replays_equal_eager(pkg, data, mode, kw):
    import torch
    fbf = pkg.fbf
    (qa, ma), (qb, mb), ds = data
    batches = [(qa, ma), (qb, mb), (qa, ma)]
    refs = []
    for j, (q, m) in enumerate(batches[:2]):
        _say(f"case: N={N} nq={NQ} mode={mode} {kw} eager batch {j}")
        refs.append(_eager(ds, q, m, mode, kw))
    refs.append(refs[0])
    assert not np.array_equal(refs[0][0], refs[1][0])
    qd = torch.from_numpy(np.ascontiguousarray(qa)).cuda()
    md = torch.from_numpy(np.ascontiguousarray(ma)).cuda()
    _say(f"case: mode={mode} {kw} capture")
    g = fbf.SearchGraph(ds, qd, md, k=100, mode=mode, **kw)
    try:
        for r, ((q, m), (ri, rd)) in enumerate(zip(batches, refs)):
            qd.copy_(torch.from_numpy(np.ascontiguousarray(q)))
            md.copy_(torch.from_numpy(np.ascontiguousarray(m)))
            _say(f"case: mode={mode} {kw} replay {r}")
            ids, d = g.launch()
            torch.cuda.synchronize()
            info = g.plan_info()
            assert not info["overflow"] and info["items_main"] > 0, info
            assert np.array_equal(_host(ids), ri), (mode, r)
            assert np.array_equal(_host(d), rd), (mode, r)
            # eager work of the program between replays
            ds.search(torch.from_numpy(np.ascontiguousarray(qb[:50])).cuda(), mb[:50], k=10, mode=mode)
            torch.cuda.synchronize()
    finally:
        g.close()


def test_any_batch_unknown_type_pads(pkg, data):
    import torch
    fbf = pkg.fbf
    (qa, ma), _, ds = data
    ref_i, ref_d = _eager(ds, qa, ma, "bf16", {})
    bad = ma.copy()
    bad[[3, 77], 0] = 7.0  # no such filter type
    qd = torch.from_numpy(np.ascontiguousarray(qa)).cuda()
    md = torch.from_numpy(bad).cuda()
    _say(f"case: N={N} nq={NQ} bf16 unknown filter types, capture")
    g = fbf.SearchGraph(ds, qd, md, k=100, mode="bf16")
    try:
        _say("case: bf16 unknown filter types, replay")
        ids, d = g.launch()
        torch.cuda.synchronize()
        ids, d = _host(ids), _host(d)
        keep = np.ones(NQ, bool)
        keep[[3, 77]] = False
        assert np.array_equal(ids[keep], ref_i[keep])
        assert np.array_equal(d[keep], ref_d[keep])
        assert (ids[~keep] == -1).all() and np.isinf(d[~keep]).all()
    finally:
        g.close()


def test_plan_info_rejects_host_planned_graph(pkg, data):
    import torch
    fbf = pkg.fbf
    (qa, ma), _, ds = data
    qd = torch.from_numpy(np.ascontiguousarray(qa[:20])).cuda()
    g = fbf.SearchGraph(ds, qd, ma[:20], k=10, mode="bf16")
    try:
        with pytest.raises(fbf.FbfError):
            g.plan_info()
    finally:
        g.close()
