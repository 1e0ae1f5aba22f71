"""Captured searches (fbf_search_graph_create / _launch / _destroy): a graph
replay must give exactly the eager search's ids and distances, on every
replay, with other eager searches and synchronisations between the replays,
and again after the query vectors are refilled in place (the filters, and so
the plan, stay fixed).  The batch is large enough for the seeded prefix pass,
its verification and the masked rescan to be part of the captured sequence.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def data(pkg):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    dg, fbf = pkg.datagen, pkg.fbf
    cfg = dg.CONFIGS[0]
    vecs, labels, ts = dg.make_base(cfg, n=200_000)
    q, metas = dg.make_queries(cfg, nq=300)
    metas[:60, 0] = 0  # unfiltered: 200000 rows, seeded prefix pass
    q2 = (q + 0.3 * np.random.default_rng(7).standard_normal(q.shape)).astype(np.float32)
    ds = fbf.Dataset(vecs, labels, ts, repr_flags=fbf.REPR_F32 | fbf.REPR_BF16 | fbf.REPR_SQ8 | fbf.REPR_SQ8SYM)
    yield q, q2, metas, ds
    ds.close()


def _host(t):
    return t.cpu().numpy()


@pytest.mark.parametrize("mode,kw", [("bf16", {}), ("bf16", {"prepass_margin": 10}),
                                     ("bf16", {"group_queries": 256}), ("sq8", {}),
                                     ("sq8sym", {}), ("exact", {})])
def test_graph_replays_equal_eager(pkg, data, mode, kw):
    import torch
    fbf = pkg.fbf
    q, q2, metas, ds = data
    qd = torch.from_numpy(np.ascontiguousarray(q)).cuda()
    _say(f"case: N=200000 nq=300 mode={mode} {kw} eager reference")
    ref_i, ref_d = ds.search(qd, metas, k=100, mode=mode, **kw)
    torch.cuda.synchronize()
    ref_i, ref_d = _host(ref_i), _host(ref_d)
    _say(f"case: mode={mode} capture")
    g = fbf.SearchGraph(ds, qd, metas, k=100, mode=mode, **kw)
    try:
        for r in range(3):
            _say(f"case: mode={mode} replay {r}")
            ids, d = g.launch()
            torch.cuda.synchronize()
            assert np.array_equal(_host(ids), ref_i), (mode, r)
            assert np.array_equal(_host(d), ref_d), (mode, r)
            # eager work of the program between replays
            other = torch.from_numpy(np.ascontiguousarray(q2[:50])).cuda()
            ds.search(other, metas[:50], k=10, mode=mode)
            torch.cuda.synchronize()
        # new query vectors, same filters
        _say(f"case: mode={mode} refilled queries")
        qd.copy_(torch.from_numpy(np.ascontiguousarray(q2)))
        e_i, e_d = ds.search(qd, metas, k=100, mode=mode, **kw)
        torch.cuda.synchronize()
        e_i, e_d = _host(e_i), _host(e_d)
        ids, d = g.launch()
        torch.cuda.synchronize()
        assert np.array_equal(_host(ids), e_i)
        assert np.array_equal(_host(d), e_d)
        assert not np.array_equal(e_i, ref_i)
    finally:
        g.close()
