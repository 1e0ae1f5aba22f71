"""Seeded MFMA scans: a prefix pass over the first 1/prepass_div of each large
range keeps a few keys whose last one seeds the threshold of the main pass;
the merge verifies each seeded query and rescans the failures.

The top-k' by approximate key is unique (keys carry the row position), so a
seeded search must return exactly what the unseeded search (prepass_div = 1)
returns, bit for bit, whatever the margin: a tiny margin only makes most
seeds fail and sends those queries through the rescan.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MODES = ["bf16", "sq8", "fp8", "sq8sym"]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def big(pkg):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    dg, fbf = pkg.datagen, pkg.fbf
    cfg = dg.CONFIGS[0]
    vecs, labels, ts = dg.make_base(cfg, n=200_000)
    q, metas = dg.make_queries(cfg, nq=160)
    metas[:40, 0] = 0  # unfiltered: 200000 rows, prefix 25000 rows at prepass_div 8
    ds = fbf.Dataset(vecs, labels, ts, repr_flags=fbf.REPR_BF16 | fbf.REPR_SQ8 | fbf.REPR_FP8 | fbf.REPR_SQ8SYM)
    yield vecs, labels, ts, q, metas, ds
    ds.close()


def _search(ds, q, metas, **kw):
    import torch
    qd = torch.from_numpy(np.ascontiguousarray(q)).cuda()
    ids, d = ds.search(qd, metas, **kw)
    torch.cuda.synchronize()
    return ids.cpu().numpy().view(np.uint32), d.cpu().numpy()


@pytest.mark.parametrize("mode", MODES)
def test_seeded_equals_unseeded(pkg, big, mode):
    vecs, labels, ts, q, metas, ds = big
    _say(f"case: N=200000 nq=160 mode={mode} unseeded reference")
    ref_i, ref_d = _search(ds, q, metas, k=100, mode=mode, prepass_div=1)
    assert ds.last_rescans() == 0
    # group 0 (default) and 256: main-pass items run as two waves of four
    # 32-query blocks, prefix items as two waves of two; group 128: both
    # passes two waves of two blocks
    for div, margin, persistent, group in [(0, 0, 0, 0), (8, 10, 0, 0), (8, 10, 2, 0), (4, 400, 1, 0), (8, 0, 2, 0),
                                           (0, 0, 0, 256), (8, 10, 0, 256), (0, 0, 0, 128), (8, 10, 0, 128)]:
        _say(f"case: mode={mode} prepass_div={div} margin={margin} persistent={persistent} group={group}")
        ids, d = _search(ds, q, metas, k=100, mode=mode, prepass_div=div, prepass_margin=margin,
                         persistent=persistent, group_queries=group)
        resc = ds.last_rescans()
        _say(f"  rescanned {resc}")
        assert np.array_equal(ids, ref_i), (mode, div, margin)
        assert np.array_equal(d, ref_d), (mode, div, margin)
        if margin == 10:
            assert resc > 0  # the tiny margin must exercise the rescan path


def test_seeded_small_kprime(pkg, big):
    vecs, labels, ts, q, metas, ds = big
    _say("case: N=200000 bf16 k=10 k'=10 seeded vs unseeded")
    ref_i, ref_d = _search(ds, q, metas, k=10, kprime=10, mode="bf16", prepass_div=1)
    ids, d = _search(ds, q, metas, k=10, kprime=10, mode="bf16", prepass_div=8, group_queries=32)
    assert np.array_equal(ids, ref_i) and np.array_equal(d, ref_d)
