"""Certified exact mode (FBF_MODE_EXACT_BF16): the BF16 MFMA scan with an FP32
rerank of k' candidates, a per-query error certificate, and the FP32 scan for
the queries the certificate leaves open.  Its ids and distances must equal
the exact FP32 mode's bit for bit (the exact mode is itself pinned to the
oracle and to the reference's Bruteforce output), whatever k' is: a small k'
only sends more queries down the fallback path.
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
    ds = fbf.Dataset(vecs, labels, ts, repr_flags=fbf.REPR_F32 | fbf.REPR_BF16)
    yield q, metas, ds
    ds.close()


def _search(ds, q, metas, **kw):
    import torch
    qd = torch.from_numpy(np.ascontiguousarray(q)).cuda()
    ids, d = ds.search(qd, metas, **kw)
    torch.cuda.synchronize()
    return ids.cpu().numpy().view(np.uint32), d.cpu().numpy()


@pytest.mark.parametrize("k,kprime", [(100, 0), (100, 100), (100, 300), (10, 0), (10, 10)])
def test_exact_bf16_equals_exact(pkg, data, k, kprime):
    q, metas, ds = data
    _say(f"case: N=200000 nq=300 exact reference k={k}")
    ref_i, ref_d = _search(ds, q, metas, k=k, mode="exact")
    _say(f"case: exact_bf16 k={k} kprime={kprime}")
    ids, d = _search(ds, q, metas, k=k, mode="exact_bf16", kprime=kprime)
    fb = ds.last_fallbacks()
    _say(f"  fallbacks {fb} of {len(q)}")
    assert np.array_equal(ids, ref_i)
    assert np.array_equal(d, ref_d)
    if kprime == k:
        assert fb > 0  # k' = k leaves almost every query uncertified: the fallback path runs
    if kprime == 0:
        assert fb < len(q) // 2


def test_exact_bf16_host_and_errors(pkg, data):
    fbf = pkg.fbf
    q, metas, ds = data
    _say("case: exact_bf16 through fbf_search_host")
    ref_i, ref_d = ds.search_host(q[:40], metas[:40], k=100, mode="exact")
    ids, d = ds.search_host(q[:40], metas[:40], k=100, mode="exact_bf16")
    assert np.array_equal(ids, ref_i) and np.array_equal(d, ref_d)
    _say("case: exact_bf16 cannot be captured")
    import torch
    qd = torch.from_numpy(np.ascontiguousarray(q[:40])).cuda()
    with pytest.raises(fbf.FbfError):
        fbf.SearchGraph(ds, qd, metas[:40], k=100, mode="exact_bf16")
    _say("case: exact_bf16 needs both representations")
    dg = pkg.datagen
    v2, l2, t2 = dg.make_base(dg.CONFIGS[0], n=5000)
    ds2 = fbf.Dataset(v2, l2, t2, repr_flags=fbf.REPR_BF16)
    try:
        with pytest.raises(fbf.FbfError):
            ds2.search_host(q[:4], metas[:4], k=10, mode="exact_bf16")
    finally:
        ds2.close()


@pytest.mark.parametrize("mode", ["exact", "bf16", "exact_bf16"])
def test_max_batch_split_equals_whole(pkg, data, mode):
    """A batch larger than max_batch runs as consecutive sub-batches (bounded
    list memory); the results must not change."""
    q, metas, ds = data
    _say(f"case: {mode} whole batch vs max_batch=70")
    ref_i, ref_d = _search(ds, q, metas, k=100, mode=mode)
    ids, d = _search(ds, q, metas, k=100, mode=mode, max_batch=70)
    assert np.array_equal(ids, ref_i)
    assert np.array_equal(d, ref_d)
