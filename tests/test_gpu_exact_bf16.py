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


@pytest.mark.parametrize("k", [1, 100, 512])
def test_edge_filters_all_modes(pkg, k):
    """Absent category, inverted and single-timestamp windows, ragged N, k up to
    512: the certified mode equals the exact mode bit for bit; the BF16 mode
    equals it on every query with no more rows than k' (all rows reranked) and
    pads the same way; an empty batch is a no-op in every mode."""
    import torch
    fbf, dg = pkg.fbf, pkg.datagen
    cfg = dg.CONFIGS[0]
    vecs, labels, ts = dg.make_base(cfg, n=3037)
    q, metas = dg.make_queries(cfg, nq=120)
    metas[0] = [1, 99999, -1, -1]
    metas[1] = [2, -1, 0.6, 0.5]
    metas[2] = [2, -1, ts[11], ts[11]]
    metas[3] = [3, labels[5], ts[5], ts[5]]
    _say(f"case: N=3037 edge filters, k={k}, exact / exact_bf16 / bf16")
    ds = fbf.Dataset(vecs, labels, ts, repr_flags=fbf.REPR_F32 | fbf.REPR_BF16)
    try:
        ref_i, ref_d = _search(ds, q, metas, k=k, mode="exact")
        ids, d = _search(ds, q, metas, k=k, mode="exact_bf16")
        assert np.array_equal(ids, ref_i) and np.array_equal(d, ref_d)
        kp = min(896, max(k, (3 * k + 1) // 2))
        bi, bd = _search(ds, q, metas, k=k, mode="bf16")
        for i in range(len(q)):
            c, _ = ds.filter_count(metas[i])
            if c <= kp:
                assert np.array_equal(bi[i], ref_i[i]) and np.array_equal(bd[i], ref_d[i]), i
        assert np.all(bi[0] == fbf.PAD_ID) and np.all(bi[1] == fbf.PAD_ID)
        empty_q = torch.zeros((0, 100), dtype=torch.float32).cuda()
        for mode in ("exact", "bf16", "exact_bf16"):
            i0, d0 = ds.search(empty_q, np.zeros((0, 4), np.float32), k=k, mode=mode)
            assert i0.shape == (0, k) and d0.shape == (0, k)
    finally:
        ds.close()
