"""The type-2 sub-interval path on the GPU (fbf_search_subinterval: the
router's SMALL time buckets, candidate pool and global refine,
hybrid_graph.cpp:91-134, 317-494) against the oracle (orc_subinterval) bit
for bit -- ids, canonical FP32 distances, bucket classes -- and against the
reference route's own outputs (tests/golden/ref_subinterval.npz)."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _type2(rng, nq, lo=0.002, hi=0.6):
    m = np.full((nq, 4), -1.0, np.float32)
    m[:, 0] = 2
    l = rng.random(nq).astype(np.float32)
    m[:, 2] = l
    m[:, 3] = np.minimum(l + np.exp(rng.uniform(np.log(lo), np.log(hi), nq)).astype(np.float32), np.float32(1))
    return m


class _Case:
    def __init__(self, pkg, orc, n):
        import torch
        assert torch.cuda.is_available(), "GPU tests need a HIP device"
        fbf, dg = pkg.fbf, pkg.datagen
        self.vecs, labels, ts = dg.make_base(dg.CONFIGS[0], n=n)
        self.L = orc.Layouts(labels.astype(np.int32), ts)
        self.tb = orc.TimeBuckets(self.vecs, self.L)
        self.ds = fbf.Dataset(self.vecs, labels, ts, repr_flags=fbf.REPR_F32 | fbf.REPR_BF16 | fbf.REPR_TBUCKETS)

    def gpu(self, q, metas, **kw):
        import torch
        qd = torch.from_numpy(np.ascontiguousarray(q, np.float32)).cuda()
        ids, d, types = self.ds.search_subinterval(qd, metas, **kw)
        torch.cuda.synchronize()
        return ids.cpu().numpy().view(np.uint32), d.cpu().numpy(), types


@pytest.fixture(scope="module")
def golden_case(pkg, orc, golden_dir):
    z = np.load(os.path.join(golden_dir, "ref_subinterval.npz"))
    c = _Case(pkg, orc, int(z["n"]))
    yield c, z
    c.ds.close()


@pytest.fixture(scope="module")
def big_case(pkg, orc):
    c = _Case(pkg, orc, 400_000)
    yield c
    c.ds.close()


def test_subinterval_golden(orc, golden_case):
    c, z = golden_case
    q, m = z["queries"], z["metas"]
    _say(f"case: N=20000 {q.shape[0]} type-2 queries, GPU vs oracle vs reference route")
    ids, d, types = c.gpu(q, m)
    oid, od, ot = orc.subinterval(c.vecs, c.L, c.tb, q, m, 100)
    assert np.array_equal(types, ot)
    assert np.array_equal(ids, oid)
    assert np.array_equal(d, od)
    assert np.array_equal(types.astype(np.int32), z["types"])
    same = np.array([set(a.tolist()) == set(b.tolist()) for a, b in zip(ids, z["ids"])])
    assert same.all(), np.nonzero(~same)[0]


@pytest.mark.parametrize("kw", [{}, {"bucket_k": 50, "pool_k": 80}, {"small_thr": 0.9}, {"k": 10, "pool_k": 30}])
def test_subinterval_large(orc, big_case, kw):
    c = big_case
    rng = np.random.default_rng(53)
    nq = 3000
    q = (c.vecs[rng.integers(0, c.vecs.shape[0], nq)] + 0.3 * rng.standard_normal((nq, 100))).astype(np.float32)
    m = _type2(rng, nq)
    # edges: a whole bucket, everything, one instant, bucket boundaries, the last bucket
    m[:6, 2:] = [[0.3, 0.4], [0.0, 1.0], [0.55, 0.55], [0.1, 0.1], [0.95, 1.0], [0.0, 0.0]]
    k = kw.get("k", 100)
    _say(f"case: N=400000 {nq} type-2 queries {kw}")
    ids, d, types = c.gpu(q, m, **kw)
    oid, od, ot = orc.subinterval(c.vecs, c.L, c.tb, q, m, k, small_thr=kw.get("small_thr", 0.0),
                                  bucket_k=kw.get("bucket_k", 140), pool_k=kw.get("pool_k", 150))
    assert np.array_equal(types, ot)
    assert np.array_equal(ids, oid)
    assert np.array_equal(d, od)
    _say(f"  {int((types == 0).sum())} SMALL slices, {int((ids[:, 0] != 0xFFFFFFFF).sum())} queries with a pool")


def test_subinterval_external_and_interleaved(pkg, orc, big_case):
    import torch
    c = big_case
    rng = np.random.default_rng(59)
    nq, n_ext = 400, 33
    q = (c.vecs[rng.integers(0, c.vecs.shape[0], nq)] + 0.3 * rng.standard_normal((nq, 100))).astype(np.float32)
    m = _type2(rng, nq)
    ep = rng.integers(0, c.vecs.shape[0], (nq, n_ext)).astype(np.uint32)
    ep[:, ::7] = 0xFFFFFFFF
    ep[0, 1] = c.vecs.shape[0] + 5  # beyond the layout: ignored
    ed = rng.uniform(1e4, 5e6, (nq, n_ext)).astype(np.float32)
    _say(f"case: N=400000 {nq} type-2 queries with {n_ext} external candidates each, between other searches")
    qd = torch.from_numpy(q).cuda()
    before, _ = c.ds.search(qd, m, k=100, mode="bf16")
    ids, d, types = c.ds.search_subinterval(qd, m, ext=(torch.from_numpy(ep.view(np.int32)).cuda(),
                                                       torch.from_numpy(ed).cuda()))
    after, _ = c.ds.search(qd, m, k=100, mode="bf16")
    torch.cuda.synchronize()
    ids = ids.cpu().numpy().view(np.uint32)
    oid, od, ot = orc.subinterval(c.vecs, c.L, c.tb, q, m, 100, ext=(ep, ed))
    assert np.array_equal(ids, oid)
    assert np.array_equal(d.cpu().numpy(), od)
    assert np.array_equal(before.cpu().numpy(), after.cpu().numpy())


def test_subinterval_rejects(pkg, big_case):
    import torch
    c = big_case
    m = np.array([[1, 3, -1, -1]], np.float32)
    with pytest.raises(pkg.fbf.FbfError, match="type-2"):
        c.ds.search_subinterval(torch.zeros((1, 100), device="cuda"), m)
