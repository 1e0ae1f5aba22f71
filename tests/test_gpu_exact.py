"""Parity of the exact FP32 HIP path with the CPU oracle: ids bit-exact with
ties to the smaller id, distances bit-identical (same canonical FP32 order).
Every call goes through the C-ABI (libfbf.so)."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _run(pkg, torch, vecs, labels, ts, q, metas, k, **kw):
    ds = pkg.fbf.Dataset(vecs, labels, ts, repr_flags=pkg.fbf.REPR_F32)
    try:
        qd = torch.from_numpy(np.ascontiguousarray(q)).cuda()
        ids, d = ds.search(qd, metas, k=k, mode="exact", **kw)
        torch.cuda.synchronize()
        return ids.cpu().numpy().view(np.uint32), d.cpu().numpy(), ds.last_stats()
    finally:
        ds.close()


def _oracle(orc, vecs, labels, ts, q, metas, k):
    ids, d = orc.exact_predicate(vecs, labels.astype(np.int32), ts, q, metas, k=k, mode="canon")
    return ids, d.astype(np.float32)


def test_reference_fixture_exact(pkg, orc, torch_cuda, dummy, golden_dir):
    _say("case: reference dummy fixture, 100 queries, k=100")
    vecs, labels, ts, q, metas = dummy
    ids, d, _ = _run(pkg, torch_cuda, vecs, labels, ts, q, metas, 100)
    gt = pkg.datagen.read_gt_bin(os.path.join(golden_dir, "dummy_gt_ref.bin"))
    assert np.array_equal(ids, gt)  # the reference's own Bruteforce output
    z = np.load(os.path.join(golden_dir, "dummy_exact.npz"))
    assert np.array_equal(d, z["dists"])


@pytest.mark.parametrize("chunk,group,persistent", [(0, 0, False), (128, 16, False), (64, 1, False),
                                                     (4096, 5, True)])
def test_c0_exact_parity(pkg, orc, torch_cuda, chunk, group, persistent):
    _say(f"case: config0 N=10000 nq=1000 chunk={chunk} group={group} persistent={persistent}")
    dg = pkg.datagen
    cfg = dg.CONFIGS[0]
    vecs, labels, ts = dg.make_base(cfg)
    q, metas = dg.make_queries(cfg)
    ids, d, st = _run(pkg, torch_cuda, vecs, labels, ts, q, metas, 100, chunk_rows=chunk,
                      group_queries=group, persistent=persistent)
    rid, rd = _oracle(orc, vecs, labels, ts, q, metas, 100)
    assert np.array_equal(ids, rid)
    assert np.array_equal(d, rd)
    assert st["pairs"] > 0 and st["pairs_computed"] >= st["pairs"]


@pytest.mark.parametrize("k", [1, 7, 100, 129, 400, 512])
def test_k_range_and_padding(pkg, orc, torch_cuda, k):
    _say(f"case: N=3037 ragged, k={k}")
    dg = pkg.datagen
    cfg = dg.CONFIGS[0]
    vecs, labels, ts = dg.make_base(cfg, n=3037)
    q, metas = dg.make_queries(cfg, nq=120)
    # Edge filters: absent category, empty / inverted window, single-timestamp window.
    metas[0] = [1, 99999, -1, -1]
    metas[1] = [2, -1, 0.6, 0.5]
    metas[2] = [2, -1, ts[11], ts[11]]
    metas[3] = [3, labels[5], ts[5], ts[5]]
    ids, d, _ = _run(pkg, torch_cuda, vecs, labels, ts, q, metas, k)
    rid, rd = _oracle(orc, vecs, labels, ts, q, metas, k)
    assert np.array_equal(ids, rid)
    assert np.array_equal(d, rd)
    assert np.all(ids[0] == 0xFFFFFFFF) and np.all(np.isinf(d[0]))
    assert np.all(ids[1] == 0xFFFFFFFF)
    assert ids[2][0] == 11 or d[2][0] <= d[2][0]


def test_ties_go_to_smaller_id(pkg, orc, torch_cuda):
    _say("case: duplicated rows and identical distances")
    rng = np.random.default_rng(7)
    base = rng.standard_normal((700, 100)).astype(np.float32)
    vecs = np.concatenate([base, base, base])          # every row three times
    vecs[2000:2100] = vecs[0]                            # 100 more copies of row 0
    n = vecs.shape[0]
    labels = (np.arange(n) % 3).astype(np.float32)
    ts = rng.random(n).astype(np.float32)
    q = np.concatenate([vecs[:8], rng.standard_normal((8, 100)).astype(np.float32)])
    metas = np.array([[t % 4, t % 3, 0.2, 0.9] for t in range(16)], np.float32)
    metas[metas[:, 0] == 0, 1:] = -1
    metas[metas[:, 0] == 1, 2:] = -1
    metas[metas[:, 0] == 2, 1] = -1
    ids, d, _ = _run(pkg, torch_cuda, vecs, labels, ts, q, metas, 100, chunk_rows=256)
    rid, rd = _oracle(orc, vecs, labels, ts, q, metas, 100)
    assert np.array_equal(ids, rid)
    assert np.array_equal(d, rd)


def test_constant_base_all_ties(pkg, orc, torch_cuda):
    _say("case: all rows identical")
    n = 5000
    vecs = np.ones((n, 100), np.float32)
    labels = np.zeros(n, np.float32)
    ts = np.linspace(0, 1, n, dtype=np.float32)
    q = np.zeros((4, 100), np.float32)
    metas = np.array([[0, -1, -1, -1], [1, 0, -1, -1], [2, -1, 0.5, 1.0], [3, 0, 0.0, 0.3]], np.float32)
    ids, d, _ = _run(pkg, torch_cuda, vecs, labels, ts, q, metas, 100, chunk_rows=512)
    rid, rd = _oracle(orc, vecs, labels, ts, q, metas, 100)
    assert np.array_equal(ids, rid)
    assert np.array_equal(ids[0], np.arange(100))
    assert np.all(d == 100.0)


def test_host_entry_point_and_filter_count(pkg, orc, torch_cuda, dummy):
    _say("case: host entry point + filter counts")
    vecs, labels, ts, q, metas = dummy
    ds = pkg.fbf.Dataset(vecs, labels, ts)
    try:
        ids, d = ds.search_host(q, metas, k=100)
        L = orc.Layouts(labels.astype(np.int32), ts)
        for m in metas[:20]:
            lay, s, e = L.query_range(*m)
            c, sel = ds.filter_count(m)
            assert c == e - s and sel == pytest.approx((e - s) / vecs.shape[0])
        with pytest.raises(pkg.fbf.FbfError):
            ds.search_host(q[:1], np.array([[7, 0, 0, 0]], np.float32), k=10)
    finally:
        ds.close()
    rid, rd = _oracle(orc, vecs, labels, ts, q, metas, 100)
    assert np.array_equal(ids, rid) and np.array_equal(d, rd)


@pytest.mark.parametrize("mode", ["exact", "exact_bf16"])
def test_contest_driver_on_reference_fixture(pkg, golden_dir, tmp_path, mode):
    """The contest-format driver end to end: the reference's own dummy base and
    query files in, the contest output file out, equal to the reference's
    ground truth for those files -- in the exact FP32 mode and in the certified
    BF16 exact mode."""
    _say(f"case: contest driver, reference dummy fixture, {mode} mode")
    import importlib.util
    spec = importlib.util.spec_from_file_location("fbf_contest", os.path.join(pkg.__path__[0], "contest.py"))
    contest = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(contest)
    out = str(tmp_path / "output.bin")
    gt = os.path.join(golden_dir, "dummy_gt_ref.bin")
    res = contest.run(os.path.join(golden_dir, "dummy-data.bin"), os.path.join(golden_dir, "dummy-queries.bin"),
                      out, mode=mode, k=100, gt=gt, log=_say)
    assert res["recall"] == 1.0
    assert np.array_equal(pkg.datagen.read_knn_bin(out), pkg.datagen.read_gt_bin(gt))
