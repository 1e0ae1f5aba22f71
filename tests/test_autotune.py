"""Autotuner: its sampling and recall helpers on the CPU, the tuner itself on
the GPU (it must return options whose recall on the sample meets the target,
and those options must be the fastest such candidate it tried)."""
import numpy as np
import pytest


def test_stratified_sample_covers_types(pkg):
    at = pkg.autotune
    metas = np.zeros((1000, 4), np.float32)
    metas[:, 0] = np.arange(1000) % 4
    idx = at.stratified_sample(metas, 100)
    assert idx.size == 100 and np.all(np.diff(idx) > 0)
    for t in range(4):
        assert np.sum(metas[idx, 0] == t) == 25


def test_recall_ignores_padding(pkg):
    at = pkg.autotune
    pad = pkg.fbf.PAD_ID
    ref = np.array([[1, 2, 3, pad], [pad, pad, pad, pad]], np.uint32)
    got = np.array([[3, 2, 9, 7], [5, 6, 7, 8]], np.uint32)
    assert at.recall_at_k(got, ref) == pytest.approx(2 / 3)
    assert at.recall_at_k(ref, ref) == 1.0


@pytest.mark.gpu
def test_tune_meets_target(pkg):
    import sys
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    print("case: autotune bf16 N=10000 nq=400", flush=True)
    sys.stdout.flush()
    fbf, dg = pkg.fbf, pkg.datagen
    cfg = dg.CONFIGS[0]
    vecs, labels, ts = dg.make_base(cfg)
    q, metas = dg.make_queries(cfg, nq=400)
    ds = fbf.Dataset(vecs, labels, ts, repr_flags=fbf.REPR_BF16 | fbf.REPR_F32)
    try:
        qd = torch.from_numpy(q).cuda()
        res = pkg.autotune.tune(ds, qd, metas, k=100, mode="bf16", recall_target=0.99, sample=128,
                                groups=(64, 128), chunks=(4096,), kprime_factors=(1.0, 1.5, 4.0))
        assert res.recall >= 0.99
        ok = [r for r in res.table if r[2] >= 0.99]
        assert res.ms <= 1.03 * min(r[1] for r in ok)
    finally:
        ds.close()
