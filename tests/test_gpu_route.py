"""The device router (fbf_route) on edge filters: unknown filter types,
labels absent from the base, empty and inverted time ranges, -0.0 bounds,
repeated filters (the query index breaks ties in the order), and searches
whose metas live in device memory (fbf_search_opts.metas_on_device)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_route_edges_and_device_metas(pkg, orc):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    fbf, dg = pkg.fbf, pkg.datagen
    cfg = dg.CONFIGS[0]
    vecs, labels, ts = dg.make_base(cfg)
    q, m = dg.make_queries(cfg, nq=500)
    m[10:20] = m[5]                                   # repeated filters
    m[20] = [1, 5000, -1, -1]                         # label absent from the base
    m[21] = [2, -1, 0.7, 0.3]                         # inverted range
    m[22] = [3, 0, -0.0, 0.0]                         # -0.0 bound
    m[23] = [2, -1, 0.0, 1.0]                         # everything
    m[24] = [0, -1, -1, -1]
    L = orc.Layouts(labels.astype(np.int32), ts)
    ds = fbf.Dataset(vecs, labels, ts, repr_flags=fbf.REPR_F32 | fbf.REPR_BF16)
    try:
        _say("case: N=10000 nq=500 device router edges")
        md = torch.from_numpy(m).cuda()
        ranges, order = ds.route(md)
        torch.cuda.synchronize()
        r = ranges.cpu().numpy().view(np.uint32)
        assert np.array_equal(r[:, :3].astype(np.int64), L.ranges(m))
        counts = (r[:, 2] - r[:, 1]).astype(np.int64)
        sel = (counts.astype(np.float32) / np.float32(vecs.shape[0])).astype(np.float32)
        sel[m[:, 0] == 0] = 1
        bf = (sel < np.float32(0.045)) | ((m[:, 0] == 3) & (sel < np.float32(0.08)))
        assert np.array_equal(r[:, 3].astype(bool), bf)
        ref_order = np.lexsort((np.arange(500), m[:, 3], m[:, 2], m[:, 1].astype(np.int32),
                                m[:, 0].astype(np.uint32)))
        assert np.array_equal(order.cpu().numpy(), ref_order.astype(np.int32))
        _say("case: searches with device metas equal host metas (exact, bf16)")
        qd = torch.from_numpy(q).cuda()
        for mode in ("exact", "bf16"):
            a, ad = ds.search(qd, m, k=100, mode=mode)
            b, bd = ds.search(qd, md, k=100, mode=mode)
            torch.cuda.synchronize()
            assert torch.equal(a, b) and torch.equal(ad, bd), mode
        _say("case: unknown filter types")
        bad = m[:8].copy()
        bad[1, 0] = 5
        bad[3, 0] = 1.5
        bad[6, 0] = -1
        rb, _ = ds.route(torch.from_numpy(bad).cuda(), order=False)
        torch.cuda.synchronize()
        rb = rb.cpu().numpy().view(np.uint32)
        assert list(np.nonzero(rb[:, 0] == 0xFFFFFFFF)[0]) == [1, 3, 6]
        assert np.all(rb[[1, 3, 6], 1:] == 0)
        with pytest.raises(fbf.FbfError, match="query 1: bad filter type"):
            ds.search(qd[:8], torch.from_numpy(bad).cuda(), k=10, mode="bf16")
    finally:
        ds.close()
