"""Device quantiser codes and the router's brute-force route, pinned.

* The codes the dataset build writes on the device (symmetric int8,
  SQ8SymmetricQuantizer::encode, sq8_sym_quant.hpp:63-74; per-dimension uint8,
  SQ8Quantizer::encode, sq8_quant.hpp:51-63) are read back through the C-ABI
  and compared byte for byte with the oracle's encoders, on the config-0 base
  plus rows built to hit the rounding edges: exact half-way ties of x / alpha *
  127 (half away from zero), the absmax element (+-127), each dimension's
  maximum (255), truncation just below an integer, and a constant dimension
  (dif = 0, code 0).
* FBF_MODE_SQ8SYM at k' = 140 is the router's brute-force branch
  (hybrid_graph.cpp:203-228: symmetric int8 top-140 of the range, id map,
  FP32 rerank to 100): its ids and distances equal the oracle's restatement
  (orc_router_bf) exactly, and its id sets equal the reference route's own
  outputs (tests/golden/ref_router_bf.npz) except where the 140th int
  distance is tied or the 100th FP32 distance nearly so.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _edge_rows(rng, d=100):
    """Rows that put x / alpha * 127 exactly on k + 0.5 (alpha = 8, set by
    one element of 8.0), the absmax at +-8, and dimension 7 constant."""
    alpha = np.float32(8.0)
    rows = []
    ties = []
    for k in range(0, 127, 3):
        target = np.float32(k + 0.5)
        y = np.float32((k + 0.5) / 127.0)
        for _ in range(64):  # walk a few ulps until fl(y * 127) is the tie
            p = np.float32(y * np.float32(127.0))
            if p == target:
                break
            y = np.nextafter(y, np.float32(np.inf) if p < target else np.float32(-np.inf))
        if np.float32(y * np.float32(127.0)) == target:
            ties.append(np.float32(y * alpha))  # x / alpha is exactly y (power-of-two alpha)
    ties = np.array(ties, np.float32)
    for sign in (1, -1):
        r = np.zeros(d, np.float32)
        r[: min(d, ties.size)] = sign * ties[:d]
        rows.append(r)
    r = rng.standard_normal(d).astype(np.float32)
    r[0], r[1] = alpha, -alpha
    rows.append(r)
    out = np.stack(rows)
    return out, ties.size


@pytest.fixture(scope="module")
def edge_ds(pkg):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    fbf, dg = pkg.fbf, pkg.datagen
    cfg = dg.CONFIGS[0]
    vecs, labels, ts = dg.make_base(cfg, n=6000)
    rng = np.random.default_rng(17)
    extra, n_ties = _edge_rows(rng)
    vecs = np.concatenate([vecs, extra]).astype(np.float32)
    labels = np.concatenate([labels, np.array([3, 3, 5], np.float32)])
    ts = np.concatenate([ts, np.array([0.25, 0.5, 0.75], np.float32)])
    vecs[:, 7] = np.float32(0.375)  # constant dimension: dif = 0 -> code 0
    # truncation edges: (x - mi) / dif * 255 just below an integer in dimension
    # 11, on rows that hold neither its minimum nor its maximum
    mi, mx = vecs[:, 11].min(), vecs[:, 11].max()
    dif = np.float32(mx - mi)
    trunc_rows = np.argsort(vecs[:, 11])[[100, 2000, 4000]]
    for j, c in zip(trunc_rows, (17, 100, 254)):
        # the largest x whose (x - mi) / dif * 255 rounds below c
        x = np.float32(mi + dif * np.float32(c) / np.float32(255.0))
        while np.float32(np.float32(np.float32(x - mi) / dif) * np.float32(255.0)) >= np.float32(c):
            x = np.nextafter(x, np.float32(-np.inf))
        vecs[j, 11] = x
    assert vecs[:, 11].min() == mi and vecs[:, 11].max() == mx
    ds = fbf.Dataset(vecs, labels, ts, repr_flags=fbf.REPR_F32 | fbf.REPR_SQ8 | fbf.REPR_SQ8SYM)
    yield vecs, labels, ts, ds, n_ties, trunc_rows
    ds.close()


def test_device_codes_equal_oracle_encoders(pkg, orc, edge_ds):
    fbf = pkg.fbf
    vecs, labels, ts, ds, n_ties, trunc_rows = edge_ds
    _say(f"case: N={vecs.shape[0]} device codes vs oracle encoders ({n_ties} half-way ties)")
    assert n_ties >= 30
    alpha, mi, dif = ds.quant_params()
    assert alpha == orc.sq8sym_train(vecs) == np.float32(8.0)
    mi2, dif2 = orc.sq8_train(vecs)
    assert np.array_equal(mi, mi2) and np.array_equal(dif, dif2)
    L = orc.Layouts(labels.astype(np.int32), ts)
    want_sym = orc.sq8sym_encode(vecs, alpha)
    want_u8 = orc.sq8_encode(vecs, mi, dif)
    for layout, perm in ((0, L.perm_both), (1, L.perm_time)):
        ids, sym = ds.layout_rows(layout, repr_flag=fbf.REPR_SQ8SYM)
        assert np.array_equal(ids, perm)
        assert np.array_equal(sym, want_sym[perm]), layout
        _, u8 = ds.layout_rows(layout, repr_flag=fbf.REPR_SQ8)
        assert np.array_equal(u8, want_u8[perm]), layout
        # a slice that starts and ends inside 32-row fragment blocks
        ids_s, u8_s = ds.layout_rows(layout, 45, 77, fbf.REPR_SQ8)
        assert np.array_equal(ids_s, perm[45:122]) and np.array_equal(u8_s, want_u8[perm][45:122])
    # The edges themselves, as the closed forms fix them.
    n = vecs.shape[0]
    cols = np.array([c for c in range(n_ties) if c != 7])  # dimension 7 was made constant
    tie_rows = want_sym[n - 3: n - 1][:, cols].astype(np.int32)
    k = np.arange(0, 127, 3)[cols]
    assert np.array_equal(tie_rows[0], k + 1)          # +(k + 0.5) -> k + 1
    assert np.array_equal(tie_rows[1], -(k + 1))       # -(k + 0.5) -> -(k + 1)
    assert want_sym[n - 1, 0] == 127 and want_sym[n - 1, 1] == -127
    assert np.all(want_sym[:, 100:] == 0)
    assert np.all(want_u8[:, 7] == 0)                  # constant dimension
    assert np.all(want_u8[np.argmax(vecs, 0), np.arange(100)][np.arange(100) != 7] == 255)
    assert np.all(want_u8[np.argmin(vecs, 0), np.arange(100)] == 0)
    assert list(want_u8[trunc_rows, 11]) == [16, 99, 253]  # truncation just below 17, 100, 254


def _router_case(orc, fbf, vecs, labels, ts, q, metas):
    import torch
    ds = fbf.Dataset(vecs, labels, ts, repr_flags=fbf.REPR_F32 | fbf.REPR_SQ8SYM)
    try:
        qd = torch.from_numpy(np.ascontiguousarray(q)).cuda()
        ids, d = ds.search(qd, metas, k=100, mode="sq8sym", kprime=140)
        torch.cuda.synchronize()
        ids = ids.cpu().numpy().view(np.uint32)
        d = d.cpu().numpy()
        alpha, _, _ = ds.quant_params()
    finally:
        ds.close()
    L = orc.Layouts(labels.astype(np.int32), ts)
    codes = orc.sq8sym_encode(vecs, alpha)
    cb = np.ascontiguousarray(codes[L.perm_both]); ct = np.ascontiguousarray(codes[L.perm_time])
    rid, rd = orc.router_bf(vecs, L, cb, ct, alpha, q, metas, 140, 100)
    return ids, d, rid, rd, L, cb, ct, alpha


def test_sq8sym_k140_is_the_router_route_dummy(pkg, orc, dummy, golden_dir):
    _say("case: reference fixture, 100 queries, sq8sym k'=140 vs router route")
    vecs, labels, ts, q, metas = dummy
    ids, d, rid, rd, L, cb, ct, alpha = _router_case(orc, pkg.fbf, vecs, labels, ts, q, metas)
    assert np.array_equal(ids, rid)
    assert np.array_equal(d, rd)
    z = np.load(os.path.join(golden_dir, "ref_router_bf.npz"))
    same = np.array([set(a.tolist()) == set(b.tolist()) for a, b in zip(ids, z["ids"])])
    qc = orc.sq8sym_encode(q, alpha)
    tied = orc.bf_boundary_tied(L, cb, ct, qc, metas, 140)
    _, d101 = orc.router_bf(vecs, L, cb, ct, alpha, q, metas, 140, 101)
    near = np.abs(d101[:, 100] - d101[:, 99]) <= 2e-5 * np.abs(d101[:, 99])
    _say(f"  {same.sum()} of {q.shape[0]} id sets equal the reference route's; {tied.sum()} tied at 140, "
         f"{near.sum()} near-tied at 100")
    assert np.all(same | tied | near)


def test_sq8sym_k140_is_the_router_route_c0(pkg, orc):
    _say("case: config0 N=10000 nq=1000 sq8sym k'=140 vs router route")
    dg = pkg.datagen
    cfg = dg.CONFIGS[0]
    vecs, labels, ts = dg.make_base(cfg)
    q, metas = dg.make_queries(cfg)
    ids, d, rid, rd, *_ = _router_case(orc, pkg.fbf, vecs, labels, ts, q, metas)
    assert np.array_equal(ids, rid)
    assert np.array_equal(d, rd)
