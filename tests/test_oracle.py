"""Pins the CPU oracle (oracle/fbf_oracle.c) before it is trusted as the
checker of the HIP kernels: golden vectors from the reference's own code and
files, independent algorithms, closed forms and invariants."""
import json
import os

import numpy as np
import pytest


def _canon_emulated(x, q):
    """Canonical order emulated in float64: t = fl32(x - q) (exact in f64, then
    rounded), acc = fl32(t*t + acc) with t*t exact in f64 -- an FMA with one
    final rounding (double rounding is possible only in astronomically rare
    ties)."""
    acc = np.float32(0)
    for xi, qi in zip(x, q):
        t = np.float32(np.float64(xi) - np.float64(qi))
        acc = np.float32(np.float64(t) * np.float64(t) + np.float64(acc))
    return acc


def test_kat_int8_l2(orc, golden_dir):
    kat = json.load(open(os.path.join(golden_dir, "kat.json")))
    x = np.zeros(112, np.int8); y = np.zeros(112, np.int8)
    x[:32] = kat["a0"]; y[:32] = kat["a1"]
    assert kat["l2_sq8sym"] == 210929  # written by the reference's own L2SqrSQ8_sym
    assert orc.l2_sq8sym(x, y) == 210929
    # closed form
    assert int(((x.astype(np.int64) - y) ** 2).sum()) == 210929


def test_int8_l2_closed_form_and_bounds(orc):
    rng = np.random.default_rng(1)
    for _ in range(50):
        x = rng.integers(-127, 128, 112).astype(np.int8)
        y = rng.integers(-127, 128, 112).astype(np.int8)
        v = orc.l2_sq8sym(x, y)
        assert v == int(((x.astype(np.int64) - y) ** 2).sum())
        assert 0 <= v <= 112 * 254 ** 2
    x = np.full(112, 127, np.int8); y = np.full(112, -127, np.int8)
    assert orc.l2_sq8sym(x, y) == 112 * 254 ** 2


def test_canonical_l2_order_and_fma(orc):
    rng = np.random.default_rng(2)
    for _ in range(200):
        x = rng.standard_normal(100).astype(np.float32) * 3
        q = rng.standard_normal(100).astype(np.float32) * 3
        c = orc.l2_canon(x, q)
        assert np.float32(c) == _canon_emulated(x, q)
        f = orc.l2_f64(x, q)
        assert abs(c - f) <= 1e-5 * f
    # A dropped, reordered or unfused term would change these bits.
    x = np.array([1e8, 1.0, -1e8], np.float32); q = np.zeros(3, np.float32)
    assert orc.l2_canon(x, q) == _canon_emulated(x, q)
    assert orc.l2_canon(np.array([3.0], np.float32), np.array([1.0], np.float32)) == 4.0


def test_normal_l2_matches_reference(orc):
    """The reference's normal_l2 / L2SqrRef bits depend on its compiler (g++
    -O3 -mfma vectorises the subtractions and contracts the tail into FMAs),
    so the pin is closeness to FP64, for the reference and for every oracle
    form."""
    R = orc.ref()
    if R is None:
        pytest.skip("reference kernels not built here")
    rng = np.random.default_rng(3)
    for _ in range(50):
        x = rng.standard_normal(100).astype(np.float32)
        q = rng.standard_normal(100).astype(np.float32)
        f = orc.l2_f64(x, q)
        for v in (orc.l2_normal(x, q), orc.l2_canon(x, q), R.ref_normal_l2(orc._p(x), orc._p(q), 100),
                  R.ref_l2_ref(orc._p(x), orc._p(q), 100)):
            assert abs(v - f) <= 1e-5 * f


def test_sq8sym_quantiser(orc):
    x = np.array([[0.5, -1.0, 0.25, 1.0 / 254]], np.float32)
    a = orc.sq8sym_train(x)
    assert a == 1.0
    c = orc.sq8sym_encode(x, a)[0]
    assert c[1] == -127 and c[0] == 64 and c[2] == 32   # 63.5 -> 64, 31.75 -> 32
    assert c[3] == 1                                      # 0.5 -> 1 (half away from zero)
    assert np.all(c[4:] == 0)                             # zero padding to 112
    # half away from zero on negatives
    assert orc.sq8sym_encode(np.array([[-1.0 / 254]], np.float32), 1.0)[0][0] == -1


def test_sq8sym_matches_reference(orc):
    R = orc.ref()
    if R is None:
        pytest.skip("reference kernels not built here")
    rng = np.random.default_rng(4)
    x = rng.standard_normal((500, 100)).astype(np.float32)
    codes = np.zeros((500, 112), np.int8)
    alpha = R.ref_sq8sym_train_encode(orc._p(x), 500, 100, orc._p(codes))
    assert alpha == orc.sq8sym_train(x)
    assert np.array_equal(codes, orc.sq8sym_encode(x, alpha))
    absmax = np.unravel_index(np.argmax(np.abs(x)), x.shape)
    assert abs(int(codes[absmax])) == 127


def test_sq8_perdim_quantiser(orc):
    x = np.array([[0.0, 1.0, 5.0], [1.0, 3.0, 5.0], [0.5, 2.0, 5.0]], np.float32)
    mi, dif = orc.sq8_train(x)
    assert np.array_equal(mi, [0, 1, 5]) and np.array_equal(dif, [1, 2, 0])
    c = orc.sq8_encode(x, mi, dif)
    assert c[0, 0] == 0 and c[1, 0] == 255          # x = mi -> 0, x = mx -> 255
    assert c[2, 0] == 127 and c[2, 1] == 127        # 127.5 truncates
    assert np.all(c[:, 2] == 0)                     # constant dimension -> 0


def test_sq8_perdim_matches_reference(orc):
    R = orc.ref()
    if R is None:
        pytest.skip("reference kernels not built here")
    rng = np.random.default_rng(5)
    x = rng.standard_normal((300, 100)).astype(np.float32)
    codes = np.zeros((300, 112), np.uint8)
    mi = np.zeros(112, np.float32); dif = np.zeros(112, np.float32)
    R.ref_sq8_train_encode(orc._p(x), 300, 100, orc._p(mi), orc._p(dif), orc._p(codes))
    mi2, dif2 = orc.sq8_train(x)
    assert np.array_equal(mi[:100], mi2) and np.array_equal(dif[:100], dif2)
    assert np.array_equal(codes[:, :100], orc.sq8_encode(x, mi2, dif2))
    # L2SqrSQ8_ext (reference, FP32 AVX-512) against the FP64 closed form
    qx = np.zeros(112, np.float32); qx[:100] = rng.standard_normal(100)
    for i in range(20):
        r = R.ref_l2_sq8ext(orc._p(qx), orc._p(np.ascontiguousarray(codes[i])), orc._p(mi), orc._p(dif))
        f = orc.l2_sq8ext_f64(qx[:100], codes[i, :100], mi[:100], dif[:100])
        assert abs(r - f) <= 1e-5 * f


def test_topk_two_algorithms_with_ties(orc):
    rng = np.random.default_rng(6)
    for n, k in [(0, 5), (3, 5), (100, 1), (1000, 100), (5000, 400)]:
        d = rng.integers(0, 50, n).astype(np.float32)          # heavy ties
        ids = rng.permutation(n).astype(np.uint32)
        keys = orc.make_keys(d, ids)
        h = orc.topk_heap(keys, k)
        f = orc.topk_fullsort(keys, k)
        assert np.array_equal(h, f)
        o = np.lexsort((ids, d))[:k]                             # numpy: by dist then id
        assert np.array_equal(h & 0xFFFFFFFF, ids[o].astype(np.uint64))


def test_key_order(orc):
    v = np.array([-np.inf, -2.5, -0.0, 0.0, 1e-30, 1.0, 3.0, np.inf], np.float32)
    o = orc.f2ord(v)
    assert np.all(np.diff(o.astype(np.int64)) >= 0)
    assert o[2] < o[3]  # -0 before +0


def test_ranges_equal_predicate(orc, dummy):
    vecs, labels, ts, q, metas = dummy
    lab = labels.astype(np.int32)
    L = orc.Layouts(lab, ts)
    extra = np.array([[1, 999, -1, -1], [2, -1, 0.5, 0.4], [3, 0, ts[5], ts[5]], [2, -1, ts[7], ts[7]],
                      [3, -7, 0, 1], [2, -1, -1.0, 2.0]], np.float32)
    allm = np.concatenate([metas, extra])
    for m in allm:
        lay, s, e = L.query_range(*m)
        perm = L.perm_both if lay == 0 else L.perm_time
        got = np.sort(perm[s:e])
        want = np.array([j for j in range(len(lab)) if orc.match(m[0], m[1], m[2], m[3], lab[j], ts[j])],
                        np.uint32)
        assert np.array_equal(got, want), m


def test_exact_ranges_equal_predicate(orc, dummy):
    vecs, labels, ts, q, metas = dummy
    lab = labels.astype(np.int32)
    L = orc.Layouts(lab, ts)
    i1, d1 = orc.exact_predicate(vecs, lab, ts, q, metas, k=100)
    i2, d2 = orc.exact_ranges(vecs, L, q, metas, k=100)
    assert np.array_equal(i1, i2)
    assert np.array_equal(d1.astype(np.float32), d2)


def test_exact_ranges_f64_equal_predicate(orc, dummy):
    """The FP64 range search (recall reference of the bench and the full-scale
    tests) equals the FP64 predicate scan on every fixture query."""
    vecs, labels, ts, q, metas = dummy
    lab = labels.astype(np.int32)
    L = orc.Layouts(lab, ts)
    vb = np.ascontiguousarray(vecs[L.perm_both]); vt = np.ascontiguousarray(vecs[L.perm_time])
    ids, d = orc.exact_ranges_f64_prepared(vb, vt, L, q, metas, 100)
    pid, pd = orc.exact_predicate(vecs, lab, ts, q, metas, k=100, mode="f64")
    assert np.array_equal(ids, pid) and np.array_equal(d, pd)
    # every returned distance is the FP64 distance of its row
    for i in range(0, q.shape[0], 7):
        for j in range(0, 100, 13):
            if ids[i, j] != 0xFFFFFFFF:
                assert d[i, j] == orc.l2_f64(vecs[ids[i, j]], q[i])


def test_l2_canon_rows(orc, dummy):
    vecs, labels, ts, q, metas = dummy
    rng = np.random.default_rng(3)
    ids = rng.integers(0, vecs.shape[0], (20, 9)).astype(np.uint32)
    ids[3, 4] = 0xFFFFFFFF
    out = orc.l2_canon_rows(vecs, q[:20], ids)
    for i in range(20):
        for j in range(9):
            want = np.inf if ids[i, j] == 0xFFFFFFFF else np.float32(orc.l2_canon(vecs[ids[i, j]], q[i]))
            assert out[i, j] == want


def test_golden_reference_ground_truth(orc, pkg, dummy, golden_dir):
    """The reference's Bruteforce output (its code, its normal_l2) equals the
    oracle's canonical exact search on the reference's fixture files."""
    vecs, labels, ts, q, metas = dummy
    gt = pkg.datagen.read_gt_bin(os.path.join(golden_dir, "dummy_gt_ref.bin"))
    ids, d = orc.exact_predicate(vecs, labels.astype(np.int32), ts, q, metas, k=100)
    assert np.array_equal(ids, gt)
    z = np.load(os.path.join(golden_dir, "dummy_exact.npz"))
    assert np.array_equal(ids, z["ids"]) and np.array_equal(d.astype(np.float32), z["dists"])
    i64, d64 = orc.exact_predicate(vecs, labels.astype(np.int32), ts, q, metas, k=100, mode="f64")
    assert np.array_equal(i64, z["ids_f64"])
    assert np.all(np.abs(d - d64) <= 1e-5 * d64)  # several queries are base rows: d = 0


def test_bf_heap_restatement_matches_reference(orc, golden_dir):
    """orc_bf_heap_ref restates maxPQIFCS: identical positions, in the reference's
    heap-array order, to the reference bruteforce()/bruteforce_subgraph()."""
    z = np.load(os.path.join(golden_dir, "ref_bruteforce.npz"))
    codes, k, alpha = z["codes"], int(z["k"]), float(z["alpha"])
    s0, s1 = z["slice"]
    for j in range(z["qcodes"].shape[0]):
        ids, dd = orc.bf_heap_ref(codes, z["qcodes"][j], k)
        assert np.array_equal(ids, z["bf_full"][j])
        ids, dd = orc.bf_heap_ref(codes[s0:s1], z["qcodes"][j], k)
        assert np.array_equal(ids, z["bf_slice"][j])
        assert np.array_equal((dd.astype(np.float32) * np.float32(alpha) * np.float32(alpha)),
                              z["bf_slice_dist"][j])
        # The sorted contract selects the same set unless the k-th distance ties.
        sid, sd = orc.bf_sorted(codes, z["qcodes"][j], k)
        full = np.sort([orc.l2_sq8sym(z["qcodes"][j], c) for c in codes])
        if full[k - 1] != full[k]:
            assert set(sid.tolist()) == set(z["bf_full"][j].tolist())
        else:
            strict = sd < full[k - 1]
            assert set(sid[strict].tolist()) <= set(z["bf_full"][j].tolist())


def test_recall_closed_form(orc):
    gt = np.arange(200, dtype=np.uint32).reshape(2, 100)
    knn = gt.copy()
    knn[0, :10] = 1000 + np.arange(10)
    knn[1] = knn[1][::-1]
    r, per = orc.recall(knn, gt)
    assert per.tolist() == [90, 100]
    assert r == pytest.approx(0.95)


def test_router_bf_port_vs_reference(orc, dummy):
    """The CPU port of the router's brute-force branch selects the same ids as
    the reference kernels driven the same way (set equality per query)."""
    R = orc.ref()
    if R is None:
        pytest.skip("reference kernels not built here")
    vecs, labels, ts, q, metas = dummy
    lab = labels.astype(np.int32)
    L = orc.Layouts(lab, ts)
    alpha = orc.sq8sym_train(vecs)
    codes = orc.sq8sym_encode(vecs, alpha)
    cb = np.ascontiguousarray(codes[L.perm_both]); ct = np.ascontiguousarray(codes[L.perm_time])
    qc = orc.sq8sym_encode(q, alpha)
    rng = L.ranges(metas)
    out = np.zeros((q.shape[0], 100), np.uint32)
    R.ref_router_bf(orc._p(cb), orc._p(ct), orc._p(L.perm_both), orc._p(L.perm_time), orc._p(vecs), 100,
                    alpha, orc._p(q), orc._p(qc), orc._p(rng), q.shape[0], 140, 100, orc._p(out))
    ids = np.zeros((q.shape[0], 100), np.uint32); dd = np.zeros((q.shape[0], 100), np.float32)
    orc.lib().orc_router_bf(orc._p(cb), orc._p(ct), orc._p(L.perm_both), orc._p(L.lab_b), orc._p(L.ts_b),
                            orc._p(L.perm_time), orc._p(L.ts_t), orc._p(vecs), vecs.shape[0], 100, alpha,
                            orc._p(q), orc._p(metas), q.shape[0], 140, 100, orc._p(ids), orc._p(dd))
    same = np.array([set(a.tolist()) == set(b.tolist()) for a, b in zip(ids, out)])
    # Only a tie at the 140th int distance can change the candidate set, or
    # a near tie at the 100th FP32 distance the rerank order (L2SqrRef's
    # compiler-chosen summation against the canonical order).
    tied = orc.bf_boundary_tied(L, cb, ct, qc, metas, 140)
    ids101, d101 = orc.router_bf(vecs, L, cb, ct, alpha, q, metas, 140, 101)
    near = np.abs(d101[:, 100] - d101[:, 99]) <= 2e-5 * np.abs(d101[:, 99])
    assert np.all(same | tied | near), np.nonzero(~(same | tied | near))[0]
    assert same.mean() >= 0.95


def test_router_bf_golden(orc, dummy, golden_dir):
    """The same check against the committed outputs of the reference route
    (tests/golden/ref_router_bf.npz, written by the reference kernels)."""
    z = np.load(os.path.join(golden_dir, "ref_router_bf.npz"))
    vecs, labels, ts, q, metas = dummy
    L = orc.Layouts(labels.astype(np.int32), ts)
    assert np.array_equal(L.ranges(metas), z["ranges"])
    alpha = orc.sq8sym_train(vecs)
    assert np.float32(alpha) == z["alpha"]
    codes = orc.sq8sym_encode(vecs, alpha)
    cb = np.ascontiguousarray(codes[L.perm_both]); ct = np.ascontiguousarray(codes[L.perm_time])
    qc = orc.sq8sym_encode(q, alpha)
    ids, _ = orc.router_bf(vecs, L, cb, ct, alpha, q, metas, 140, 100)
    _, d101 = orc.router_bf(vecs, L, cb, ct, alpha, q, metas, 140, 101)
    same = np.array([set(a.tolist()) == set(b.tolist()) for a, b in zip(ids, z["ids"])])
    tied = orc.bf_boundary_tied(L, cb, ct, qc, metas, 140)
    near = np.abs(d101[:, 100] - d101[:, 99]) <= 2e-5 * np.abs(d101[:, 99])
    assert np.all(same | tied | near), np.nonzero(~(same | tied | near))[0]
    # A plausible slip (ties to the larger position, a dropped rerank) fails here.
    assert same.sum() >= q.shape[0] - (tied | near).sum()
