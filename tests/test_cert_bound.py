"""The certificate of FBF_MODE_EXACT_BF16, checked on the CPU.

1. Its rounding-error bound: a BF16 score as the MFMA scan forms it --
   |x|^2 (FP32, sequential FMA) split into hi + lo bf16, rows and queries
   rounded to bf16 (round to nearest even), -2 bf16(q) . bf16(x) plus hi and lo
   summed in FP32 in any order -- lies within |q||x| 0.00784 + |x|^2 1.7e-5 of
   |x|^2 - 2 q.x.  Checked for several summation orders on random and on
   adversarial inputs (aligned vectors, values on bf16 rounding midpoints).
2. The certificate itself: on random batches with a small k', every query the
   test certifies has the same top-k as a brute-force search over all rows
   with canonical FP32 distances (the exact mode's), ties to the smaller id.
"""
import numpy as np

BOUND_QX, BOUND_XX = 0.00784, 1.7e-5


def bf16(a):
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def fma_seq_norm(x):
    """|x|^2 per row by sequential FP32 FMA (the dataset build's row norms)."""
    acc = np.zeros(x.shape[0], np.float32)
    for d in range(x.shape[1]):
        acc = (x[:, d].astype(np.float64) ** 2 + acc).astype(np.float32)
    return acc


def bf16_terms(x, q):
    """The 102 exact FP32 products the BF16 tile sums per row."""
    n = fma_seq_norm(x)
    hi = bf16(n)
    lo = bf16((n - hi).astype(np.float32))
    prod = bf16(x).astype(np.float64) * (-2.0 * bf16(q).astype(np.float64))[None, :]
    return np.concatenate([prod, hi[:, None].astype(np.float64), lo[:, None].astype(np.float64)], axis=1)


def sum_f32(t, order):
    """FP32 sum of each row of t (exact float64 terms) in a given order."""
    t = t[:, order]
    acc = np.zeros(t.shape[0], np.float32)
    for j in range(t.shape[1]):
        acc = (acc.astype(np.float64) + t[:, j]).astype(np.float32)
    return acc


def sum_f32_pairwise(t):
    v = [t[:, j].astype(np.float32) for j in range(t.shape[1])]
    while len(v) > 1:
        nxt = [(v[i].astype(np.float64) + v[i + 1]).astype(np.float32) for i in range(0, len(v) - 1, 2)]
        if len(v) % 2:
            nxt.append(v[-1])
        v = nxt
    return v[0]


def check_bound(x, q):
    t = bf16_terms(x, q)
    xr, qr = x.astype(np.float64), q.astype(np.float64)
    real = (xr * xr).sum(1) - 2.0 * xr @ qr
    nx = np.sqrt((xr * xr).sum(1))
    bound = BOUND_QX * np.linalg.norm(qr) * nx + BOUND_XX * nx * nx
    rng = np.random.default_rng(3)
    orders = [np.arange(102), np.arange(102)[::-1], rng.permutation(102)]
    worst = 0.0
    for s in [sum_f32(t, o) for o in orders] + [sum_f32_pairwise(t)]:
        err = np.abs(s.astype(np.float64) - real)
        assert np.all(err <= bound), float(np.max(err / bound))
        worst = max(worst, float(np.max(err / bound)))
    return worst


def test_bound_random():
    rng = np.random.default_rng(0)
    x = (rng.standard_normal((4000, 100)) * rng.uniform(0.01, 30, (4000, 1))).astype(np.float32)
    q = (rng.standard_normal(100) * 7).astype(np.float32)
    assert check_bound(x, q) < 1.0


def test_bound_adversarial():
    rng = np.random.default_rng(1)
    # Rows aligned with the query (every product positive), values on bf16
    # rounding midpoints (low 16 bits 0x8000: the worst relative rounding).
    base = np.abs(rng.standard_normal(100)).astype(np.float32) + 1.0
    mid = (base.view(np.uint32) & np.uint32(0xFFFF0000)) | np.uint32(0x8000)
    q = mid.view(np.float32)
    x = np.repeat(q[None, :], 64, axis=0) * rng.uniform(0.5, 2.0, (64, 1)).astype(np.float32)
    xm = (x.view(np.uint32) & np.uint32(0xFFFF0000)) | np.uint32(0x7FFF)
    x = np.concatenate([x, xm.view(np.float32), -x], axis=0)
    worst = check_bound(x, q)
    assert 0.2 < worst < 1.0  # the bound is not loose by orders of magnitude here


def canon_dist(x, q):
    """Canonical FP32 distance: t = x - q (FP32), acc = fma(t, t, acc) over d in order."""
    acc = np.zeros(x.shape[0], np.float32)
    for d in range(x.shape[1]):
        t = (x[:, d] - q[d]).astype(np.float32).astype(np.float64)
        acc = (t * t + acc).astype(np.float32)
    return acc


def test_certificate_implies_exact_topk():
    rng = np.random.default_rng(2)
    n, k, kp = 6000, 10, 24
    centers = rng.standard_normal((16, 100)).astype(np.float32)
    x = (centers[rng.integers(0, 16, n)] + 0.3 * rng.standard_normal((n, 100))).astype(np.float32)
    xm = np.float32(np.sqrt(float(fma_seq_norm(x).max())) * (1 + 1e-5))
    n_cert = 0
    for qi in range(40):
        q = (centers[qi % 16] + 0.3 * rng.standard_normal(100)).astype(np.float32)
        s = sum_f32(bf16_terms(x, q), np.arange(102))
        order = np.lexsort((np.arange(n), s))           # by (score, position)
        cand = order[:kp]
        ts = s[order[kp - 1]]
        d = canon_dist(x, q)
        cd = d[cand]
        o2 = np.lexsort((cand, cd))
        dk = cd[o2[k - 1]]
        qn = np.float32(0)
        for v in q:
            qn = np.float32(np.float64(v) * v + qn)
        eps = np.float32(1.01) * (np.float32(np.sqrt(qn)) * xm * np.float32(BOUND_QX) + xm * xm * np.float32(BOUND_XX))
        lower = (ts + qn * np.float32(1 - 2e-5) - eps) * np.float32(1 - 3e-5)
        if lower > dk:
            n_cert += 1
            full = np.lexsort((np.arange(n), d))[:k]
            assert np.array_equal(cand[o2[:k]], full), qi
    assert 0 < n_cert < 40  # both outcomes occur at this k'
