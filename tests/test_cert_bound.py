"""The certificate of FBF_MODE_EXACT_BF16, checked on the CPU.

1. Its rounding-error bound.  The BF16 tile forms a row's score as |x|^2
   (FP32, sequential FMA) split into hi + lo bf16, plus sum_i bf16(x_i)
   (-2 bf16(q_i)), summed in FP32.  With qb = bf16(q), xb = bf16(x) and
   rho_x = |xb - x| / |x|, the bound the merge kernel applies is
       2 |x| (|qb - q| (1 + rho) + |q| rho) + 1.4e-5 (2.02 |q||x| + 1.01 |x|^2)
       + 2.2e-5 |x|^2
   (the kernel uses the dataset's largest rho and |x|, so checking it per row
   with the row's own rho_x and |x| is the stronger statement).  Checked for
   several FP32 summation orders on random inputs and on the worst case of
   bf16 rounding: aligned vectors whose every element lies on the bf16 tie
   just above a power of two (relative error 2^-8 in both operands, all of
   one sign), where the error comes within 10% of the bound (and exceeds the
   constant a 2^-9 unit roundoff would give).
2. The certificate itself: on random batches with a small k', every query the
   test certifies has the same top-k as a brute-force search over all rows
   with canonical FP32 distances (the exact mode's), ties to the smaller id.
"""
import numpy as np


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


def row_bound(x, q):
    xr, qr = x.astype(np.float64), q.astype(np.float64)
    nx = np.sqrt((xr * xr).sum(1))
    rho = np.sqrt(((bf16(x).astype(np.float64) - xr) ** 2).sum(1)) / nx
    nq = np.linalg.norm(qr)
    dq = np.linalg.norm(bf16(q).astype(np.float64) - qr)
    return 2 * nx * (dq * (1 + rho) + nq * rho) + 1.4e-5 * (2.02 * nq * nx + 1.01 * nx * nx) + 2.2e-5 * nx * nx


def check_bound(x, q):
    t = bf16_terms(x, q)
    xr, qr = x.astype(np.float64), q.astype(np.float64)
    real = (xr * xr).sum(1) - 2.0 * xr @ qr
    bound = row_bound(x, q)
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


def test_bound_worst_case_rounding():
    rng = np.random.default_rng(1)
    # Every element 2^e (1 + 2^-8): a bf16 tie, rounded to even (down) by a
    # relative 2^-8, the largest relative rounding error of bf16.
    e = rng.integers(-3, 4, 100)
    q = (np.ldexp(1.0 + 2.0 ** -8, e)).astype(np.float32)
    assert np.all(bf16(q) < q)
    x = np.repeat(q[None, :], 8, axis=0) * np.ldexp(1.0, rng.integers(-2, 3, (8, 1))).astype(np.float32)
    worst = check_bound(x.astype(np.float32), q)
    assert worst > 0.9  # the bound is tight here: rounding errors all of one sign
    # The constant a 2^-9 unit roundoff would give falls short on these rows.
    t = bf16_terms(x.astype(np.float32), q)
    xr, qr = x.astype(np.float64), q.astype(np.float64)
    err = np.abs(sum_f32(t, np.arange(102)).astype(np.float64) - ((xr * xr).sum(1) - 2 * xr @ qr))
    naive = 0.00784 * np.linalg.norm(qr) * np.sqrt((xr * xr).sum(1)) + 1.7e-5 * (xr * xr).sum(1)
    assert np.any(err > naive)


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
    f32 = np.float32
    xm = f32(np.sqrt(float(fma_seq_norm(x).max())) * (1 + 1e-5))
    xr = x.astype(np.float64)
    rho = f32(np.max(np.sqrt(((bf16(x).astype(np.float64) - xr) ** 2).sum(1) / (xr * xr).sum(1))) * 1.001 + 1e-7)
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
        qn = f32(0)
        dq2 = f32(0)
        for v in q:
            qn = f32(np.float64(v) * v + qn)
            dv = np.float64(bf16(np.array([v]))[0]) - v
            dq2 = f32(dv * dv + dq2)
        qa, dq = f32(np.sqrt(qn) * 1.0001), f32(np.sqrt(dq2) * 1.0001)
        eps = f32(1.01) * (f32(2) * xm * (dq * (f32(1) + rho) + qa * rho)
                           + f32(1.4e-5) * (f32(2.02) * qa * xm + f32(1.01) * xm * xm) + f32(2.2e-5) * xm * xm)
        lower = (ts + qn * f32(1 - 2e-5) - eps) * f32(1 - 3e-5)
        if lower > dk:
            n_cert += 1
            full = np.lexsort((np.arange(n), d))[:k]
            assert np.array_equal(cand[o2[:k]], full), qi
    assert 0 < n_cert < 40  # both outcomes occur at this k'
