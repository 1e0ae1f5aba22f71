"""Reference-shaped bruteforce() / bruteforce_subgraph() (symmetric int8) on
the GPU against the oracle and the reference's own outputs."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ref_bf(golden_dir):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return np.load(os.path.join(golden_dir, "ref_bruteforce.npz"))


def _check_against_reference(orc, codes, qc, k, got_pos, ref_heap):
    """Same set as the reference heap unless the k-th distance ties; always
    the oracle's sorted (dist, position) list."""
    sid, sd = orc.bf_sorted(codes, qc, k)
    assert np.array_equal(got_pos, sid)
    full = np.sort([orc.l2_sq8sym(qc, c) for c in codes])
    if k < len(full) and full[k - 1] == full[k]:
        strict = sd < full[k - 1]
        assert set(got_pos[strict].tolist()) <= set(ref_heap.tolist())
    else:
        assert set(got_pos.tolist()) == set(ref_heap.tolist())


def test_bruteforce_unregistered_and_registered(pkg, orc, ref_bf):
    fbf = pkg.fbf
    codes = np.ascontiguousarray(ref_bf["codes"])
    k = int(ref_bf["k"])
    alpha = float(ref_bf["alpha"])
    _say("case: bruteforce unregistered full array, 12 queries, k=140")
    for j in range(ref_bf["qcodes"].shape[0]):
        pos = fbf.bruteforce(k, codes.shape[0], codes, alpha, ref_bf["qcodes"][j])
        _check_against_reference(orc, codes, ref_bf["qcodes"][j], k, pos, ref_bf["bf_full"][j])
    _say("case: bruteforce_subgraph registered slice")
    fbf.register_codes(codes)
    try:
        s0, s1 = ref_bf["slice"]
        for j in range(ref_bf["qcodes"].shape[0]):
            pos, dist = fbf.bruteforce_subgraph(k, int(s1 - s0), codes, alpha, ref_bf["qcodes"][j],
                                                offset_rows=int(s0))
            _check_against_reference(orc, codes[s0:s1], ref_bf["qcodes"][j], k, pos, ref_bf["bf_slice"][j])
            sid, sd = orc.bf_sorted(codes[s0:s1], ref_bf["qcodes"][j], k)
            assert np.array_equal(dist, sd.astype(np.float32) * np.float32(alpha) * np.float32(alpha))
    finally:
        fbf.unregister_codes(codes)


def test_bruteforce_padding_and_small(pkg, orc, ref_bf):
    _say("case: bruteforce k > npoints pads with -1")
    fbf = pkg.fbf
    codes = np.ascontiguousarray(ref_bf["codes"][:37])
    pos = fbf.bruteforce(50, 37, codes, 1.0, ref_bf["qcodes"][0])
    sid, sd = orc.bf_sorted(codes, ref_bf["qcodes"][0], 50)
    assert np.array_equal(pos, sid)
    assert np.all(pos[37:] == -1)


def test_bruteforce_batch(pkg, orc, ref_bf):
    _say("case: batched slices of one registered array")
    fbf = pkg.fbf
    codes = np.ascontiguousarray(ref_bf["codes"])
    rng = np.random.default_rng(11)
    nq = 40
    starts = rng.integers(0, 9000, nq).astype(np.uint64)
    counts = np.minimum(rng.integers(1, 3000, nq), codes.shape[0] - starts).astype(np.uint64)
    counts[0] = 0
    qc = orc.sq8sym_encode(rng.standard_normal((nq, 100)).astype(np.float32), float(ref_bf["alpha"]))
    fbf.register_codes(codes)
    try:
        pos, dist = fbf.bruteforce_batch(codes, starts, counts, qc, 140)
    finally:
        fbf.unregister_codes(codes)
    for i in range(nq):
        s, c = int(starts[i]), int(counts[i])
        sid, sd = orc.bf_sorted(codes[s:s + c], qc[i], 140)
        assert np.array_equal(pos[i], sid), i
        valid = sid >= 0
        assert np.array_equal(dist[i][valid], sd[valid])


def test_bruteforce_many_threads(pkg, orc, ref_bf):
    """The router calls bruteforce() from its OpenMP threads at once
    (hybrid_graph.cpp:168-211): 16 host threads, each with its own stream and
    scratch, call it concurrently on slices of one registered array; every
    result must equal the oracle's."""
    import time
    from concurrent.futures import ThreadPoolExecutor
    fbf = pkg.fbf
    codes = np.ascontiguousarray(ref_bf["codes"])
    alpha = float(ref_bf["alpha"])
    rng = np.random.default_rng(5)
    n_calls = 320
    starts = rng.integers(0, codes.shape[0] - 1, n_calls)
    counts = np.minimum(rng.integers(1, 4000, n_calls), codes.shape[0] - starts)
    qc = orc.sq8sym_encode(rng.standard_normal((n_calls, 100)).astype(np.float32), alpha)
    _say(f"case: {n_calls} bruteforce calls from 16 threads, registered array")
    fbf.register_codes(codes)
    try:
        def one(i):
            return fbf.bruteforce(140, int(counts[i]), codes, alpha, qc[i], offset_rows=int(starts[i]))
        one(0)
        t0 = time.perf_counter()
        with ThreadPoolExecutor(16) as ex:
            got = list(ex.map(one, range(n_calls)))
        dt = time.perf_counter() - t0
    finally:
        fbf.unregister_codes(codes)
    _say(f"  {n_calls / dt:.0f} calls/s")
    for i in range(n_calls):
        s, c = int(starts[i]), int(counts[i])
        sid, _ = orc.bf_sorted(codes[s:s + c], qc[i], 140)
        assert np.array_equal(got[i], sid), i


def test_bruteforce_combined_mixed(pkg, orc, ref_bf):
    """Concurrent calls are combined into shared launches: calls on two
    registered arrays, with different k (1 to 300, each answered from the
    batch's largest-k list), slices shorter than k, and bruteforce_subgraph
    calls among them, from 16 threads -- every result equals the oracle's."""
    from concurrent.futures import ThreadPoolExecutor
    fbf = pkg.fbf
    codes_a = np.ascontiguousarray(ref_bf["codes"])
    rng = np.random.default_rng(17)
    codes_b = rng.integers(-127, 128, (30_000, 112), dtype=np.int8)
    alpha = float(ref_bf["alpha"])
    n_calls = 400
    which = rng.integers(0, 2, n_calls)
    ks = rng.choice([1, 7, 100, 140, 300], n_calls)
    qc = orc.sq8sym_encode(rng.standard_normal((n_calls, 100)).astype(np.float32), alpha)
    arrs = (codes_a, codes_b)
    starts = np.array([rng.integers(0, arrs[w].shape[0] - 1) for w in which])
    counts = np.array([min(int(rng.integers(1, 6000)), arrs[w].shape[0] - int(s)) for w, s in zip(which, starts)])
    counts[::17] = np.minimum(counts[::17], 5)  # shorter than k
    sub = rng.random(n_calls) < 0.3
    _say(f"case: {n_calls} combined calls from 16 threads, 2 registered arrays, k in 1..300")
    fbf.register_codes(codes_a)
    fbf.register_codes(codes_b)
    try:
        def one(i):
            arr = arrs[which[i]]
            if sub[i]:
                return fbf.bruteforce_subgraph(int(ks[i]), int(counts[i]), arr, alpha, qc[i], offset_rows=int(starts[i]))
            return fbf.bruteforce(int(ks[i]), int(counts[i]), arr, alpha, qc[i], offset_rows=int(starts[i])), None
        with ThreadPoolExecutor(16) as ex:
            got = list(ex.map(one, range(n_calls)))
    finally:
        fbf.unregister_codes(codes_a)
        fbf.unregister_codes(codes_b)
    for i in range(n_calls):
        s, c, k = int(starts[i]), int(counts[i]), int(ks[i])
        sid, sd = orc.bf_sorted(arrs[which[i]][s:s + c], qc[i], k)
        pos, dist = got[i]
        assert np.array_equal(pos, sid), (i, k, c)
        if dist is not None:
            valid = sid >= 0
            assert np.array_equal(dist[valid], sd[valid].astype(np.float32) * np.float32(alpha) * np.float32(alpha)), i
