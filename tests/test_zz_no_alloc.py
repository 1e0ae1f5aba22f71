"""The allocation-free entry point (fbf_bruteforce_batch_noalloc), through the
C-ABI, against fbf_bruteforce_batch on the same inputs: two calls in a row
(the second smaller, with another k, so stale contents of the shared scratch
would show), a call on a side stream, and the host-side rejection of a batch
larger than the scratch."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def codes(golden_dir, pkg):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    z = np.load(os.path.join(golden_dir, "ref_bruteforce.npz"))
    c = np.ascontiguousarray(z["codes"])
    pkg.fbf.register_codes(c)
    yield c, float(z["alpha"])
    pkg.fbf.unregister_codes(c)


def _slices(rng, n_rows, nq, max_len):
    starts = rng.integers(0, n_rows - 1, nq).astype(np.uint64)
    counts = np.minimum(rng.integers(0, max_len, nq), n_rows - starts).astype(np.uint64)
    return starts, counts


def test_noalloc_equals_batch_two_calls(pkg, orc, codes):
    fbf = pkg.fbf
    c, alpha = codes
    rng = np.random.default_rng(23)
    for call, (nq, k, max_len) in enumerate(((48, 140, 9000), (9, 100, 2500))):
        starts, counts = _slices(rng, c.shape[0], nq, max_len)
        counts[0] = 0
        qc = orc.sq8sym_encode(rng.standard_normal((nq, 100)).astype(np.float32), alpha)
        _say(f"case: call {call}: nq={nq} k={k} slices up to {max_len} rows")
        pos, dist = fbf.bruteforce_batch_noalloc(c, starts, counts, qc, k)
        ref_pos, ref_dist = fbf.bruteforce_batch(c, starts, counts, qc, k)
        assert np.array_equal(pos, ref_pos), call
        assert np.array_equal(dist, ref_dist), call
        for i in (1, nq - 1):
            s, n = int(starts[i]), int(counts[i])
            sid, sd = orc.bf_sorted(c[s:s + n], qc[i], k)
            assert np.array_equal(pos[i], sid)


def test_noalloc_side_stream_and_limit(pkg, orc, codes):
    import torch
    fbf = pkg.fbf
    c, alpha = codes
    rng = np.random.default_rng(29)
    starts, counts = _slices(rng, c.shape[0], 20, 6000)
    qc = orc.sq8sym_encode(rng.standard_normal((20, 100)).astype(np.float32), alpha)
    _say("case: noalloc on a torch side stream")
    side = torch.cuda.Stream()
    pos, dist = fbf.bruteforce_batch_noalloc(c, starts, counts, qc, 140, stream=side)
    ref = fbf.bruteforce_batch(c, starts, counts, qc, 140)
    assert np.array_equal(pos, ref[0]) and np.array_equal(dist, ref[1])
    _say("case: noalloc batch beyond the scratch is rejected on the host")
    big = 8000  # 8000 x 1024-key lists alone need 64 MiB
    st = np.zeros(big, np.uint64)
    cn = np.full(big, c.shape[0], np.uint64)
    qb = np.repeat(qc[:1], big, 0)
    with pytest.raises(fbf.FbfError, match="scratch"):
        fbf.bruteforce_batch_noalloc(c, st, cn, qb, 896)
    # and the next call still works
    pos2, _ = fbf.bruteforce_batch_noalloc(c, starts[:3], counts[:3], qc[:3], 140)
    assert np.array_equal(pos2, ref[0][:3])
