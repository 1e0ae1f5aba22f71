"""Lifecycle of the library's device state.

* A search after fbf_search_host (which creates, uses and destroys its own
  stream) on another stream: the dataset's shared scratch hands over through
  an event, never through the destroyed stream's handle.
* k at the list limit (896) of the reference-shaped bruteforce() equals the
  oracle; the C-ABI rejects larger k on the host (tests/test_abi.py).
* A process that uses a dataset, the per-thread engines of bruteforce() (from
  the main thread and from a worker thread) and a captured graph, and then
  exits without releasing anything, exits with status 0: no destructor or
  finaliser calls HIP while the runtime tears down.
"""
import os
import subprocess
import sys
import textwrap

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_search_after_search_host_on_side_stream(pkg, orc):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    fbf, dg = pkg.fbf, pkg.datagen
    cfg = dg.CONFIGS[0]
    vecs, labels, ts = dg.make_base(cfg, n=30_000)
    q, metas = dg.make_queries(cfg, nq=200)
    _say("case: N=30000 nq=200 search_host, then search on a torch side stream, both modes")
    ds = fbf.Dataset(vecs, labels, ts, repr_flags=fbf.REPR_F32 | fbf.REPR_BF16)
    try:
        eid, ed = orc.exact_predicate(vecs, labels.astype(np.int32), ts, q, metas, k=100, mode="canon")
        for mode in ("exact", "bf16"):
            hi, hd = ds.search_host(q, metas, k=100, mode=mode)
            side = torch.cuda.Stream()
            qd = torch.from_numpy(q).cuda()
            torch.cuda.synchronize()
            with torch.cuda.stream(side):
                ids, d = ds.search(qd, metas, k=100, mode=mode, stream=side)
            side.synchronize()
            ids = ids.cpu().numpy().view(np.uint32)
            assert np.array_equal(ids, hi) and np.array_equal(d.cpu().numpy(), hd), mode
            if mode == "exact":
                assert np.array_equal(ids, eid)
            # and back on the default stream after the side stream is gone
            del side
            ids2, _ = ds.search(qd, metas, k=100, mode=mode)
            torch.cuda.synchronize()
            assert np.array_equal(ids2.cpu().numpy().view(np.uint32), hi), mode
    finally:
        ds.close()


def test_bruteforce_k_at_list_limit(pkg, orc, golden_dir):
    fbf = pkg.fbf
    z = np.load(os.path.join(golden_dir, "ref_bruteforce.npz"))
    codes = np.ascontiguousarray(z["codes"][:3000])
    _say("case: bruteforce k=896 over 3000 codes, registered and not")
    for j in range(3):
        qc = z["qcodes"][j]
        pos = fbf.bruteforce(896, codes.shape[0], codes, 1.0, qc)
        sid, _ = orc.bf_sorted(codes, qc, 896)
        assert np.array_equal(pos, sid), j
    fbf.register_codes(codes)
    try:
        starts = np.array([0, 100, 2000], np.uint64)
        counts = np.array([3000, 1500, 1000], np.uint64)
        pos, dist = fbf.bruteforce_batch(codes, starts, counts, z["qcodes"][:3], 896)
        for i in range(3):
            s, c = int(starts[i]), int(counts[i])
            sid, sd = orc.bf_sorted(codes[s:s + c], z["qcodes"][i], 896)
            assert np.array_equal(pos[i], sid), i
        with pytest.raises(fbf.FbfError):
            fbf.bruteforce_batch(codes, starts, counts, z["qcodes"][:3], 897)
    finally:
        fbf.unregister_codes(codes)


_CHILD = textwrap.dedent("""
    import sys, threading
    import numpy as np
    import torch
    sys.path.insert(0, {root!r})
    import __graft_entry__ as ge
    pkg = ge.load_pkg()
    fbf, dg = pkg.fbf, pkg.datagen
    cfg = dg.CONFIGS[0]
    vecs, labels, ts = dg.make_base(cfg, n=20000)
    q, metas = dg.make_queries(cfg, nq=64)
    ds = fbf.Dataset(vecs, labels, ts, repr_flags=fbf.REPR_F32 | fbf.REPR_BF16 | fbf.REPR_SQ8SYM)
    qd = torch.from_numpy(q).cuda()
    ds.search(qd, metas, k=100, mode="bf16")
    g = fbf.SearchGraph(ds, qd, metas, k=100, mode="bf16")
    g.launch()
    codes = np.zeros((5000, 112), np.int8)
    codes[:, :100] = np.random.default_rng(1).integers(-127, 128, (5000, 100))
    fbf.register_codes(codes)
    out = []
    def worker():
        out.append(fbf.bruteforce(140, 4000, codes, 1.0, codes[7], offset_rows=500))
    t = threading.Thread(target=worker); t.start(); t.join()
    out.append(fbf.bruteforce(140, 5000, codes, 1.0, codes[9]))
    torch.cuda.synchronize()
    assert out[0].size == 140 and out[1][0] == 9  # the query is row 9 itself: distance 0
    print("child done", flush=True)
    # exit with the dataset, graph, registered codes and both engines alive
    """)


@pytest.mark.parametrize("release", [False, True])
def test_process_exit_is_clean(release):
    script = _CHILD.format(root=ROOT)
    if release:
        script += "g.close(); ds.close(); fbf.unregister_codes(codes)\n"
    _say(f"case: child process exits (release={release})")
    r = subprocess.run([sys.executable, "-c", script], cwd=ROOT, capture_output=True, text=True, timeout=110)
    assert "child done" in r.stdout, r.stdout + r.stderr
    assert r.returncode == 0, f"rc={r.returncode}\n{r.stdout}\n{r.stderr[-3000:]}"
