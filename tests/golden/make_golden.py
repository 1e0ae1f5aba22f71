"""Writes the golden fixtures under tests/golden/ (run in a container that has
the reference sources; the fixtures are committed, this script is how they
were made).

* dummy-data.bin / dummy-queries.bin: the reference's own fixture files,
  copied byte for byte (contest formats, utils.h:257-276 and :550-566).
* dummy_gt_ref.bin: ground truth written by the reference's Bruteforce
  (utils.h:80-158, normal_l2 from distance.h:62-71, compiled by oracle/Makefile
  target "ref") in the SaveGroundTruth format (utils.h:168-188).
* dummy_exact.npz: this project's canonical-FP32 exact top-100 (ids, dists)
  and the FP64 top-100 for the same queries.
* ref_bruteforce.npz: symmetric-int8 codes of the dummy base (reference
  SQ8SymmetricQuantizer train/encode), and for 12 queries the reference
  bruteforce() and bruteforce_subgraph() outputs (heap-array order) at
  k = 140 over the whole code array and over a slice.
* kat.json: the int8 known answer of test_simd_distance.cpp:199-203.
* ref_router_bf.npz: the reference's brute-force route (hybrid_graph.cpp:
  203-228: bruteforce() top-140 of the symmetric int8 codes, id map, L2SqrRef
  rerank to 100 in maxPQIFCS) run by the reference kernels over every dummy
  query's filtered range of the sorted layouts; ids in heap-array order.
* ref_subinterval.npz: the reference's type-2 sub-interval route for the
  SMALL buckets (its SplitInterval, per-bucket quantisers, bruteforce_subgraph,
  candidate pool and L2SqrRef refine, oracle/ref_shim.cpp
  ref_subinterval_small) on the config-0 base at 20,000 rows (datagen seed)
  for 400 type-2 queries; ids in heap-array order, RFType per bucket, the
  bucket alphas.
"""
from __future__ import annotations

import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("FBF_REF_DIR", "/root/reference")
sys.path.insert(0, ROOT)


def main() -> None:
    import __graft_entry__ as ge
    orc = ge.load_oracle()
    orc.build(ref=True)
    R = orc.ref()
    assert R is not None, "reference shim not built"
    dg = ge.load_pkg().datagen

    for f in ("dummy-data.bin", "dummy-queries.bin"):
        shutil.copyfile(os.path.join(REF, f), os.path.join(HERE, f))
    vecs, labels, ts = dg.read_base_bin(os.path.join(HERE, "dummy-data.bin"))
    q, metas = dg.read_query_bin(os.path.join(HERE, "dummy-queries.bin"))
    n, d = vecs.shape
    nq = q.shape[0]

    # Reference ground truth (its own Bruteforce, threads and all).
    nodes = np.ascontiguousarray(np.concatenate([labels[:, None], ts[:, None], vecs], 1), np.float32)
    qrec = np.ascontiguousarray(np.concatenate([metas, q], 1), np.float32)
    gt = np.zeros((nq, 100), np.uint32)
    R.ref_ground_truth(orc._p(nodes), n, orc._p(qrec), nq, d, orc._p(gt))
    dg.write_gt_bin(os.path.join(HERE, "dummy_gt_ref.bin"), gt)

    ids, dists = orc.exact_predicate(vecs, labels.astype(np.int32), ts, q, metas, k=100, mode="canon")
    ids64, d64 = orc.exact_predicate(vecs, labels.astype(np.int32), ts, q, metas, k=100, mode="f64")
    np.savez_compressed(os.path.join(HERE, "dummy_exact.npz"), ids=ids, dists=dists.astype(np.float32),
                        ids_f64=ids64, dists_f64=d64)

    # Reference symmetric-int8 quantiser and bruteforce() outputs.
    codes = np.zeros((n, 112), np.int8)
    alpha = R.ref_sq8sym_train_encode(orc._p(np.ascontiguousarray(vecs)), n, d, orc._p(codes))
    sel = np.arange(0, nq, max(1, nq // 12))[:12]
    k = 140
    qcodes = np.zeros((sel.size, 112), np.int8)
    bf_full = np.zeros((sel.size, k), np.int32)
    bf_slice = np.zeros((sel.size, k), np.int32)
    bf_slice_d = np.zeros((sel.size, k), np.float32)
    s0, s1 = 1234, 1234 + 3001
    for j, i in enumerate(sel):
        R.ref_sq8sym_encode(orc._p(np.ascontiguousarray(q[i])), d, alpha, orc._p(qcodes[j]))
        R.ref_bruteforce(k, orc._p(bf_full[j]), n, orc._p(codes), alpha, orc._p(qcodes[j]))
        R.ref_bruteforce_subgraph(k, orc._p(bf_slice[j]), orc._p(bf_slice_d[j]), s1 - s0,
                                  orc._p(codes[s0:s1]), alpha, orc._p(qcodes[j]))
    np.savez_compressed(os.path.join(HERE, "ref_bruteforce.npz"), codes=codes, alpha=np.float32(alpha),
                        query_index=sel, qcodes=qcodes, k=k, bf_full=bf_full, slice=np.array([s0, s1]),
                        bf_slice=bf_slice, bf_slice_dist=bf_slice_d)

    # The reference's brute-force route over every dummy query's range.
    L = orc.Layouts(labels.astype(np.int32), ts)
    ranges = L.ranges(metas).astype(np.int64)
    codes_b = np.ascontiguousarray(codes[L.perm_both])
    codes_t = np.ascontiguousarray(codes[L.perm_time])
    qcodes_all = np.zeros((nq, 112), np.int8)
    for i in range(nq):
        R.ref_sq8sym_encode(orc._p(np.ascontiguousarray(q[i])), d, alpha, orc._p(qcodes_all[i]))
    rr_ids = np.zeros((nq, 100), np.uint32)
    R.ref_router_bf(orc._p(codes_b), orc._p(codes_t), orc._p(L.perm_both), orc._p(L.perm_time),
                    orc._p(np.ascontiguousarray(vecs)), d, alpha, orc._p(np.ascontiguousarray(q)),
                    orc._p(qcodes_all), orc._p(np.ascontiguousarray(ranges)), nq, 140, 100, orc._p(rr_ids))
    np.savez_compressed(os.path.join(HERE, "ref_router_bf.npz"), ranges=ranges, ids=rr_ids, kq=140, k=100,
                        alpha=np.float32(alpha))

    # The type-2 sub-interval route on a uniform-timestamp base.
    sv, sl, st = dg.make_base(dg.CONFIGS[0], n=20000)
    Ls = orc.Layouts(sl.astype(np.int32), st)
    rng = np.random.default_rng(41)
    nq2 = 400
    q2, _ = dg.make_queries(dg.CONFIGS[0], nq=nq2)
    l2 = rng.random(nq2).astype(np.float32)
    w2 = np.exp(rng.uniform(np.log(0.005), np.log(0.5), nq2)).astype(np.float32)
    m2 = np.full((nq2, 4), -1.0, np.float32)
    m2[:, 0] = 2
    m2[:, 2] = l2
    m2[:, 3] = np.minimum(l2 + w2, np.float32(1.0))
    m2[:4, 2:] = [[0.1, 0.2], [0.0, 1.0], [0.25, 0.26], [0.5, 0.5]]  # whole bucket, everything, tiny, one instant
    rid, rtypes, ralpha = orc.ref_subinterval(sv, Ls, q2, m2, 100)
    np.savez_compressed(os.path.join(HERE, "ref_subinterval.npz"), n=20000, queries=q2, metas=m2, ids=rid,
                        types=rtypes, alphas=ralpha)

    a0 = [127, 127, 127, -4, 5, 6, 7, 8, 9, -127, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, -24,
          25, 26, 27, 28, 29, 30, 31, 32]
    a1 = [-127, -127, 4, 5, 6, 7, 8, 9, 10, 127, 12, -13, 14, 15, 16, 17, 18, 19, -20, 21, 22, 23, 24, -25,
          26, 27, 28, 29, 30, 31, 32, 33]
    x = np.zeros(112, np.int8); y = np.zeros(112, np.int8)
    x[:32] = a0; y[:32] = a1
    kat = {"source": "test_simd_distance.cpp:199-203", "a0": a0, "a1": a1,
           "l2_sq8sym": int(R.ref_l2_sq8sym(orc._p(x), orc._p(y)))}
    with open(os.path.join(HERE, "kat.json"), "w") as f:
        json.dump(kat, f, indent=1)
    print("golden fixtures written; reference int8 KAT =", kat["l2_sq8sym"])


if __name__ == "__main__":
    main()
