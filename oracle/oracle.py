"""ctypes front end of the CPU oracle (oracle/fbf_oracle.c) and of the
reference-kernel shim (oracle/_ref/libfbf_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the cpu_baseline leg of bench.py, always as the checker.  The product
library never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libfbf_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libfbf_ref.so")
PAD_ID = 0xFFFFFFFF
CODE_DIM = 112

_P = C.c_void_p
_U64 = C.c_uint64
_I = C.c_int
_F = C.c_float


def build(ref: bool | None = None) -> None:
    """Compile the oracle (always) and the reference shim (when the reference
    sources are present in this container)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    ref_dir = os.environ.get("FBF_REF_DIR", "/root/reference")
    if ref is None:
        ref = os.path.isdir(ref_dir)
    if ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref", f"REF_DIR={ref_dir}"], check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build(ref=False)
        L = C.CDLL(ORACLE_SO)
        L.orc_f2ord.restype = C.c_uint32
        L.orc_f2ord.argtypes = [_F]
        L.orc_ord2f.restype = _F
        L.orc_ord2f.argtypes = [C.c_uint32]
        L.orc_l2_canon.restype = _F
        L.orc_l2_canon.argtypes = [_P, _P, _I]
        L.orc_l2_normal.restype = _F
        L.orc_l2_normal.argtypes = [_P, _P, _I]
        L.orc_l2_f64.restype = C.c_double
        L.orc_l2_f64.argtypes = [_P, _P, _I]
        L.orc_l2_sq8sym.restype = C.c_int32
        L.orc_l2_sq8sym.argtypes = [_P, _P]
        L.orc_l2_sq8ext_f64.restype = C.c_double
        L.orc_l2_sq8ext_f64.argtypes = [_P, _P, _P, _P, _I]
        L.orc_sq8sym_train.restype = _F
        L.orc_sq8sym_train.argtypes = [_P, _U64, _I]
        L.orc_sq8sym_encode.argtypes = [_P, _U64, _I, _F, _P]
        L.orc_sq8_train.argtypes = [_P, _U64, _I, _P, _P]
        L.orc_sq8_encode.argtypes = [_P, _U64, _I, _P, _P, _P]
        L.orc_topk_heap.argtypes = [_P, _U64, _I, _P, C.POINTER(_I)]
        L.orc_topk_fullsort.argtypes = [_P, _U64, _I, _P, C.POINTER(_I)]
        L.orc_match.restype = _I
        L.orc_match.argtypes = [_I, C.c_int32, _F, _F, C.c_int32, _F]
        L.orc_exact_predicate.argtypes = [_P, _P, _P, _U64, _I, _P, _P, _U64, _I, _I, _P, _P]
        L.orc_build_layouts.argtypes = [_P, _P, _U64, _P, _P]
        L.orc_query_range.argtypes = [_I, C.c_int32, _F, _F, _P, _P, _P, _U64,
                                      C.POINTER(_I), C.POINTER(_U64), C.POINTER(_U64)]
        L.orc_exact_ranges.argtypes = [_P, _P, _P, _P, _P, _P, _P, _U64, _I, _P, _P, _U64, _I, _P, _P]
        L.orc_exact_ranges_f64.argtypes = [_P, _P, _P, _P, _P, _P, _P, _U64, _I, _P, _P, _U64, _I, _P, _P]
        L.orc_l2_canon_rows.argtypes = [_P, _I, _P, _P, _U64, _I, _P]
        L.orc_bf_heap_ref.argtypes = [_I, _U64, _P, _P, _P, _P]
        L.orc_time_buckets.argtypes = [_P, _U64, _P, _P]
        L.orc_split_interval.argtypes = [_F, _F, _P, _U64, _U64, _P, _P, _F, _P, _P, _P]
        L.orc_subinterval.argtypes = [_P, _P, _P, _U64, _I, _P, _P, _P, _P, _P, _P, _P, _U64, _U64, _F, _I, _I,
                                      _I, _P, _P, _I, _P, _P, _P]
        L.orc_bf_sorted.argtypes = [_I, _U64, _P, _P, _P, _P]
        L.orc_router_bf.argtypes = [_P, _P, _P, _P, _P, _P, _P, _P, _U64, _I, _F, _P, _P, _U64,
                                    _I, _I, _P, _P]
        L.orc_recall.restype = C.c_double
        L.orc_recall.argtypes = [_P, _P, _U64, _I, _P]
        L.orc_num_threads.restype = _I
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO) and _cpu_has_avx512()


def _cpu_has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            flags = f.read()
    except OSError:
        return False
    return all(f" {x}" in flags for x in ("avx512f", "avx512vl", "avx512bw", "avx512dq"))


def ref():
    """The reference's own kernels (oracle/_ref), or None when not built or the
    host CPU lacks the AVX-512 subsets they were compiled for."""
    global _ref
    if _ref is None and ref_available():
        R = C.CDLL(REF_SO)
        R.ref_l2_sq8sym.restype = C.c_int
        R.ref_l2_sq8sym.argtypes = [_P, _P]
        R.ref_l2_ref.restype = _F
        R.ref_l2_ref.argtypes = [_P, _P, _I]
        R.ref_normal_l2.restype = _F
        R.ref_normal_l2.argtypes = [_P, _P, _I]
        R.ref_l2_sq8ext.restype = _F
        R.ref_l2_sq8ext.argtypes = [_P, _P, _P, _P]
        R.ref_bruteforce.argtypes = [_U64, _P, _U64, _P, _F, _P]
        R.ref_bruteforce_subgraph.argtypes = [_U64, _P, _P, _U64, _P, _F, _P]
        R.ref_sq8sym_train_encode.restype = _F
        R.ref_sq8sym_train_encode.argtypes = [_P, _I, _I, _P]
        R.ref_sq8sym_encode.argtypes = [_P, _I, _F, _P]
        R.ref_sq8_train_encode.argtypes = [_P, _I, _I, _P, _P, _P]
        R.ref_ground_truth.argtypes = [_P, C.c_uint32, _P, C.c_uint32, _I, _P]
        R.ref_router_bf.argtypes = [_P, _P, _P, _P, _P, _I, _F, _P, _P, _P, C.c_uint32,
                                    C.c_uint32, C.c_uint32, _P]
        R.ref_subinterval_small.argtypes = [_P, _P, _P, C.c_uint32, _I, _P, _P, C.c_uint32, C.c_uint32,
                                            C.c_uint32, _P, _P, _P]
        _ref = R
    return _ref


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"], "arrays passed to the oracle must be C-contiguous"
    return a.ctypes.data_as(_P)


# ---------------------------------------------------------------- keys
def f2ord(d: np.ndarray) -> np.ndarray:
    u = np.ascontiguousarray(d, dtype=np.float32).view(np.uint32)
    return np.where(u & 0x80000000, ~u, u | 0x80000000).astype(np.uint32)


def make_keys(dists: np.ndarray, ids: np.ndarray) -> np.ndarray:
    return (f2ord(dists).astype(np.uint64) << np.uint64(32)) | ids.astype(np.uint64)


# ---------------------------------------------------------------- distances
def l2_canon(x: np.ndarray, q: np.ndarray) -> float:
    x = np.ascontiguousarray(x, np.float32); q = np.ascontiguousarray(q, np.float32)
    return lib().orc_l2_canon(_p(x), _p(q), x.size)


def l2_normal(x, q) -> float:
    x = np.ascontiguousarray(x, np.float32); q = np.ascontiguousarray(q, np.float32)
    return lib().orc_l2_normal(_p(x), _p(q), x.size)


def l2_f64(x, q) -> float:
    x = np.ascontiguousarray(x, np.float32); q = np.ascontiguousarray(q, np.float32)
    return lib().orc_l2_f64(_p(x), _p(q), x.size)


def l2_sq8sym(a: np.ndarray, b: np.ndarray) -> int:
    a = np.zeros(CODE_DIM, np.int8) if a is None else np.ascontiguousarray(a, np.int8)
    b = np.ascontiguousarray(b, np.int8)
    assert a.size == CODE_DIM and b.size == CODE_DIM
    return lib().orc_l2_sq8sym(_p(a), _p(b))


def l2_sq8ext_f64(x, code, mi, dif) -> float:
    x = np.ascontiguousarray(x, np.float32); code = np.ascontiguousarray(code, np.uint8)
    mi = np.ascontiguousarray(mi, np.float32); dif = np.ascontiguousarray(dif, np.float32)
    return lib().orc_l2_sq8ext_f64(_p(x), _p(code), _p(mi), _p(dif), x.size)


# ---------------------------------------------------------------- quantisers
def sq8sym_train(x: np.ndarray) -> float:
    x = np.ascontiguousarray(x, np.float32)
    return lib().orc_sq8sym_train(_p(x), x.shape[0], x.shape[1])


def sq8sym_encode(x: np.ndarray, alpha: float) -> np.ndarray:
    x = np.ascontiguousarray(np.atleast_2d(x), np.float32)
    out = np.zeros((x.shape[0], CODE_DIM), np.int8)
    lib().orc_sq8sym_encode(_p(x), x.shape[0], x.shape[1], alpha, _p(out))
    return out


def sq8_train(x: np.ndarray):
    x = np.ascontiguousarray(x, np.float32)
    mi = np.zeros(x.shape[1], np.float32); dif = np.zeros(x.shape[1], np.float32)
    lib().orc_sq8_train(_p(x), x.shape[0], x.shape[1], _p(mi), _p(dif))
    return mi, dif


def sq8_encode(x: np.ndarray, mi, dif) -> np.ndarray:
    x = np.ascontiguousarray(np.atleast_2d(x), np.float32)
    mi = np.ascontiguousarray(mi, np.float32); dif = np.ascontiguousarray(dif, np.float32)
    out = np.zeros(x.shape, np.uint8)
    lib().orc_sq8_encode(_p(x), x.shape[0], x.shape[1], _p(mi), _p(dif), _p(out))
    return out


# ---------------------------------------------------------------- selection
def topk_heap(keys: np.ndarray, k: int) -> np.ndarray:
    keys = np.ascontiguousarray(keys, np.uint64)
    out = np.zeros(max(k, 1), np.uint64); n = C.c_int(0)
    lib().orc_topk_heap(_p(keys), keys.size, k, _p(out), C.byref(n))
    return out[: n.value]


def topk_fullsort(keys: np.ndarray, k: int) -> np.ndarray:
    keys = np.ascontiguousarray(keys, np.uint64)
    out = np.zeros(max(k, 1), np.uint64); n = C.c_int(0)
    lib().orc_topk_fullsort(_p(keys), keys.size, k, _p(out), C.byref(n))
    return out[: n.value]


def match(type_, v, l, r, label, ts) -> bool:
    return bool(lib().orc_match(int(type_), int(v), float(l), float(r), int(label), float(ts)))


# ---------------------------------------------------------------- exact search
def exact_predicate(base, labels, ts, queries, metas, k=100, mode="canon"):
    """Bruteforce (utils.h:80-158) restated; returns (ids uint32 nq x k,
    dists float64 nq x k) ascending by (dist, id), padded."""
    base = np.ascontiguousarray(base, np.float32)
    labels = np.ascontiguousarray(labels, np.int32)
    ts = np.ascontiguousarray(ts, np.float32)
    queries = np.ascontiguousarray(queries, np.float32)
    metas = np.ascontiguousarray(metas, np.float32)
    nq = queries.shape[0]
    ids = np.zeros((nq, k), np.uint32); dists = np.zeros((nq, k), np.float64)
    m = {"canon": 0, "f64": 1, "normal": 2}[mode]
    lib().orc_exact_predicate(_p(base), _p(labels), _p(ts), base.shape[0], base.shape[1],
                              _p(queries), _p(metas), nq, k, m, _p(ids), _p(dists))
    return ids, dists


class Layouts:
    """Sorted layouts of SortDataset (hybrid_graph.cpp:505-524) with ties broken
    by original id: perm_both by (label, ts, id), perm_time by (ts, id)."""

    def __init__(self, labels: np.ndarray, ts: np.ndarray):
        labels = np.ascontiguousarray(labels, np.int32); ts = np.ascontiguousarray(ts, np.float32)
        n = labels.size
        self.n = n
        self.perm_both = np.zeros(n, np.uint32); self.perm_time = np.zeros(n, np.uint32)
        lib().orc_build_layouts(_p(labels), _p(ts), n, _p(self.perm_both), _p(self.perm_time))
        self.lab_b = np.ascontiguousarray(labels[self.perm_both])
        self.ts_b = np.ascontiguousarray(ts[self.perm_both])
        self.ts_t = np.ascontiguousarray(ts[self.perm_time])

    def query_range(self, type_, v, l, r):
        lay = C.c_int(0); s = _U64(0); e = _U64(0)
        lib().orc_query_range(int(type_), int(v), float(l), float(r), _p(self.lab_b), _p(self.ts_b),
                              _p(self.ts_t), self.n, C.byref(lay), C.byref(s), C.byref(e))
        return lay.value, s.value, e.value

    def ranges(self, metas: np.ndarray) -> np.ndarray:
        out = np.zeros((metas.shape[0], 3), np.int64)
        for i, m in enumerate(metas):
            out[i] = self.query_range(int(m[0]), int(m[1]), m[2], m[3])
        return out


def exact_ranges(base, lay: Layouts, queries, metas, k=100):
    """Canonical-FP32 exact top-k over contiguous ranges of permuted copies (the
    CPU path timed beside the GPU).  Returns (ids, dists float32)."""
    base = np.ascontiguousarray(base, np.float32)
    vb = np.ascontiguousarray(base[lay.perm_both]); vt = np.ascontiguousarray(base[lay.perm_time])
    return exact_ranges_prepared(vb, vt, lay, queries, metas, k)


def exact_ranges_prepared(vb, vt, lay: Layouts, queries, metas, k=100):
    queries = np.ascontiguousarray(queries, np.float32); metas = np.ascontiguousarray(metas, np.float32)
    nq = queries.shape[0]
    ids = np.zeros((nq, k), np.uint32); dists = np.zeros((nq, k), np.float32)
    lib().orc_exact_ranges(_p(vb), _p(lay.perm_both), _p(lay.lab_b), _p(lay.ts_b), _p(vt),
                           _p(lay.perm_time), _p(lay.ts_t), lay.n, vb.shape[1], _p(queries),
                           _p(metas), nq, k, _p(ids), _p(dists))
    return ids, dists


def exact_ranges_f64_prepared(vb, vt, lay: Layouts, queries, metas, k=100):
    """FP64 exact top-k over the sorted ranges (check time only).  Returns
    (ids uint32, dists float64), ascending by (distance, id), padded."""
    queries = np.ascontiguousarray(queries, np.float32); metas = np.ascontiguousarray(metas, np.float32)
    nq = queries.shape[0]
    ids = np.zeros((nq, k), np.uint32); dists = np.zeros((nq, k), np.float64)
    lib().orc_exact_ranges_f64(_p(vb), _p(lay.perm_both), _p(lay.lab_b), _p(lay.ts_b), _p(vt),
                               _p(lay.perm_time), _p(lay.ts_t), lay.n, vb.shape[1], _p(queries),
                               _p(metas), nq, k, _p(ids), _p(dists))
    return ids, dists


def l2_canon_rows(base, queries, ids) -> np.ndarray:
    """Canonical FP32 distance of each (query, returned id); +inf for the pad id."""
    base = np.ascontiguousarray(base, np.float32); queries = np.ascontiguousarray(queries, np.float32)
    ids = np.ascontiguousarray(ids, np.uint32)
    out = np.zeros(ids.shape, np.float32)
    lib().orc_l2_canon_rows(_p(base), base.shape[1], _p(queries), _p(ids), ids.shape[0], ids.shape[1], _p(out))
    return out


# ---------------------------------------------------------------- reference bf semantics
def bf_heap_ref(codes: np.ndarray, qcode: np.ndarray, k: int):
    """maxPQIFCS-driven bruteforce()/bruteforce_subgraph() (bruteforce.h:244-289):
    positions in heap-array order and their int distances."""
    codes = np.ascontiguousarray(codes, np.int8); qcode = np.ascontiguousarray(qcode, np.int8)
    ids = np.zeros(k, np.int32); d = np.zeros(k, np.int32)
    lib().orc_bf_heap_ref(k, codes.shape[0], _p(codes), _p(qcode), _p(ids), _p(d))
    return ids, d


def bf_sorted(codes: np.ndarray, qcode: np.ndarray, k: int):
    codes = np.ascontiguousarray(codes, np.int8); qcode = np.ascontiguousarray(qcode, np.int8)
    ids = np.zeros(k, np.int32); d = np.zeros(k, np.int32)
    lib().orc_bf_sorted(k, codes.shape[0], _p(codes), _p(qcode), _p(ids), _p(d))
    return ids, d


def router_bf(base, lay: Layouts, codes_b, codes_t, alpha, queries, metas, kq=140, k=100):
    """The router's brute-force branch (hybrid_graph.cpp:203-228) restated
    (orc_router_bf): symmetric int8 top-kq of each query's range, ties to the
    smaller position, id map, canonical FP32 rerank to k.  Returns (ids, dists)."""
    base = np.ascontiguousarray(base, np.float32)
    queries = np.ascontiguousarray(queries, np.float32); metas = np.ascontiguousarray(metas, np.float32)
    nq = queries.shape[0]
    ids = np.zeros((nq, k), np.uint32); dists = np.zeros((nq, k), np.float32)
    lib().orc_router_bf(_p(codes_b), _p(codes_t), _p(lay.perm_both), _p(lay.lab_b), _p(lay.ts_b),
                        _p(lay.perm_time), _p(lay.ts_t), _p(base), base.shape[0], base.shape[1], alpha,
                        _p(queries), _p(metas), nq, kq, k, _p(ids), _p(dists))
    return ids, dists


def bf_boundary_tied(lay: Layouts, codes_b, codes_t, qcodes, metas, kq=140) -> np.ndarray:
    """True for queries whose range holds more than kq rows and whose kq-th and
    (kq+1)-th smallest symmetric int8 distances are equal: there the
    reference's heap may keep a different row than the smaller position."""
    out = np.zeros(metas.shape[0], bool)
    for i, m in enumerate(metas):
        layout, s, e = lay.query_range(int(m[0]), int(m[1]), m[2], m[3])
        if e - s <= kq:
            continue
        c = codes_b if layout == 0 else codes_t
        _, d = bf_sorted(c[s:e], qcodes[i], kq + 1)
        out[i] = d[kq - 1] == d[kq]
    return out


# ---------------------------------------------------------------- type-2 sub-interval path
class TimeBuckets:
    """The ten 0.1-wide time buckets of the timestamp layout (utils.h:470-478)
    with each bucket's symmetric int8 quantiser trained on its own rows
    (SQ8SymmetricQuantizer via the timestamp searchers' SetData,
    baseline.cpp:130-141): alpha[t], and codes of bucket t at rows
    [off[t], off[t] + count[t]) of one array."""

    def __init__(self, base: np.ndarray, lay: Layouts):
        self.start = np.zeros(10, np.uint64); self.count = np.zeros(10, np.uint64)
        lib().orc_time_buckets(_p(lay.ts_t), lay.n, _p(self.start), _p(self.count))
        self.off = np.concatenate([[0], np.cumsum(self.count)[:-1]]).astype(np.uint64)
        self.alpha = np.zeros(10, np.float32)
        codes = []
        for t in range(10):
            s, c = int(self.start[t]), int(self.count[t])
            rows = np.ascontiguousarray(base[lay.perm_time[s:s + c]])
            if c:
                self.alpha[t] = sq8sym_train(rows)
                codes.append(sq8sym_encode(rows, float(self.alpha[t])))
        self.codes = np.ascontiguousarray(np.concatenate(codes) if codes else np.zeros((0, CODE_DIM), np.int8))

    def split(self, lay: Layouts, l, r, nb, small_thr=0.0):
        types = np.zeros(10, np.uint8); ls = np.zeros(10, np.uint64); ln = np.zeros(10, np.uint64)
        lib().orc_split_interval(float(l), float(r), _p(lay.ts_t), lay.n, nb, _p(self.start), _p(self.count),
                                 float(small_thr), _p(types), _p(ls), _p(ln))
        return types, ls, ln


def subinterval(base, lay: Layouts, tb: TimeBuckets, queries, metas, k=100, nb=None, small_thr=0.0,
                bucket_k=140, pool_k=150, ext=None):
    """The type-2 SMALL route + candidate pool + global refine restated
    (orc_subinterval); ext: optional (positions uint32 [nq, n_ext], pool
    distances float32 [nq, n_ext]).  Returns (ids, dists, types nq x 10)."""
    base = np.ascontiguousarray(base, np.float32)
    queries = np.ascontiguousarray(queries, np.float32); metas = np.ascontiguousarray(metas, np.float32)
    nq = queries.shape[0]
    ids = np.zeros((nq, k), np.uint32); dists = np.zeros((nq, k), np.float32); types = np.zeros((nq, 10), np.uint8)
    ep = np.zeros(1, np.uint32); ed = np.zeros(1, np.float32); n_ext = 0
    if ext is not None:
        ep = np.ascontiguousarray(ext[0], np.uint32); ed = np.ascontiguousarray(ext[1], np.float32)
        n_ext = ep.shape[1]
    lib().orc_subinterval(_p(base), _p(lay.perm_time), _p(lay.ts_t), lay.n, base.shape[1], _p(tb.codes), _p(tb.off),
                          _p(tb.start), _p(tb.count), _p(tb.alpha), _p(queries), _p(metas), nq,
                          lay.n if nb is None else nb, float(small_thr), bucket_k, pool_k, k, _p(ep), _p(ed), n_ext,
                          _p(ids), _p(dists), _p(types))
    return ids, dists, types


def ref_subinterval(base, lay: Layouts, queries, metas, k=100, nb=None):
    """The reference's own route (oracle/_ref ref_subinterval_small): ids in
    heap-array order, RFType per bucket (255 none), bucket alphas; None when
    the reference kernels are not available."""
    R = ref()
    if R is None:
        return None
    base = np.ascontiguousarray(base, np.float32)
    vt = np.ascontiguousarray(base[lay.perm_time])
    queries = np.ascontiguousarray(queries, np.float32); metas = np.ascontiguousarray(metas, np.float32)
    nq = queries.shape[0]
    ids = np.zeros((nq, k), np.uint32); types = np.zeros((nq, 10), np.int32); alphas = np.zeros(10, np.float32)
    R.ref_subinterval_small(_p(vt), _p(lay.ts_t), _p(lay.perm_time), lay.n, base.shape[1], _p(queries), _p(metas), nq,
                            k, lay.n if nb is None else nb, _p(ids), _p(types), _p(alphas))
    return ids, types, alphas


def recall(knn: np.ndarray, gt: np.ndarray):
    knn = np.ascontiguousarray(knn, np.uint32); gt = np.ascontiguousarray(gt, np.uint32)
    per = np.zeros(knn.shape[0], np.int32)
    r = lib().orc_recall(_p(knn), _p(gt), knn.shape[0], knn.shape[1], _p(per))
    return r, per


def num_threads() -> int:
    return lib().orc_num_threads()
