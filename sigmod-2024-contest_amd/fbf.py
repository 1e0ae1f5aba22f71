"""Python front end of libfbf.so (the C-ABI in include/fbf.h).

Mirrors the reference's call surface for the brute-force path:

* ``Dataset`` -- the device-resident counterpart of the sorted layouts and
  codes that HybridGraph keeps (hybrid_graph.h, SortDataset at
  hybrid_graph.cpp:505-524);
* ``Dataset.search`` -- the batched brute-force branch of BatchSearch
  (hybrid_graph.cpp:168-228) and the exact Bruteforce of utils.h:80-158;
* ``bruteforce`` / ``bruteforce_subgraph`` -- bruteforce.h:244-289.

The library is required: every call goes to the HIP kernels and there is no
CPU path behind this module.  A missing or stale libfbf.so raises.
"""
from __future__ import annotations

import ctypes as C
import os
import sys as _sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libfbf.so")
PAD_ID = 0xFFFFFFFF
CODE_BYTES = 112

MODE_EXACT_F32, MODE_SQ8, MODE_BF16, MODE_FP8, MODE_SQ8SYM, MODE_EXACT_BF16 = range(6)
MODES = {"exact": MODE_EXACT_F32, "sq8": MODE_SQ8, "bf16": MODE_BF16, "fp8": MODE_FP8,
         "sq8sym": MODE_SQ8SYM, "exact_bf16": MODE_EXACT_BF16}
REPR_F32, REPR_SQ8, REPR_BF16, REPR_FP8, REPR_SQ8SYM, REPR_TBUCKETS = 0x1, 0x2, 0x4, 0x8, 0x10, 0x20
REPR_ALL = 0x3F
MODE_REPR = {MODE_EXACT_F32: REPR_F32, MODE_SQ8: REPR_SQ8, MODE_BF16: REPR_BF16,
             MODE_FP8: REPR_FP8, MODE_SQ8SYM: REPR_SQ8SYM, MODE_EXACT_BF16: REPR_BF16 | REPR_F32}

# Every symbol include/fbf.h declares (checked by the CPU test suite).
EXPORTS = [
    "fbf_dataset_create", "fbf_dataset_create_device", "fbf_dataset_destroy", "fbf_filter_count",
    "fbf_search", "fbf_search_host", "fbf_dataset_quant_params", "fbf_register_codes",
    "fbf_unregister_codes", "fbf_bruteforce", "fbf_bruteforce_subgraph", "fbf_bruteforce_batch",
    "fbf_profile_enable", "fbf_profile_collect", "fbf_last_search_stats", "fbf_last_search_rescans",
    "fbf_last_error", "fbf_version", "fbf_search_graph_create", "fbf_search_graph_launch",
    "fbf_search_graph_destroy", "fbf_last_search_fallbacks", "fbf_dataset_layout_rows",
    "fbf_last_search_pass_pairs", "fbf_bruteforce_batch_noalloc", "fbf_search_subinterval", "fbf_route",
    "fbf_search_graph_plan_info",
]


class FbfError(RuntimeError):
    pass


class SubintervalOpts(C.Structure):
    _fields_ = [("small_thr", C.c_float), ("bucket_k", C.c_uint32), ("pool_k", C.c_uint32),
                ("reserved", C.c_uint32)]


class SearchOpts(C.Structure):
    _fields_ = [("mode", C.c_uint32), ("kprime", C.c_uint32), ("chunk_rows", C.c_uint32),
                ("group_queries", C.c_uint32), ("persistent", C.c_uint32), ("prepass_div", C.c_uint32),
                ("prepass_margin", C.c_uint32), ("max_batch", C.c_uint32), ("metas_on_device", C.c_uint32)]


_lib = None
_P = C.c_void_p


def lib() -> C.CDLL:
    """Load libfbf.so (raises FbfError when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FbfError(f"{LIB_PATH} is not built: run __graft_entry__.build() or make -C {HERE}")
    L = C.CDLL(LIB_PATH)
    L.fbf_last_error.restype = C.c_char_p
    L.fbf_version.restype = C.c_char_p
    i = C.c_int
    L.fbf_dataset_create.argtypes = [_P, _P, _P, C.c_uint64, C.c_uint32, C.c_uint32, C.POINTER(_P)]
    L.fbf_dataset_create_device.argtypes = [_P, _P, _P, C.c_uint64, C.c_uint32, C.c_uint32, C.POINTER(_P)]
    L.fbf_dataset_destroy.argtypes = [_P]
    L.fbf_filter_count.argtypes = [_P, _P, C.POINTER(C.c_uint64), C.POINTER(C.c_float)]
    L.fbf_search.argtypes = [_P, _P, _P, C.c_uint64, C.c_uint32, C.POINTER(SearchOpts), _P, _P, _P]
    L.fbf_search_host.argtypes = [_P, _P, _P, C.c_uint64, C.c_uint32, C.POINTER(SearchOpts), _P, _P]
    L.fbf_route.argtypes = [_P, _P, C.c_uint64, C.c_float, C.c_float, _P, _P, _P]
    L.fbf_dataset_quant_params.argtypes = [_P, _P, _P, _P]
    L.fbf_register_codes.argtypes = [_P, C.c_uint64]
    L.fbf_unregister_codes.argtypes = [_P]
    L.fbf_bruteforce.argtypes = [C.c_size_t, C.c_size_t, _P, C.c_size_t, _P, C.c_float, C.c_size_t, _P]
    L.fbf_bruteforce_subgraph.argtypes = [C.c_size_t, C.c_size_t, _P, _P, C.c_size_t, _P, C.c_float,
                                          C.c_size_t, _P]
    L.fbf_bruteforce_batch.argtypes = [_P, _P, _P, _P, C.c_uint64, C.c_uint32, _P, _P]
    L.fbf_bruteforce_batch_noalloc.argtypes = [_P, _P, _P, _P, C.c_uint64, C.c_uint32, _P, _P, _P]
    L.fbf_search_subinterval.argtypes = [_P, _P, _P, C.c_uint64, C.c_uint32, C.POINTER(SubintervalOpts), _P, _P,
                                         C.c_uint32, _P, _P, _P, _P]
    L.fbf_profile_enable.argtypes = [_P, i]
    L.fbf_profile_collect.argtypes = [_P, C.c_char_p, C.POINTER(C.c_uint64), C.POINTER(C.c_double)]
    L.fbf_last_search_stats.argtypes = [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                        C.POINTER(C.c_uint64)]
    L.fbf_last_search_rescans.argtypes = [_P, C.POINTER(C.c_uint64)]
    L.fbf_last_search_fallbacks.argtypes = [_P, C.POINTER(C.c_uint64)]
    L.fbf_last_search_pass_pairs.argtypes = [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.fbf_search_graph_create.argtypes = [_P, _P, _P, C.c_uint64, C.c_uint32, C.POINTER(SearchOpts), _P, _P,
                                          C.POINTER(_P)]
    L.fbf_search_graph_launch.argtypes = [_P, _P]
    L.fbf_search_graph_destroy.argtypes = [_P]
    L.fbf_search_graph_plan_info.argtypes = [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                             C.POINTER(C.c_uint32)]
    L.fbf_dataset_layout_rows.argtypes = [_P, C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint32, _P, _P]
    for name in EXPORTS:
        if name not in ("fbf_last_error", "fbf_version"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _check(status: int, what: str) -> None:
    if status != 0:
        msg = lib().fbf_last_error().decode(errors="replace")
        raise FbfError(f"{what} failed with status {status}: {msg}")


def _np_ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_P)


def _dev_ptr(t) -> int:
    import torch  # device memory plumbing only
    assert isinstance(t, torch.Tensor) and t.is_cuda and t.is_contiguous()
    return t.data_ptr()


def version() -> str:
    return lib().fbf_version().decode()


class Dataset:
    """Device-resident filtered dataset (both sorted layouts + representations)."""

    def __init__(self, vecs, labels: np.ndarray, timestamps: np.ndarray, repr_flags: int = REPR_F32):
        L = lib()
        labels = np.ascontiguousarray(labels, np.float32)
        timestamps = np.ascontiguousarray(timestamps, np.float32)
        h = _P()
        if isinstance(vecs, np.ndarray):
            vecs = np.ascontiguousarray(vecs, np.float32)
            n, d = vecs.shape
            st = L.fbf_dataset_create(_np_ptr(vecs), _np_ptr(labels), _np_ptr(timestamps), n, d,
                                      repr_flags, C.byref(h))
        else:  # torch CUDA tensor
            n, d = vecs.shape
            st = L.fbf_dataset_create_device(_P(_dev_ptr(vecs)), _np_ptr(labels), _np_ptr(timestamps),
                                             n, d, repr_flags, C.byref(h))
        _check(st, "fbf_dataset_create")
        self.h = h
        self.n, self.d = int(n), int(d)
        self.repr_flags = repr_flags

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().fbf_dataset_destroy(self.h)
            self.h = None

    def __del__(self):
        # At interpreter shutdown the HIP runtime may already be tearing down:
        # no device calls then (the process exit releases the memory).
        if _sys is None or _sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def filter_count(self, meta) -> tuple[int, float]:
        m = np.ascontiguousarray(meta, np.float32)
        c = C.c_uint64(0); s = C.c_float(0)
        _check(lib().fbf_filter_count(self.h, _np_ptr(m), C.byref(c), C.byref(s)), "fbf_filter_count")
        return c.value, s.value

    @staticmethod
    def _opts(mode, kprime, chunk_rows, group_queries, persistent, prepass_div=0,
              prepass_margin=0, max_batch=0) -> SearchOpts:
        o = SearchOpts()
        o.mode = MODES[mode] if isinstance(mode, str) else int(mode)
        o.kprime = kprime
        o.chunk_rows = chunk_rows
        o.group_queries = group_queries
        o.persistent = int(persistent)  # 0 default, 1 (True) persistent, 2 one workgroup per item
        o.prepass_div = prepass_div
        o.prepass_margin = prepass_margin
        o.max_batch = max_batch
        return o

    def search(self, queries, metas: np.ndarray, k: int = 100, mode="exact", kprime: int = 0,
               chunk_rows: int = 0, group_queries: int = 0, persistent: bool = False,
               out=None, stream=None, prepass_div: int = 0, prepass_margin: int = 0, max_batch: int = 0):
        """queries: torch CUDA float32 [nq, d]; metas: numpy float32 [nq, 4].
        Returns torch (ids int32 view of uint32 [nq, k], dists float32 [nq, k])
        on the queries' device; asynchronous on ``stream`` (torch stream or None
        for torch's current stream)."""
        import torch
        on_dev = isinstance(metas, torch.Tensor)
        if on_dev:  # device metas: routed on the device (fbf_route)
            assert metas.is_cuda and metas.dtype == torch.float32 and metas.is_contiguous()
        else:
            metas = np.ascontiguousarray(metas, np.float32)
        nq = metas.shape[0]
        assert queries.shape == (nq, self.d) and queries.dtype == torch.float32
        if out is None:
            ids = torch.empty((nq, k), dtype=torch.int32, device=queries.device)
            dists = torch.empty((nq, k), dtype=torch.float32, device=queries.device)
        else:
            ids, dists = out
        st = (stream or torch.cuda.current_stream(queries.device)).cuda_stream
        o = self._opts(mode, kprime, chunk_rows, group_queries, persistent, prepass_div, prepass_margin,
                       max_batch)
        o.metas_on_device = int(on_dev)
        mp = _P(_dev_ptr(metas)) if on_dev else _np_ptr(metas)
        _check(lib().fbf_search(self.h, _P(_dev_ptr(queries)), mp, nq, k, C.byref(o),
                                _P(_dev_ptr(ids)), _P(_dev_ptr(dists)), _P(st)), "fbf_search")
        return ids, dists

    def route(self, metas_dev, bf_thr: float = 0.0, bf_thr3: float = 0.0, order: bool = True, stream=None):
        """Device router (fbf_route): metas torch CUDA float32 [nq, 4] ->
        (ranges int32 [nq, 4] = layout, s, e, bf; order int32 [nq] or None)."""
        import torch
        assert metas_dev.is_cuda and metas_dev.dtype == torch.float32 and metas_dev.is_contiguous()
        nq = metas_dev.shape[0]
        ranges = torch.empty((nq, 4), dtype=torch.int32, device=metas_dev.device)
        ordr = torch.empty(nq, dtype=torch.int32, device=metas_dev.device) if order else None
        st = (stream or torch.cuda.current_stream(metas_dev.device)).cuda_stream
        _check(lib().fbf_route(self.h, _P(_dev_ptr(metas_dev)), nq, bf_thr, bf_thr3, _P(_dev_ptr(ranges)),
                               _P(_dev_ptr(ordr)) if order else None, _P(st)), "fbf_route")
        return ranges, ordr

    def search_host(self, queries: np.ndarray, metas: np.ndarray, k: int = 100, mode="exact",
                    kprime: int = 0, chunk_rows: int = 0, group_queries: int = 0,
                    persistent: bool = False, prepass_div: int = 0, prepass_margin: int = 0):
        """Host-memory form: returns numpy (ids uint32 [nq, k], dists float32 [nq, k])."""
        queries = np.ascontiguousarray(queries, np.float32)
        metas = np.ascontiguousarray(metas, np.float32)
        nq = queries.shape[0]
        ids = np.zeros((nq, k), np.uint32)
        dists = np.zeros((nq, k), np.float32)
        o = self._opts(mode, kprime, chunk_rows, group_queries, persistent, prepass_div, prepass_margin)
        _check(lib().fbf_search_host(self.h, _np_ptr(queries), _np_ptr(metas), nq, k, C.byref(o),
                                     _np_ptr(ids), _np_ptr(dists)), "fbf_search_host")
        return ids, dists

    def search_subinterval(self, queries, metas: np.ndarray, k: int = 100, small_thr: float = 0.0,
                           bucket_k: int = 0, pool_k: int = 0, ext=None, stream=None):
        """Type-2 sub-interval search (fbf_search_subinterval): queries torch
        CUDA float32 [nq, d], metas numpy [nq, 4] (type 2).  ext: optional
        (positions uint32/int32 CUDA [nq, n_ext], distances float32 CUDA
        [nq, n_ext]) candidates of the graph-searched buckets.  Returns torch
        (ids, dists) and numpy bucket classes [nq, 10] (255 none, 0 SMALL,
        1 MEDIUM, 3 FULL)."""
        import torch
        metas = np.ascontiguousarray(metas, np.float32)
        nq = metas.shape[0]
        assert queries.shape == (nq, self.d) and queries.dtype == torch.float32
        ids = torch.empty((nq, k), dtype=torch.int32, device=queries.device)
        dists = torch.empty((nq, k), dtype=torch.float32, device=queries.device)
        types = np.zeros((nq, 10), np.uint8)
        o = SubintervalOpts()
        o.small_thr, o.bucket_k, o.pool_k = small_thr, bucket_k, pool_k
        n_ext, ep, ed = 0, None, None
        if ext is not None:
            ep, ed = ext
            assert ep.shape == ed.shape and ep.shape[0] == nq
            n_ext = ep.shape[1]
            ep, ed = _P(_dev_ptr(ep)), _P(_dev_ptr(ed))
        st = (stream or torch.cuda.current_stream(queries.device)).cuda_stream
        _check(lib().fbf_search_subinterval(self.h, _P(_dev_ptr(queries)), _np_ptr(metas), nq, k, C.byref(o), ep, ed,
                                            n_ext, _P(_dev_ptr(ids)), _P(_dev_ptr(dists)), _np_ptr(types), _P(st)),
               "fbf_search_subinterval")
        return ids, dists, types

    def quant_params(self):
        """(alpha, mi, dif): the symmetric int8 scale (None without
        REPR_SQ8SYM) and the per-dimension uint8 min / range (None without
        REPR_SQ8)."""
        a = C.c_float(0)
        has_sym, has_u8 = bool(self.repr_flags & REPR_SQ8SYM), bool(self.repr_flags & REPR_SQ8)
        mi = np.zeros(self.d, np.float32) if has_u8 else None
        dif = np.zeros(self.d, np.float32) if has_u8 else None
        _check(lib().fbf_dataset_quant_params(self.h, C.byref(a) if has_sym else None,
                                              _np_ptr(mi) if has_u8 else None, _np_ptr(dif) if has_u8 else None),
               "fbf_dataset_quant_params")
        return (a.value if has_sym else None), mi, dif

    def layout_rows(self, layout: int, row0: int = 0, nrows: int | None = None, repr_flag: int = 0):
        """Original ids and (repr_flag REPR_SQ8SYM / REPR_SQ8) the device codes of
        positions [row0, row0 + nrows) of one sorted layout."""
        nrows = self.n - row0 if nrows is None else nrows
        ids = np.zeros(nrows, np.uint32)
        codes = None
        if repr_flag == REPR_SQ8SYM:
            codes = np.zeros((nrows, CODE_BYTES), np.int8)
        elif repr_flag == REPR_SQ8:
            codes = np.zeros((nrows, self.d), np.uint8)
        _check(lib().fbf_dataset_layout_rows(self.h, layout, row0, nrows, repr_flag, _np_ptr(ids),
                                             None if codes is None else _np_ptr(codes)),
               "fbf_dataset_layout_rows")
        return ids, codes

    # ---------------------------------------------------------- measurement
    def profile(self, enable: bool = True) -> None:
        _check(lib().fbf_profile_enable(self.h, int(enable)), "fbf_profile_enable")

    def profile_collect(self, kernel: str = "all") -> tuple[int, float]:
        n = C.c_uint64(0); ms = C.c_double(0)
        _check(lib().fbf_profile_collect(self.h, kernel.encode(), C.byref(n), C.byref(ms)),
               "fbf_profile_collect")
        return n.value, ms.value

    def last_stats(self) -> dict:
        p = C.c_uint64(0); pc = C.c_uint64(0); wi = C.c_uint64(0)
        _check(lib().fbf_last_search_stats(self.h, C.byref(p), C.byref(pc), C.byref(wi)),
               "fbf_last_search_stats")
        return {"pairs": p.value, "pairs_computed": pc.value, "work_items": wi.value}

    def last_pass_pairs(self) -> tuple[int, int]:
        """(prefix-pass pairs, main-pass pairs) of the last search."""
        a = C.c_uint64(0); b = C.c_uint64(0)
        _check(lib().fbf_last_search_pass_pairs(self.h, C.byref(a), C.byref(b)), "fbf_last_search_pass_pairs")
        return a.value, b.value

    def last_rescans(self) -> int:
        """Queries of the last search rescanned after their seed failed."""
        n = C.c_uint64(0)
        _check(lib().fbf_last_search_rescans(self.h, C.byref(n)), "fbf_last_search_rescans")
        return n.value

    def last_fallbacks(self) -> int:
        """Queries of the last exact_bf16 search that the certificate left to the FP32 scan."""
        n = C.c_uint64(0)
        _check(lib().fbf_last_search_fallbacks(self.h, C.byref(n)), "fbf_last_search_fallbacks")
        return n.value


# ---------------------------------------------------------------- reference-shaped calls
def register_codes(codes: np.ndarray) -> None:
    """Upload a host array of 112-byte codes once (fbf_register_codes)."""
    assert codes.flags["C_CONTIGUOUS"]
    _check(lib().fbf_register_codes(_np_ptr(codes), codes.nbytes), "fbf_register_codes")


def unregister_codes(codes: np.ndarray) -> None:
    _check(lib().fbf_unregister_codes(_np_ptr(codes)), "fbf_unregister_codes")


def bruteforce(k: int, npoints: int, codes_in: np.ndarray, alpha: float, query_code: np.ndarray,
               offset_rows: int = 0) -> np.ndarray:
    """bruteforce (bruteforce.h:244-265): positions (relative to codes_in +
    offset_rows*112) of the k smallest symmetric-int8 distances, ascending."""
    out = np.zeros(k, np.int32)
    q = np.ascontiguousarray(query_code, np.int8)
    ptr = codes_in.ctypes.data + offset_rows * CODE_BYTES
    _check(lib().fbf_bruteforce(100, k, _np_ptr(out), npoints, _P(ptr), alpha, 1, _np_ptr(q)),
           "fbf_bruteforce")
    return out


def bruteforce_subgraph(k: int, npoints: int, codes_in: np.ndarray, alpha: float,
                        query_code: np.ndarray, offset_rows: int = 0):
    """bruteforce_subgraph (bruteforce.h:267-289): positions and int distances
    scaled by alpha^2, ascending."""
    out = np.zeros(k, np.int32)
    dist = np.zeros(k, np.float32)
    q = np.ascontiguousarray(query_code, np.int8)
    ptr = codes_in.ctypes.data + offset_rows * CODE_BYTES
    _check(lib().fbf_bruteforce_subgraph(100, k, _np_ptr(out), _np_ptr(dist), npoints, _P(ptr), alpha, 1,
                                         _np_ptr(q)), "fbf_bruteforce_subgraph")
    return out, dist


def bruteforce_batch(codes_base: np.ndarray, starts, counts, qcodes: np.ndarray, k: int):
    starts = np.ascontiguousarray(starts, np.uint64); counts = np.ascontiguousarray(counts, np.uint64)
    qcodes = np.ascontiguousarray(qcodes, np.int8)
    nq = starts.size
    pos = np.zeros((nq, k), np.int32); dist = np.zeros((nq, k), np.int32)
    _check(lib().fbf_bruteforce_batch(_np_ptr(codes_base), _np_ptr(starts), _np_ptr(counts), _np_ptr(qcodes),
                                      nq, k, _np_ptr(pos), _np_ptr(dist)), "fbf_bruteforce_batch")
    return pos, dist


def bruteforce_batch_noalloc(codes_base: np.ndarray, starts, counts, qcodes: np.ndarray, k: int, stream=None):
    """fbf_bruteforce_batch_noalloc: bruteforce_batch with its scratch in the
    library's fixed device array (no device allocation; codes_base must be
    registered).  stream: a torch stream or None (default stream)."""
    starts = np.ascontiguousarray(starts, np.uint64); counts = np.ascontiguousarray(counts, np.uint64)
    qcodes = np.ascontiguousarray(qcodes, np.int8)
    nq = starts.size
    pos = np.zeros((nq, k), np.int32); dist = np.zeros((nq, k), np.int32)
    st = None if stream is None else _P(stream.cuda_stream)
    _check(lib().fbf_bruteforce_batch_noalloc(_np_ptr(codes_base), _np_ptr(starts), _np_ptr(counts),
                                              _np_ptr(qcodes), nq, k, _np_ptr(pos), _np_ptr(dist), st),
           "fbf_bruteforce_batch_noalloc")
    return pos, dist


class SearchGraph:
    """One batch's search captured as a HIP graph (fbf_search_graph_create):
    ``launch()`` replays it, reading the query tensor given at creation and
    writing ``ids`` / ``dists``.  With metas as a numpy array the plan is made
    once on the host (the query vectors may change between replays, the
    filters may not); with metas as a torch CUDA tensor the graph routes and
    plans on the device at every replay, so both tensors may be refilled
    (BF16, FP8, SQ8, SQ8SYM modes)."""

    def __init__(self, ds: "Dataset", queries, metas, k: int = 100, mode="bf16", kprime: int = 0,
                 chunk_rows: int = 0, group_queries: int = 0, persistent: int = 0, prepass_div: int = 0,
                 prepass_margin: int = 0):
        import torch
        self.any_batch = isinstance(metas, torch.Tensor)
        if self.any_batch:
            assert metas.is_cuda and metas.dtype == torch.float32 and metas.is_contiguous()
            self.metas = metas  # kept alive: every replay reads it
            mp = _P(_dev_ptr(metas))
        else:
            metas = np.ascontiguousarray(metas, np.float32)
            mp = _np_ptr(metas)
        nq = metas.shape[0]
        assert queries.shape == (nq, ds.d) and queries.dtype == torch.float32 and queries.is_contiguous()
        self.queries = queries  # kept alive: the graph reads it
        self.ids = torch.empty((nq, k), dtype=torch.int32, device=queries.device)
        self.dists = torch.empty((nq, k), dtype=torch.float32, device=queries.device)
        o = Dataset._opts(mode, kprime, chunk_rows, group_queries, persistent, prepass_div, prepass_margin)
        o.metas_on_device = int(self.any_batch)
        h = _P()
        _check(lib().fbf_search_graph_create(ds.h, _P(_dev_ptr(queries)), mp, nq, k, C.byref(o),
                                             _P(_dev_ptr(self.ids)), _P(_dev_ptr(self.dists)), C.byref(h)),
               "fbf_search_graph_create")
        self.h = h

    def plan_info(self) -> dict:
        """Any-batch graphs, after a replay completed: the device plan's work
        items per pass and the overflow flag."""
        a, b, f = C.c_uint64(), C.c_uint64(), C.c_uint32()
        _check(lib().fbf_search_graph_plan_info(self.h, C.byref(a), C.byref(b), C.byref(f)),
               "fbf_search_graph_plan_info")
        return {"items_prefix": a.value, "items_main": b.value, "overflow": bool(f.value)}

    def launch(self, stream=None):
        import torch
        st = (stream or torch.cuda.current_stream(self.queries.device)).cuda_stream
        _check(lib().fbf_search_graph_launch(self.h, _P(st)), "fbf_search_graph_launch")
        return self.ids, self.dists

    def close(self) -> None:
        if getattr(self, "h", None):
            _check(lib().fbf_search_graph_destroy(self.h), "fbf_search_graph_destroy")
            self.h = None

    def __del__(self):
        if _sys is None or _sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass
