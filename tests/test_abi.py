"""CPU checks of the C-ABI library: it loads, exports every entry point the
public header declares, and its host-side argument checks reject bad input
without touching a device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", f)).read()
            names |= set(re.findall(r"^\s*(?:fbf_status|const char \*)\s*(fbf_\w+)\s*\(", txt, re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol(pkg):
    L = pkg.fbf.lib()
    declared = _declared()
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(L, name), name
    assert set(declared) == set(pkg.fbf.EXPORTS)


def test_version_and_error_channel(pkg):
    assert "gfx950" in pkg.fbf.version()
    L = pkg.fbf.lib()
    # null dataset pointer: rejected on the host
    assert L.fbf_dataset_create(None, None, None, 0, 100, 1, C.byref(C.c_void_p())) == 1
    assert b"null" in L.fbf_last_error()


def test_bruteforce_argument_checks(pkg):
    L = pkg.fbf.lib()
    out = np.zeros(4, np.int32)
    q = np.zeros(112, np.int8)
    # k = 0 is a no-op, null query is rejected -- neither reaches a device
    assert L.fbf_bruteforce(100, 0, out.ctypes.data_as(C.c_void_p), 10, None, 1.0, 1,
                            q.ctypes.data_as(C.c_void_p)) == 0
    assert L.fbf_bruteforce(100, 4, out.ctypes.data_as(C.c_void_p), 10, None, 1.0, 1,
                            q.ctypes.data_as(C.c_void_p)) == 1
    assert L.fbf_unregister_codes(C.c_void_p(0x1234)) == 3
    # k beyond the 1024-key lists' room (k + one tile of appends): rejected
    # before any device work
    big = np.zeros(1000, np.int32)
    codes = np.zeros((10, 112), np.int8)
    assert L.fbf_bruteforce(100, 1000, big.ctypes.data_as(C.c_void_p), 10, codes.ctypes.data_as(C.c_void_p),
                            1.0, 1, q.ctypes.data_as(C.c_void_p)) == 1
    assert b"896" in L.fbf_last_error()
    st = np.zeros(1, np.uint64); cn = np.ones(1, np.uint64)
    assert L.fbf_bruteforce_batch(codes.ctypes.data_as(C.c_void_p), st.ctypes.data_as(C.c_void_p),
                                  cn.ctypes.data_as(C.c_void_p), q.ctypes.data_as(C.c_void_p), 1, 897,
                                  big.ctypes.data_as(C.c_void_p), big.ctypes.data_as(C.c_void_p)) == 1


def test_shim_header_compiles_against_abi(tmp_path):
    """The header-only C++ shims (bruteforce / bruteforce_subgraph with the
    reference signatures) compile and link against libfbf.so."""
    src = tmp_path / "t.cpp"
    src.write_text('#include "fbf_bruteforce.hpp"\n'
                   'int main(){ int ids[4]; float d[4]; int8_t q[112] = {0}; char c[112*4] = {0};\n'
                   '  (void)&fbf_ref::bruteforce; (void)&fbf_ref::bruteforce_subgraph; (void)ids; (void)d;\n'
                   '  (void)q; (void)c; return 0; }\n')
    lib_dir = os.path.join(ROOT, "sigmod-2024-contest_amd", "lib")
    r = os.system(f"g++ -std=c++17 -I{ROOT}/include {src} -L{lib_dir} -lfbf -Wl,-rpath,{lib_dir} "
                  f"-o {tmp_path}/t 2>{tmp_path}/err.txt")
    assert r == 0, open(f"{tmp_path}/err.txt").read()


def test_datagen_formats_roundtrip(pkg, tmp_path):
    dg = pkg.datagen
    cfg = dg.CONFIGS[0]
    v, lab, ts = dg.make_base(cfg, n=300)
    q, m = dg.make_queries(cfg, nq=40)
    dg.write_base_bin(str(tmp_path / "b.bin"), v, lab, ts)
    dg.write_query_bin(str(tmp_path / "q.bin"), q, m)
    v2, l2, t2 = dg.read_base_bin(str(tmp_path / "b.bin"))
    q2, m2 = dg.read_query_bin(str(tmp_path / "q.bin"))
    assert np.array_equal(v, v2) and np.array_equal(lab, l2) and np.array_equal(ts, t2)
    assert np.array_equal(q, q2) and np.array_equal(m, m2)
    assert set(np.unique(m[:, 0]).tolist()) <= {0, 1, 2, 3}
    assert np.all(m[m[:, 0] == 0, 1:] == -1)
    assert np.all((m[:, 3] >= m[:, 2]) | (m[:, 2] == -1))
