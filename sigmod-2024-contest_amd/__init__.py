"""MI355X filtered brute-force top-k (drop-in for the reference's brute-force
path).  The compute lives in lib/libfbf.so (HIP for gfx950, C-ABI in
include/fbf.h); ``fbf`` is its ctypes front end and ``datagen`` the seeded
synthetic workloads and contest file formats.

The directory name is not a Python identifier; load it with ``load()`` from
the repository root helpers (bench.py, tests/conftest.py), which registers it
as ``fbf_amd``.
"""
from . import autotune, datagen, fbf  # noqa: F401

__all__ = ["fbf", "datagen", "autotune"]
