"""Seeded synthetic workloads for the filtered top-k benchmark configurations.

Each configuration mirrors the contest data model (utils.h:257-276 base rows
[label, timestamp, 100 floats]; utils.h:550-566 queries [type, v, l, r, 100
floats]):

* base vectors: a mixture of isotropic Gaussians, centres ~ N(0, I), every row
  x = centre + 0.3 * N(0, I), component chosen uniformly;
* labels: integer categories with a Zipf(s) law over ``n_cat`` values
  (category 0 is the most frequent), stored as float32 like the contest files;
* timestamps: float32 uniform in [0, 1];
* queries: same mixture; type uniform over {0, 1, 2, 3}; v drawn from the
  category law; l ~ U[0, 1]; width from the configuration's law; r = min(l +
  width, 1); unused fields are -1 (as in the contest query file).

Seeds: base 1000 + config, queries 2000 + config (numpy PCG64).
"""
from __future__ import annotations

import dataclasses
import os
import struct

import numpy as np

DIM = 100


@dataclasses.dataclass(frozen=True)
class Config:
    name: str
    n: int
    nq: int
    centres: int
    n_cat: int
    zipf_s: float
    width_lo: float
    width_hi: float
    width_log: bool = False
    seed: int = 0
    uneven_batches: bool = False  # queries arrive in batches of log-uniform size (batch_sizes)


CONFIGS = {
    # Plain exact FP32, CPU-runnable (N=10k, 1k queries).
    0: Config("c0_n10k_q1k", 10_000, 1_000, 64, 100, 1.1, 0.05, 0.5, False, 0),
    # Full scale: N=10M, 100k queries (exact FP32 / SQ8 / BF16 / FP8 share it).
    1: Config("c1_n10m_q100k", 10_000_000, 100_000, 1024, 1000, 1.1, 0.01, 0.5, False, 1),
    # Uneven persistent batch: 10k Zipf(1.3) categories, log-uniform widths.
    4: Config("c4_n10m_zipf13", 10_000_000, 100_000, 1024, 10_000, 1.3, 0.001, 0.8, True, 4, True),
}


def zipf_pmf(n_cat: int, s: float) -> np.ndarray:
    p = np.arange(1, n_cat + 1, dtype=np.float64) ** (-s)
    return p / p.sum()


def _draw_categories(rng: np.random.Generator, pmf: np.ndarray, size: int) -> np.ndarray:
    cdf = np.cumsum(pmf)
    cdf[-1] = 1.0
    return np.searchsorted(cdf, rng.random(size), side="right").astype(np.int32)


def make_base(cfg: Config, n: int | None = None, chunk: int = 1 << 20):
    """Return (vecs float32 [n, 100], labels float32 [n], timestamps float32 [n])."""
    n = cfg.n if n is None else n
    rng = np.random.Generator(np.random.PCG64(1000 + cfg.seed))
    centres = rng.standard_normal((cfg.centres, DIM), dtype=np.float32)
    pmf = zipf_pmf(cfg.n_cat, cfg.zipf_s)
    vecs = np.empty((n, DIM), np.float32)
    labels = np.empty(n, np.float32)
    ts = np.empty(n, np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        m = e - s
        comp = rng.integers(0, cfg.centres, m)
        noise = rng.standard_normal((m, DIM), dtype=np.float32)
        noise *= np.float32(0.3)
        noise += centres[comp]
        vecs[s:e] = noise
        labels[s:e] = _draw_categories(rng, pmf, m).astype(np.float32)
        ts[s:e] = rng.random(m, dtype=np.float32)
    return vecs, labels, ts


def make_queries(cfg: Config, nq: int | None = None):
    """Return (queries float32 [nq, 100], metas float32 [nq, 4])."""
    nq = cfg.nq if nq is None else nq
    rng_c = np.random.Generator(np.random.PCG64(1000 + cfg.seed))
    centres = rng_c.standard_normal((cfg.centres, DIM), dtype=np.float32)
    rng = np.random.Generator(np.random.PCG64(2000 + cfg.seed))
    pmf = zipf_pmf(cfg.n_cat, cfg.zipf_s)
    comp = rng.integers(0, cfg.centres, nq)
    q = centres[comp] + np.float32(0.3) * rng.standard_normal((nq, DIM), dtype=np.float32)
    types = rng.integers(0, 4, nq)
    v = _draw_categories(rng, pmf, nq)
    l = rng.random(nq)
    if cfg.width_log:
        w = np.exp(rng.uniform(np.log(cfg.width_lo), np.log(cfg.width_hi), nq))
    else:
        w = rng.uniform(cfg.width_lo, cfg.width_hi, nq)
    r = np.minimum(l + w, 1.0)
    metas = np.full((nq, 4), -1.0, np.float32)
    metas[:, 0] = types
    has_v = (types == 1) | (types == 3)
    has_t = (types == 2) | (types == 3)
    metas[has_v, 1] = v[has_v]
    metas[has_t, 2] = l[has_t]
    metas[has_t, 3] = r[has_t]
    return q.astype(np.float32), metas


def batch_sizes(cfg: Config, nq: int, seed: int = 0) -> list[int]:
    """Uneven batch sizes, log-uniform in [1, 4096], summing to nq."""
    rng = np.random.Generator(np.random.PCG64(3000 + cfg.seed + seed))
    out, left = [], nq
    while left > 0:
        b = int(np.exp(rng.uniform(0.0, np.log(4096.0))))
        b = max(1, min(b, left))
        out.append(b)
        left -= b
    return out


# ---------------------------------------------------------------- contest file formats
def write_base_bin(path: str, vecs, labels, ts) -> None:
    """utils.h:257-276: u32 N, then N x [f32 label, f32 ts, 100 f32]."""
    n = vecs.shape[0]
    rec = np.empty((n, DIM + 2), np.float32)
    rec[:, 0] = labels
    rec[:, 1] = ts
    rec[:, 2:] = vecs
    with open(path, "wb") as f:
        f.write(struct.pack("<I", n))
        f.write(rec.tobytes())


def read_base_bin(path: str):
    with open(path, "rb") as f:
        n = struct.unpack("<I", f.read(4))[0]
        rec = np.frombuffer(f.read(), np.float32).reshape(n, DIM + 2)
    return np.ascontiguousarray(rec[:, 2:]), rec[:, 0].copy(), rec[:, 1].copy()


def write_query_bin(path: str, queries, metas) -> None:
    """utils.h:550-566: u32 N, then N x [f32 type, f32 v, f32 l, f32 r, 100 f32]."""
    n = queries.shape[0]
    rec = np.empty((n, DIM + 4), np.float32)
    rec[:, :4] = metas
    rec[:, 4:] = queries
    with open(path, "wb") as f:
        f.write(struct.pack("<I", n))
        f.write(rec.tobytes())


def read_query_bin(path: str):
    with open(path, "rb") as f:
        n = struct.unpack("<I", f.read(4))[0]
        rec = np.frombuffer(f.read(), np.float32).reshape(n, DIM + 4)
    return np.ascontiguousarray(rec[:, 4:]), np.ascontiguousarray(rec[:, :4])


def write_knn_bin(path: str, ids: np.ndarray) -> None:
    """io.h:22-33 SaveKNN: nq x 100 uint32, no header."""
    np.ascontiguousarray(ids, np.uint32).tofile(path)


def read_knn_bin(path: str, k: int = 100) -> np.ndarray:
    return np.fromfile(path, np.uint32).reshape(-1, k)


def write_gt_bin(path: str, gt: np.ndarray) -> None:
    """utils.h:168-188 SaveGroundTruth: u32 N, then N x 100 uint32."""
    with open(path, "wb") as f:
        f.write(struct.pack("<I", gt.shape[0]))
        f.write(np.ascontiguousarray(gt, np.uint32).tobytes())


def read_gt_bin(path: str, k: int = 100) -> np.ndarray:
    with open(path, "rb") as f:
        n = struct.unpack("<I", f.read(4))[0]
        return np.frombuffer(f.read(), np.uint32).reshape(n, k).copy()

