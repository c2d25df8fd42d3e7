"""Seeded synthetic workloads for the five search configurations.

Every generator returns ``(base, queries)`` as C-contiguous float32 arrays.
The base stream uses ``numpy.random.default_rng(seed)`` and the query stream
``default_rng(seed + 1)``, so the queries come from the same process as the
base but are not copies of it.

==========  =========================================================
name        process
==========  =========================================================
uniform     i.i.d. U[0,1) (the reference README example, README.md:54-89)
int_mix     integer 0..255: round(clip(c + N(0, sigma^2), 0, 255)),
            centres U[0,255]^D
pos_mix     max(c + N(0, sigma^2), 0), centres U[0,1]^D
unit_mix    normalise(c + N(0, sigma^2)); centres normalised N(0, I)
            (centre_law="unit", the default) or N(0, I) as drawn
            (centre_law="normal")
==========  =========================================================

The unit-norm configurations name only sigma "before normalisation", not the
centres' law.  With unit-length centres, noise of sigma 0.25 per dimension
has norm ~0.25 sqrt(D) (6.9 at D = 768), so the points are close to isotropic
and a graph search reaches recall@10 ~0.2 at ef 200; the 96-d and 768-d
configurations therefore use centres ~ N(0, I) (per-dimension scale 1, the
scale sigma is quoted against), which gives the clustered unit vectors a
mixture is meant to produce.
"""
from __future__ import annotations

import numpy as np

# name -> (N, D, degree, build EF, number of queries, generator, kwargs)
CONFIGS = {
    "c1_uniform_100k_d64": dict(n=100_000, d=64, deg=32, ef_build=200, nq=1_000,
                                kind="uniform", kw={}),
    "c2_intmix_1m_d128": dict(n=1_000_000, d=128, deg=32, ef_build=400, nq=10_000,
                              kind="int_mix", kw=dict(centres=1024, sigma=20.0)),
    "c3_posmix_1m_d960": dict(n=1_000_000, d=960, deg=64, ef_build=500, nq=1_000,
                              kind="pos_mix", kw=dict(centres=512, sigma=0.05)),
    "c4_unitmix_10m_d96": dict(n=10_000_000, d=96, deg=32, ef_build=300, nq=10_000,
                               kind="unit_mix",
                               kw=dict(centres=4096, sigma=0.3, centre_law="normal")),
    "c5_unitmix_1m_d768": dict(n=1_000_000, d=768, deg=32, ef_build=None, nq=None,
                               kind="unit_mix",
                               kw=dict(centres=2048, sigma=0.25, centre_law="normal")),
}


def _chunks(n, step=65536):
    for s in range(0, n, step):
        yield s, min(n, s + step)


def uniform(n: int, d: int, rng: np.random.Generator) -> np.ndarray:
    return rng.random((n, d), dtype=np.float32)


def int_mix(n: int, d: int, rng: np.random.Generator, centres_arr: np.ndarray,
            sigma: float) -> np.ndarray:
    out = np.empty((n, d), np.float32)
    lab = rng.integers(0, centres_arr.shape[0], size=n)
    for s, e in _chunks(n):
        x = centres_arr[lab[s:e]] + rng.normal(0.0, sigma, size=(e - s, d))
        out[s:e] = np.rint(np.clip(x, 0.0, 255.0)).astype(np.float32)
    return out


def pos_mix(n: int, d: int, rng: np.random.Generator, centres_arr: np.ndarray,
            sigma: float) -> np.ndarray:
    out = np.empty((n, d), np.float32)
    lab = rng.integers(0, centres_arr.shape[0], size=n)
    for s, e in _chunks(n):
        x = centres_arr[lab[s:e]] + rng.normal(0.0, sigma, size=(e - s, d))
        out[s:e] = np.maximum(x, 0.0).astype(np.float32)
    return out


def unit_mix(n: int, d: int, rng: np.random.Generator, centres_arr: np.ndarray,
             sigma: float) -> np.ndarray:
    out = np.empty((n, d), np.float32)
    lab = rng.integers(0, centres_arr.shape[0], size=n)
    for s, e in _chunks(n):
        x = centres_arr[lab[s:e]] + rng.normal(0.0, sigma, size=(e - s, d))
        x /= np.linalg.norm(x, axis=1, keepdims=True)
        out[s:e] = x.astype(np.float32)
    return out


def make(kind: str, n: int, d: int, nq: int, seed: int = 0, **kw):
    """Generate ``(base[n, d], queries[nq, d])`` for one process kind."""
    rng = np.random.default_rng(seed)
    qrng = np.random.default_rng(seed + 1)
    if kind == "uniform":
        return uniform(n, d, rng), uniform(nq, d, qrng)
    c = int(kw.get("centres", 16))
    sigma = float(kw.get("sigma", 1.0))
    if kind == "int_mix":
        cen = rng.uniform(0.0, 255.0, size=(c, d))
        return int_mix(n, d, rng, cen, sigma), int_mix(nq, d, qrng, cen, sigma)
    if kind == "pos_mix":
        cen = rng.uniform(0.0, 1.0, size=(c, d))
        return pos_mix(n, d, rng, cen, sigma), pos_mix(nq, d, qrng, cen, sigma)
    if kind == "unit_mix":
        law = kw.get("centre_law", "unit")
        if law not in ("unit", "normal"):
            raise ValueError(f"centre_law must be 'unit' or 'normal', got {law!r}")
        cen = rng.normal(size=(c, d))
        if law == "unit":
            cen /= np.linalg.norm(cen, axis=1, keepdims=True)
        return unit_mix(n, d, rng, cen, sigma), unit_mix(nq, d, qrng, cen, sigma)
    raise ValueError(f"unknown process kind {kind!r}")


def config(name: str, seed: int = 0, n: int | None = None, nq: int | None = None):
    """Generate a named configuration, optionally shrunk to ``n`` / ``nq``."""
    c = CONFIGS[name]
    return make(c["kind"], n or c["n"], c["d"], nq or c["nq"] or 1000, seed, **c["kw"])
