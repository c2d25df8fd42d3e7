import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
GOLDEN_INDEXES = ("g_uniform_n2000_d64", "g_intmix_n1200_d128", "g_unitmix_n1000_d96_deg64",
                  "g_ties_n1500_d64")
# fixtures whose evaluated distances never tie (ties need equal float sums;
# the ties fixture is built to have them everywhere)
TIE_FREE = ("g_uniform_n2000_d64", "g_intmix_n1200_d128", "g_unitmix_n1000_d96_deg64")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def load_golden(name):
    from oracle import oracle as O

    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    n, d, deg = (int(x) for x in z["shape"])
    ix = O.read_index(os.path.join(GOLDEN, name + ".index"), n, d, deg)
    return ix, z


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O

    O.build()
    return O


def make_random_index(n, d, deg, seed=0, kind="uniform", knn=False, dup_rate=0.0):
    """A small index with a random (or exact-kNN) graph, through the file writer."""
    from oracle import oracle as O

    rng = np.random.default_rng(seed)
    if kind == "int":
        base = rng.integers(0, 256, size=(n, d)).astype(np.float32)
    else:
        base = rng.random((n, d), dtype=np.float32)
    if knn:
        sq = (base.astype(np.float64) ** 2).sum(1)
        nbr = np.empty((n, deg), np.uint32)
        for s in range(0, n, 1024):
            e = min(n, s + 1024)
            dd = sq[s:e, None] + sq[None, :] - 2.0 * base[s:e].astype(np.float64) @ base.T.astype(np.float64)
            dd[np.arange(e - s), np.arange(s, e)] = np.inf
            nbr[s:e] = np.argsort(dd, axis=1)[:, :deg]
        # a few long-range edges so every node is reachable
        nbr[:, -4:] = rng.integers(0, n, size=(n, 4))
    else:
        nbr = rng.integers(0, n, size=(n, deg)).astype(np.uint32)
    if dup_rate > 0:
        m = rng.random((n, deg)) < dup_rate
        nbr[:, 1:][m[:, 1:]] = nbr[:, :1].repeat(deg - 1, 1)[m[:, 1:]]
    codes, params = O.sq8_quantize(base)
    ep = int(rng.integers(0, n))
    return O.IndexArrays(ep, base, nbr, codes, params), base


def gpu_available():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False
