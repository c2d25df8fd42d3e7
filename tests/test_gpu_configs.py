"""Device search at the shapes of the benchmark configurations, on graphs the
device builder made from each configuration's data process, against the
oracle.

Bar: ids and squared-distance bits identical to the oracle's, per query:
orc_search_canon (the SQ8 path, qg.hpp:298-344, both output orders) and
orc_search_h2 (the FastScan path, qg.hpp:252-287) over the device's own
sidecar.  Shapes: 96-d unit-norm mixture at degree 32 with ef 50 and 200
(the 10M x 96 workload), 960-d non-negative mixture at degree 64 with
per-query ef in 64..1024 and k in {10, 100} in one launch (the 1M x 960
workload), 128-d integer mixture with k 100 (the 1M x 128 workload).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _built(tmp_path, kind, n, d, deg, ef_build, nq, kw, seed=0, iters=2):
    from liteqg_amd import datasets
    from liteqg_amd.symphonyqg import Index
    from oracle import oracle as O

    base, Q = datasets.make(kind, n, d, nq, seed=seed, **kw)
    ix = Index("QG", "L2", n, d, deg)
    ix.build_index(base, ef_build, num_iter=iters)
    p = tmp_path / "b.index"
    ix.save(str(p))
    return ix, O.read_index(str(p), n, d, deg), Q


def _same(ids, dis, e_ids, e_dis):
    return np.array_equal(ids, e_ids) and np.array_equal(
        np.ascontiguousarray(dis, np.float32).view(np.uint32),
        np.ascontiguousarray(e_dis, np.float32).view(np.uint32))


def _per_query(O, arr, H, ix, Q, ef, k, order):
    """Device batch with per-query ef and k against one oracle call per query."""
    ids, dis = ix.search_batch(Q, k, ef=ef, order=order, return_distances=True)
    kmax = int(k.max())
    for i in range(Q.shape[0]):
        ki = int(k[i])
        if H is None:
            e_ids, e_dis, _ = O.search_canon_batch(arr, Q[i:i + 1], int(ef[i]), ki, order)
        else:
            e_ids, e_dis, _ = O.search_h2_batch(H, Q[i:i + 1], int(ef[i]), ki)
        assert _same(ids[i, :ki], dis[i, :ki], e_ids[0], e_dis[0]), (i, int(ef[i]), ki)
        assert np.all(ids[i, ki:kmax] == 0xFFFFFFFF)


def _h2(O, arr, ix):
    s, codes, fac = ix.get_sidecar()
    return O.H2Arrays(arr, s, codes, fac, 1)


def test_unit_mixture_d96_degree32(oracle_mod, tmp_path):
    O = oracle_mod
    ix, arr, Q = _built(tmp_path, "unit_mix", 20000, 96, 32, 300, 200,
                        dict(centres=64, sigma=0.3, centre_law="normal"))
    for ef in (50, 200):
        for order in ("reference", "nearest"):
            ids, dis = ix.search_batch(Q, 10, ef=ef, order=order, return_distances=True)
            e_ids, e_dis, _ = O.search_canon_batch(arr, Q, ef, 10, order)
            assert _same(ids, dis, e_ids, e_dis), (ef, order)
    ix.build_sidecar(seed=3)
    ix.set_mode("fastscan")
    H = _h2(O, arr, ix)
    for ef in (50, 200):
        ids, dis = ix.search_batch(Q, 10, ef=ef, return_distances=True)
        e_ids, e_dis, _ = O.search_h2_batch(H, Q, ef, 10)
        assert _same(ids, dis, e_ids, e_dis), ef


def test_positive_mixture_d960_degree64_uneven(oracle_mod, tmp_path):
    O = oracle_mod
    ix, arr, Q = _built(tmp_path, "pos_mix", 3000, 960, 64, 500, 120,
                        dict(centres=16, sigma=0.05))
    rng = np.random.default_rng(9)
    ef = rng.integers(64, 1025, Q.shape[0]).astype(np.uint32)
    ef[:4] = (64, 1024, 513, 1000)
    k = rng.choice([10, 100], Q.shape[0]).astype(np.uint32)
    for order in ("nearest", "reference"):
        _per_query(O, arr, None, ix, Q, ef, k, order)
    ix.build_sidecar(seed=4)
    ix.set_mode("fastscan")
    _per_query(O, arr, _h2(O, arr, ix), ix, Q, ef, k, "nearest")


def test_integer_mixture_d128_k100(oracle_mod, tmp_path):
    """The integer-valued 128-d process (the integer kernel's exact dot
    products) at k 100 and ef 100 / 500, and FastScan beside it."""
    O = oracle_mod
    ix, arr, Q = _built(tmp_path, "int_mix", 20000, 128, 32, 400, 200,
                        dict(centres=32, sigma=20.0))
    for ef in (100, 500):
        ids, dis = ix.search_batch(Q, 100, ef=ef, order="nearest", return_distances=True)
        e_ids, e_dis, _ = O.search_canon_batch(arr, Q, ef, 100, "nearest")
        assert _same(ids, dis, e_ids, e_dis), ef
    ix.build_sidecar(seed=5)
    ix.set_mode("fastscan")
    H = _h2(O, arr, ix)
    ids, dis = ix.search_batch(Q, 100, ef=300, return_distances=True)
    e_ids, e_dis, _ = O.search_h2_batch(H, Q, 300, 100)
    assert _same(ids, dis, e_ids, e_dis)
