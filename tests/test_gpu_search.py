"""Device search against the oracle and the reference's golden outputs.

Bar: ids and squared distances bit-identical to the total-order oracle
(orc_search_canon) on every input; ids identical to the reference's own
search() output on every golden fixture whose distances do not tie.
"""
import numpy as np
import pytest

from conftest import GOLDEN, GOLDEN_INDEXES, TIE_FREE, load_golden, make_random_index

pytestmark = pytest.mark.gpu


def device_index(name):
    from liteqg_amd.symphonyqg import Index

    z = np.load(f"{GOLDEN}/{name}.npz")
    n, d, deg = (int(x) for x in z["shape"])
    ix = Index("QG", "L2", n, d, deg)
    ix.load(f"{GOLDEN}/{name}.index")
    return ix


def device_from_arrays(arr, tmp_path, tag="x"):
    from liteqg_amd.symphonyqg import Index
    from oracle import oracle as O

    p = tmp_path / f"{tag}.index"
    O.write_index(str(p), arr)
    ix = Index("QG", "L2", arr.n, arr.dim, arr.deg)
    ix.load(str(p))
    return ix


def oracle_batch(O, arr, Q, ef, k, order):
    res, dis, st = O.search_canon_batch(arr, Q, ef, k, order)
    return res, dis, st


def assert_same(got_ids, got_d, exp_ids, exp_d):
    assert np.array_equal(got_ids, exp_ids)
    assert np.array_equal(got_d.view(np.uint32), exp_d.astype(np.float32).view(np.uint32))


@pytest.mark.parametrize("name", GOLDEN_INDEXES)
def test_golden_reference_order(oracle_mod, name):
    arr, z = load_golden(name)
    ix = device_index(name)
    for ef in z["efs"]:
        ids, dis = ix.search_batch(z["queries"], 10, ef=int(ef), order="reference",
                                   return_distances=True)
        exp, exp_d, _ = oracle_batch(oracle_mod, arr, z["queries"], int(ef), 10, "reference")
        assert_same(ids, dis, exp, exp_d)
        if name in TIE_FREE:
            assert np.array_equal(ids, z[f"ids_ef{ef}"]), f"ef={ef} differs from reference"


@pytest.mark.parametrize("name", GOLDEN_INDEXES)
def test_golden_plain_search(oracle_mod, name):
    """Index.search (one query per launch) == reference search() on tie-free data."""
    arr, z = load_golden(name)
    ix = device_index(name)
    for ef in (10, 50):
        ix.set_ef(ef)
        got = np.stack([ix.search(q, 10) for q in z["queries"][:40]])
        exp, _, _ = oracle_batch(oracle_mod, arr, z["queries"][:40], ef, 10, "reference")
        assert np.array_equal(got, exp)
        if name in TIE_FREE:
            assert np.array_equal(got, z[f"ids_ef{ef}"][:40])


@pytest.mark.parametrize("name", TIE_FREE)
def test_golden_nearest_order(oracle_mod, name):
    arr, z = load_golden(name)
    ix = device_index(name)
    ids, dis = ix.search_batch(z["queries"], 10, ef=100, order="nearest", return_distances=True)
    exp, exp_d, _ = oracle_batch(oracle_mod, arr, z["queries"], 100, 10, "nearest")
    assert_same(ids, dis, exp, exp_d)
    # nearest order at ef=k is the reference output (the whole pool), reversed order not needed
    ids10 = ix.search_batch(z["queries"], 10, ef=10, order="nearest")
    assert np.array_equal(ids10, z["ids_ef10"])


CASES = [
    # n, d, deg, ef, k, kind, knn, dup_rate
    (20000, 128, 32, 100, 10, "int", True, 0.0),
    (20000, 64, 32, 200, 10, "uniform", True, 0.0),
    (20000, 128, 32, 150, 10, "int", True, 0.0),      # 192 class, integer kernel
    (5000, 64, 32, 192, 10, "uniform", True, 0.0),    # 192 class at its boundary
    (5000, 64, 32, 129, 10, "uniform", True, 0.0),    # 192 class just above 128
    (5000, 96, 64, 64, 10, "uniform", True, 0.0),
    (3000, 80, 32, 50, 10, "uniform", True, 0.0),      # D % 16 != 0: FMA tail
    (3000, 100, 32, 50, 10, "uniform", True, 0.0),     # D % 4 != 0: byte loads
    (3000, 32, 32, 1, 10, "uniform", True, 0.0),       # ef = 1
    (3000, 32, 32, 5, 20, "uniform", True, 0.0),       # k > ef: tail filled
    (4000, 64, 32, 1024, 100, "uniform", True, 0.0),   # largest ef class
    (3000, 64, 32, 300, 10, "uniform", False, 0.0),    # random graph
    (3000, 64, 32, 100, 10, "uniform", True, 0.3),     # repeated ids in lists
    (2000, 48, 96, 70, 10, "uniform", True, 0.0),      # deg > 64: two id chunks
    (3000, 768, 32, 300, 100, "uniform", True, 0.0),   # 768-d rows (several chunk loads)
    (2000, 960, 64, 500, 100, "uniform", True, 0.0),   # 960-d rows, degree 64
]


@pytest.mark.parametrize("case", CASES, ids=[f"n{c[0]}_d{c[1]}_g{c[2]}_ef{c[3]}_k{c[4]}"
                                              f"{'_dup' if c[7] else ''}{'' if c[6] else '_rand'}"
                                              for c in CASES])
def test_random_index_parity(oracle_mod, tmp_path, case):
    n, d, deg, ef, k, kind, knn, dup = case
    arr, base = make_random_index(n, d, deg, seed=n + d, kind=kind, knn=knn, dup_rate=dup)
    ix = device_from_arrays(arr, tmp_path)
    rng = np.random.default_rng(7)
    Q = (base[rng.integers(0, n, 64)] + rng.normal(0, 0.05 if kind == "uniform" else 10,
                                                   size=(64, d))).astype(np.float32)
    for order in ("reference", "nearest"):
        ids, dis, st = ix.search_batch(Q, k, ef=ef, order=order, return_distances=True,
                                       return_stats=True)
        exp, exp_d, est = oracle_batch(oracle_mod, arr, Q, ef, k, order)
        assert_same(ids, dis, exp, exp_d)
        assert st["expansions"] == int(est[:, 0].sum())
        assert st["evaluations"] >= int(est[:, 1].sum())


@pytest.mark.parametrize("prefetch", [1, 2, 3, 4])
@pytest.mark.parametrize("kind,d,deg,ef", [("int", 128, 32, 50), ("uniform", 96, 64, 300),
                                           ("uniform", 768, 32, 100)])
def test_latency_mode_matches_oracle(oracle_mod, tmp_path, prefetch, kind, d, deg, ef):
    """Latency mode (one query per workgroup, a second wave loading ahead, or
    at depth 4 evaluating ahead): batches of up to 256 queries (1,024 at
    D <= 256) and the plain per-query search give the oracle's ids and
    distance bits in both orders; a larger batch runs the batch kernel with
    the same results."""
    arr, base = make_random_index(3000, d, deg, seed=d + prefetch, kind=kind, knn=True)
    ix = device_from_arrays(arr, tmp_path)
    ix.set_launch(ef, waves_per_group=1, cache_bits=10, prefetch=prefetch)
    assert ix.get_launch(ef)["prefetch"] == prefetch
    rng = np.random.default_rng(5)
    nmax = 1025 if d <= 256 else 300
    Q = (base[rng.integers(0, 3000, nmax)] + rng.normal(0, 0.05 if kind == "uniform" else 10,
                                                        size=(nmax, d))).astype(np.float32)
    if kind == "int":
        Q = np.clip(np.rint(Q), 0, 255).astype(np.float32)
    for order in ("reference", "nearest"):
        for nq in (1, 37, 256, 300) + ((1024, 1025) if d <= 256 else ()):
            ids, dis = ix.search_batch(Q[:nq], 10, ef=ef, order=order, return_distances=True)
            exp, exp_d, _ = oracle_batch(oracle_mod, arr, Q[:nq], ef, 10, order)
            assert_same(ids, dis, exp, exp_d)
    ix.set_ef(ef)
    ix.set_order("nearest")
    exp, _, _ = oracle_batch(oracle_mod, arr, Q[:20], ef, 10, "nearest")
    assert np.array_equal(np.stack([ix.search(q, 10) for q in Q[:20]]), exp)


@pytest.mark.parametrize("split", [2, 4])
@pytest.mark.parametrize("kind,d,deg", [("int", 128, 32), ("uniform", 96, 64),
                                        ("uniform", 768, 32), ("uniform", 960, 64),
                                        ("uniform", 100, 32)])
def test_split_evaluation_matches_oracle(oracle_mod, tmp_path, split, kind, d, deg):
    """Split evaluation (one query per workgroup of 2 or 4 waves sharing each
    pass's distance evaluation): ids and distance bits equal the oracle's in
    both orders at ef 1, 50, 150, 300 and 600, with per-query ef and k in one
    launch, and for the plain search.  The integer index runs it with its
    integer kernel disabled; D = 100 takes the generic row loop."""
    n = 3000 if d <= 128 else 1500
    arr, base = make_random_index(n, d, deg, seed=d + split, kind=kind, knn=True)
    ix = device_from_arrays(arr, tmp_path)
    if kind == "int":
        ix.set_sq8_integer(False)
    ix.set_launch(0, waves_per_group=1, cache_bits=10, split=split)
    assert ix.get_launch(300)["split"] == split
    rng = np.random.default_rng(d)
    nq = 120
    Q = (base[rng.integers(0, n, nq)] + rng.normal(0, 0.05 if kind == "uniform" else 10,
                                                   size=(nq, d))).astype(np.float32)
    if kind == "int":
        Q = np.clip(np.rint(Q), 0, 255).astype(np.float32)
    for order in ("reference", "nearest"):
        for ef in (1, 50, 150, 300, 600):
            for m in (1, 37, nq):
                ids, dis = ix.search_batch(Q[:m], 10, ef=ef, order=order, return_distances=True)
                exp, exp_d, _ = oracle_batch(oracle_mod, arr, Q[:m], ef, 10, order)
                assert_same(ids, dis, exp, exp_d)
    efs = rng.integers(1, 1025, nq).astype(np.uint32)
    ks = rng.choice([1, 10, 100], nq).astype(np.uint32)
    ids, dis = ix.search_batch(Q, ks, ef=efs, order="nearest", return_distances=True)
    for i in range(nq):
        e_ids, e_d, _ = oracle_batch(oracle_mod, arr, Q[i:i + 1], int(efs[i]), int(ks[i]),
                                     "nearest")
        kk = int(ks[i])
        assert np.array_equal(ids[i, :kk], e_ids[0]), i
        assert np.array_equal(dis[i, :kk].view(np.uint32), e_d[0].view(np.uint32)), i
        assert (ids[i, kk:] == 0xFFFFFFFF).all()
    ix.set_ef(150)
    ix.set_order("nearest")
    exp, _, _ = oracle_batch(oracle_mod, arr, Q[:10], 150, 10, "nearest")
    assert np.array_equal(np.stack([ix.search(q, 10) for q in Q[:10]]), exp)


@pytest.mark.parametrize("kind,d,dup", [("int", 128, 0.0), ("uniform", 96, 0.1),
                                         ("uniform", 100, 0.0), ("uniform", 960, 0.0)])
def test_wide_pass_matches_oracle(oracle_mod, tmp_path, kind, d, dup):
    """The 1024-entry pool class takes a degree-64 node's neighbours in one
    pass of 64 ids (up to 64 fresh candidates evaluated together, one lane
    each past 32; the integer path in two calls): ids and distance bits equal
    the oracle's in both orders at ef 513..1024, k 10 and 100, with duplicate
    ids in the rows, in the batch kernel, the split kernel, latency mode
    (prefetch 2) and speculative latency mode (prefetch 4)."""
    n = 2500 if d <= 128 else 1200
    arr, base = make_random_index(n, d, 64, seed=d + 11, kind=kind, knn=True, dup_rate=dup)
    ix = device_from_arrays(arr, tmp_path)
    rng = np.random.default_rng(d + 12)
    nq = 40
    Q = (base[rng.integers(0, n, nq)] + rng.normal(0, 0.05 if kind == "uniform" else 10,
                                                   size=(nq, d))).astype(np.float32)
    if kind == "int":
        Q = np.clip(np.rint(Q), 0, 255).astype(np.float32)
    settings = [dict(waves_per_group=1, cache_bits=10), dict(waves_per_group=4, cache_bits=12),
                dict(waves_per_group=1, cache_bits=10, prefetch=2),
                dict(waves_per_group=1, cache_bits=10, prefetch=4)]
    if kind != "int":
        settings.append(dict(waves_per_group=1, cache_bits=11, split=2))
    for cfg in settings:
        ix.set_launch(1024, **cfg)
        for ef, k, order in ((600, 10, "nearest"), (1024, 100, "reference"), (777, 10, "reference")):
            for m in (1, nq):
                ids, dis = ix.search_batch(Q[:m], k, ef=ef, order=order, return_distances=True)
                exp, exp_d, _ = oracle_batch(oracle_mod, arr, Q[:m], ef, k, order)
                assert_same(ids, dis, exp, exp_d)
    ix.set_ef(1024)
    ix.set_order("nearest")
    exp, _, _ = oracle_batch(oracle_mod, arr, Q[:5], 1024, 10, "nearest")
    assert np.array_equal(np.stack([ix.search(q, 10) for q in Q[:5]]), exp)


def test_split_is_default_at_high_dimension(oracle_mod, tmp_path):
    """An index of D >= 512 splits its float-path evaluation over two waves
    by default; autotuning on it considers split launches and keeps results."""
    arr, base = make_random_index(1500, 512, 32, seed=3, knn=True)
    ix = device_from_arrays(arr, tmp_path)
    assert ix.get_launch(50)["split"] == 2
    rng = np.random.default_rng(4)
    Q = (base[rng.integers(0, 1500, 64)] + rng.normal(0, 0.05, (64, 512))).astype(np.float32)
    best = ix.autotune(Q, 50, 10)
    assert best["split"] in (1, 2, 4)
    exp, exp_d, _ = oracle_batch(oracle_mod, arr, Q, 50, 10, "nearest")
    ids, dis = ix.search_batch(Q, 10, ef=50, order="nearest", return_distances=True)
    assert_same(ids, dis, exp, exp_d)


def test_autotune_and_prefetch_depth(oracle_mod, tmp_path):
    """The default prefetch depth of small launches survives tuning on a large
    batch (above the 1,024 queries of a latency-mode launch at D <= 256, so
    it times the batch kernel only); tuning on a small set picks a depth
    among the latency-mode ones or the batch kernel; results stay the
    oracle's throughout."""
    arr, base = make_random_index(3000, 64, 32, seed=8, knn=True)
    ix = device_from_arrays(arr, tmp_path)
    rng = np.random.default_rng(9)
    Q = (base[rng.integers(0, 3000, 1100)] + rng.normal(0, 0.05, (1100, 64))).astype(np.float32)
    assert ix.get_launch(50)["prefetch"] == 2
    big = ix.autotune(Q, 50, 10)
    assert big["prefetch"] == 2 and ix.get_launch(50)["prefetch"] == 2
    small = ix.autotune(Q[:64], 50, 10)
    assert small["prefetch"] in (0, 1, 2, 3, 4)
    exp, exp_d, _ = oracle_batch(oracle_mod, arr, Q, 50, 10, "nearest")
    for nq in (1, 64, 400):
        ids, dis = ix.search_batch(Q[:nq], 10, ef=50, order="nearest", return_distances=True)
        assert_same(ids, dis, exp[:nq], exp_d[:nq])


def test_uneven_ef_and_k_in_one_launch(oracle_mod, tmp_path):
    arr, base = make_random_index(8000, 64, 32, seed=5, knn=True)
    ix = device_from_arrays(arr, tmp_path)
    rng = np.random.default_rng(3)
    nq = 300
    Q = (base[rng.integers(0, 8000, nq)] + rng.normal(0, 0.05, size=(nq, 64))).astype(np.float32)
    efs = rng.integers(1, 700, nq).astype(np.uint32)
    ks = rng.choice([1, 10, 100], nq).astype(np.uint32)
    ids, dis = ix.search_batch(Q, ks, ef=efs, order="nearest", return_distances=True)
    assert ids.shape == (nq, 100)
    for i in range(nq):
        e_ids, e_d, _ = oracle_batch(oracle_mod, arr, Q[i:i + 1], int(efs[i]), int(ks[i]), "nearest")
        kk = int(ks[i])
        assert np.array_equal(ids[i, :kk], e_ids[0]), i
        assert np.array_equal(dis[i, :kk].view(np.uint32), e_d[0].view(np.uint32)), i
        assert (ids[i, kk:] == 0xFFFFFFFF).all()


@pytest.mark.parametrize("wpg", [1, 4])
def test_batch_sizes_around_the_xcd_split(oracle_mod, tmp_path, wpg):
    """Batches from 1 to 257 queries: below 8 workgroups one queue serves the
    grid, from 8 up each XCD owns nq/8 queries and drains the others' ranges
    when its own is empty.  Every query is searched once, with the oracle's
    results, on the integer and on the float path."""
    for kind, d in (("int", 64), ("uniform", 48)):
        arr, base = make_random_index(4000, d, 32, seed=d, kind=kind, knn=True)
        ix = device_from_arrays(arr, tmp_path, tag=kind)
        ix.set_launch(0, waves_per_group=wpg, cache_bits=10)
        rng = np.random.default_rng(d)
        noise = 4.0 if kind == "int" else 0.05
        Qall = base[rng.integers(0, 4000, 257)] + rng.normal(0, noise, size=(257, d))
        if kind == "int":
            Qall = np.clip(np.rint(Qall), 0, 255)
        Qall = Qall.astype(np.float32)
        for nq in (1, 7, 8, 9, 33, 64, 257):
            Q = Qall[:nq]
            ids, dis, st = ix.search_batch(Q, 10, ef=40, order="nearest", return_distances=True,
                                           return_stats=True)
            exp, exp_d, est = oracle_batch(oracle_mod, arr, Q, 40, 10, "nearest")
            assert_same(ids, dis, exp, exp_d)
            assert st["queries"] == nq and st["expansions"] == int(est[:, 0].sum()), (kind, nq)
        ix.reset_launch(0)


def test_repeated_calls_identical(tmp_path):
    arr, base = make_random_index(6000, 128, 32, seed=9, kind="int", knn=True)
    ix = device_from_arrays(arr, tmp_path)
    Q = base[:500] + 1.0
    a = ix.search_batch(Q, 10, ef=64, order="nearest")
    b = ix.search_batch(Q, 10, ef=64, order="nearest")
    c = ix.search_batch(Q[:37], 10, ef=64, order="nearest")
    assert np.array_equal(a, b)
    assert np.array_equal(a[:37], c)


def test_recall_small_knn_graph(oracle_mod, tmp_path):
    """With the nearest order a kNN graph reaches high recall against FP64 truth."""
    arr, base = make_random_index(20000, 32, 32, seed=1, knn=True)
    ix = device_from_arrays(arr, tmp_path)
    rng = np.random.default_rng(2)
    Q = rng.random((200, 32), dtype=np.float32)
    gt, _ = oracle_mod.bruteforce_fp64(base, Q, 10)
    ids = ix.search_batch(Q, 10, ef=128, order="nearest")
    assert oracle_mod.recall_at_k(ids, gt, 10) > 0.9


@pytest.mark.parametrize("occ,bits,wpg", [(0, 4, 1), (0, 10, 4), (8, 9, 4), (8, 4, 2), (0, 14, 1)])
def test_launch_settings_do_not_change_results(oracle_mod, tmp_path, occ, bits, wpg):
    """Occupancy variant, cache size and group shape change speed only.  A tiny
    cache (16 entries) forgets almost everything, so most revisits are caught by
    the pool test instead of the cache."""
    arr, base = make_random_index(6000, 64, 32, seed=11, knn=True)
    ix = device_from_arrays(arr, tmp_path)
    rng = np.random.default_rng(4)
    Q = (base[rng.integers(0, 6000, 200)] + rng.normal(0, 0.05, size=(200, 64))).astype(np.float32)
    ix.set_launch(0, waves_per_group=wpg, cache_bits=bits, occupancy=occ)
    assert ix.get_launch(100)["cache_bits"] == bits
    for ef in (10, 100, 300):
        ids, dis, st = ix.search_batch(Q, 10, ef=ef, order="nearest", return_distances=True,
                                       return_stats=True)
        exp, exp_d, est = oracle_batch(oracle_mod, arr, Q, ef, 10, "nearest")
        assert_same(ids, dis, exp, exp_d)
        assert st["expansions"] == int(est[:, 0].sum())


@pytest.mark.parametrize("bits,prefetch", [(4, 0), (12, 0), (4, 2)])
def test_large_pool_ranked_admission(oracle_mod, tmp_path, bits, prefetch):
    """Pool class 1024 admits a pass's candidates by binary search over the
    pool's staging copy and a prefix sum of per-position counters.  With a
    16-entry visited cache the kernel re-evaluates many nodes (more
    evaluations than the oracle's distinct ones), and those still in the pool
    reach admission again (the duplicate drop); ids, distance bits and
    expansion counts equal the oracle's at ef 600 and 1024 in both orders, in
    the batch kernel and in latency mode."""
    arr, base = make_random_index(6000, 64, 32, seed=31, knn=True)
    ix = device_from_arrays(arr, tmp_path)
    rng = np.random.default_rng(6)
    Q = (base[rng.integers(0, 6000, 120)] + rng.normal(0, 0.05, size=(120, 64))).astype(np.float32)
    ix.set_launch(0, waves_per_group=2, cache_bits=bits, occupancy=0, prefetch=prefetch)
    for ef in (600, 1024):
        for order in ("nearest", "reference"):
            ids, dis, st = ix.search_batch(Q, 10, ef=ef, order=order, return_distances=True,
                                           return_stats=True)
            exp, exp_d, est = oracle_batch(oracle_mod, arr, Q, ef, 10, order)
            assert_same(ids, dis, exp, exp_d)
            assert st["expansions"] == int(est[:, 0].sum())
            if bits == 4:
                assert st["evaluations"] > int(est[:, 1].sum())


def test_autotune_keeps_results(oracle_mod, tmp_path):
    arr, base = make_random_index(8000, 128, 32, seed=12, kind="int", knn=True)
    ix = device_from_arrays(arr, tmp_path)
    Q = base[:400] + 3.0
    before = ix.search_batch(Q, 10, ef=80, order="nearest")
    best = ix.autotune(Q, 80, 10)
    assert ix.get_launch(80) == best
    assert np.array_equal(ix.search_batch(Q, 10, ef=80, order="nearest"), before)


@pytest.mark.parametrize("d", [128, 64, 256, 272])
def test_integer_path_matches_oracle(oracle_mod, tmp_path, d):
    """Integer data with min 0 / scale 1 SQ8 params and integer queries take the
    v_dot4 path (d <= 256); the result must be bit-identical to the FP32 order
    of the reference.  Queries with a fractional or out-of-range value in the
    same batch take the FP32 path."""
    arr, base = make_random_index(12000, d, 32, seed=d, kind="int", knn=True)
    assert (arr.params[:, 0] == 0).all() and (arr.params[:, 1] == 1).all()
    ix = device_from_arrays(arr, tmp_path)
    rng = np.random.default_rng(d)
    Q = np.clip(np.rint(base[rng.integers(0, 12000, 96)] + rng.normal(0, 30, size=(96, d))),
                0, 255).astype(np.float32)
    Q[5, 3] += 0.5          # fractional: FP32 path
    Q[9, 0] = 256.0         # out of byte range: FP32 path
    Q[11, :] = 0.0          # all zeros: integer path
    Q[12, :] = 255.0
    for ef in (16, 64, 200):
        for order in ("nearest", "reference"):
            ids, dis, st = ix.search_batch(Q, 10, ef=ef, order=order, return_distances=True,
                                           return_stats=True)
            exp, exp_d, est = oracle_batch(oracle_mod, arr, Q, ef, 10, order)
            assert_same(ids, dis, exp, exp_d)
            assert st["expansions"] == int(est[:, 0].sum())


@pytest.mark.parametrize("name", GOLDEN_INDEXES)
def test_count_visits_equals_reference_visited_set(oracle_mod, name):
    """E and the number of distinct nodes evaluated equal the oracle's exact
    visited-set counts (the reference's own E and V on tie-free inputs)."""
    arr, z = load_golden(name)
    ix = device_index(name)
    for ef in (10, 50, 100):
        got = ix.count_visits(z["queries"], ef, 10).astype(np.int64)
        _, _, est = oracle_batch(oracle_mod, arr, z["queries"], ef, 10, "nearest")
        assert np.array_equal(got, est.astype(np.int64)), ef
        if name in TIE_FREE:
            _, _, _, (E, V) = oracle_mod.search_heap(arr, z["queries"][0], ef, 10)
            assert (got[0, 0], got[0, 1]) == (E, V)


def test_integer_index_full_device_batch(oracle_mod, tmp_path):
    """An integer index (exact SQ8 decode) with a batch large enough to fill
    the device: integer-valued queries take the integer kernel, the others the
    float kernel that follows it; results, and the exact visit counts (one
    bitmap per wave of the larger of the two grids), equal the oracle's."""
    arr, base = make_random_index(3000, 64, 32, seed=17, kind="int", knn=True)
    assert np.all(arr.params[:, 0] == 0) and np.all(arr.params[:, 1] == 1)
    ix = device_from_arrays(arr, tmp_path)
    rng = np.random.default_rng(4)
    nq = 12000
    Q = np.clip(base[rng.integers(0, 3000, nq)] + rng.integers(-3, 4, (nq, 64)), 0, 255)
    Q = Q.astype(np.float32)
    Q[::7] += 0.25  # every 7th query is not integer valued
    ids, dis = ix.search_batch(Q, 10, ef=50, order="nearest", return_distances=True)
    exp, exp_d, est = oracle_batch(oracle_mod, arr, Q, 50, 10, "nearest")
    assert_same(ids, dis, exp, exp_d)
    got = ix.count_visits(Q, 50, 10).astype(np.int64)
    assert np.array_equal(got, est.astype(np.int64))


def test_count_visits_with_tiny_cache(oracle_mod, tmp_path):
    """With a 16-entry cache the kernel re-evaluates many nodes; the exact count
    still equals the oracle's and is below the evaluation count."""
    arr, base = make_random_index(6000, 64, 32, seed=21, knn=True)
    ix = device_from_arrays(arr, tmp_path)
    Q = base[:150] + 0.01
    ix.set_launch(0, waves_per_group=4, cache_bits=4, occupancy=0)
    got = ix.count_visits(Q, 100, 10).astype(np.int64)
    _, _, est = oracle_batch(oracle_mod, arr, Q, 100, 10, "nearest")
    assert np.array_equal(got, est.astype(np.int64))
    _, st = ix.search_batch(Q, 10, ef=100, order="nearest", return_stats=True)
    assert st["evaluations"] > int(got[:, 1].sum())


def test_plain_search_on_integer_index(oracle_mod, tmp_path):
    """One query per call on an index with exact SQ8 decode: an integer-valued
    query runs the integer launch alone and any other the float launch alone
    (the host makes the kernels' test); both equal the batched search and the
    oracle, in both orders."""
    arr, base = make_random_index(4000, 128, 32, seed=23, kind="int", knn=True)
    ix = device_from_arrays(arr, tmp_path)
    rng = np.random.default_rng(8)
    Q = np.clip(np.rint(base[rng.integers(0, 4000, 24)] + rng.normal(0, 15, (24, 128))), 0, 255)
    Q = Q.astype(np.float32)
    Q[1::3, 5] += 0.5      # fractional
    Q[2::6, 0] = 256.0     # out of byte range
    for order in ("nearest", "reference"):
        ix.set_order(order)
        for ef in (10, 64):
            ix.set_ef(ef)
            got = np.stack([ix.search(q, 10) for q in Q])
            want = ix.search_batch(Q, 10, ef=ef, order=order)
            exp, _, _ = oracle_batch(oracle_mod, arr, Q, ef, 10, order)
            assert np.array_equal(got, want), (order, ef)
            assert np.array_equal(got, exp), (order, ef)
