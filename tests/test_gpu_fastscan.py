"""Device FastScan RaBitQ path (search_qg, qg.hpp:252-287) against the
reference's outputs and the oracle.

Bar: query preparation (LUT bytes, lo, width, sumq), exact distances and
FastScan estimates bit-identical to the reference's on the h2_* fixtures
(which hold the reference's own sidecar); result ids identical to the
reference's search_qg on every fixture query and ef; on random indexes, ids,
distances and (expansions, pops) identical to the oracle's orc_search_h2 over
the same sidecar; the device-built sidecar equal to the oracle's (codes bit
for bit, factors within FACTOR_TOL of their magnitude).
"""
import numpy as np
import pytest

from conftest import GOLDEN, TIE_FREE, load_golden, make_random_index

pytestmark = pytest.mark.gpu

H2_NAMES = TIE_FREE
# Factors are FP64 sums rounded to float; the device and the oracle sum in the
# same order, so they are expected to agree exactly.  Allowed: 2e-6 of the
# largest magnitude of each factor array (the oracle-vs-reference bound).
FACTOR_TOL = 2e-6


def _index_from_golden(name):
    from liteqg_amd.symphonyqg import Index

    z = np.load(f"{GOLDEN}/{name}.npz")
    n, d, deg = (int(x) for x in z["shape"])
    ix = Index("QG", "L2", n, d, deg)
    ix.load(f"{GOLDEN}/{name}.index")
    return ix


def _index_from_arrays(arr, tmp_path, tag):
    from liteqg_amd.symphonyqg import Index
    from oracle import oracle as O

    p = tmp_path / f"{tag}.index"
    O.write_index(str(p), arr)
    ix = Index("QG", "L2", arr.n, arr.dim, arr.deg)
    ix.load(str(p))
    return ix


def _signs(p, seed):
    rng = np.random.default_rng(seed)
    return ((rng.integers(0, 2, p) * 2 - 1) / np.sqrt(np.float32(p))).astype(np.float32)


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("name", H2_NAMES)
def test_probe_matches_reference(name):
    """LUT, lo, width, sumq, sqr_y and every estimate of the entry point's row,
    computed on the device from the reference's sidecar, equal the
    reference's (qg_query.hpp:28-43, qg_scanner.hpp:99-155)."""
    arr, _ = load_golden(name)
    h = np.load(f"{GOLDEN}/h2_{name}.npz")
    ix = _index_from_golden(name)
    ix.set_sidecar(h["signs"], h["codes"], h["factors"], layout="blocks")
    nprep = h["prep_lut"].shape[0]
    out = ix.fastscan_probe(h["queries"][:nprep], nodes=np.full(nprep, arr.ep, np.uint32))
    assert np.array_equal(out["lut"], h["prep_lut"])
    assert np.array_equal(_bits(out["lo"]), _bits(h["prep_lo"]))
    assert np.array_equal(_bits(out["width"]), _bits(h["prep_width"]))
    assert np.array_equal(out["sumq"], h["prep_sumq"])
    assert np.array_equal(_bits(out["sqr_y"]), _bits(h["prep_sqr_y"]))
    assert np.array_equal(_bits(out["est"]), _bits(h["prep_est"]))


@pytest.mark.parametrize("name", H2_NAMES)
def test_search_matches_reference(name):
    """search_qg ids for every fixture query at every ef, reference sidecar."""
    h = np.load(f"{GOLDEN}/h2_{name}.npz")
    ix = _index_from_golden(name)
    ix.set_sidecar(h["signs"], h["codes"], h["factors"], layout="blocks")
    ix.set_mode("fastscan")
    k = int(h["k"])
    for ef in h["efs"]:
        ids = ix.search_batch(h["queries"], k, ef=int(ef))
        assert np.array_equal(ids, h[f"ids_ef{ef}"]), f"ef={ef}"
    # the plain per-query entry point runs the same path
    ix.set_ef(int(h["efs"][1]))
    for i in range(5):
        assert np.array_equal(ix.search(h["queries"][i], k), h[f"ids_ef{h['efs'][1]}"][i])


@pytest.mark.parametrize("name", H2_NAMES)
def test_update_results_matches_reference(oracle_mod, name):
    """Searches that expand fewer than k nodes (ef 1-4 at k 10, ef 16/64 at
    k 100): the update_results tail (qg.hpp:395-418), including the probe of
    node 0 through the zero-filled id snapshot (buffer.hpp:81-102).  Ids equal
    the reference's search_qg (h2u_*.npz) and ids and distance bits equal
    the oracle's, in one launch per case and in one launch holding every
    case."""
    arr, _ = load_golden(name)
    h = np.load(f"{GOLDEN}/h2_{name}.npz")
    u = np.load(f"{GOLDEN}/h2u_{name}.npz")
    ix = _index_from_golden(name)
    ix.set_sidecar(h["signs"], h["codes"], h["factors"], layout="blocks")
    ix.set_mode("fastscan")
    H = oracle_mod.H2Arrays(arr, h["signs"], h["codes"], h["factors"], 0)
    q = u["queries"]
    efs, ks = [], []
    for ef, k in u["cases"]:
        ids, dis = ix.search_batch(q, int(k), ef=int(ef), return_distances=True)
        assert np.array_equal(ids, u[f"ids_ef{ef}_k{k}"]), f"ef={ef} k={k}"
        e_ids, e_dis, _ = oracle_mod.search_h2_batch(H, q, int(ef), int(k))
        assert np.array_equal(ids, e_ids)
        assert np.array_equal(_bits(dis), _bits(e_dis))
        efs.append(np.full(q.shape[0], ef, np.uint32))
        ks.append(np.full(q.shape[0], k, np.uint32))
    # every case in one launch: per-query ef and k
    Q = np.concatenate([q] * len(u["cases"]))
    ids = ix.search_batch(Q, np.concatenate(ks), ef=np.concatenate(efs))
    for c, (ef, k) in enumerate(u["cases"]):
        got = ids[c * q.shape[0]:(c + 1) * q.shape[0], :k]
        assert np.array_equal(got, u[f"ids_ef{ef}_k{k}"]), f"mixed launch ef={ef} k={k}"


@pytest.mark.parametrize("name", H2_NAMES)
def test_expansions_match_oracle(oracle_mod, name):
    """The expanded-node lists the FastScan-candidate graph build takes
    (find_candidates, qg.hpp:439-473) equal the oracle's, in order, on the
    reference's sidecar; a cap below the count keeps the first entries."""
    arr, _ = load_golden(name)
    h = np.load(f"{GOLDEN}/h2_{name}.npz")
    ix = _index_from_golden(name)
    ix.set_sidecar(h["signs"], h["codes"], h["factors"], layout="blocks")
    H = oracle_mod.H2Arrays(arr, h["signs"], h["codes"], h["factors"], 0)
    for ef in (10, 50, 100, 400):
        got = ix.fastscan_expansions(h["queries"], ef, 1024)
        assert np.array_equal(got, oracle_mod.h2_expansions_batch(H, h["queries"], ef, 1024)), ef
    assert np.array_equal(ix.fastscan_expansions(h["queries"], 50, 9),
                          oracle_mod.h2_expansions_batch(H, h["queries"], 50, 9))


@pytest.mark.parametrize("d,deg", [(128, 32), (960, 64)])
def test_expansions_match_oracle_device_sidecar(oracle_mod, tmp_path, d, deg):
    arr, base = make_random_index(1200 if d < 900 else 500, d, deg, seed=d, knn=True)
    ix = _index_from_arrays(arr, tmp_path, "e")
    ix.build_sidecar(seed=2)
    H = _h2arr(oracle_mod, arr, ix)
    rng = np.random.default_rng(4)
    Q = (base[rng.integers(0, arr.n, 40)] + rng.normal(0, 0.02, (40, d))).astype(np.float32)
    for ef in (20, 200):
        assert np.array_equal(ix.fastscan_expansions(Q, ef, 960),
                              oracle_mod.h2_expansions_batch(H, Q, ef, 960)), ef


def _h2arr(O, arr, ix):
    s, codes, fac = ix.get_sidecar()
    return O.H2Arrays(arr, s, codes, fac, 1)


@pytest.mark.parametrize("name", H2_NAMES)
def test_device_sidecar_matches_oracle_and_reference(oracle_mod, name):
    """The device-built sidecar equals the oracle's; searching with it gives
    the reference's ids on the fixtures."""
    O = oracle_mod
    arr, _ = load_golden(name)
    h = np.load(f"{GOLDEN}/h2_{name}.npz")
    ix = _index_from_golden(name)
    ix.build_sidecar(signs=h["signs"])
    s, codes, fac = ix.get_sidecar()
    assert np.array_equal(s, h["signs"])
    oc, of = O.sidecar_build(arr, h["signs"])
    assert np.array_equal(codes, oc)
    g = arr.deg
    a, b = fac.reshape(-1, 3, g), of.reshape(-1, 3, g)
    for j in range(3):
        assert np.abs(a[:, j] - b[:, j]).max() <= FACTOR_TOL * np.abs(b[:, j]).max()
    ix.set_mode("fastscan")
    for ef in h["efs"]:
        assert np.array_equal(ix.search_batch(h["queries"], int(h["k"]), ef=int(ef)),
                              h[f"ids_ef{ef}"]), f"ef={ef}"


def _oracle_h2(O, arr, ix, Q, ef, k):
    s, codes, fac = ix.get_sidecar()
    H = O.H2Arrays(arr, s, codes, fac, 1)
    return O.search_h2_batch(H, Q, ef, k)


# (n, d, deg, kind, knn, dup_rate): covers code words per lane 1, 2, 4, 8, 16
# and 32, a dimension that is not a multiple of 16 (tail FMA), duplicate ids
# inside neighbour rows, and random as well as kNN graphs
SHAPES = [
    (1500, 64, 32, "uniform", True, 0.0),
    (1500, 100, 32, "uniform", True, 0.05),
    (1000, 96, 32, "uniform", True, 0.05),
    (1200, 128, 32, "int", True, 0.0),
    (1000, 96, 64, "uniform", False, 0.1),
    (800, 200, 32, "uniform", True, 0.0),
    (800, 256, 64, "uniform", True, 0.0),
    (500, 768, 32, "uniform", True, 0.0),
    (400, 960, 64, "uniform", True, 0.0),
]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"n{s[0]}_d{s[1]}_deg{s[2]}")
def test_search_matches_oracle(oracle_mod, tmp_path, shape):
    O = oracle_mod
    n, d, deg, kind, knn, dup = shape
    arr, _ = make_random_index(n, d, deg, seed=d + deg, kind=kind, knn=knn, dup_rate=dup)
    ix = _index_from_arrays(arr, tmp_path, "r")
    p = ix.padded_dim
    ix.build_sidecar(signs=_signs(p, d))
    _, codes, fac = ix.get_sidecar()
    oc, of = O.sidecar_build(arr, _signs(p, d))
    assert np.array_equal(codes, oc)
    ix.set_mode("fastscan")
    rng = np.random.default_rng(d)
    Q = (arr.raw[rng.integers(0, n, 40)] + rng.normal(0, 0.05, (40, d))).astype(np.float32)
    for ef, k in ((16, 10), (64, 10), (150, 10), (192, 20), (200, 20), (700, 10)):
        ids, dis, st = ix.search_batch(Q, k, ef=ef, return_distances=True, return_stats=True)
        e_ids, e_dis, e_st = _oracle_h2(O, arr, ix, Q, ef, k)
        assert np.array_equal(ids, e_ids), f"ef={ef}"
        assert np.array_equal(_bits(dis), _bits(e_dis)), f"ef={ef}"
        assert st["expansions"] == int(e_st[:, 0].sum())
        assert st["evaluations"] == int(e_st[:, 1].sum())


@pytest.mark.parametrize("n,d,deg", [(1500, 128, 32), (1500, 100, 32), (800, 256, 64),
                                     (800, 200, 32), (500, 768, 32), (400, 960, 64)])
def test_bit_plane_scan_and_lut_wrap_match_oracle(oracle_mod, tmp_path, n, d, deg):
    """From P = 256 on, the general kernel sums a neighbour's LUT entries by
    the query's bit planes, except for a query with a codebook whose four
    quantised values add up to 256 or more: its uint8 LUT entry wraps, and
    the query keeps the LUT scan (as every query at P <= 128 does).  Scaled unit vectors c e_j rotate to vectors whose entries
    are all +-c s_j, so half or all of them quantise to 64 (or 63): they
    reach the wrap.  Estimates (probe kernel) and search results equal the
    oracle's for queries on both sides; D = 128 forces the general kernel
    over the specialised one."""
    O = oracle_mod
    arr, _ = make_random_index(n, d, deg, seed=d + 7, knn=True)
    ix = _index_from_arrays(arr, tmp_path, "w")
    p = ix.padded_dim
    signs = _signs(p, d + 1)
    ix.build_sidecar(signs=signs)
    ix.set_generic_kernel(True)
    ix.set_mode("fastscan")
    rng = np.random.default_rng(d + 2)
    def wraps(q):
        # the single-bit entries are the codebook's q values (exact); the
        # all-bits entry is their uint8 sum
        lut = O.query_prepare(signs, q, p)[5].reshape(p // 4, 16).astype(np.int64)
        return bool(((lut[:, 1] + lut[:, 2] + lut[:, 4] + lut[:, 8]) >= 256).any())

    # columns 0, 4 and 64 of the Hadamard matrix have runs of equal signs at
    # least a codebook long; whether the top entries quantise to 64 depends on
    # the scale's rounding, so scales are tried until some queries wrap
    Q = [(arr.raw[rng.integers(0, n, 20)] + rng.normal(0, 0.05, (20, d))).astype(np.float32)]
    picked = {True: 0, False: 0}
    for j in (0, 4, min(64, d - 1)):
        for c in np.geomspace(0.1, 10, 40).astype(np.float32):
            for sg in (1.0, -1.0):
                e = np.zeros((1, d), np.float32)
                e[0, j] = sg * c
                w = wraps(e[0])
                if picked[w] < 8:
                    picked[w] += 1
                    Q.append(e)
    assert picked[True] >= 4 and picked[False] >= 4
    Q = np.concatenate(Q)
    wrap = [wraps(q) for q in Q]
    nodes = rng.integers(0, n, Q.shape[0]).astype(np.uint32)
    pr = ix.fastscan_probe(Q, nodes)
    _, codes, fac = ix.get_sidecar()
    H = O.H2Arrays(arr, signs, codes, fac, 1)
    for i, q in enumerate(Q):
        _, lo, width, sumq, _, lut = O.query_prepare(signs, q, p)
        node = int(nodes[i])
        _, est = O.scan_estimates(codes[node], 1, fac[node], deg, p, lut, float(pr["sqr_y"][i]),
                                  lo, width, sumq)
        assert np.array_equal(_bits(pr["est"][i]), _bits(est)), (i, wrap[i])
    for ef, k in ((32, 10), (200, 20)):
        ids, dis = ix.search_batch(Q, k, ef=ef, return_distances=True)
        e_ids, e_dis, _ = O.search_h2_batch(H, Q, ef, k)
        assert np.array_equal(ids, e_ids), f"ef={ef}"
        assert np.array_equal(_bits(dis), _bits(e_dis)), f"ef={ef}"


@pytest.mark.parametrize("shape", [s for s in SHAPES if s[1] <= 256],
                         ids=lambda s: f"n{s[0]}_d{s[1]}_deg{s[2]}")
def test_occupancy_variants_match_oracle(oracle_mod, tmp_path, shape):
    """The uncapped and the register-capped kernels (6 and 8 waves per SIMD,
    code words per lane <= 4) give the oracle's ids, distances and counts, also
    with an LDS visited table small enough (64 entries at ef 256) that queries
    go to the overflow kernel and its device table; P/lanes above 128 bits
    rejects the cap."""
    from liteqg_amd._lib import LqgError

    O = oracle_mod
    n, d, deg, kind, knn, dup = shape
    arr, _ = make_random_index(n, d, deg, seed=d + deg, kind=kind, knn=knn, dup_rate=dup)
    ix = _index_from_arrays(arr, tmp_path, "c")
    ix.build_sidecar(signs=_signs(ix.padded_dim, d))
    ix.set_mode("fastscan")
    rng = np.random.default_rng(d + 1)
    Q = (arr.raw[rng.integers(0, n, 40)] + rng.normal(0, 0.05, (40, d))).astype(np.float32)
    cw = ix.padded_dim // 8 // (1 if deg >= 64 else 2) // 4
    if cw > 4:
        with pytest.raises(LqgError) as e:
            ix.set_launch(0, waves_per_group=2, cache_bits=10, occupancy=6)
        assert e.value.status == 6
        return
    # degree 32 at D = 64, 96, 128 has a kernel specialised for the record
    # shape; the general kernel is run beside it
    kernels = (False, True) if deg == 32 and d in (64, 96, 128) else (False,)
    for generic in kernels:
        ix.set_generic_kernel(generic)
        for occ in (0, 6, 8):
            for ef, k, bits in ((64, 10, 10), (150, 10, 10), (200, 20, 10), (256, 10, 6),
                                (100, 100, 10)):
                ix.set_launch(0, waves_per_group=4, cache_bits=bits, occupancy=occ)
                assert ix.get_launch(ef)["occupancy"] == occ
                if occ == 8 and k > 64:  # the 8-wave cap has no k > 64 kernel
                    with pytest.raises(LqgError) as e:
                        ix.search_batch(Q, k, ef=ef)
                    assert e.value.status == 6
                    continue
                ids, dis, st = ix.search_batch(Q, k, ef=ef, return_distances=True,
                                               return_stats=True)
                e_ids, e_dis, e_st = _oracle_h2(O, arr, ix, Q, ef, k)
                assert np.array_equal(ids, e_ids), (generic, occ, ef, k)
                assert np.array_equal(_bits(dis), _bits(e_dis)), (generic, occ, ef, k)
                assert st["expansions"] == int(e_st[:, 0].sum()), (generic, occ, ef, k)
                assert st["evaluations"] == int(e_st[:, 1].sum()), (generic, occ, ef, k)
    ix.set_generic_kernel(False)
    with pytest.raises(LqgError) as e:
        ix.set_launch(0, waves_per_group=2, cache_bits=10, occupancy=5)
    assert e.value.status == 1
    ix.reset_launch(0)


# Max |GEMM - FP64| over max |FP64| of each rotated query.  FP32 sums of D
# terms, each at most max|q| / sqrt(P): the rounding stays within a small
# multiple of sqrt(D) ulps of the largest value, far below this bound; a wrong
# operand layout or sign pattern gives errors of order 1.
GEMM_ROT_TOL = 1e-5


@pytest.mark.parametrize("shape", [(1500, 64, 32), (1200, 128, 32), (800, 200, 32),
                                   (500, 768, 32), (400, 960, 64)],
                         ids=lambda s: f"n{s[0]}_d{s[1]}_deg{s[2]}")
def test_rotation_fht_and_gemm(oracle_mod, tmp_path, shape):
    """Query rotation: the device FHT equals the oracle's (the reference's
    FHTRotator, bit for bit); the MFMA GEMM is within GEMM_ROT_TOL of an FP64
    rotation H_P diag(signs) q; a FastScan search with GEMM-rotated queries
    finds at least 98% of the FHT search's ids (mean per-query overlap)."""
    import scipy.linalg

    O = oracle_mod
    n, d, deg = shape
    arr, _ = make_random_index(n, d, deg, seed=d + 7, knn=True)
    ix = _index_from_arrays(arr, tmp_path, "rot")
    p = ix.padded_dim
    signs = _signs(p, d + 1)
    ix.build_sidecar(signs=signs)
    rng = np.random.default_rng(d)
    Q = (arr.raw[rng.integers(0, n, 37)] + rng.normal(0, 0.05, (37, d))).astype(np.float32)
    fht = ix.rotate_queries(Q, "fht")
    gemm = ix.rotate_queries(Q, "gemm")
    assert fht.shape == (37, p) and gemm.shape == (37, p)
    for i in range(Q.shape[0]):
        assert np.array_equal(_bits(fht[i]), _bits(O.fht_rotate(signs, Q[i], p)))
    H = scipy.linalg.hadamard(p).astype(np.float64)
    qpad = np.zeros((Q.shape[0], p))
    qpad[:, :d] = Q.astype(np.float64) * signs[:d].astype(np.float64)
    exact = qpad @ H.T
    scale = np.abs(exact).max(axis=1, keepdims=True)
    err_gemm = float((np.abs(gemm - exact) / scale).max())
    err_fht = float((np.abs(fht - exact) / scale).max())
    assert err_gemm <= GEMM_ROT_TOL, (err_gemm, err_fht)
    ix.set_mode("fastscan")
    ids_fht = ix.search_batch(Q, 10, ef=128)
    ix.set_rotation("gemm")
    assert ix.get_rotation() == "gemm"
    ids_gemm = ix.search_batch(Q, 10, ef=128)
    ix.set_rotation("fht")
    overlap = np.mean([len(set(a) & set(b)) / len(a) for a, b in zip(ids_fht, ids_gemm)])
    assert overlap >= 0.98, (overlap, err_gemm, err_fht)


def test_update_results_and_uneven_ef_k(oracle_mod, tmp_path):
    """ef below k runs update_results (qg.hpp:395-418), including rows of node
    0 from the zero-filled id snapshot; per-query ef and k in one launch."""
    O = oracle_mod
    arr, _ = make_random_index(1200, 64, 32, seed=11, knn=True, dup_rate=0.05)
    ix = _index_from_arrays(arr, tmp_path, "u")
    ix.build_sidecar(seed=5)
    ix.set_mode("fastscan")
    rng = np.random.default_rng(3)
    Q = rng.random((60, 64), dtype=np.float32)
    ef = rng.integers(1, 300, 60).astype(np.uint32)
    ef[:10] = 1
    k = rng.integers(1, 40, 60).astype(np.uint32)
    k[:10] = 30
    ids, dis = ix.search_batch(Q, k, ef=ef, return_distances=True)
    kmax = int(k.max())
    for i in range(60):
        e_ids, e_dis, _ = _oracle_h2(O, arr, ix, Q[i:i + 1], int(ef[i]), int(k[i]))
        assert np.array_equal(ids[i, :k[i]], e_ids[0]), i
        assert np.array_equal(_bits(dis[i, :k[i]]), _bits(e_dis[0])), i
        assert np.all(ids[i, k[i]:kmax] == 0xFFFFFFFF)


@pytest.mark.parametrize("prefetch", [1, 2, 3])
@pytest.mark.parametrize("n,d,deg,ef", [(1500, 128, 32, 128), (1000, 96, 64, 300),
                                        (400, 960, 64, 200)])
def test_latency_mode_matches_oracle(oracle_mod, tmp_path, prefetch, n, d, deg, ef):
    """FastScan latency mode (one query per workgroup, a second wave loading
    the records of the best unexpanded entries ahead; the specialised kernel
    at D 128 / degree 32, the general one otherwise): batches of up to 256
    queries and the plain per-query search give the oracle's ids and distance
    bits; a larger batch runs the batch kernel with the same results."""
    arr, base = make_random_index(n, d, deg, seed=d + prefetch, knn=True)
    ix = _index_from_arrays(arr, tmp_path, "lat")
    ix.build_sidecar(seed=3)
    ix.set_mode("fastscan")
    ix.set_launch(ef, waves_per_group=1, cache_bits=10, prefetch=prefetch)
    assert ix.get_launch(ef)["prefetch"] == prefetch
    rng = np.random.default_rng(6)
    Q = (base[rng.integers(0, n, 300)] + rng.normal(0, 0.02, (300, d))).astype(np.float32)
    H = _h2arr(oracle_mod, arr, ix)
    exp, exp_d, _ = oracle_mod.search_h2_batch(H, Q, ef, 10)
    for nq in (1, 37, 256, 300):
        ids, dis = ix.search_batch(Q[:nq], 10, ef=ef, return_distances=True)
        assert np.array_equal(ids, exp[:nq]), nq
        assert np.array_equal(_bits(dis), _bits(exp_d[:nq])), nq
    ix.set_ef(ef)
    assert np.array_equal(np.stack([ix.search(q, 10) for q in Q[:10]]), exp[:10])


def test_visited_overflow_table(oracle_mod, tmp_path):
    """A 64-entry LDS visited table moves to device memory after 48
    expansions; results do not change."""
    O = oracle_mod
    arr, _ = make_random_index(1500, 64, 32, seed=2, knn=True)
    ix = _index_from_arrays(arr, tmp_path, "o")
    ix.build_sidecar(seed=9)
    ix.set_mode("fastscan")
    rng = np.random.default_rng(4)
    Q = rng.random((30, 64), dtype=np.float32)
    ix.set_launch(0, waves_per_group=2, cache_bits=6)
    ids, dis, st = ix.search_batch(Q, 10, ef=256, return_distances=True, return_stats=True)
    e_ids, e_dis, e_st = _oracle_h2(O, arr, ix, Q, 256, 10)
    assert int(e_st[:, 0].max()) > 48  # the overflow path ran
    assert np.array_equal(ids, e_ids)
    assert np.array_equal(_bits(dis), _bits(e_dis))
    ix.reset_launch(0)


def test_sidecar_file_roundtrip(tmp_path):
    from liteqg_amd._lib import LqgError
    from liteqg_amd.symphonyqg import Index

    arr, _ = make_random_index(900, 64, 32, seed=7, knn=True)
    ix = _index_from_arrays(arr, tmp_path, "s")
    ix.build_sidecar(seed=1)
    ix.set_mode("fastscan")
    Q = np.random.default_rng(0).random((20, 64), dtype=np.float32)
    want = ix.search_batch(Q, 10, ef=64)
    ix.save_sidecar(str(tmp_path / "s.side"))
    ix2 = Index("QG", "L2", arr.n, arr.dim, arr.deg)
    ix2.load(str(tmp_path / "s.index"))
    ix2.load_sidecar(str(tmp_path / "s.side"))
    ix2.set_mode("fastscan")
    assert np.array_equal(ix2.search_batch(Q, 10, ef=64), want)
    other, _ = make_random_index(900, 64, 32, seed=8, knn=True)
    ix3 = _index_from_arrays(other, tmp_path, "t")
    with pytest.raises(LqgError):
        ix3.load_sidecar(str(tmp_path / "s.side"))


def test_mode_switch_keeps_sq8_results(tmp_path):
    arr, _ = make_random_index(1000, 64, 32, seed=3, knn=True)
    ix = _index_from_arrays(arr, tmp_path, "m")
    Q = np.random.default_rng(1).random((30, 64), dtype=np.float32)
    before = ix.search_batch(Q, 10, ef=50, order="nearest")
    ix.build_sidecar(seed=2)
    ix.set_mode("fastscan")
    assert ix.get_mode() == "fastscan"
    ix.search_batch(Q, 10, ef=50)
    ix.set_mode("sq8")
    assert np.array_equal(ix.search_batch(Q, 10, ef=50, order="nearest"), before)


@pytest.mark.parametrize("rotation", ["fht", "gemm"])
def test_plain_search_equals_batch(oracle_mod, tmp_path, rotation):
    """One query per call in FastScan mode (the kernels read the query from
    pinned host memory and write the ids back there) equals the batched
    search with either rotation, and with the FHT also the oracle."""
    arr, _ = make_random_index(1500, 96, 32, seed=31, knn=True)
    ix = _index_from_arrays(arr, tmp_path, "p")
    ix.build_sidecar(signs=_signs(ix.padded_dim, 96))
    ix.set_mode("fastscan")
    ix.set_rotation(rotation)
    rng = np.random.default_rng(12)
    Q = (arr.raw[rng.integers(0, 1500, 20)] + rng.normal(0, 0.05, (20, 96))).astype(np.float32)
    for ef in (16, 100):
        ix.set_ef(ef)
        got = np.stack([ix.search(q, 10) for q in Q])
        want = ix.search_batch(Q, 10, ef=ef)
        assert np.array_equal(got, want), (rotation, ef)
        if rotation == "fht":
            exp, _, _ = _oracle_h2(oracle_mod, arr, ix, Q, ef, 10)
            assert np.array_equal(got, exp), ef
