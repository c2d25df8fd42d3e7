"""The FastScan RaBitQ oracle (search_qg, qg.hpp:252-287) against the reference.

The h2_*.npz fixtures were written by oracle/_ref/ref_h2 (the reference's
QuantizedGraph, rabitq_codes, FHTRotator, SearchBuffer and ResultBuffer
compiled from /root/reference with the reference's -Ofast flags) over the
g_*.index fixtures; buffers.npz holds buffer operation traces replayed by the
same binary.  These tests only read them.
"""
import numpy as np
import pytest
from scipy.linalg import hadamard

from conftest import GOLDEN, TIE_FREE, load_golden, make_random_index

H2_NAMES = TIE_FREE


def _h2(name):
    ix, _ = load_golden(name)
    return ix, np.load(f"{GOLDEN}/h2_{name}.npz")


def test_buffers_match_reference_trace(oracle_mod):
    """SearchBuffer / ResultBuffer (utils/buffer.hpp:11-127) on a trace with
    many equal distances: the popped id sequence and the final results."""
    z = np.load(f"{GOLDEN}/buffers.npz")
    popped, res = oracle_mod.buffers_replay(z["ops"], int(z["cap"]))
    assert np.array_equal(popped, z["popped"])
    assert np.array_equal(res, z["results"])


def test_buffers_small_cases(oracle_mod):
    """Closed cases: a full buffer drops the largest; equal keys insert before
    existing ones, so among equals the newest pops first; pop marks only."""
    def ops(*t):
        return np.array([(k, i, np.float32(d).view(np.uint32)) for k, i, d in t],
                        np.uint32).reshape(-1)
    # cap 2: 3 (d 3), 1 (d 1), 2 (d 2) -> holds 1, 2; pops 1 then 2 then none
    popped, res = oracle_mod.buffers_replay(ops((0, 3, 3), (0, 1, 1), (0, 2, 2), (1, 0, 0),
                                                (1, 0, 0), (1, 0, 0)), 2)
    assert list(popped) == [1, 2, 0xFFFFFFFF]
    # equal distances: the later insert sits first
    popped, _ = oracle_mod.buffers_replay(ops((0, 5, 1), (0, 6, 1), (1, 0, 0)), 4)
    assert list(popped) == [6]
    # ResultBuffer keeps the cap smallest, ascending
    _, res = oracle_mod.buffers_replay(ops((2, 9, 9), (2, 4, 4), (2, 7, 7), (2, 1, 1)), 3)
    assert list(res) == [1, 4, 7]


@pytest.mark.parametrize("name", H2_NAMES)
def test_fht_rotation_matches_reference(oracle_mod, name):
    """FHTRotator::rotate (utils/rotator.hpp:88-114): bit-exact rotated queries."""
    ix, h = _h2(name)
    p = oracle_mod.padded_dim(ix.dim)
    for i in range(h["prep_rq"].shape[0]):
        rq = oracle_mod.fht_rotate(h["signs"], h["queries"][i], p)
        assert np.array_equal(rq.view(np.uint32), h["prep_rq"][i].view(np.uint32))


@pytest.mark.parametrize("d", (64, 96, 100, 128, 200))
def test_fht_rotation_is_sylvester_hadamard(oracle_mod, d):
    """rotate(x) = H_P diag(s) [x, 0]: the Sylvester Hadamard matrix (scipy) in
    FP64; with |s| = 1/sqrt(P) it is orthonormal, so norms are kept."""
    rng = np.random.default_rng(d)
    p = oracle_mod.padded_dim(d)
    s = ((rng.integers(0, 2, p) * 2 - 1) / np.sqrt(p)).astype(np.float32)
    for _ in range(4):
        x = rng.standard_normal(d).astype(np.float32)
        got = oracle_mod.fht_rotate(s, x, p).astype(np.float64)
        xp = np.zeros(p)
        xp[:d] = x
        want = hadamard(p) @ (s.astype(np.float64) * xp)
        assert np.allclose(got, want, rtol=0, atol=1e-5 * np.abs(want).max())
        assert abs(np.linalg.norm(got) - np.linalg.norm(x)) < 1e-5 * np.linalg.norm(x)


@pytest.mark.parametrize("name", H2_NAMES)
def test_query_prepare_matches_reference(oracle_mod, name):
    """QGQuery::query_prepare (qg_query.hpp:28-43): lo, width, sumq and the
    4P-byte LUT equal the reference's, byte for byte."""
    ix, h = _h2(name)
    p = oracle_mod.padded_dim(ix.dim)
    for i in range(h["prep_rq"].shape[0]):
        rq, lo, w, sq, qb, lut = oracle_mod.query_prepare(h["signs"], h["queries"][i], p)
        assert np.float32(lo) == h["prep_lo"][i]
        assert np.float32(w) == h["prep_width"][i]
        assert sq == int(h["prep_sumq"][i])
        assert np.array_equal(lut, h["prep_lut"][i])


def test_query_prepare_reciprocal_is_decided_by_goldens(oracle_mod):
    """The exact reciprocal 1/width quantises some golden query differently
    from the reference build's 62.999996f/(hi - lo), so the fixtures decide."""
    f32 = np.float32
    differs = 0
    for name in H2_NAMES:
        _, h = _h2(name)
        for i in range(h["prep_rq"].shape[0]):
            rq = h["prep_rq"][i]
            lo = f32(min(rq.min(), 0))
            prod = ((rq - lo).astype(f32) * (f32(1) / h["prep_width"][i])).astype(f32)
            q = np.floor(prod.astype(np.float64) + 1.0)  # lround(x + 0.5), x >= 0
            differs += int(q.sum()) != int(h["prep_sumq"][i])
    assert differs > 0


@pytest.mark.parametrize("d", (64, 100, 256))
def test_query_prepare_closed_forms(oracle_mod, d):
    """q in [0, 64] within one step of (v - lo)/width + 0.5; sumq = sum q; each
    LUT entry is the uint8 sum of the q bytes its bit pattern selects."""
    rng = np.random.default_rng(d + 1)
    p = oracle_mod.padded_dim(d)
    s = ((rng.integers(0, 2, p) * 2 - 1) / np.sqrt(p)).astype(np.float32)
    q = rng.standard_normal(d).astype(np.float32) * 3
    rq, lo, w, sq, qb, lut = oracle_mod.query_prepare(s, q, p)
    lo64 = min(0.0, float(rq.min()))
    hi64 = max(0.0, float(rq.max()))
    assert abs(w - (hi64 - lo64) / 63) <= 1e-6 * (hi64 - lo64)
    t = (rq.astype(np.float64) - lo64) / w + 0.5
    assert np.all(np.abs(qb - t) <= 0.5 + 1e-4)
    assert qb.max() <= 64 and sq == int(qb.astype(np.int64).sum())
    for c in range(p // 4):
        for j in range(16):
            want = sum(int(qb[4 * c + 3 - b]) for b in range(4) if j >> b & 1) & 0xFF
            assert lut[16 * c + j] == want


@pytest.mark.parametrize("name", H2_NAMES)
def test_scan_estimates_match_reference(oracle_mod, name):
    """QGScanner::scan_neighbors (qg_scanner.hpp:99-155) on the entry point's
    row of the reference's own sidecar: every estimate bit-exact."""
    ix, h = _h2(name)
    p = oracle_mod.padded_dim(ix.dim)
    for i in range(h["prep_rq"].shape[0]):
        lut = h["prep_lut"][i]
        sqr_y = oracle_mod.l2_sqr(h["queries"][i], ix.raw[ix.ep])
        assert np.float32(sqr_y) == h["prep_sqr_y"][i]
        _, est = oracle_mod.scan_estimates(h["codes"][ix.ep], 0, h["factors"][ix.ep], ix.deg, p,
                                           lut, sqr_y, float(h["prep_lo"][i]),
                                           float(h["prep_width"][i]), int(h["prep_sumq"][i]))
        assert np.array_equal(est.view(np.uint32), h["prep_est"][i].view(np.uint32))


def test_estimate_order_is_decided_by_goldens(oracle_mod):
    """The source order (fac_dq * width) * f differs from the reference build's
    (f * fac_dq) * width on some golden estimate, so the fixtures decide."""
    f32 = np.float32
    differs = 0
    for name in H2_NAMES:
        ix, h = _h2(name)
        p = oracle_mod.padded_dim(ix.dim)
        fac = h["factors"][ix.ep].reshape(3, ix.deg)
        for i in range(h["prep_rq"].shape[0]):
            sums, est = oracle_mod.scan_estimates(
                h["codes"][ix.ep], 0, h["factors"][ix.ep], ix.deg, p, h["prep_lut"][i],
                float(h["prep_sqr_y"][i]), float(h["prep_lo"][i]), float(h["prep_width"][i]),
                int(h["prep_sumq"][i]))
            f = (sums.astype(np.uint16).view(np.int16).astype(np.int32) * 2 -
                 int(h["prep_sumq"][i])).astype(f32)
            w, lo = h["prep_width"][i], h["prep_lo"][i]
            a = (fac[0] + h["prep_sqr_y"][i]).astype(f32)
            # fma(fac_vq, lo, a): the product is exact in double
            b = (fac[2].astype(np.float64) * np.float64(lo) + a.astype(np.float64)).astype(f32)
            built = ((f * fac[1]).astype(f32) * w).astype(f32) + b
            alt = ((fac[1] * w).astype(f32) * f).astype(f32) + b
            assert np.array_equal(built.view(np.uint32), est.view(np.uint32))
            differs += int((alt.view(np.uint32) != h["prep_est"][i].view(np.uint32)).sum())
    assert differs > 0


@pytest.mark.parametrize("name", H2_NAMES)
def test_scan_layouts_agree(oracle_mod, name):
    """The vector-major layout read by this build's kernels holds the same
    nibbles as the reference's FastScan blocks: identical uint16 sums."""
    ix, h = _h2(name)
    p = oracle_mod.padded_dim(ix.dim)
    vm = oracle_mod.codes_to_vector_major(h["codes"][:50], ix.deg, p)
    rng = np.random.default_rng(0)
    lut = rng.integers(0, 256, 4 * p).astype(np.uint8)
    for v in range(50):
        a, _ = oracle_mod.scan_estimates(h["codes"][v], 0, h["factors"][v], ix.deg, p, lut,
                                         1.0, -1.0, 0.1, 100)
        b, _ = oracle_mod.scan_estimates(vm[v], 1, h["factors"][v], ix.deg, p, lut,
                                         1.0, -1.0, 0.1, 100)
        assert np.array_equal(a, b)


@pytest.mark.parametrize("name", H2_NAMES)
def test_search_qg_matches_reference(oracle_mod, name):
    """search_qg (qg.hpp:252-287) with the reference's sidecar: ids for ids at
    every ef of the fixture, 100 queries each."""
    ix, h = _h2(name)
    arr = oracle_mod.H2Arrays(ix, h["signs"], h["codes"], h["factors"], 0)
    for ef in h["efs"]:
        ids, dist, _ = oracle_mod.search_h2_batch(arr, h["queries"], int(ef), int(h["k"]))
        assert np.array_equal(ids, h[f"ids_ef{ef}"]), f"ef={ef}"


@pytest.mark.parametrize("name", H2_NAMES)
def test_update_results_matches_reference(oracle_mod, name):
    """update_results (qg.hpp:395-418): at ef 1-4 with k 10 and ef 16/64 with
    k 100 most searches expand fewer than k nodes, so the result list is
    filled from the neighbours of the id snapshot, whose unfilled slots are
    zero (buffer.hpp:81-102) and probe node 0.  The reference's search_qg ids
    (h2u_*.npz), 100 queries per case."""
    ix, h = _h2(name)
    u = np.load(f"{GOLDEN}/h2u_{name}.npz")
    arr = oracle_mod.H2Arrays(ix, h["signs"], h["codes"], h["factors"], 0)
    short = 0
    for ef, k in u["cases"]:
        ids, dist, st = oracle_mod.search_h2_batch(arr, u["queries"], int(ef), int(k))
        assert np.array_equal(ids, u[f"ids_ef{ef}_k{k}"]), f"ef={ef} k={k}"
        short += int((st[:, 0] < k).sum())
        assert np.all(np.diff(dist, axis=1) >= 0)
    assert short >= 400  # the tail runs in most of the 500 searches


@pytest.mark.parametrize("name", H2_NAMES)
def test_expansion_list_invariants(oracle_mod, name):
    """find_candidates (qg.hpp:439-473) keeps the nodes search_qg's traversal
    expands, in order.  The list's length is the expansion count of the
    pinned search (same traversal), its ids are distinct, it starts at the
    entry point, and every later id is a neighbour of an earlier one (a node
    enters the buffer only as a neighbour of an expanded node)."""
    ix, h = _h2(name)
    arr = oracle_mod.H2Arrays(ix, h["signs"], h["codes"], h["factors"], 0)
    q = h["queries"][:40]
    for ef in (10, 50, 100):
        ex = oracle_mod.h2_expansions_batch(arr, q, ef, 1024)
        _, _, st = oracle_mod.search_h2_batch(arr, q, ef, 10)
        for r in range(q.shape[0]):
            e = ex[r][ex[r] != 0xFFFFFFFF]
            assert e.size == st[r, 0]
            assert np.all(ex[r][e.size:] == 0xFFFFFFFF)
            assert len(set(e.tolist())) == e.size
            assert e[0] == ix.ep
            seen = {int(e[0])}
            nb = set(ix.nbr[e[0]].tolist())
            for x in e[1:]:
                assert int(x) in nb
                seen.add(int(x))
                nb |= set(ix.nbr[x].tolist())
    # a cap below the expansion count keeps the first cap of them
    full = oracle_mod.h2_expansions_batch(arr, q, 50, 1024)
    assert np.array_equal(oracle_mod.h2_expansions_batch(arr, q, 50, 7), full[:, :7])


@pytest.mark.parametrize("name", H2_NAMES)
def test_sidecar_codes_equal_reference(oracle_mod, name):
    """rabitq_codes (quantization/rabitq.hpp:23-59): this build's sidecar has
    the reference's sign codes bit for bit (in vector-major layout)."""
    ix, h = _h2(name)
    p = oracle_mod.padded_dim(ix.dim)
    codes, _ = oracle_mod.sidecar_build(ix, h["signs"])
    assert np.array_equal(codes, oracle_mod.codes_to_vector_major(h["codes"], ix.deg, p))


# The factors are sums over P terms: the reference sums in float with its own
# SIMD order, this build in FP64.  Tolerance: 2e-6 of the largest magnitude of
# each factor array (observed <= 1.1e-6).
FACTOR_TOL = 2e-6


@pytest.mark.parametrize("name", H2_NAMES)
def test_sidecar_factors_within_tolerance(oracle_mod, name):
    ix, h = _h2(name)
    _, fac = oracle_mod.sidecar_build(ix, h["signs"])
    ref = h["factors"].reshape(-1, 3, ix.deg)
    got = fac.reshape(-1, 3, ix.deg)
    for j in range(3):
        scale = np.abs(ref[:, j]).max()
        assert np.abs(ref[:, j] - got[:, j]).max() <= FACTOR_TOL * scale


@pytest.mark.parametrize("name", H2_NAMES)
def test_search_with_own_sidecar_matches_reference(oracle_mod, name):
    """Searching with this build's sidecar returns the reference's ids on the
    fixtures (factor differences never reorder a decision there)."""
    ix, h = _h2(name)
    codes, fac = oracle_mod.sidecar_build(ix, h["signs"])
    arr = oracle_mod.H2Arrays(ix, h["signs"], codes, fac, 1)
    for ef in h["efs"]:
        ids, _, _ = oracle_mod.search_h2_batch(arr, h["queries"], int(ef), int(h["k"]))
        assert np.array_equal(ids, h[f"ids_ef{ef}"]), f"ef={ef}"


def test_search_qg_invariants(oracle_mod):
    """On a random index: results are distinct, in ascending exact distance,
    their distances are the exact l2_sqr, and at ef 256 the 1-bit estimator
    still steers to the true neighbours (recall@10 > 0.9; 0.98 observed)."""
    ix, _ = make_random_index(1500, 64, 32, seed=5, knn=True)
    p = oracle_mod.padded_dim(64)
    rng = np.random.default_rng(2)
    s = ((rng.integers(0, 2, p) * 2 - 1) / np.sqrt(p)).astype(np.float32)
    codes, fac = oracle_mod.sidecar_build(ix, s)
    arr = oracle_mod.H2Arrays(ix, s, codes, fac, 1)
    q = rng.random((20, 64), dtype=np.float32)
    ids, dist, st = oracle_mod.search_h2_batch(arr, q, 256, 10)
    for r in range(q.shape[0]):
        assert len(set(ids[r].tolist())) == 10
        ex = np.array([oracle_mod.l2_sqr(q[r], ix.raw[i]) for i in ids[r]], np.float32)
        assert np.array_equal(ex.view(np.uint32), dist[r].view(np.uint32))
        assert np.all(np.diff(dist[r]) >= 0)
    gt, _ = oracle_mod.bruteforce_fp64(ix.raw, q, 10)
    assert oracle_mod.recall_at_k(ids, gt, 10) > 0.9
