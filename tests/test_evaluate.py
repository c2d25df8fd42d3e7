"""The evaluation harness (liteqg_amd/evaluate.py): vector file formats, the ef
sequence, the stopping rule, recall and distance-ratio metrics.  CPU tests use
a stand-in index; the GPU test runs the harness on a fixture index."""
import numpy as np
import pytest

from conftest import GOLDEN


def test_vecs_files_roundtrip_and_layout(tmp_path):
    from liteqg_amd import evaluate as E

    rng = np.random.default_rng(0)
    f = rng.random((7, 5), dtype=np.float32)
    i = rng.integers(-5, 1 << 30, (4, 3)).astype(np.int32)
    E.fvecs_write(str(tmp_path / "a.fvecs"), f)
    E.ivecs_write(str(tmp_path / "b.ivecs"), i)
    raw = np.fromfile(tmp_path / "a.fvecs", np.int32).reshape(7, 6)
    assert (raw[:, 0] == 5).all()
    assert np.array_equal(raw[:, 1:].view(np.float32), f)
    assert np.array_equal(E.fvecs_read(str(tmp_path / "a.fvecs")), f)
    assert np.array_equal(E.ivecs_read(str(tmp_path / "b.ivecs")), i)
    assert np.array_equal(E.read_vectors(str(tmp_path / "b.ivecs")), i)
    # a row with another dimension, and a truncated file
    bad = np.fromfile(tmp_path / "b.ivecs", np.int32)
    bad[4] = 2
    bad.tofile(tmp_path / "bad.ivecs")
    with pytest.raises(ValueError):
        E.ivecs_read(str(tmp_path / "bad.ivecs"))
    np.fromfile(tmp_path / "b.ivecs", np.int32)[:-1].tofile(tmp_path / "short.ivecs")
    with pytest.raises(ValueError):
        E.ivecs_read(str(tmp_path / "short.ivecs"))
    with pytest.raises(ValueError):
        E.read_vectors(str(tmp_path / "x.npy"))


def test_bin_files_roundtrip_and_slices(tmp_path):
    from liteqg_amd import evaluate as E

    rng = np.random.default_rng(1)
    f = rng.random((9, 4), dtype=np.float32)
    E.fbin_write(str(tmp_path / "a.fbin"), f)
    hdr = np.fromfile(tmp_path / "a.fbin", np.int32, count=2)
    assert hdr.tolist() == [9, 4]
    assert np.array_equal(E.fbin_read(str(tmp_path / "a.fbin")), f)
    assert np.array_equal(E.fbin_read(str(tmp_path / "a.fbin"), start=3, count=4), f[3:7])
    assert np.array_equal(E.fbin_read(str(tmp_path / "a.fbin"), start=7), f[7:])
    i = rng.integers(0, 100, (3, 2)).astype(np.int32)
    E.ibin_write(str(tmp_path / "b.ibin"), i)
    assert np.array_equal(E.ibin_read(str(tmp_path / "b.ibin")), i)
    with pytest.raises(ValueError):
        E.fbin_read(str(tmp_path / "a.fbin"), start=10)


def test_beam_sizes_sequence():
    from itertools import islice

    from liteqg_amd import evaluate as E

    assert list(islice(E.beam_sizes(10), 12)) == [10, 15, 20, 25, 30, 40, 50, 60, 70, 80, 100, 150]
    assert list(islice(E.beam_sizes(1), 3)) == [10, 15, 20]
    assert list(islice(E.beam_sizes(99), 1)) == [10]
    assert list(islice(E.beam_sizes(100), 2)) == [100, 150]
    assert list(islice(E.beam_sizes(1000), 1)) == [1000]
    with pytest.raises(ValueError):
        next(E.beam_sizes(0))


def test_recall_and_distance_ratio():
    from liteqg_amd import evaluate as E

    gt = np.array([[0, 1, 2], [3, 4, 5]])
    assert E.recall_percent(gt, gt, 3) == 100.0
    assert E.recall_percent(np.array([[2, 9, 9], [5, 4, 3]]), gt, 3) == pytest.approx(100 * 4 / 6)
    assert E.recall_percent(np.array([[1, 0], [3, 9]]), gt, 2) == 75.0
    # points on a line; query at 0.5
    base = np.array([[0.0], [1.0], [2.0], [3.0], [10.0]], np.float32)
    q = np.array([[0.5]], np.float32)
    g = np.array([[0, 1]])
    assert E.avg_distance_ratio(base, q, g, g, 2) == 1.0
    # found {1, 2}: sorted distances (0.5, 1.5) over (0.5, 0.5) -> (1 + 3) / 2
    assert E.avg_distance_ratio(base, q, np.array([[2, 1]]), g, 2) == pytest.approx(2.0)
    # a true distance below 1e-4 is skipped; a query with none left counts 1
    q0 = np.array([[0.0]], np.float32)
    assert E.avg_distance_ratio(base, q0, np.array([[1, 2]]), np.array([[0, 1]]), 2) == \
        pytest.approx(2.0 / 1.0)
    assert E.avg_distance_ratio(base, q0, np.array([[4]]), np.array([[0]]), 1) == 1.0
    # unfilled slots are left out
    assert E.avg_distance_ratio(base, q, np.array([[1, 0xFFFFFFFF]]), g, 2) == 1.0


class _FakeIndex:
    """search_batch returns the first r(ef) true ids of each query, then misses."""

    def __init__(self, gt, recall_of_ef):
        self.gt = gt
        self.r = recall_of_ef
        self.calls = []

    def search_batch(self, Q, k, ef, order):
        self.calls.append(ef)
        hits = self.r(ef)
        out = np.full((len(Q), k), 10**6, np.uint32)
        out[:, :hits] = self.gt[:len(Q), :hits]
        return out


def test_find_efs_stopping_rule_and_sweep_columns(tmp_path):
    from liteqg_amd import evaluate as E

    gt = np.arange(40).reshape(4, 10)
    Q = np.zeros((4, 3), np.float32)
    # perfect at the first ef: one ef
    assert E.find_efs(_FakeIndex(gt, lambda ef: 10), Q, gt, 10) == [10]
    # recall 10%, 20%, ..., stops when the gain falls below 0.05 points
    steps = {10: 1, 15: 2, 20: 3, 25: 3}
    assert E.find_efs(_FakeIndex(gt, lambda ef: steps[ef]), Q, gt, 10) == [10, 15, 20, 25]
    cols = E.sweep(_FakeIndex(gt, lambda ef: steps[ef]), Q, gt, [10, 15], 10, rounds=2,
                   method="symphonyqg32", memory_mb=12.5)
    assert cols["Recall"] == [10.0, 20.0]
    assert cols["EFS"] == [10, 15] and cols["Method"] == ["symphonyqg32"] * 2
    assert len(cols["QPS"]) == 2 and all(v > 0 for v in cols["QPS"])
    E.write_csv(str(tmp_path / "s.csv"), cols)
    lines = open(tmp_path / "s.csv").read().splitlines()
    assert lines[0] == "QPS,Recall,EFS,Method,Memory" and len(lines) == 3


@pytest.mark.gpu
def test_harness_on_fixture_index(oracle_mod, tmp_path):
    """The sweep's recall per ef equals the recall of the batched search; the
    plain per-query mode returns the same ids; the CLI writes the CSVs."""
    from liteqg_amd import evaluate as E
    from liteqg_amd.symphonyqg import Index

    name = "g_uniform_n2000_d64"
    arr = oracle_mod.read_index(f"{GOLDEN}/{name}.index", 2000, 64, 32)
    Q = np.load(f"{GOLDEN}/{name}.npz")["queries"][:60]
    gt = E.ground_truth(arr.raw, Q, 10)
    ref_gt, _ = oracle_mod.bruteforce_fp64(arr.raw, Q, 10)
    assert np.array_equal(gt, ref_gt.astype(np.int32))
    ix = Index("QG", "L2", 2000, 64, 32)
    ix.load(f"{GOLDEN}/{name}.index")
    efs = E.find_efs(ix, Q, gt, 10, order="nearest")
    assert efs[0] == 10 and len(efs) >= 1
    cols = E.sweep(ix, Q, gt, efs, 10, rounds=1, order="nearest")
    for ef, rec in zip(efs, cols["Recall"]):
        assert rec == E.recall_percent(ix.search_batch(Q, 10, ef=ef, order="nearest"), gt, 10)
    ids_plain, _ = E._run(ix, Q[:8], 10, 50, False, "reference")
    assert np.array_equal(ids_plain, ix.search_batch(Q[:8], 10, ef=50, order="reference"))
    E.fvecs_write(str(tmp_path / "b.fvecs"), arr.raw)
    E.fvecs_write(str(tmp_path / "q.fvecs"), Q)
    E.ivecs_write(str(tmp_path / "g.ivecs"), gt)
    E.main(["--base", str(tmp_path / "b.fvecs"), "--query", str(tmp_path / "q.fvecs"),
            "--gt", str(tmp_path / "g.ivecs"), "--index", f"{GOLDEN}/{name}.index",
            "--rounds", "1", "--ratio", "--out", str(tmp_path / "res")])
    assert (tmp_path / "res" / "symphonyqg32_10.csv").exists()
    lines = open(tmp_path / "res" / "ratio32_10.csv").read().splitlines()
    assert lines[0] == "QPS,RATIO,EFS,Method"
