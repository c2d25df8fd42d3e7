"""Evaluation harness of the reference's reproduce/ scripts, on the device.

What the reference's scripts do, for users with real datasets:

- vector files: ``.fvecs`` / ``.ivecs`` (every row is an int32 dimension d
  followed by d values; reproduce/utils/io.py:3-27) and ``.fbin`` / ``.ibin``
  (an int32 header (n, d), then n*d values; reproduce/utils/io.py:30-75);
- ``beam_sizes(k)``: the ef sequence of reproduce/utils/beam_size.py (10, 15,
  20, 25, 30, 40, 50, 60, 70, 80 times a power of ten, starting one decade
  below k);
- ``find_efs``: the stopping rule of reproduce/run.py:14-44 (stop once recall
  passes 99.8%, gains less than 0.05 points, or QPS drops below 10);
- ``sweep``: QPS and recall (percent) per ef averaged over rounds, the rows
  reproduce/run.py writes (QPS, Recall, EFS, Method, Memory);
- ``avg_distance_ratio``: reproduce/utils/avg_distance_ratio.py with
  reproduce/ratio.py's averaging (sorted found distances over sorted true
  distances, true distances below 1e-4 skipped, a query with none counts 1);
- ``ground_truth``: exact top-k by L2 (reproduce/compute_gt.py), here in FP64
  on the device with ties broken by id.

Searches run through ``symphonyqg.Index``: one query per ``search`` call as the
reference times it (``batched=False``), or one ``search_batch`` launch per ef.

Command line (writes the reference's CSV names into --out):
  python -m liteqg_amd.evaluate --base sift_base.fvecs --query sift_query.fvecs \\
      --gt sift_groundtruth.ivecs --index symphonyqg_32.index --degree 32 --out results/sift
"""
from __future__ import annotations

import argparse
import csv
import os
import time

import numpy as np

BASES = (10, 15, 20, 25, 30, 40, 50, 60, 70, 80)


# -- vector files -------------------------------------------------------------
def _vecs_read(path: str, dtype) -> np.ndarray:
    raw = np.fromfile(path, dtype=np.int32)
    if raw.size == 0:
        return np.empty((0, 0), dtype)
    d = int(raw[0])
    if d <= 0 or raw.size % (d + 1) != 0:
        raise ValueError(f"{path}: not a vecs file (row header {d}, {raw.size} words)")
    rows = raw.reshape(-1, d + 1)
    if not np.all(rows[:, 0] == d):
        raise ValueError(f"{path}: rows of different dimensions")
    return np.ascontiguousarray(rows[:, 1:]).view(dtype)


def _vecs_write(path: str, m: np.ndarray, dtype) -> None:
    m = np.ascontiguousarray(m, dtype=dtype)
    if m.ndim != 2:
        raise ValueError("expected a 2-D array")
    out = np.empty((m.shape[0], m.shape[1] + 1), np.int32)
    out[:, 0] = m.shape[1]
    out[:, 1:] = m.view(np.int32)
    out.tofile(path)


def ivecs_read(path: str) -> np.ndarray:
    return _vecs_read(path, np.int32)


def fvecs_read(path: str) -> np.ndarray:
    return _vecs_read(path, np.float32)


def ivecs_write(path: str, m) -> None:
    _vecs_write(path, m, np.int32)


def fvecs_write(path: str, m) -> None:
    _vecs_write(path, m, np.float32)


def _bin_read(path: str, dtype, start: int, count: int | None) -> np.ndarray:
    with open(path, "rb") as f:
        n, d = (int(x) for x in np.fromfile(f, dtype=np.int32, count=2))
        if start < 0 or start > n:
            raise ValueError(f"{path}: start {start} outside [0, {n}]")
        m = n - start if count is None else min(count, n - start)
        a = np.fromfile(f, dtype=dtype, count=m * d, offset=start * d * 4)
    if a.size != m * d:
        raise ValueError(f"{path}: file shorter than its header says")
    return a.reshape(m, d)


def _bin_write(path: str, m, dtype) -> None:
    m = np.ascontiguousarray(m, dtype=dtype)
    with open(path, "wb") as f:
        np.array(m.shape, np.int32).tofile(f)
        m.tofile(f)


def fbin_read(path: str, start: int = 0, count: int | None = None) -> np.ndarray:
    return _bin_read(path, np.float32, start, count)


def ibin_read(path: str, start: int = 0, count: int | None = None) -> np.ndarray:
    return _bin_read(path, np.int32, start, count)


def fbin_write(path: str, m) -> None:
    _bin_write(path, m, np.float32)


def ibin_write(path: str, m) -> None:
    _bin_write(path, m, np.int32)


def read_vectors(path: str) -> np.ndarray:
    """Any of the four formats, by extension."""
    ext = os.path.splitext(path)[1]
    readers = {".fvecs": fvecs_read, ".ivecs": ivecs_read, ".fbin": fbin_read, ".ibin": ibin_read}
    if ext not in readers:
        raise ValueError(f"unknown vector file type {ext!r}")
    return readers[ext](path)


# -- metrics ------------------------------------------------------------------
def beam_sizes(k: int):
    """The ef sequence: 10, 15, ..., 80 times 10^e with e from max(0, digits(k) - 2)."""
    if k < 1:
        raise ValueError("k must be positive")
    e = max(0, len(str(int(k))) - 2)
    while True:
        for b in BASES:
            yield b * 10 ** e
        e += 1


def recall_percent(found, gt, k: int) -> float:
    """Percent of the first k true neighbours of each query found among its results."""
    found = np.asarray(found)
    gt = np.asarray(gt)
    hit = sum(len(set(gt[i, :k].tolist()) & set(found[i].tolist())) for i in range(found.shape[0]))
    return 100.0 * hit / (found.shape[0] * k)


def avg_distance_ratio(base, queries, found, gt, k: int) -> float:
    """Mean over queries of the mean ratio sorted(found distance) / sorted(true
    distance), skipping true distances below 1e-4 (a query with none counts 1);
    invalid ids (unfilled slots) are left out."""
    base = np.asarray(base, np.float64)
    total = 0.0
    for i in range(len(queries)):
        q = np.asarray(queries[i], np.float64)
        ids = np.asarray(found[i], np.int64)
        ids = ids[(ids >= 0) & (ids < base.shape[0])]
        df = np.sort(np.linalg.norm(base[ids] - q, axis=1))
        dg = np.sort(np.linalg.norm(base[np.asarray(gt[i][:k], np.int64)] - q, axis=1))
        r = [a / b for a, b in zip(df, dg) if b >= 1e-4]
        total += float(np.mean(r)) if r else 1.0
    return total / len(queries)


def ground_truth(base, queries, k: int) -> np.ndarray:
    """Exact top-k by L2 in FP64 on the device, ties broken by id (int32 ids)."""
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("ground_truth runs on the GPU")
    B = torch.from_numpy(np.ascontiguousarray(base, np.float32)).cuda().double()
    bn = (B * B).sum(1)
    out = np.empty((len(queries), k), np.int32)
    kk = min(B.shape[0], k + 16)
    for s in range(0, len(queries), 256):
        q = torch.from_numpy(np.ascontiguousarray(queries[s:s + 256], np.float32)).cuda().double()
        d = (q * q).sum(1)[:, None] + bn[None, :] - 2.0 * (q @ B.T)
        vals, idx = torch.topk(d, kk, largest=False)
        vals, idx = vals.cpu().numpy(), idx.cpu().numpy()
        for r in range(vals.shape[0]):
            out[s + r] = idx[r][np.lexsort((idx[r], vals[r]))[:k]]
    return out


# -- search sweeps ------------------------------------------------------------
def _run(index, queries, k: int, ef: int, batched: bool, order: str):
    """(results, seconds) for all queries at one ef."""
    if batched:
        index.search_batch(queries[:1], k, ef=ef, order=order)  # warm-up
        t = time.perf_counter()
        ids = index.search_batch(queries, k, ef=ef, order=order)
        return ids, time.perf_counter() - t
    index.set_ef(ef)
    index.set_order(order)
    ids = np.empty((len(queries), k), np.uint32)
    total = 0.0
    for i in range(len(queries)):
        t = time.perf_counter()
        ids[i] = index.search(queries[i], k)
        total += time.perf_counter() - t
    return ids, total


def find_efs(index, queries, gt, k: int = 10, batched: bool = True, order: str = "reference",
             max_steps: int = 60) -> list:
    """ef values of the sweep, by the reference's stopping rule."""
    efs, prev = [], 0.0
    for ef in beam_sizes(k):
        if ef > 1024 or len(efs) >= max_steps:
            break
        efs.append(ef)
        ids, sec = _run(index, queries, k, ef, batched, order)
        rec = recall_percent(ids, gt, k)
        qps = len(queries) / sec
        if rec > 99.8 or rec - prev < 0.05 or qps < 10:
            break
        prev = rec
    return efs


def sweep(index, queries, gt, efs, k: int = 10, rounds: int = 5, batched: bool = True,
          order: str = "reference", method: str = "", memory_mb: float = 0.0) -> dict:
    """QPS and recall (percent) per ef, averaged over rounds: the columns of the
    reference's per-degree CSV."""
    qps = np.zeros(len(efs))
    rec = np.zeros(len(efs))
    for _ in range(rounds):
        for j, ef in enumerate(efs):
            ids, sec = _run(index, queries, k, ef, batched, order)
            qps[j] += len(queries) / sec
            rec[j] += recall_percent(ids, gt, k)
    return {"QPS": (qps / rounds).tolist(), "Recall": (rec / rounds).tolist(), "EFS": list(efs),
            "Method": [method] * len(efs), "Memory": [memory_mb] * len(efs)}


def ratio_sweep(index, base, queries, gt, efs, k: int = 10, order: str = "reference") -> list:
    """Average distance ratio per ef (reproduce/ratio.py), plain per-query search."""
    out = []
    for ef in efs:
        ids, _ = _run(index, queries, k, ef, False, order)
        out.append(avg_distance_ratio(base, queries, ids, gt, k))
    return out


def write_csv(path: str, cols: dict) -> None:
    keys = list(cols)
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(keys)
        for row in zip(*(cols[c] for c in keys)):
            w.writerow(row)


def _rss_mb() -> float:
    try:
        import psutil

        return psutil.Process().memory_info().rss / 1024 / 1024
    except ImportError:
        return 0.0


def main(argv=None) -> None:
    from .symphonyqg import Index

    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--base", required=True)
    ap.add_argument("--query", required=True)
    ap.add_argument("--gt", help="ground-truth ids (.ivecs / .ibin); computed on the device if absent")
    ap.add_argument("--index", required=True, help="index file written by the reference or by this library")
    ap.add_argument("--degree", type=int, default=32)
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--rounds", type=int, default=5)
    ap.add_argument("--plain", action="store_true", help="one query per search() call, as the reference times it")
    ap.add_argument("--order", default="reference", choices=("reference", "nearest"))
    ap.add_argument("--normalize", action="store_true", help="unit-normalise the queries (angular datasets)")
    ap.add_argument("--ratio", action="store_true", help="also write the distance-ratio CSV")
    ap.add_argument("--out", default=".")
    a = ap.parse_args(argv)
    base = read_vectors(a.base)
    queries = read_vectors(a.query).astype(np.float32)
    if a.normalize:
        queries = queries / np.linalg.norm(queries, axis=1, keepdims=True)
    gt = read_vectors(a.gt) if a.gt else ground_truth(base, queries, a.k)
    n, d = base.shape
    m0 = _rss_mb()
    index = Index("QG", "L2", n, d, a.degree)
    index.load(a.index)
    mem = _rss_mb() - m0
    method = f"symphonyqg{a.degree}"
    efs = find_efs(index, queries, gt, a.k, batched=not a.plain, order=a.order)
    cols = sweep(index, queries, gt, efs, a.k, a.rounds, batched=not a.plain, order=a.order,
                 method=method, memory_mb=mem)
    os.makedirs(a.out, exist_ok=True)
    write_csv(os.path.join(a.out, f"{method}_{a.k}.csv"), cols)
    if a.ratio:
        r = ratio_sweep(index, base, queries, gt, efs, a.k, a.order)
        write_csv(os.path.join(a.out, f"ratio{a.degree}_{a.k}.csv"),
                  {"QPS": cols["QPS"], "RATIO": r, "EFS": efs, "Method": [method] * len(efs)})


if __name__ == "__main__":
    main()
