"""Benchmark of the query path on one MI355X: QPS at recall@10 >= 0.95.

Workload (the 1M x 128 degree-32 configuration): 1,000,000 float32 vectors,
integer valued 0..255, from a 1,024-component isotropic Gaussian mixture
(centres U[0,255]^128, sigma 20, clipped and rounded); 10,000 queries from
the same mixture; graph built on the device with build EF 400, 3 iterations;
k = 10; ef swept over {20, 50, 100, 200, 500} and the smallest ef whose
recall@10 against FP64 brute force reaches 0.95 is the operating point.

A step is one batched search launch over the 10,000 device-resident queries.
``value`` is queries per second over the K timed steps (HIP events on the
stream the kernel runs on, inputs already in HBM).  The plain path (one
query per Index.search call, host buffers) is timed beside it for the
batched/plain ratio and p50/p99 latency.

``steady_state`` (beside ``value``, not in it): the same 10,000-query batch
replayed back to back from four HIP graphs on four streams, so each batch's
tail overlaps the next batch's start; ids are checked equal to the single
launch's.  It shows the kernel's rate without the tail of a lone batch.

``parity`` (the bench exits with status 3 after printing its line if any
check fails): at full size, on a 2,000-query sample, the device's ids and
distance bits against the oracle at every swept ef and both output orders,
and the reference's own search() against the oracle's replica of its heap
order; the FastScan leg checks its ids and distance bits against the oracle
and, with the reference's own sidecar installed, its ids against the
reference's search_qg on all 10,000 queries.

``cpu_baseline``: the reference's search() (and, in the FastScan leg, its
search_qg) compiled from its own headers, on every host core this process
may use (affinity capped by the cgroup quota; count and CPU model recorded).

``other_workloads``: the 100k x 64 and 10M x 96 configurations at full size
(recall, QPS and roofline per ef); skipped with --no-configs.

Usage:  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                        [--no-parity] [--no-configs] [--no-serve] [--no-fastscan]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
EF_SWEEP = (20, 50, 100, 200, 500)
FS_EF_SWEEP = (50, 100, 128, 150, 200, 300, 500, 700, 1000)
SERVE_RATES = (100_000.0, 1_000_000.0, 4_000_000.0)
SERVE_LANES, SERVE_BATCH, SERVE_SLOTS = 4, 2048, 16384
SERVE_MIX_RATES = (5_000.0, 20_000.0, 50_000.0)
SERVE_MIX_LANES, SERVE_MIX_BATCH = 16, 2048
TARGET_RECALL = 0.95
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "h1_search_kernel_c2_traffic.md")
TRAFFIC_FILE_FS = os.path.join(ROOT, "profiles", "h2_fast_kernel_c2_traffic.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpus():
    """Host cores this process may use: its CPU affinity, capped by the
    cgroup's CPU quota where one is set; and the CPU model."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = max(1, int(int(q) // int(period)))
    except (OSError, ValueError):
        pass
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    threads = min(aff, quota) if quota else aff
    return {"threads": threads, "affinity": aff, "cgroup_quota": quota, "model": model}


def ground_truth(base: np.ndarray, Q: np.ndarray, k: int):
    """FP64 brute force on the device (exact for integer-valued data), ties by id."""
    import torch

    from liteqg_amd.evaluate import ground_truth as gt_fp64

    out = gt_fp64(base, Q, k)
    torch.cuda.empty_cache()
    return out


def recall_at_k(found, gt, k):
    return float(np.mean([len(set(f[:k].tolist()) & set(g[:k].tolist())) / k
                          for f, g in zip(found, gt)]))


class DeviceBatch:
    """Device-resident queries/results and the raw C ABI call on torch's stream."""

    def __init__(self, ix, Q, k, ef_per_query=None, k_per_query=None):
        import torch

        from liteqg_amd import _lib

        self.L = _lib.lib()
        self.check = _lib.check
        self.h = ix._h
        self.nq = Q.shape[0]
        self.k = k
        self.q = torch.from_numpy(Q).cuda()
        # optional per-query ef and k (then launch's ef is their maximum and
        # k the largest k)
        self.d_ef = (None if ef_per_query is None else
                     torch.from_numpy(np.ascontiguousarray(ef_per_query, np.int32)).cuda())
        self.d_k = (None if k_per_query is None else
                    torch.from_numpy(np.ascontiguousarray(k_per_query, np.int32)).cuda())
        self.ids = torch.empty((self.nq, k), dtype=torch.int32, device="cuda")
        self.stats = torch.empty((self.nq, 2), dtype=torch.int32, device="cuda")
        # a stream of our own (default priority): its handle is never NULL, so
        # the kernel runs on exactly the stream the timing events are recorded on
        self.stream = torch.cuda.Stream()
        torch.cuda.synchronize()

    def launch(self, ef, with_stats=False):
        self.check(self.L.lqg_search_batch_device(
            self.h, C.c_void_p(self.q.data_ptr()), self.nq,
            None if self.d_ef is None else C.c_void_p(self.d_ef.data_ptr()),
            None if self.d_k is None else C.c_void_p(self.d_k.data_ptr()), ef, self.k, 1,
            C.c_void_p(self.ids.data_ptr()), None,
            C.c_void_p(self.stats.data_ptr()) if with_stats else None,
            C.c_void_p(self.stream.cuda_stream)))

    def results(self):
        return self.ids.cpu().numpy().view(np.uint32)

    def counters(self):
        st = self.stats.cpu().numpy().view(np.uint32).astype(np.int64)
        return st[:, 0], st[:, 1]


def time_launches(db, ef, steps, warmup):
    import torch

    for _ in range(warmup):
        db.launch(ef)
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    t0.record(db.stream)
    for _ in range(steps):
        db.launch(ef)
    t1.record(db.stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    return t0.elapsed_time(t1) / steps, wall


def graph_leg(ix, db, ef, args):
    """K replays of the batch launch captured into a HIP graph (HIP events on
    the replay stream)."""
    import torch

    from liteqg_amd.serving import BatchGraph

    g = BatchGraph(ix, db.q.data_ptr(), db.nq, db.ids.data_ptr(), db.k, ef)
    for _ in range(args.warmup):
        g.launch(db.stream.cuda_stream)
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(db.stream)
    for _ in range(args.steps):
        g.launch(db.stream.cuda_stream)
    t1.record(db.stream)
    torch.cuda.synchronize()
    g.close()
    return t0.elapsed_time(t1) / args.steps


def steady_state_leg(ix, db, ef, args, depth=4):
    """Batches back to back: `depth` HIP-graph replays of the same batch
    launch, each with its own result buffer and work-queue scratch, taken in
    turn on `depth` streams, so the tail of one batch (its last queries, when
    most waves have no query left) overlaps the start of the next.  K steps
    of one batch each are timed between events on the first stream; every
    replay's ids must equal the single launch's.  HIP maps streams onto its
    few hardware queues (4 per process here), and two streams that share one
    run their launches one after the other; with four streams at least two
    queues take part."""
    import torch

    from liteqg_amd.serving import BatchGraph

    db.launch(ef)
    torch.cuda.synchronize()
    ref = db.results().copy()
    outs = [torch.full((db.nq, db.k), -1, dtype=torch.int32, device="cuda")
            for _ in range(depth)]
    graphs = [BatchGraph(ix, db.q.data_ptr(), db.nq, o.data_ptr(), db.k, ef) for o in outs]
    streams = [torch.cuda.Stream() for _ in range(depth)]

    def run(steps):
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record(streams[0])
        for s in streams[1:]:
            s.wait_event(t0)
        for i in range(steps):
            graphs[i % depth].launch(streams[i % depth].cuda_stream)
        for s in streams[1:]:
            e = torch.cuda.Event()
            e.record(s)
            streams[0].wait_event(e)
        t1.record(streams[0])
        torch.cuda.synchronize()
        return t0.elapsed_time(t1) / steps

    try:
        run(max(args.warmup, depth))
        ms = run(args.steps)
        same = all(np.array_equal(o.cpu().numpy().view(np.uint32), ref) for o in outs)
    finally:
        for g in graphs:
            g.close()
    if not same:
        raise RuntimeError("a back-to-back batch returned ids other than the single launch's")
    return {"streams": depth, "ms_per_batch": round(ms, 4),
            "qps": round(db.nq / (ms / 1e3), 1), "ids_equal_single_launch": same}


def serving_leg(ix, Q, ef, k, args):
    """The serving engine under Poisson arrivals of single-query requests at
    fixed rates, plus one unthrottled run (every request at t = 0).  Latency
    is completion minus scheduled arrival; throughput is requests over the
    span from the first arrival to the last completion."""
    from liteqg_amd.serving import Server, poisson_arrivals

    out = {"requests_per_rate": int(Q.shape[0]), "ef": ef, "k": k, "rates": []}
    with Server(ix, lanes=SERVE_LANES, max_batch=SERVE_BATCH, slots=SERVE_SLOTS, ef_max=ef,
                kmax=k) as srv:
        srv.replay(Q[:2000], ef, k, np.zeros(2000, np.uint64))  # warm-up
        for rate in SERVE_RATES + (None,):
            arr = (np.zeros(Q.shape[0], np.uint64) if rate is None
                   else poisson_arrivals(Q.shape[0], rate, seed=args.seed + 2))
            lat, _ = srv.replay(Q, ef, k, arr)
            span = float((arr.astype(np.float64) + lat).max() - float(arr[0])) / 1e9
            row = {"rate_per_s": rate if rate is not None else "unthrottled",
                   "p50_ms": round(float(np.percentile(lat, 50)) / 1e6, 4),
                   "p99_ms": round(float(np.percentile(lat, 99)) / 1e6, 4),
                   "achieved_qps": round(Q.shape[0] / span, 1)}
            out["rates"].append(row)
            log(f"serving {row}")
        out["engine"] = dict(srv.stats(), lanes=SERVE_LANES, max_batch=SERVE_BATCH,
                             slots=SERVE_SLOTS)
    return out


def serving_mix_leg(args):
    """The serving workload of its own configuration: 1,000,000 unit-norm
    768-d vectors from a 2,048-centre mixture (centres N(0, I), sigma 0.25
    before normalisation), degree 32, graph built on the device (build EF 400, 3
    iterations); Poisson arrivals of single-query requests at 5k, 20k and
    50k per second with per-request ef uniform in {64..512} and k in {10, 100},
    served in both search modes: the SQ8 path (order=nearest) and the FastScan
    path (1-bit estimates, exact re-rank).  Reports p50/p99 latency and
    recall@10 (first 10 ids of each request against FP64 brute force) per
    rate, plus an unthrottled run for the engine's ceiling."""
    from liteqg_amd import datasets
    from liteqg_amd.serving import Server, poisson_arrivals
    from liteqg_amd.symphonyqg import Index

    n, d, deg, nq = args.serve_n, 768, 32, args.serve_nq
    t = time.perf_counter()
    base, Q = datasets.make("unit_mix", n, d, nq, seed=args.seed, centres=2048, sigma=0.25,
                            centre_law="normal")
    ix = Index("QG", "L2", n, d, deg)
    ix.build_index(base, 400, num_iter=3)
    build_s = time.perf_counter() - t
    gt = ground_truth(base, Q, 10)
    del base
    rng = np.random.default_rng(args.seed + 3)
    ef = rng.integers(64, 513, nq).astype(np.uint32)
    k = np.where(rng.random(nq) < 0.5, 10, 100).astype(np.uint32)
    out = {"workload": f"unit_mix N={n} D={d} deg={deg} (2048 centres N(0, I), sigma 0.25, "
                       f"normalised), "
                       f"{nq} requests per rate, ef ~ U{{64..512}}, k ~ {{10, 100}}",
           "data_and_build_s": round(build_s, 1), "rates": []}
    log(f"serving mix: data + build {build_s:.1f}s")

    def serve(key):
        rows = []
        with Server(ix, lanes=SERVE_MIX_LANES, max_batch=SERVE_MIX_BATCH, slots=SERVE_SLOTS,
                    ef_max=512, kmax=100) as srv:
            srv.replay(Q[:1000], ef[:1000], k[:1000], np.zeros(1000, np.uint64))  # warm-up
            for rate in SERVE_MIX_RATES + (None,):
                arr = (np.zeros(nq, np.uint64) if rate is None
                       else poisson_arrivals(nq, rate, seed=args.seed + 2))
                lat, ids = srv.replay(Q, ef, k, arr)
                span = float((arr.astype(np.float64) + lat).max() - float(arr[0])) / 1e9
                row = {"rate_per_s": rate if rate is not None else "unthrottled",
                       "p50_ms": round(float(np.percentile(lat, 50)) / 1e6, 4),
                       "p99_ms": round(float(np.percentile(lat, 99)) / 1e6, 4),
                       "achieved_qps": round(nq / span, 1),
                       "recall_at_10": round(recall_at_k(ids[:, :10], gt, 10), 4)}
                rows.append(row)
                log(f"serving mix ({key}) {row}")
            out[key + "_engine"] = dict(srv.stats(), lanes=SERVE_MIX_LANES,
                                        max_batch=SERVE_MIX_BATCH, slots=SERVE_SLOTS)
        return rows

    # the SQ8 path (what the reference's search() runs), then the FastScan
    # path with its exact re-rank on a device-built sidecar of the same graph
    out["rates"] = serve("sq8")
    t = time.perf_counter()
    ix.build_sidecar(seed=args.seed)
    ix.set_mode("fastscan")
    out["fastscan_sidecar_build_s"] = round(time.perf_counter() - t, 2)
    out["fastscan_rates"] = serve("fastscan")
    ix.set_mode("sq8")
    return out


def wide_leg(args, host):
    """The 960-d workload at full size: 1,000,000 x 960 non-negative mixture
    (512 centres U[0,1]^960, sigma 0.05, clipped at 0), degree 64, graph built
    on the device (build EF 500, 3 iterations); 1,000 queries from the same
    process with per-query ef ~ U{64..1024} and k ~ {10, 100}, all of them in
    one persistent launch (pool class of ef 1024, launch settings autotuned
    for that class), in both search modes: SQ8 (order nearest) and FastScan
    with its exact re-rank.  Recall@10 (first 10 ids of each row) against
    FP64 brute force, QPS, and the roofline on algorithmic bytes:
      SQ8:      4D + 8D + 4*deg*E + D*V + 4k  (E, V exact per query)
      FastScan: 4D + E*(4D + deg*P/8 + 16*deg) + 4k
    plus per-query parity against the oracle on a sample (ids and distance
    bits)."""
    import torch

    from liteqg_amd import datasets
    from liteqg_amd.symphonyqg import Index

    n, d, deg, nq, efb = args.wide_n, 960, 64, 1000, 500
    t = time.perf_counter()
    base, Q = datasets.make("pos_mix", n, d, nq, seed=args.seed, centres=512, sigma=0.05)
    ix = Index("QG", "L2", n, d, deg)
    ix.build_index(base, efb, num_iter=3)
    build_s = time.perf_counter() - t
    log(f"wide: data + build {build_s:.1f}s")
    gt = ground_truth(base, Q, 10)
    rng = np.random.default_rng(args.seed + 4)
    ef = rng.integers(64, 1025, nq).astype(np.uint32)
    k = np.where(rng.random(nq) < 0.5, 10, 100).astype(np.uint32)
    ef_max, kmax = int(ef.max()), int(k.max())
    db = DeviceBatch(ix, Q, kmax, ef_per_query=ef, k_per_query=k)
    p = ix.padded_dim
    out = {"workload": f"pos_mix N={n} D={d} deg={deg} (512 centres U[0,1]^D, sigma 0.05, "
                       f"clipped at 0), {nq} queries, ef ~ U{{64..1024}}, k ~ {{10, 100}}, "
                       f"one launch; graph built on device with build EF {efb} x3 iterations",
           "data_and_build_s": round(build_s, 1)}

    def timed(mode):
        tuned = ix.autotune(Q, ef_max, kmax)
        db.launch(ef_max, with_stats=True)
        torch.cuda.synchronize()
        ids = db.results()
        E, V = db.counters()
        ms, _ = time_launches(db, ef_max, args.steps, args.warmup)
        rec = recall_at_k(ids[:, :10], gt, 10)
        return dict(mode=mode, launch=tuned, ms_per_step=round(ms, 4),
                    qps=round(nq / ms * 1e3, 1), recall_at_10=round(rec, 4)), ids, E, V, ms

    # SQ8: E and V of the reference's visited set from the exact counting
    # launch, one call per distinct ef (the traversal does not depend on k)
    row, ids_sq8, E, _, ms = timed("sq8")
    ev = np.zeros((nq, 2), np.int64)
    for e in np.unique(ef):
        sel = np.nonzero(ef == e)[0]
        ev[sel] = ix.count_visits(Q[sel], int(e), 10)
    if not np.array_equal(ev[:, 0], E):
        raise RuntimeError("wide: expansion counts differ between the timed and counting launch")
    alg = float(np.sum(4 * d + 8 * d + 4 * deg * ev[:, 0] + d * ev[:, 1] + 4 * k.astype(np.int64)))
    gbs = alg / (ms / 1e3) / 1e9
    row["roofline"] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                       "bytes_per_query": round(alg / nq, 1),
                       "expansions_per_query": round(float(ev[:, 0].mean()), 2),
                       "distinct_evaluations_per_query": round(float(ev[:, 1].mean()), 2)}
    out["sq8"] = row
    log(f"wide sq8: {row}")

    ix.build_sidecar(seed=args.seed)
    ix.set_mode("fastscan")
    row, ids_fs, E, _, ms = timed("fastscan")
    rec_bytes = 4 * d + deg * p // 8 + 16 * deg
    alg = float(np.sum(4 * d + E.astype(np.int64) * rec_bytes + 4 * k.astype(np.int64)))
    gbs = alg / (ms / 1e3) / 1e9
    row["roofline"] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                       "bytes_per_query": round(alg / nq, 1), "record_bytes": rec_bytes,
                       "expansions_per_query": round(float(E.mean()), 2)}
    out["fastscan"] = row
    log(f"wide fastscan: {row}")

    if not args.no_parity:
        from oracle import oracle as O

        sample = np.arange(0, nq, nq // args.wide_parity_nq)[:args.wide_parity_nq]
        with tempfile.TemporaryDirectory() as tmp:
            path = os.path.join(tmp, "w.index")
            ix.save(path)
            arr = O.read_index(path, n, d, deg)
        s_, codes, fac = ix.get_sidecar()
        H = O.H2Arrays(arr, s_, codes, fac, 1)
        dis = {}
        for mode in ("sq8", "fastscan"):
            ix.set_mode(mode)
            _, dis[mode] = ix.search_batch(Q, k, ef=ef, order="nearest", return_distances=True)
        ok = {"sq8": True, "fastscan": True}
        for i in sample:
            ki, ei = int(k[i]), int(ef[i])
            e_ids, e_d, _ = O.search_canon_batch(arr, Q[i:i + 1], ei, ki, "nearest")
            ok["sq8"] &= bool(np.array_equal(ids_sq8[i, :ki], e_ids[0]) and np.array_equal(
                dis["sq8"][i, :ki].view(np.uint32), e_d[0].view(np.uint32)))
            e_ids, e_d, _ = O.search_h2_batch(H, Q[i:i + 1], ei, ki)
            ok["fastscan"] &= bool(np.array_equal(ids_fs[i, :ki], e_ids[0]) and np.array_equal(
                dis["fastscan"][i, :ki].view(np.uint32), e_d[0].view(np.uint32)))
        out["parity_vs_oracle"] = {"queries": int(sample.size), "ids_and_dist_bits_equal": ok,
                                   "ok": ok["sq8"] and ok["fastscan"]}
        log(f"wide parity: {out['parity_vs_oracle']}")
        del arr, H, codes, fac
    ix.set_mode("sq8")
    del ix, db, base
    torch.cuda.empty_cache()
    return out


def wide_probe(args):
    """--pmc-probe on the 960-d workload of wide_leg (pos_mix --wide-n x 960,
    degree 64, 1,000 queries at per-query ef U{64..1024} and k {10, 100}, the
    same draws): --probe-launches launches of the probe's mode (sq8-wide or
    fastscan-wide) with the --launch settings, then a 1 GiB device copy."""
    import torch

    from liteqg_amd import datasets
    from liteqg_amd.symphonyqg import Index

    n, d, deg, nq = args.wide_n, 960, 64, 1000
    base, Q = datasets.make("pos_mix", n, d, nq, seed=args.seed, centres=512, sigma=0.05)
    ix = Index("QG", "L2", n, d, deg)
    ix.build_index(base, 500, num_iter=3)
    rng = np.random.default_rng(args.seed + 4)
    ef = rng.integers(64, 1025, nq).astype(np.uint32)
    k = np.where(rng.random(nq) < 0.5, 10, 100).astype(np.uint32)
    db = DeviceBatch(ix, Q, int(k.max()), ef_per_query=ef, k_per_query=k)
    if args.probe_mode == "fastscan-wide":
        ix.build_sidecar(seed=args.seed)
        ix.set_mode("fastscan")
    ef_max = int(ef.max())
    if args.launch:
        ix.set_launch(ef_max, *[int(x) for x in args.launch.split(",")])
    else:
        ix.autotune(Q, ef_max, int(k.max()))
    ms, _ = time_launches(db, ef_max, args.probe_launches, 2)
    db.launch(ef_max, with_stats=True)
    torch.cuda.synchronize()
    n_exp, _ = db.counters()
    log(f"wide probe counters: {int(n_exp.sum())} expansions, {n_exp.mean():.2f} per query, "
        f"max {int(n_exp.max())}")
    x = torch.zeros(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    log(f"pmc probe: {args.probe_launches} {args.probe_mode} launches of {nq} queries "
        f"(N={n}), launch {ix.get_launch(ef_max)}, {ms:.4f} ms per launch (HIP events), "
        f"then a 1 GiB copy")


def configs_leg(args, host):
    """Two more workloads at full size, SQ8 path, order=nearest, graphs built on
    the device (3 iterations), FP64 ground truth, 10 timed launches per ef,
    with the roofline of each launch (algorithmic bytes from the exact visit
    counts, as for the headline workload):
      100,000 x 64 i.i.d. U[0,1) (the reference README's example), degree 32,
        build EF 200, 1,000 queries, ef {50, 100, 200}, k 10; the plain path,
        the reference's CPU search at ef 100 and oracle parity beside it;
      10,000,000 x 96 unit-norm, 4,096 centres N(0, I), sigma 0.3 before
        normalisation, degree 32, build EF 300, 10,000 queries, ef {50, 200},
        k 10; the batch also replayed as a HIP graph."""
    import torch

    from liteqg_amd import datasets
    from liteqg_amd.symphonyqg import Index

    out = {}
    for name, n, d, efb, nq, kind, kw, efs in (
            ("uniform_100k_d64", 100_000, 64, 200, 1_000, "uniform", {}, (50, 100, 200)),
            ("unitmix_10m_d96", 10_000_000, 96, 300, 10_000, "unit_mix",
             dict(centres=4096, sigma=0.3, centre_law="normal"), (50, 200))):
        t = time.perf_counter()
        base, Q = datasets.make(kind, n, d, nq, seed=args.seed, **kw)
        ix = Index("QG", "L2", n, d, 32)
        ix.build_index(base, efb, num_iter=3)
        build_s = time.perf_counter() - t
        gt = ground_truth(base, Q, 10)
        db = DeviceBatch(ix, Q, 10)
        rows = []
        for ef in efs:
            db.launch(ef)
            torch.cuda.synchronize()
            rec = recall_at_k(db.results(), gt, 10)
            ms, _ = time_launches(db, ef, 10, 2)
            ev = ix.count_visits(Q, ef, 10).astype(np.int64)
            alg = float(np.sum(4 * d + 8 * d + 4 * 32 * ev[:, 0] + d * ev[:, 1] + 4 * 10))
            gbs = alg / (ms / 1e3) / 1e9
            row = {"ef": ef, "recall": round(rec, 4), "qps": round(nq / ms * 1e3, 1),
                   "ms_per_step": round(ms, 4),
                   "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                                "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                                "bytes_per_query": round(alg / nq, 1)}}
            if name.startswith("unitmix"):
                row["graph_replay_ms"] = round(graph_leg(ix, db, ef, args), 4)
            rows.append(row)
        res = {"data_and_build_s": round(build_s, 1), "queries": nq, "ef_sweep": rows}
        if name.startswith("uniform"):
            ix.set_ef(100)
            ix.set_order("nearest")
            t = time.perf_counter()
            for q in Q[:300]:
                ix.search(q, 10)
            res["plain_qps_ef100"] = round(300 / (time.perf_counter() - t), 1)
            with tempfile.TemporaryDirectory() as tmp:
                p = os.path.join(tmp, "c.index")
                ix.save(p)
                if not args.no_parity:
                    from oracle import oracle as O

                    arr = O.read_index(p, n, d, 32)
                    par = []
                    for ef in efs:
                        ids, dis = ix.search_batch(Q, 10, ef=ef, order="nearest",
                                                   return_distances=True)
                        e_ids, e_dis, _ = O.search_canon_batch(arr, Q, ef, 10, "nearest",
                                                               threads=host["threads"])
                        par.append(dict(ef=ef, ids_equal=bool(np.array_equal(ids, e_ids)),
                                        dist_bits_equal=bool(np.array_equal(
                                            dis.view(np.uint32), e_dis.view(np.uint32)))))
                    res["parity_vs_oracle"] = par
                if not args.no_cpu:
                    res["cpu_reference_ef100"] = cpu_baseline(p, n, d, 32, Q, 100, 10, host)
        out[name] = res
        log(f"config {name}: {res}")
        del ix, db, base
        torch.cuda.empty_cache()
    return out


def fastscan_leg(ix, db, Q, gt, k, args, ctx):
    """The FastScan RaBitQ path (search_qg) on the same index and queries:
    device sidecar build, ef sweep to recall@10 >= 0.95, batched timing, and
    its roofline on algorithmic bytes per query
      4D (query) + E * (4D + deg*P/8 + 16*deg) (node records) + 4k (ids out)
    with E the exact expansion count the kernel reports.  Then parity at full
    size (the oracle's search_h2 on a sample; the reference's own search_qg on
    its own sidecar for all queries) and the reference's search_qg timed on
    the host cores as this path's CPU baseline."""
    import torch

    d, deg = ix.dimension, ix.degree_bound
    p = ix.padded_dim
    t = time.perf_counter()
    ix.build_sidecar(seed=args.seed)
    torch.cuda.synchronize()
    side_s = time.perf_counter() - t
    ix.set_mode("fastscan")
    sweep, chosen = [], None
    for ef in FS_EF_SWEEP:
        db.launch(ef)
        torch.cuda.synchronize()
        rec = recall_at_k(db.results(), gt, k)
        ms, _ = time_launches(db, ef, 2, 0)
        sweep.append(dict(ef=ef, recall=round(rec, 4), qps=round(args.nq / ms * 1e3, 1)))
        log(f"fastscan ef={ef} recall={rec:.4f} qps={args.nq / ms * 1e3:.0f}")
        if rec >= TARGET_RECALL:
            chosen = ef
            break
    if chosen is None:
        chosen = sweep[-1]["ef"]
    rec = next(s["recall"] for s in sweep if s["ef"] == chosen)
    tuned = ix.autotune(Q, chosen, k)
    ms_step, _ = time_launches(db, chosen, args.steps, args.warmup)
    qps = args.nq / (ms_step / 1e3)
    db.launch(chosen, with_stats=True)
    torch.cuda.synchronize()
    E, pops = db.counters()
    rec_bytes = 4 * d + deg * p // 8 + 16 * deg
    alg = float(np.sum(4 * d + E * rec_bytes + 4 * k))
    achieved = alg / (ms_step / 1e3) / 1e9
    steady = steady_state_leg(ix, db, chosen, args)
    steady_gbs = alg / (steady["ms_per_batch"] / 1e3) / 1e9
    steady.update({"algorithmic_GBps": round(steady_gbs, 1),
                   "frac_of_peak": round(steady_gbs / HBM_PEAK_GBS, 4)})
    log(f"fastscan batches back to back: {steady}")
    # the matrix-core rotation: its error against FP64 (the FHT's beside it),
    # then the same search with GEMM-rotated queries
    import scipy.linalg

    qs = Q[:1000]
    signs = ix.get_sidecar(signs_only=True)
    qpad = np.zeros((qs.shape[0], p))
    qpad[:, :d] = qs.astype(np.float64) * signs[:d].astype(np.float64)
    exact = qpad @ scipy.linalg.hadamard(p).astype(np.float64).T
    scale = np.abs(exact).max(axis=1, keepdims=True)
    rot_err = {m: float((np.abs(ix.rotate_queries(qs, m) - exact) / scale).max())
               for m in ("fht", "gemm")}
    ids_fht = db.results()
    ix.set_rotation("gemm")
    ms_gemm, _ = time_launches(db, chosen, args.steps, args.warmup)
    ids_gemm = db.results()
    ix.set_rotation("fht")
    gemm = {"rel_err_vs_fp64": rot_err["gemm"], "fht_rel_err_vs_fp64": rot_err["fht"],
            "ms_per_step": round(ms_gemm, 4), "qps": round(args.nq / ms_gemm * 1e3, 1),
            "recall_at_10": round(recall_at_k(ids_gemm, gt, k), 4),
            "ids_overlap_with_fht": round(float(np.mean(
                [len(set(a) & set(b)) / len(a) for a, b in zip(ids_fht, ids_gemm)])), 5)}
    log(f"fastscan gemm rotation: {gemm}")
    # parity: the oracle on the device's sidecar, at the operating point and
    # the ef below it in the sweep
    par_efs = sorted({chosen, sweep[max(0, len(sweep) - 2)]["ef"]})
    par = {"queries": int(ctx["sample"].shape[0]),
           "vs_oracle": parity_fastscan(ix, ctx["arr"], ctx["sample"], par_efs, k, ctx["host"])}
    log(f"fastscan parity vs oracle: {par['vs_oracle']}")
    cpu = None
    if not args.no_cpu:
        cpu, r_ids, side = cpu_baseline_fastscan(ctx["path"], ix.num_elements, d, deg, signs, Q,
                                                 chosen, k, ctx["host"], ctx["tmp"], True)
        log(f"fastscan cpu baseline {cpu}")
        if side is not None:
            # the reference's own codes and factors on the device: its
            # search_qg and this kernel then read the same bytes
            ix.set_sidecar(signs, side[0], side[1], layout="blocks")
            del side
            ref_dev = ix.search_batch(Q, k, ef=chosen)
            par["vs_reference_search_qg"] = {
                "queries": int(Q.shape[0]), "ef": chosen,
                "sidecar": "the reference's rabitq_codes rows, installed on the device",
                "rows_identical": round(float(np.mean(np.all(ref_dev == r_ids, 1))), 5),
                "rows_equal_up_to_ties": round(float(np.mean(
                    same_up_to_ties(ctx["base"], Q, ref_dev, r_ids))), 5)}
            log(f"fastscan parity vs reference: {par['vs_reference_search_qg']}")
            ix.build_sidecar(seed=args.seed)  # back to the device-built sidecar
    par["ok"] = all(r["ids_equal"] and r["dist_bits_equal"] for r in par["vs_oracle"]) and \
        par.get("vs_reference_search_qg", {}).get("rows_equal_up_to_ties", 1.0) == 1.0
    # the plain path in this mode: one Index.search call per query
    ix.set_ef(chosen)
    lat = []
    for q in Q[:args.plain]:
        t0 = time.perf_counter()
        ix.search(q, k)
        lat.append((time.perf_counter() - t0) * 1e3)
    lat = np.array(lat)
    plain = {"qps": round(len(lat) / (lat.sum() / 1e3), 1),
             "p50_ms": round(float(np.median(lat)), 4),
             "p99_ms": round(float(np.percentile(lat, 99)), 4), "queries": len(lat)}
    log(f"fastscan plain path: {plain}")
    ix.set_mode("sq8")
    log(f"fastscan timed ef={chosen}: {ms_step:.3f} ms/step, {qps:.0f} QPS, "
        f"{achieved:.0f} GB/s")
    return {"path": "FastScan RaBitQ search_qg (1-bit codes, exact re-rank)", "ef": chosen,
            "recall_at_10": rec, "qps": round(qps, 1), "ms_per_step": round(ms_step, 4),
            "ef_sweep": sweep, "sidecar_build_s": round(side_s, 2), "launch": tuned,
            "gemm_rotation": gemm, "steady_state": steady, "parity": par,
            "plain_gpu": plain, "batched_over_plain": round(qps / plain["qps"], 1),
            "cpu_baseline": cpu,
            "gpu_over_cpu": round(qps / cpu["value"], 1) if cpu else None,
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                         "traffic": read_traffic(TRAFFIC_FILE_FS),
                         "bytes_per_query": round(alg / args.nq, 1),
                         "record_bytes": rec_bytes,
                         "expansions_per_query": round(float(E.mean()), 2),
                         "pops_per_query": round(float(pops.mean()), 2)}}


def exact_dists(base, Q, ids):
    """FP64 squared L2 from each query to each returned id (inf for
    0xFFFFFFFF slots)."""
    ok = ids != 0xFFFFFFFF
    B = base[np.where(ok, ids, 0).astype(np.int64)].astype(np.float64)
    d = ((B - Q[:, None, :].astype(np.float64)) ** 2).sum(-1)
    return np.where(ok, d, np.inf)


def same_up_to_ties(base, Q, a, b):
    """Per query: the two id lists hold the same multiset of exact distances
    (they may differ only where distances are equal)."""
    da, db = np.sort(exact_dists(base, Q, a), 1), np.sort(exact_dists(base, Q, b), 1)
    return np.all(da == db, axis=1)


def ref_search_ids(ix_path, n, d, deg, Q, ef, k, tmp):
    """The reference's own Index.search (set_ef, then search per query:
    qg.hpp:226-243, k farthest of the final pool) via oracle/_ref/ref_driver."""
    drv = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    if not os.path.exists(drv):
        return None
    qp, op = os.path.join(tmp, "qpar.f32"), os.path.join(tmp, "ref_ids.u32")
    np.ascontiguousarray(Q, np.float32).tofile(qp)
    r = subprocess.run([drv, "search", ix_path, str(n), str(d), str(deg), qp, str(Q.shape[0]),
                        str(ef), str(k), op], capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        log("ref_driver search failed:", r.returncode, r.stderr[-300:])
        return None
    return np.fromfile(op, np.uint32).reshape(Q.shape[0], k)


def parity_sq8(ix, arr, base, sample, efs, k, host, ix_path, tmp):
    """Full-size parity of the SQ8 path on a query sample:
    - against the oracle's canonical search (orc_search_canon, keys ordered by
      (distance, id)) at every swept ef, both output orders: ids and distance
      bits must be identical;
    - against the reference's own search() at every swept ef.  The device
      orders equal distances by id, the reference by its heap's sift order
      (libstdc++ push_heap / pop_heap), and where two distances tie the walk
      itself can take another path.  The reference's ids must equal those of
      the oracle's replica of its heap order (orc_search_heap, the canonical
      search with the heap's tie order) on every query; the device's equal
      the canonical search's (above).  The two oracles differ in nothing but
      the order of equal keys, so every device/reference difference is a tie
      difference.  Reported beside it: rows identical, and rows that hold the
      same multiset of exact distances (exact here: the data are integers
      0..255, so SQ8 is lossless and every FP32 sum is exact)."""
    from oracle import oracle as O

    oracle_rows, ref_rows = [], []
    for ef in efs:
        for order in ("reference", "nearest"):
            ids, dis = ix.search_batch(sample, k, ef=ef, order=order, return_distances=True)
            e_ids, e_dis, _ = O.search_canon_batch(arr, sample, ef, k, order,
                                                   threads=host["threads"])
            oracle_rows.append(dict(
                ef=ef, order=order, ids_equal=bool(np.array_equal(ids, e_ids)),
                dist_bits_equal=bool(np.array_equal(dis.view(np.uint32),
                                                    e_dis.view(np.uint32)))))
        r_ids = ref_search_ids(ix_path, arr.n, arr.dim, arr.deg, sample, ef, k, tmp)
        if r_ids is not None:
            ids = ix.search_batch(sample, k, ef=ef, order="reference")
            h_ids = O.search_heap_batch(arr, sample, ef, k, threads=host["threads"])
            ref_rows.append(dict(
                ef=ef,
                reference_equals_heap_replica=round(float(np.mean(np.all(h_ids == r_ids, 1))),
                                                    5),
                rows_identical=round(float(np.mean(np.all(ids == r_ids, 1))), 5),
                rows_equal_up_to_ties=round(float(np.mean(same_up_to_ties(base, sample, ids,
                                                                          r_ids))), 5)))
    ok = all(r["ids_equal"] and r["dist_bits_equal"] for r in oracle_rows) and \
        all(r["reference_equals_heap_replica"] == 1.0 for r in ref_rows)
    return {"queries": int(sample.shape[0]), "vs_oracle": oracle_rows,
            "vs_reference_search": ref_rows, "ok": ok}


def parity_fastscan(ix, arr, sample, efs, k, host):
    """Full-size parity of the FastScan path against the oracle's search_h2
    on the device's own sidecar: ids and distance bits at each ef."""
    from oracle import oracle as O

    signs, codes, factors = ix.get_sidecar()
    H = O.H2Arrays(arr, signs, codes, factors, 1)
    rows = []
    for ef in efs:
        ids, dis = ix.search_batch(sample, k, ef=ef, return_distances=True)
        e_ids, e_dis, _ = O.search_h2_batch(H, sample, ef, k, threads=host["threads"])
        rows.append(dict(ef=ef, ids_equal=bool(np.array_equal(ids, e_ids)),
                         dist_bits_equal=bool(np.array_equal(dis.view(np.uint32),
                                                             e_dis.view(np.uint32)))))
    return rows


def read_traffic(path=TRAFFIC_FILE):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(ix_path, n, d, deg, Q, ef, k, host, nq=2000):
    """The reference's search() (SQ8 path) timed on the host cores
    (oracle/_ref/ref_driver), else the oracle port."""
    threads = host["threads"]
    drv = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    sample = Q[:nq]
    if os.path.exists(drv) and ix_path:
        with tempfile.TemporaryDirectory() as tmp:
            qp = os.path.join(tmp, "q.f32")
            sample.tofile(qp)
            try:
                r = subprocess.run([drv, "time", ix_path, str(n), str(d), str(deg), qp,
                                    str(sample.shape[0]), str(ef), str(k), str(threads), "3"],
                                   capture_output=True, text=True, timeout=600)
                if r.returncode == 0:
                    res = json.loads(r.stdout.strip().splitlines()[-1])
                    return dict(value=round(res["qps_all"], 1), unit="queries/s", cores=threads,
                                kind="reference", host=host,
                                single_thread_p50_ms=res["p50_ms"],
                                single_thread_p99_ms=res["p99_ms"],
                                sample=f"{sample.shape[0]} queries, ef={ef}, k={k}, reference "
                                       f"search() (same traversal; its output order differs), "
                                       f"{threads} threads (every host core this process may "
                                       f"use) each with its own index copy, median of 3 rounds; "
                                       f"single-thread p50 {res['p50_ms']:.3f} ms "
                                       f"p99 {res['p99_ms']:.3f} ms")
                log("ref_driver failed:", r.returncode, r.stderr[-300:])
            except Exception as e:  # noqa: BLE001
                log("ref_driver unavailable:", e)
    from oracle import oracle as O

    arr = O.read_index(ix_path, n, d, deg)
    t0 = time.perf_counter()
    O.search_canon_batch(arr, sample, ef, k, "nearest", threads=threads)
    dt = time.perf_counter() - t0
    return dict(value=round(sample.shape[0] / dt, 1), unit="queries/s", cores=threads,
                kind="port", host=host,
                sample=f"{sample.shape[0]} queries, ef={ef}, k={k}, C restatement, "
                       f"OpenMP over queries")


def cpu_baseline_fastscan(ix_path, n, d, deg, signs, Q, ef, k, host, tmp, want_sidecar):
    """The reference's search_qg (FastScan path) timed on the host cores: the
    reference's own rows rebuilt with its rabitq_codes under the device's
    rotation signs (oracle/_ref/ref_h2 time), all 10,000 queries, one graph
    copy per thread, median of 3 rounds.  Also returns the reference's ids of
    those queries and, if asked, its sidecar (codes in its block layout,
    factors) for a search on the device with the very same codes."""
    drv = os.path.join(ROOT, "oracle", "_ref", "ref_h2")
    if not os.path.exists(drv):
        return None, None, None
    sp, qp = os.path.join(tmp, "signs.f32"), os.path.join(tmp, "qfs.f32")
    idp, scp = os.path.join(tmp, "ref_fs_ids.u32"), os.path.join(tmp, "ref_sidecar.bin")
    np.ascontiguousarray(signs, np.float32).tofile(sp)
    np.ascontiguousarray(Q, np.float32).tofile(qp)
    nq, threads = Q.shape[0], host["threads"]
    r = subprocess.run([drv, "time", ix_path, str(n), str(d), str(deg), sp, qp, str(nq), str(ef),
                        str(k), str(threads), "3", idp, scp if want_sidecar else "-"],
                       capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        log("ref_h2 time failed:", r.returncode, r.stderr[-300:])
        return None, None, None
    res = json.loads(r.stdout.strip().splitlines()[-1])
    ids = np.fromfile(idp, np.uint32).reshape(nq, k)
    side = None
    if want_sidecar:
        p = 1 << (d - 1).bit_length()
        raw = np.fromfile(scp, np.uint8).reshape(n, deg * p // 8 + 12 * deg)
        side = (np.ascontiguousarray(raw[:, :deg * p // 8]),
                np.ascontiguousarray(raw[:, deg * p // 8:]).view(np.float32))
        os.remove(scp)
    base = dict(value=round(res["qps_all"], 1), unit="queries/s", cores=threads,
                kind="reference", host=host, single_thread_p50_ms=res["p50_ms"],
                single_thread_p99_ms=res["p99_ms"],
                sample=f"{nq} queries, ef={ef}, k={k}, the reference's search_qg on its own "
                       f"rows rebuilt with rabitq_codes (rebuild {res['sidecar_s']:.1f} s, not "
                       f"timed), {threads} threads (every host core this process may use) "
                       f"each with its own graph copy, median of 3 rounds; single-thread p50 "
                       f"{res['p50_ms']:.3f} ms p99 {res['p99_ms']:.3f} ms")
    return base, ids, side


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--nq", type=int, default=10_000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--ef", type=int, default=0, help="skip the sweep and use this ef")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fastscan", action="store_true", help="skip the FastScan RaBitQ leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving legs")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the 100k x 64 and 10M x 96 workloads (about a minute)")
    ap.add_argument("--no-parity", action="store_true",
                    help="skip the full-size parity checks against the oracle and the reference")
    ap.add_argument("--parity-nq", type=int, default=2000,
                    help="queries of the full-size parity sample")
    ap.add_argument("--no-wide", action="store_true",
                    help="skip the 1M x 960 degree-64 workload (about 90 s)")
    ap.add_argument("--wide-n", type=int, default=1_000_000,
                    help="base vectors of the 960-d workload")
    ap.add_argument("--wide-parity-nq", type=int, default=50,
                    help="queries of the 960-d workload checked against the oracle")
    ap.add_argument("--serve-n", type=int, default=1_000_000,
                    help="base vectors of the 768-d serving workload")
    ap.add_argument("--serve-nq", type=int, default=10_000,
                    help="requests per arrival rate of the 768-d serving workload")
    ap.add_argument("--plain", type=int, default=500, help="queries timed on the plain path")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="for rocprofv3 runs: build, --probe-launches launches at --ef (default 50) with "
                         "the tuned launch settings, then a 1 GiB device copy (counter scale)")
    ap.add_argument("--probe-mode", choices=("sq8", "fastscan", "sq8-wide", "fastscan-wide"),
                    default="sq8",
                    help="search path of the --pmc-probe launches (fastscan: default ef 200; "
                         "*-wide: the 960-d workload at --wide-n nodes, per-query ef and k)")
    ap.add_argument("--probe-launches", type=int, default=3,
                    help="number of --pmc-probe search launches")
    ap.add_argument("--launch", default="",
                    help="pin the launch settings of the probe instead of autotuning: "
                         "waves_per_group,cache_bits,occupancy[,groups[,prefetch[,split]]]")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the query path is data-parallel over queries and "
              "this build runs on a single MI355X; --gpus above 1 is not supported.",
              file=sys.stderr)
        sys.exit(2)

    if args.pmc_probe and args.probe_mode.endswith("-wide"):
        wide_probe(args)
        return

    import torch

    from liteqg_amd import datasets
    from liteqg_amd.symphonyqg import Index

    n, d, deg, k, ef_build = args.n, 128, 32, 10, 400
    t = time.perf_counter()
    base, Q = datasets.make("int_mix", n, d, args.nq, seed=args.seed, centres=1024, sigma=20.0)
    log(f"data {time.perf_counter() - t:.1f}s")
    t = time.perf_counter()
    ix = Index("QG", "L2", n, d, deg)
    ix.build_index(base, ef_build, num_iter=3)
    build_s = time.perf_counter() - t
    log(f"build {build_s:.1f}s")
    if args.pmc_probe:
        db = DeviceBatch(ix, Q, k)
        if args.probe_mode == "fastscan":
            ef = args.ef or 200
            ix.build_sidecar(seed=args.seed)
            ix.set_mode("fastscan")
        else:
            ef = args.ef or 50
        if args.launch:
            v = [int(x) for x in args.launch.split(",")]
            ix.set_launch(ef, *v)
        else:
            ix.autotune(Q, ef, k)
        ms, _ = time_launches(db, ef, args.probe_launches, 2)
        db.launch(ef, with_stats=True)
        torch.cuda.synchronize()
        n_exp, n_pop = db.counters()
        pct = np.percentile(n_exp, [50, 90, 99, 100]).astype(int).tolist()
        log(f"probe counters: {n_exp.mean():.2f} expansions, {n_pop.mean():.2f} pops per query "
            f"(nq {db.nq}); expansions p50/p90/p99/max {pct}")
        x = torch.zeros(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)  # reads 1 GiB, writes 1 GiB
        torch.cuda.synchronize()
        log(f"pmc probe: {args.probe_launches} launches at ef={ef}, launch {ix.get_launch(ef)}, "
            f"{ms:.4f} ms per launch (HIP events), then a 1 GiB copy")
        return
    t = time.perf_counter()
    gt = ground_truth(base, Q, k)
    log(f"ground truth {time.perf_counter() - t:.1f}s")
    host = host_cpus()
    log(f"host cpus {host}")
    tmpdir = tempfile.TemporaryDirectory()
    ix_path = os.path.join(tmpdir.name, "c2.index")
    ix.save(ix_path)
    from oracle import oracle as O  # the checker: parity and CPU-baseline legs only

    arr = O.read_index(ix_path, n, d, deg)
    sample = Q[:args.parity_nq]

    db = DeviceBatch(ix, Q, k)
    sweep = []
    chosen = None
    for ef in ([args.ef] if args.ef else EF_SWEEP):
        db.launch(ef)
        torch.cuda.synchronize()
        rec = recall_at_k(db.results(), gt, k)
        ms, _ = time_launches(db, ef, 2, 0)
        sweep.append(dict(ef=ef, recall=round(rec, 4), qps=round(args.nq / ms * 1e3, 1)))
        log(f"ef={ef} recall={rec:.4f} qps={args.nq / ms * 1e3:.0f}")
        if chosen is None and rec >= TARGET_RECALL:
            chosen = ef
    if chosen is None:
        chosen = sweep[-1]["ef"]
    rec = next(s["recall"] for s in sweep if s["ef"] == chosen)
    t = time.perf_counter()
    tuned = ix.autotune(Q, chosen, k)
    log(f"autotune ef={chosen}: {tuned} ({time.perf_counter() - t:.1f}s)")

    # timed region: K launches over the resident queries
    ms_step, wall = time_launches(db, chosen, args.steps, args.warmup)
    qps = args.nq / (ms_step / 1e3)
    db.launch(chosen, with_stats=True)
    torch.cuda.synchronize()
    E, V = db.counters()
    log(f"timed ef={chosen}: {ms_step:.3f} ms/step, {qps:.0f} QPS")

    # the general float SQ8 kernel on the same workload (the integer kernel
    # above is exact only because these vectors are integers 0..255), with
    # launch settings tuned for it; the integer kernel's are restored after
    ix.set_sq8_integer(False)
    ids_int = db.results().copy()
    float_launch = ix.autotune(Q, chosen, k)
    db.launch(chosen)
    torch.cuda.synchronize()
    float_same = bool(np.array_equal(db.results(), ids_int))
    ms_float, _ = time_launches(db, chosen, args.steps, args.warmup)
    ix.set_sq8_integer(True)
    ix.set_launch(chosen, **tuned)
    db.launch(chosen)
    torch.cuda.synchronize()
    log(f"float SQ8 kernel: {ms_float:.3f} ms/step ({float_launch}), ids equal to the integer "
        f"kernel: {float_same}")

    # the same launch captured into a HIP graph and replayed
    graph_ms = graph_leg(ix, db, chosen, args)
    log(f"graph replay: {graph_ms:.3f} ms/step")
    steady = steady_state_leg(ix, db, chosen, args)
    log(f"batches back to back: {steady}")

    # plain path: one query per Index.search call
    ix.set_ef(chosen)
    ix.set_order("nearest")
    lat = []
    for q in Q[:args.plain]:
        t0 = time.perf_counter()
        ix.search(q, k)
        lat.append((time.perf_counter() - t0) * 1e3)
    lat = np.array(lat)
    plain_qps = len(lat) / (lat.sum() / 1e3)

    # roofline: algorithmic bytes per query, SQ8 path
    #   4D (query) + 8D (SQ8 params) + 4*deg*E (id rows) + D*V (codes) + 4k (ids out)
    # with E and V the reference's counts: expansions and distinct nodes
    # evaluated, from the exact counting launch (the timed kernel's own
    # evaluation count V_eval also includes re-evaluations its lossy visited
    # cache lets through).
    exact = ix.count_visits(Q, chosen, k).astype(np.int64)
    if not np.array_equal(exact[:, 0], E):
        raise RuntimeError("expansion counts differ between the timed and the counting launch")
    V_ref = exact[:, 1]
    alg_bytes = float(np.sum(4 * d + 8 * d + 4 * deg * E + d * V_ref + 4 * k))
    achieved = alg_bytes / (ms_step / 1e3) / 1e9
    traffic = read_traffic()

    parity = None
    if not args.no_parity:
        t = time.perf_counter()
        parity = parity_sq8(ix, arr, base, sample, [r["ef"] for r in sweep], k, host, ix_path,
                            tmpdir.name)
        log(f"parity (SQ8) {parity} ({time.perf_counter() - t:.1f}s)")

    serving = None if args.no_serve else serving_leg(ix, Q, chosen, k, args)

    ctx = dict(arr=arr, base=base, sample=sample, host=host, path=ix_path, tmp=tmpdir.name)
    fastscan = None if args.no_fastscan else fastscan_leg(ix, db, Q, gt, k, args, ctx)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(ix_path, n, d, deg, Q, chosen, k, host)
        log(f"cpu baseline {cpu}")
    del arr, ctx
    tmpdir.cleanup()

    serving_mix = None if args.no_serve else serving_mix_leg(args)

    configs = None if args.no_configs else configs_leg(args, host)
    wide = None if args.no_wide else wide_leg(args, host)
    if wide:
        configs = dict(configs or {}, posmix_1m_d960=wide)

    line = {
        "metric": "QPS at recall@10≥0.95 (batched vs plain GPU), 1M×128 deg-32; p99 ms",
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"int_mix N={n} D={d} deg={deg} (1024-centre Gaussian mixture, sigma 20, "
                        f"integer 0..255), {args.nq} queries, k={k}, graph built on device with "
                        f"build EF {ef_build} x3 iterations",
            "ef": chosen, "k": k, "nq": args.nq, "path": "SQ8 beam search, order=nearest",
            "recall_at_10": rec, "ef_sweep": sweep, "build_s": round(build_s, 1),
            "seed": args.seed, "launch": tuned,
        },
        "plain_gpu": {"qps": round(plain_qps, 1), "p50_ms": round(float(np.median(lat)), 4),
                      "p99_ms": round(float(np.percentile(lat, 99)), 4), "queries": len(lat)},
        "batched_over_plain": round(qps / plain_qps, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "bytes_per_query": round(alg_bytes / args.nq, 1),
                     "expansions_per_query": round(float(E.mean()), 2),
                     "distinct_evaluations_per_query": round(float(V_ref.mean()), 2),
                     "kernel_evaluations_per_query": round(float(V.mean()), 2)},
        "cpu_baseline": cpu,
    }
    line["graph_replay"] = {"ms_per_step": round(graph_ms, 4),
                            "qps": round(args.nq / (graph_ms / 1e3), 1)}
    float_gbs = alg_bytes / (ms_float / 1e3) / 1e9
    line["sq8_float_kernel"] = {
        "note": "the general SQ8 kernel (dequantise + FMA partials), integer kernel disabled",
        "launch": float_launch,
        "ms_per_step": round(ms_float, 4), "qps": round(args.nq / (ms_float / 1e3), 1),
        "ids_equal_integer_kernel": float_same,
        "roofline": {"bound": "hbm", "achieved": round(float_gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(float_gbs / HBM_PEAK_GBS, 4)}}
    if parity:
        line["parity"] = parity
    # the same algorithmic bytes per batch over the back-to-back batch time
    steady_gbs = alg_bytes / (steady["ms_per_batch"] / 1e3) / 1e9
    steady.update({"algorithmic_GBps": round(steady_gbs, 1),
                   "frac_of_peak": round(steady_gbs / HBM_PEAK_GBS, 4)})
    line["steady_state"] = steady
    if serving:
        line["serving"] = serving
    if serving_mix:
        line["serving_mix"] = serving_mix
    if configs:
        line["other_workloads"] = configs
    if fastscan:
        line["fastscan"] = fastscan
    if cpu:
        line["gpu_over_cpu"] = round(qps / cpu["value"], 1)
    print(json.dumps(line), flush=True)
    bad = [name for name, p in (("SQ8", parity), ("FastScan", fastscan and fastscan["parity"]),
                                ("960-d", wide and wide.get("parity_vs_oracle")))
           if p and not p["ok"]]
    if bad:
        log(f"PARITY FAILED: {bad}")
        sys.exit(3)


if __name__ == "__main__":
    main()
