"""Serving and graph-replay front ends over the C ABI (include/liteqg.h).

``Server`` wraps the native serving engine (lqg_serve_*): single-query
requests with their own ef and k are submitted as they arrive and batched
continuously over several launches in flight.  ``BatchGraph`` wraps a batch
launch captured into a HIP graph (lqg_graph_*).  Neither has a counterpart in
the reference, which searches one query per call (qg.hpp:235-243).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (LQG_ERR_BUSY, ORDER, LqgError, ServeConfig, ServeStats, check, lib, ptr)


class Server:
    """Continuous-batching request server over one index.

    lanes: launches in flight; max_batch: requests per launch; slots:
    outstanding requests (admission limit); ef_max, kmax: largest ef and k a
    request may ask (they fix the kernel's LDS layout).
    """

    def __init__(self, index, lanes: int = 4, max_batch: int = 1024, slots: int | None = None,
                 ef_max: int = 256, kmax: int = 100, order: str = "nearest"):
        self.index = index
        self.dim = index.dimension
        self.kmax = int(kmax)
        cfg = ServeConfig(int(lanes), int(max_batch), int(slots or 4 * lanes * max_batch),
                          int(ef_max), int(kmax), ORDER[order])
        self._s = C.c_void_p()
        self._L = lib()  # kept for close() at interpreter exit
        check(self._L.lqg_serve_start(index._h, C.byref(cfg), C.byref(self._s)))

    def submit(self, query, ef: int, k: int) -> int | None:
        """Ticket, or None when every slot is taken (retry later)."""
        q = np.ascontiguousarray(query, dtype=np.float32).reshape(-1)
        if q.size != self.dim:
            raise ValueError(f"query must have {self.dim} values")
        t = C.c_uint64()
        st = lib().lqg_serve_submit(self._s, ptr(q), int(ef), int(k), C.byref(t))
        if st == LQG_ERR_BUSY:
            return None
        check(st)
        return t.value

    def poll(self, ticket: int, k: int):
        """(ids, dists) once done, else None.  k is the request's k (the size
        of the arrays returned); a k below it raises and keeps the ticket.  A
        poll that returns or raises for a failed request releases the
        ticket."""
        ids = np.empty(int(k), np.uint32)
        dis = np.empty(int(k), np.float32)
        st = lib().lqg_serve_poll(self._s, C.c_uint64(ticket), ptr(ids), ptr(dis), int(k))
        if st == LQG_ERR_BUSY:
            return None
        check(st)
        return ids, dis

    def wait(self, ticket: int, k: int):
        ids = np.empty(int(k), np.uint32)
        dis = np.empty(int(k), np.float32)
        check(lib().lqg_serve_wait(self._s, C.c_uint64(ticket), ptr(ids), ptr(dis), int(k)))
        return ids, dis

    def stats(self) -> dict:
        st = ServeStats()
        check(lib().lqg_serve_get_stats(self._s, C.byref(st)))
        return {f: getattr(st, f) for f, _ in ServeStats._fields_}

    def replay(self, queries, ef, k, arrival_ns):
        """Native load replay: query i submitted at arrival_ns[i] (ns after the
        start).  Returns (latency_ns uint64[nq], ids uint32[nq, kmax])."""
        Q = np.ascontiguousarray(queries, dtype=np.float32).reshape(-1, self.dim)
        nq = Q.shape[0]
        ef = np.ascontiguousarray(np.broadcast_to(np.asarray(ef, np.uint32), (nq,)))
        k = np.ascontiguousarray(np.broadcast_to(np.asarray(k, np.uint32), (nq,)))
        arr = np.ascontiguousarray(arrival_ns, dtype=np.uint64).reshape(nq)
        lat = np.empty(nq, np.uint64)
        ids = np.empty((nq, self.kmax), np.uint32)
        check(lib().lqg_serve_replay(self._s, ptr(Q), nq, ptr(ef), ptr(k), ptr(arr), ptr(lat),
                                     ptr(ids)))
        return lat, ids

    def close(self) -> None:
        s = getattr(self, "_s", None)
        L = getattr(self, "_L", None)
        if s is not None and s.value and L is not None:
            L.lqg_serve_stop(s)
            self._s = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()


def poisson_arrivals(nq: int, rate_per_s: float, seed: int = 0) -> np.ndarray:
    """Arrival times (ns) of a Poisson process: exponential gaps of mean 1/rate."""
    gaps = np.random.default_rng(seed).exponential(1e9 / rate_per_s, nq)
    return np.cumsum(gaps).astype(np.uint64)


class BatchGraph:
    """A batch launch over device arrays (torch tensors or raw pointers)
    captured into a HIP graph; ``launch(stream)`` replays it."""

    def __init__(self, index, d_queries: int, nq: int, d_results: int, kmax: int, ef_max: int,
                 order: str = "nearest", d_ef: int | None = None, d_k: int | None = None,
                 d_dists: int | None = None, d_stats: int | None = None):
        self._g = C.c_void_p()
        self._L = lib()  # kept for close() at interpreter exit
        vp = lambda x: None if x is None else C.c_void_p(x)  # noqa: E731
        check(self._L.lqg_graph_create(index._h, vp(d_queries), int(nq), vp(d_ef), vp(d_k),
                                     int(ef_max), int(kmax), ORDER[order], vp(d_results),
                                     vp(d_dists), vp(d_stats), C.byref(self._g)))

    def launch(self, stream: int | None = None) -> None:
        check(lib().lqg_graph_launch(self._g, None if stream is None else C.c_void_p(stream)))

    def close(self) -> None:
        g = getattr(self, "_g", None)
        L = getattr(self, "_L", None)
        if g is not None and g.value and L is not None:
            L.lqg_graph_destroy(g)
            self._g = None

    def __del__(self):
        self.close()


__all__ = ["Server", "BatchGraph", "poisson_arrivals", "LqgError"]
