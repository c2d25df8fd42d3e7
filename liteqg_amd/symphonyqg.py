"""Drop-in for the reference's ``symphonyqg`` module (python/bindings.cpp:101-125).

``Index`` keeps the reference's constructor, method names, positional order
and keyword names; every call goes through the C ABI of liblqg.so to the
gfx950 kernels.  Differences, all deliberate:

* Invalid constructor arguments raise ``ValueError`` (the reference prints to
  stderr and leaves a null index that crashes on the next call,
  bindings.cpp:45-60).
* ``search`` before ``set_ef`` raises (the reference reads an uninitialised
  ef, qg.hpp:66).
* Result slots the pool cannot fill are ``0xFFFFFFFF`` (the reference leaves
  them uninitialised, qg.hpp:340).
* The output order is selectable: ``"reference"`` (default) returns the k
  farthest entries of the final ef-pool in ascending order, exactly as the
  shipped ``search`` does (qg.hpp:289-296,340-343); ``"nearest"`` returns the
  k nearest of the same pool.
* Additions: ``set_order``, ``search_batch`` (many queries, one launch), and
  the FastScan RaBitQ path (``search_qg``, qg.hpp:252-287, which the reference
  ships but never calls): ``build_sidecar`` / ``set_sidecar`` /
  ``load_sidecar`` give it its codes and factors, ``set_mode("fastscan")``
  routes ``search`` and ``search_batch`` to it.  Its results are the k
  nearest expanded nodes by exact distance, ascending; slots it cannot fill
  hold 0, as the reference's zero-initialised ResultBuffer does.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (BUILD_CANDIDATES, MODE, ORDER, ROTATION, Launch, LqgError, Stats, check, lib,
                   ptr)

UINT_MAX = 0xFFFFFFFF


def _order_code(order: str) -> int:
    try:
        return ORDER[order]
    except KeyError:
        raise ValueError(f"order must be 'reference' or 'nearest', got {order!r}") from None


class Index:
    """symphonyqg.Index(index_type, metric, num_elements, dimension, degree_bound=32)."""

    def __init__(self, index_type: str, metric: str, num_elements: int, dimension: int,
                 degree_bound: int = 32):
        self._h = C.c_void_p()
        # the library object is kept here so the handle can be released even at
        # interpreter exit, when module globals may already be cleared
        self._L = lib()
        st = self._L.lqg_create(str(index_type).encode(), str(metric).encode(),
                                int(num_elements), int(dimension), int(degree_bound),
                                C.byref(self._h))
        if st != 0:
            raise ValueError(lib().lqg_last_error().decode())
        self.num_elements = int(num_elements)
        self.dimension = int(dimension)
        self.degree_bound = int(degree_bound)
        self._order = "reference"

    def __del__(self):
        h = getattr(self, "_h", None)
        L = getattr(self, "_L", None)
        if h is not None and h.value and L is not None:
            L.lqg_destroy(h)
            self._h = None

    # -- reference API -----------------------------------------------------
    def load(self, filename: str) -> None:
        check(lib().lqg_load(self._h, str(filename).encode()))

    def save(self, filename: str) -> None:
        check(lib().lqg_save(self._h, str(filename).encode()))

    def set_ef(self, EF: int) -> None:  # noqa: N803 (keyword name of the reference)
        check(lib().lqg_set_ef(self._h, int(EF)))

    def build_index(self, data, EF: int, num_iter: int = 3,  # noqa: N803
                    num_thread: int = UINT_MAX, *, candidates: str | None = None) -> None:
        """Build the graph on the device (bindings.cpp:69-90).  candidates:
        ``"sq8"`` (default: the nearest entries of an SQ8 beam search) or
        ``"fastscan"`` (the reference's find_candidates: every node a FastScan
        search over the graph being built expands); None keeps the current
        setting."""
        arr = np.ascontiguousarray(data, dtype=np.float32)
        if arr.ndim == 1:
            arr = arr.reshape(1, -1)
        if arr.ndim != 2:
            raise ValueError("Input data has an incorrect shape. Data must be a 1D or 2D array.")
        if candidates is not None:
            if candidates not in BUILD_CANDIDATES:
                raise ValueError(f"candidates must be 'sq8' or 'fastscan', got {candidates!r}")
            check(lib().lqg_set_build_candidates(self._h, BUILD_CANDIDATES[candidates]))
        check(lib().lqg_build_index(self._h, ptr(arr), arr.shape[0], arr.shape[1], int(EF),
                                    int(num_iter), int(min(num_thread, UINT_MAX))))

    def search(self, query, k: int) -> np.ndarray:
        q = np.ascontiguousarray(query, dtype=np.float32).reshape(-1)
        if q.size < self.dimension:
            raise ValueError(f"query has {q.size} values, dimension is {self.dimension}")
        q = q[: self.dimension]
        out = np.empty(int(k), dtype=np.uint32)
        check(lib().lqg_search(self._h, ptr(q), int(k), ptr(out)))
        return out

    # -- additions ---------------------------------------------------------
    def set_order(self, order: str) -> None:
        check(lib().lqg_set_order(self._h, _order_code(order)))
        self._order = order

    def search_batch(self, queries, k, ef=None, order: str | None = None,
                     return_distances: bool = False, return_stats: bool = False):
        """Search many queries in one persistent launch.

        ``k`` and ``ef`` may be scalars or per-query arrays.  Returns ids
        ``uint32[nq, kmax]`` (and distances ``float32[nq, kmax]``, and a stats
        dict, when asked); row i holds k_i results then 0xFFFFFFFF.
        """
        Q = np.ascontiguousarray(queries, dtype=np.float32)
        if Q.ndim == 1:
            Q = Q.reshape(1, -1)
        if Q.ndim != 2 or Q.shape[1] != self.dimension:
            raise ValueError(f"queries must be (nq, {self.dimension})")
        nq = Q.shape[0]
        k_arr = None
        if np.ndim(k) == 0:
            kmax = int(k)
        else:
            k_arr = np.ascontiguousarray(k, dtype=np.uint32).reshape(-1)
            if k_arr.size != nq:
                raise ValueError("k array must have one entry per query")
            kmax = int(k_arr.max())
        ef_arr = None
        if ef is not None:
            ef_arr = np.ascontiguousarray(np.broadcast_to(np.asarray(ef, dtype=np.uint32), (nq,)))
        ids = np.empty((nq, kmax), dtype=np.uint32)
        dis = np.empty((nq, kmax), dtype=np.float32) if return_distances else None
        st = Stats()
        check(lib().lqg_search_batch(self._h, ptr(Q), nq, ptr(ef_arr), ptr(k_arr), kmax,
                                     _order_code(order or self._order), ptr(ids), ptr(dis),
                                     C.byref(st) if return_stats else None))
        out = [ids]
        if return_distances:
            out.append(dis)
        if return_stats:
            out.append(dict(queries=st.queries, expansions=st.expansions,
                            evaluations=st.evaluations, device_ms=st.device_ms))
        return out[0] if len(out) == 1 else tuple(out)

    def set_launch(self, ef: int = 0, waves_per_group: int = 4, cache_bits: int = 10,
                   occupancy: int = 0, groups: int = 0, prefetch: int = 0,
                   split: int = 1) -> None:
        """Kernel launch settings for ef's capacity class (ef=0: every class).
        prefetch (SQ8 launches of at most 256 queries, e.g. a lone search(), or
        1,024 on an index of at most 256 dims):
        0 = the batch kernel, 1..3 = latency mode with that prefetch depth.
        split (SQ8 float path): 1 = off, 2 or 4 waves per query sharing each
        expansion's distance evaluation (takes precedence over prefetch)."""
        cfg = Launch(waves_per_group, cache_bits, occupancy, groups, prefetch, split)
        check(lib().lqg_set_launch(self._h, int(ef), C.byref(cfg)))

    def reset_launch(self, ef: int = 0) -> None:
        check(lib().lqg_set_launch(self._h, int(ef), None))

    def get_launch(self, ef: int) -> dict:
        cfg = Launch()
        check(lib().lqg_get_launch(self._h, int(ef), C.byref(cfg)))
        return {f: getattr(cfg, f) for f, _ in Launch._fields_}

    def autotune(self, queries, ef: int, k: int = 10) -> dict:
        """Time the candidate launch settings on these queries; keep the fastest."""
        Q = np.ascontiguousarray(queries, dtype=np.float32).reshape(-1, self.dimension)
        best = Launch()
        check(lib().lqg_autotune(self._h, ptr(Q), Q.shape[0], int(ef), int(k), C.byref(best)))
        return {f: getattr(best, f) for f, _ in Launch._fields_}

    def count_visits(self, queries, ef: int, k: int = 10) -> np.ndarray:
        """Per query (expansions, distinct nodes evaluated), counted exactly."""
        Q = np.ascontiguousarray(queries, dtype=np.float32).reshape(-1, self.dimension)
        out = np.empty((Q.shape[0], 2), dtype=np.uint32)
        check(lib().lqg_count_visits(self._h, ptr(Q), Q.shape[0], int(ef), int(k), ptr(out)))
        return out

    def set_sq8_integer(self, enable: bool) -> None:
        """Diagnostic: allow (default) or forbid the exact integer SQ8 kernel;
        forbidden, every query runs the general float kernel (same results)."""
        check(lib().lqg_set_sq8_integer(self._h, 1 if enable else 0))

    # -- FastScan RaBitQ path (search_qg) ----------------------------------
    def fastscan_expansions(self, queries, ef: int, cap: int) -> np.ndarray:
        """Diagnostic: the first ``cap`` nodes each query's FastScan search
        expands at ``ef``, in order (0xFFFFFFFF past the last): the candidate
        list of the reference's find_candidates (qg.hpp:439-473)."""
        Q = np.ascontiguousarray(queries, dtype=np.float32).reshape(-1, self.dimension)
        out = np.empty((Q.shape[0], int(cap)), np.uint32)
        check(lib().lqg_h2_expansions(self._h, ptr(Q), Q.shape[0], int(ef), int(cap), ptr(out)))
        return out

    def set_overflow_bits(self, bits: int) -> None:
        """log2 entries of the per-wave device visited table a FastScan query
        moves to once its LDS table is half full (default 14); 0 = none, the
        query fails instead (diagnostic)."""
        check(lib().lqg_h2_set_overflow_bits(self._h, int(bits)))

    def set_generic_kernel(self, generic: bool) -> None:
        """Run the general FastScan kernel even where one specialised for the
        index's record shape exists (D = 64, 96, 128 at degree 32); same
        results (for comparisons)."""
        check(lib().lqg_h2_set_generic(self._h, 1 if generic else 0))

    @property
    def padded_dim(self) -> int:
        p = 1
        while p < self.dimension:
            p <<= 1
        return p

    def set_mode(self, mode: str) -> None:
        """``"sq8"`` (the reference's live search, default) or ``"fastscan"``."""
        if mode not in MODE:
            raise ValueError(f"mode must be 'sq8' or 'fastscan', got {mode!r}")
        check(lib().lqg_set_mode(self._h, MODE[mode]))

    def get_mode(self) -> str:
        m = C.c_int()
        check(lib().lqg_get_mode(self._h, C.byref(m)))
        return {v: k for k, v in MODE.items()}[m.value]

    def build_sidecar(self, signs=None, seed: int = 0) -> None:
        """RaBitQ codes and factors of every node's neighbours, built on the
        device; rotation signs given (P floats) or drawn from ``seed``."""
        s = None
        if signs is not None:
            s = np.ascontiguousarray(signs, dtype=np.float32).reshape(-1)
            if s.size != self.padded_dim:
                raise ValueError(f"signs must have {self.padded_dim} values")
        check(lib().lqg_h2_build_sidecar(self._h, ptr(s), int(seed)))

    def set_sidecar(self, signs, codes, factors, layout: str = "vector") -> None:
        """Install a sidecar: codes ``uint8[N, deg*P/8]`` in the reference's
        FastScan block layout (``layout="blocks"``) or vector-major
        (``"vector"``), factors ``float32[N, 3*deg]``."""
        p, g, n = self.padded_dim, self.degree_bound, self.num_elements
        s = np.ascontiguousarray(signs, dtype=np.float32).reshape(-1)
        c = np.ascontiguousarray(codes, dtype=np.uint8).reshape(n, g * p // 8)
        f = np.ascontiguousarray(factors, dtype=np.float32).reshape(n, 3 * g)
        if s.size != p:
            raise ValueError(f"signs must have {p} values")
        if layout not in ("blocks", "vector"):
            raise ValueError("layout must be 'blocks' or 'vector'")
        check(lib().lqg_h2_set_sidecar(self._h, ptr(s), ptr(c), 0 if layout == "blocks" else 1,
                                       ptr(f)))

    def get_sidecar(self, signs_only: bool = False):
        """(signs, codes vector-major, factors) of the installed sidecar, or
        only the signs."""
        p, g, n = self.padded_dim, self.degree_bound, self.num_elements
        s = np.empty(p, np.float32)
        if signs_only:
            check(lib().lqg_h2_get_sidecar(self._h, ptr(s), None, None))
            return s
        c = np.empty((n, g * p // 8), np.uint8)
        f = np.empty((n, 3 * g), np.float32)
        check(lib().lqg_h2_get_sidecar(self._h, ptr(s), ptr(c), ptr(f)))
        return s, c, f

    def save_sidecar(self, path: str) -> None:
        check(lib().lqg_h2_save_sidecar(self._h, str(path).encode()))

    def load_sidecar(self, path: str) -> None:
        check(lib().lqg_h2_load_sidecar(self._h, str(path).encode()))

    def set_rotation(self, method: str) -> None:
        """FastScan query rotation: ``"fht"`` (in the search kernel, the
        reference's rotation bit for bit; default) or ``"gemm"`` (one FP32 GEMM
        on the matrix cores before the search)."""
        if method not in ROTATION:
            raise ValueError(f"rotation must be 'fht' or 'gemm', got {method!r}")
        check(lib().lqg_h2_set_rotation(self._h, ROTATION[method]))

    def get_rotation(self) -> str:
        v = C.c_int()
        check(lib().lqg_h2_get_rotation(self._h, C.byref(v)))
        return {c: n for n, c in ROTATION.items()}[v.value]

    def rotate_queries(self, queries, method: str = "fht") -> np.ndarray:
        """The rotated queries (nq x P float32) by either method."""
        if method not in ROTATION:
            raise ValueError(f"rotation must be 'fht' or 'gemm', got {method!r}")
        Q = np.ascontiguousarray(queries, dtype=np.float32).reshape(-1, self.dimension)
        out = np.empty((Q.shape[0], self.padded_dim), np.float32)
        check(lib().lqg_h2_rotate(self._h, ptr(Q), Q.shape[0], ROTATION[method], ptr(out)))
        return out

    def fastscan_probe(self, queries, nodes=None):
        """Device query preparation (and, with nodes, the exact distance to
        nodes[i] and its neighbours' FastScan estimates) for parity checks.
        Returns dict(lut, lo, width, sumq[, sqr_y, est])."""
        Q = np.ascontiguousarray(queries, dtype=np.float32).reshape(-1, self.dimension)
        nq, p = Q.shape[0], self.padded_dim
        lut = np.empty((nq, 4 * p), np.uint8)
        lw = np.empty((nq, 2), np.float32)
        sq = np.empty(nq, np.int32)
        nd = sy = est = None
        if nodes is not None:
            nd = np.ascontiguousarray(nodes, dtype=np.uint32).reshape(nq)
            sy = np.empty(nq, np.float32)
            est = np.empty((nq, self.degree_bound), np.float32)
        check(lib().lqg_h2_probe(self._h, ptr(Q), nq, ptr(nd), ptr(lut), ptr(lw), ptr(sq),
                                 ptr(sy), ptr(est)))
        out = dict(lut=lut, lo=lw[:, 0].copy(), width=lw[:, 1].copy(), sumq=sq)
        if nodes is not None:
            out.update(sqr_y=sy, est=est)
        return out

    def info(self) -> dict:
        n, d, g = C.c_uint64(), C.c_uint64(), C.c_uint64()
        ep, ef = C.c_uint32(), C.c_uint32()
        check(lib().lqg_info(self._h, C.byref(n), C.byref(d), C.byref(g), C.byref(ep),
                             C.byref(ef)))
        return dict(num_elements=n.value, dimension=d.value, degree_bound=g.value,
                    entry_point=ep.value, ef=ef.value)

    def graph(self) -> np.ndarray:
        out = np.empty((self.num_elements, self.degree_bound), dtype=np.uint32)
        check(lib().lqg_get_graph(self._h, ptr(out)))
        return out


__all__ = ["Index", "LqgError", "UINT_MAX"]
