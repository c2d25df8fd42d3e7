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
* Additions: ``set_order``, ``search_batch`` (many queries, one launch).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import ORDER, Launch, LqgError, Stats, check, lib, ptr

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
        st = lib().lqg_create(str(index_type).encode(), str(metric).encode(), int(num_elements),
                              int(dimension), int(degree_bound), C.byref(self._h))
        if st != 0:
            raise ValueError(lib().lqg_last_error().decode())
        self.num_elements = int(num_elements)
        self.dimension = int(dimension)
        self.degree_bound = int(degree_bound)
        self._order = "reference"

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().lqg_destroy(h)
            self._h = None

    # -- reference API -----------------------------------------------------
    def load(self, filename: str) -> None:
        check(lib().lqg_load(self._h, str(filename).encode()))

    def save(self, filename: str) -> None:
        check(lib().lqg_save(self._h, str(filename).encode()))

    def set_ef(self, EF: int) -> None:  # noqa: N803 (keyword name of the reference)
        check(lib().lqg_set_ef(self._h, int(EF)))

    def build_index(self, data, EF: int, num_iter: int = 3,  # noqa: N803
                    num_thread: int = UINT_MAX) -> None:
        arr = np.ascontiguousarray(data, dtype=np.float32)
        if arr.ndim == 1:
            arr = arr.reshape(1, -1)
        if arr.ndim != 2:
            raise ValueError("Input data has an incorrect shape. Data must be a 1D or 2D array.")
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
                   occupancy: int = 0, groups: int = 0) -> None:
        """Kernel launch settings for ef's capacity class (ef=0: every class)."""
        cfg = Launch(waves_per_group, cache_bits, occupancy, groups)
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
