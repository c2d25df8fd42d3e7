"""ctypes front end of the CPU oracle (oracle/lqg_oracle.c) and of the index
file layout.  TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check
and by the CPU-baseline leg of bench.py, never by the library.

Index file body (qg.hpp:206-224; third/svs/array.hpp:109-116; shapes from
qg.hpp:183-193,420-436), no header:

    u32 entry point
    N x (D + deg) float32 rows: D raw floats, then deg u32 neighbour ids
    N x D uint8 SQ8 codes
    D x {float32 min, float32 scale}
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# LQG_ORACLE_LIB names another build of lqg_oracle.c (`make -C oracle san`:
# oracle/_san/liblqg_oracle.so, built with AddressSanitizer and UBSan)
LIB_PATH = os.environ.get("LQG_ORACLE_LIB") or os.path.join(HERE, "liblqg_oracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")
FILL = 0xFFFFFFFF


def build(force: bool = False) -> None:
    """Compile the C restatement (and oracle/_ref where /root/reference exists)."""
    if os.environ.get("LQG_ORACLE_LIB"):
        return
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_sq8_dequantize.argtypes = [P, C.c_uint32, P, P, C.c_int]
        L.orc_l2_sqr.argtypes = [P, P, C.c_uint32]
        L.orc_l2_sqr.restype = C.c_float
        L.orc_search_heap.argtypes = [P, P, C.c_uint32, C.c_uint32, P, P, P, P, P]
        L.orc_search_canon.argtypes = [P, P, C.c_uint32, C.c_uint32, C.c_int, P, P, P, P, P, P]
        L.orc_search_canon_batch.argtypes = [P, P, C.c_uint64, C.c_uint32, C.c_uint32, C.c_int,
                                             P, P, P, C.c_int]
        L.orc_search_heap_batch.argtypes = [P, P, C.c_uint64, C.c_uint32, C.c_uint32, P, C.c_int]
        L.orc_bruteforce_fp64.argtypes = [P, C.c_uint64, C.c_uint32, P, C.c_uint64, C.c_uint32,
                                          P, P, C.c_int]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class _OrcIndex(C.Structure):
    _fields_ = [("n", C.c_uint32), ("dim", C.c_uint32), ("deg", C.c_uint32), ("ep", C.c_uint32),
                ("nbr", C.c_void_p), ("qdata", C.c_void_p), ("params", C.c_void_p),
                ("fused", C.c_int)]


class IndexArrays:
    """The three arrays the search reads, plus the raw vectors and entry point."""

    def __init__(self, ep, raw, nbr, qdata, params, fused: bool = True):
        self.ep = int(ep)
        self.raw = np.ascontiguousarray(raw, np.float32)
        self.nbr = np.ascontiguousarray(nbr, np.uint32)
        self.qdata = np.ascontiguousarray(qdata, np.uint8)
        self.params = np.ascontiguousarray(params, np.float32).reshape(-1, 2)
        self.n, self.dim = self.qdata.shape
        self.deg = self.nbr.shape[1]
        self.fused = bool(fused)
        self._s = _OrcIndex(self.n, self.dim, self.deg, self.ep, self.nbr.ctypes.data,
                            self.qdata.ctypes.data, self.params.ctypes.data, int(self.fused))

    @property
    def ref(self):
        return C.byref(self._s)

    def with_rounding(self, fused: bool) -> "IndexArrays":
        return IndexArrays(self.ep, self.raw, self.nbr, self.qdata, self.params, fused)


def file_size(n: int, d: int, deg: int) -> int:
    return 4 + n * (d + deg) * 4 + n * d + 8 * d


def read_index(path: str, n: int, d: int, deg: int, fused: bool = True) -> IndexArrays:
    expect = file_size(n, d, deg)
    got = os.path.getsize(path)
    if got != expect:
        raise ValueError(f"{path}: {got} bytes, expected {expect} for N={n} D={d} deg={deg}")
    buf = np.fromfile(path, dtype=np.uint8)
    ep = int(buf[:4].view(np.uint32)[0])
    off = 4
    rows = buf[off:off + n * (d + deg) * 4].view(np.float32).reshape(n, d + deg)
    off += n * (d + deg) * 4
    qdata = buf[off:off + n * d].reshape(n, d)
    off += n * d
    params = buf[off:off + 8 * d].view(np.float32).reshape(d, 2)
    raw = rows[:, :d]
    nbr = rows[:, d:].view(np.uint32)
    return IndexArrays(ep, raw, nbr, qdata, params, fused)


def write_index(path: str, ix: IndexArrays) -> None:
    n, d, deg = ix.n, ix.dim, ix.deg
    rows = np.empty((n, d + deg), np.float32)
    rows[:, :d] = ix.raw
    rows[:, d:] = ix.nbr.view(np.float32)
    with open(path, "wb") as f:
        f.write(np.uint32(ix.ep).tobytes())
        f.write(rows.tobytes())
        f.write(ix.qdata.tobytes())
        f.write(ix.params.astype(np.float32).tobytes())


def sq8_quantize(data: np.ndarray):
    """Per-dimension min/scale codes as quantization/sq8.hpp:11-26 computes them
    in the reference's -Ofast build.

    ``scale = (max - min) * fl(1/255)`` (the build turns the division by the
    constant 255.0f into a multiplication by its float reciprocal), then codes
    ``round((v - min) / scale)`` with a true float32 division and
    round-half-away-from-zero (std::round).  Both facts are fixed by the
    reference-built index files in tests/golden.  A dimension with max == min
    gets scale 0 and the reference's 0/0 -> NaN -> cast; here such codes are 0.
    """
    data = np.ascontiguousarray(data, np.float32)
    mn = data.min(axis=0)
    mx = data.max(axis=0)
    scale = ((mx - mn) * (np.float32(1.0) / np.float32(255.0))).astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = (data - mn) / scale
    r = np.where(np.isfinite(r), r, 0.0).astype(np.float32)
    t = np.trunc(r)
    q = t + np.sign(r) * (np.abs(r - t) >= np.float32(0.5))
    params = np.stack([mn, scale], axis=1).astype(np.float32)
    return np.clip(q, 0, 255).astype(np.uint8), params


def sq8_dequantize(codes, params, fused: bool = True) -> np.ndarray:
    codes = np.ascontiguousarray(codes, np.uint8)
    params = np.ascontiguousarray(params, np.float32)
    out = np.empty(codes.shape, np.float32)
    L = lib()
    rows = codes.reshape(-1, codes.shape[-1])
    o = out.reshape(rows.shape)
    for i in range(rows.shape[0]):
        L.orc_sq8_dequantize(_p(rows[i]), rows.shape[1], _p(params), _p(o[i]), int(fused))
    return out


def l2_sqr(a, b) -> float:
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return float(lib().orc_l2_sqr(_p(a), _p(b), a.shape[0]))


def search_heap(ix: IndexArrays, query, ef: int, k: int):
    """Reference H1 replica: (results[k], pool_ids, pool_d, (E, V))."""
    q = np.ascontiguousarray(query, np.float32)
    res = np.empty(k, np.uint32)
    pid = np.empty(ef, np.uint32)
    pd = np.empty(ef, np.float32)
    pn = np.zeros(1, np.uint32)
    st = np.zeros(2, np.uint64)
    rc = lib().orc_search_heap(ix.ref, _p(q), ef, k, _p(res), _p(pid), _p(pd), _p(pn), _p(st))
    if rc != 0:
        raise RuntimeError(f"orc_search_heap failed ({rc})")
    n = int(pn[0])
    return res, pid[:n].copy(), pd[:n].copy(), (int(st[0]), int(st[1]))


def search_canon(ix: IndexArrays, query, ef: int, k: int, order: str = "nearest"):
    """Total-order H1: (results[k], dists[k], pool_ids, pool_d, (E, V))."""
    q = np.ascontiguousarray(query, np.float32)
    res = np.empty(k, np.uint32)
    dis = np.empty(k, np.float32)
    pid = np.empty(ef, np.uint32)
    pd = np.empty(ef, np.float32)
    pn = np.zeros(1, np.uint32)
    st = np.zeros(2, np.uint64)
    rc = lib().orc_search_canon(ix.ref, _p(q), ef, k, 0 if order == "reference" else 1,
                                _p(res), _p(dis), _p(pid), _p(pd), _p(pn), _p(st))
    if rc != 0:
        raise RuntimeError(f"orc_search_canon failed ({rc})")
    n = int(pn[0])
    return res, dis, pid[:n].copy(), pd[:n].copy(), (int(st[0]), int(st[1]))


def search_canon_batch(ix: IndexArrays, queries, ef: int, k: int, order: str = "nearest",
                       threads: int = 0):
    Q = np.ascontiguousarray(queries, np.float32)
    nq = Q.shape[0]
    res = np.empty((nq, k), np.uint32)
    dis = np.empty((nq, k), np.float32)
    st = np.zeros((nq, 2), np.uint64)
    rc = lib().orc_search_canon_batch(ix.ref, _p(Q), nq, ef, k, 0 if order == "reference" else 1,
                                      _p(res), _p(dis), _p(st), threads)
    if rc != 0:
        raise RuntimeError("orc_search_canon_batch failed")
    return res, dis, st


def search_heap_batch(ix: IndexArrays, queries, ef: int, k: int, threads: int = 0):
    Q = np.ascontiguousarray(queries, np.float32)
    res = np.empty((Q.shape[0], k), np.uint32)
    if lib().orc_search_heap_batch(ix.ref, _p(Q), Q.shape[0], ef, k, _p(res), threads) != 0:
        raise RuntimeError("orc_search_heap_batch failed")
    return res


def bruteforce_fp64(base, queries, k: int, threads: int = 0):
    B = np.ascontiguousarray(base, np.float32)
    Q = np.ascontiguousarray(queries, np.float32)
    ids = np.empty((Q.shape[0], k), np.uint32)
    dis = np.empty((Q.shape[0], k), np.float64)
    if lib().orc_bruteforce_fp64(_p(B), B.shape[0], B.shape[1], _p(Q), Q.shape[0], k, _p(ids),
                                 _p(dis), threads) != 0:
        raise RuntimeError("orc_bruteforce_fp64 failed")
    return ids, dis


def recall_at_k(found: np.ndarray, gt: np.ndarray, k: int) -> float:
    """|found[:k] & gt[:k]| / k averaged over queries (reproduce/run.py:31-40)."""
    tot = 0
    for f, g in zip(found, gt):
        tot += len(set(int(x) for x in f[:k]) & set(int(x) for x in g[:k]))
    return tot / (k * len(found))


# --------------------------------------------------------------------------
# H2: FastScan RaBitQ search (qg.hpp:252-287) over an explicit sidecar
# --------------------------------------------------------------------------
class _OrcH2(C.Structure):
    _fields_ = [("n", C.c_uint32), ("dim", C.c_uint32), ("deg", C.c_uint32), ("p", C.c_uint32),
                ("ep", C.c_uint32), ("raw", C.c_void_p), ("nbr", C.c_void_p),
                ("codes", C.c_void_p), ("factors", C.c_void_p), ("signs", C.c_void_p),
                ("layout", C.c_int)]


def padded_dim(d: int) -> int:
    """P = 1 << ceil_log2(D) (utils/tools.hpp:25-34)."""
    p = 1
    while p < d:
        p <<= 1
    return p


def _h2sigs():
    L = lib()
    if getattr(L, "_h2", False):
        return L
    P = C.c_void_p
    L.orc_fht_rotate.argtypes = [P, P, C.c_uint32, C.c_uint32, P]
    L.orc_query_prepare.argtypes = [P, P, C.c_uint32, C.c_uint32, P, P, P, P, P, P]
    L.orc_scan_estimates.argtypes = [P, C.c_int, P, C.c_uint32, C.c_uint32, P, C.c_float,
                                     C.c_float, C.c_float, C.c_int32, P, P]
    L.orc_buffers_replay.argtypes = [P, C.c_uint64, C.c_uint32, P, P, P]
    L.orc_search_h2.argtypes = [P, P, C.c_uint32, C.c_uint32, P, P, P]
    L.orc_search_h2_batch.argtypes = [P, P, C.c_uint64, C.c_uint32, C.c_uint32, P, P, P, C.c_int]
    L.orc_h2_expansions_batch.argtypes = [P, P, C.c_uint64, C.c_uint32, C.c_uint32, P, C.c_int]
    L.orc_sidecar_build.argtypes = [P, P, P, C.c_int]
    L.orc_codes_to_vector_major.argtypes = [P, C.c_uint32, C.c_uint32, P]
    L._h2 = True
    return L


def fht_rotate(signs, x, p):
    s = np.ascontiguousarray(signs, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty(p, np.float32)
    _h2sigs().orc_fht_rotate(_p(s), _p(x), x.shape[0], p, _p(out))
    return out


def query_prepare(signs, q, p):
    """(rotated P floats, lo, width, sumq, q bytes, LUT 4P bytes)."""
    s = np.ascontiguousarray(signs, np.float32)
    q = np.ascontiguousarray(q, np.float32)
    rq = np.empty(p, np.float32)
    lo = np.zeros(1, np.float32)
    wd = np.zeros(1, np.float32)
    sq = np.zeros(1, np.int32)
    qb = np.empty(p, np.uint8)
    lut = np.empty(4 * p, np.uint8)
    _h2sigs().orc_query_prepare(_p(s), _p(q), q.shape[0], p, _p(rq), _p(lo), _p(wd), _p(sq),
                                _p(qb), _p(lut))
    return rq, float(lo[0]), float(wd[0]), int(sq[0]), qb, lut


def scan_estimates(codes, layout, factors, deg, p, lut, sqr_y, lo, width, sumq):
    codes = np.ascontiguousarray(codes, np.uint8)
    factors = np.ascontiguousarray(factors, np.float32)
    lut = np.ascontiguousarray(lut, np.uint8)
    sums = np.empty(deg, np.uint16)
    est = np.empty(deg, np.float32)
    _h2sigs().orc_scan_estimates(_p(codes), layout, _p(factors), deg, p, _p(lut), sqr_y, lo,
                                 width, sumq, _p(sums), _p(est))
    return sums, est


def buffers_replay(ops, cap):
    ops = np.ascontiguousarray(ops, np.uint32)
    nops = ops.shape[0] // 3
    popped = np.empty(nops, np.uint32)
    npop = np.zeros(1, np.uint32)
    res = np.empty(cap, np.uint32)
    _h2sigs().orc_buffers_replay(_p(ops), nops, cap, _p(popped), _p(npop), _p(res))
    return popped[:int(npop[0])].copy(), res


class H2Arrays:
    """Everything the FastScan search reads: raw rows, ids, sidecar, signs."""

    def __init__(self, ix: IndexArrays, signs, codes, factors, layout: int):
        self.ix = ix
        self.p = padded_dim(ix.dim)
        self.signs = np.ascontiguousarray(signs, np.float32)
        self.codes = np.ascontiguousarray(codes, np.uint8)
        self.factors = np.ascontiguousarray(factors, np.float32)
        self.layout = int(layout)
        self._s = _OrcH2(ix.n, ix.dim, ix.deg, self.p, ix.ep, ix.raw.ctypes.data,
                         ix.nbr.ctypes.data, self.codes.ctypes.data, self.factors.ctypes.data,
                         self.signs.ctypes.data, self.layout)

    @property
    def ref(self):
        return C.byref(self._s)


def sidecar_build(ix: IndexArrays, signs, threads: int = 0):
    """This build's sidecar (vector-major codes, FP64-summed factors)."""
    p = padded_dim(ix.dim)
    codes = np.zeros((ix.n, ix.deg * p // 8), np.uint8)
    factors = np.zeros((ix.n, 3 * ix.deg), np.float32)
    h = H2Arrays(ix, signs, codes, factors, 1)
    _h2sigs().orc_sidecar_build(h.ref, _p(codes), _p(factors), threads)
    return codes, factors


def codes_to_vector_major(ref_codes, deg, p):
    ref_codes = np.ascontiguousarray(ref_codes, np.uint8).reshape(-1, deg * p // 8)
    out = np.empty_like(ref_codes)
    L = _h2sigs()
    for i in range(ref_codes.shape[0]):
        L.orc_codes_to_vector_major(_p(ref_codes[i]), deg, p, _p(out[i]))
    return out


def h2_expansions_batch(h: H2Arrays, queries, ef: int, cap: int, threads: int = 0):
    """find_candidates (qg.hpp:439-473): the first cap nodes the FastScan search
    of each query expands, in order (0xFFFFFFFF past the last)."""
    Q = np.ascontiguousarray(queries, np.float32)
    out = np.empty((Q.shape[0], cap), np.uint32)
    if _h2sigs().orc_h2_expansions_batch(h.ref, _p(Q), Q.shape[0], ef, cap, _p(out),
                                         threads) != 0:
        raise RuntimeError("orc_h2_expansions_batch failed")
    return out


def search_h2_batch(h: H2Arrays, queries, ef: int, k: int, threads: int = 0):
    Q = np.ascontiguousarray(queries, np.float32)
    res = np.empty((Q.shape[0], k), np.uint32)
    dis = np.empty((Q.shape[0], k), np.float32)
    st = np.zeros((Q.shape[0], 2), np.uint64)
    if _h2sigs().orc_search_h2_batch(h.ref, _p(Q), Q.shape[0], ef, k, _p(res), _p(dis), _p(st),
                                     threads) != 0:
        raise RuntimeError("orc_search_h2_batch failed")
    return res, dis, st
