"""Writes the golden fixtures in tests/golden/ from the reference's own code.

Runs only where the reference tree is present: it drives oracle/_ref/ref_driver
(built by ``make -C oracle ref`` from the reference's headers) and stores
inputs and outputs as data files.  Reference builds are not reproducible
(the rotator and the random initial graph draw from std::random_device,
utils/rotator.hpp:37-43, utils/tools.hpp:8-15), so the index FILES are the
fixtures, not a recipe.

Fixtures written (all small):
  <name>.index           reference-built index file (QGBuilder, save_index)
  <name>.npz             base/queries shapes, queries, and the reference's
                         search() ids for each ef in EFS at k=10
  funcs_d<D>.npz         sq8_dequantize + l2_sqr vectors for several D
  h2_<name>.npz          the reference's search_qg on the FastScan rows:
                         sidecar, query preparation, ids at H2_EFS, k=10
  h2u_<name>.npz         search_qg ids at the H2_UPDATE_CASES (ef, k) pairs

Usage:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from liteqg_amd import datasets  # noqa: E402

DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
DRIVER_H2 = os.path.join(ROOT, "oracle", "_ref", "ref_h2")
H2_EFS = (10, 50, 100)
# (ef, k) pairs whose searches mostly expand fewer than k nodes, so the
# reference's update_results tail (qg.hpp:395-418) and its zero-filled id
# snapshot (buffer.hpp:81-102) decide the output
H2_UPDATE_CASES = ((1, 10), (2, 10), (4, 10), (16, 100), (64, 100))
H2_INDEXES = ("g_uniform_n2000_d64", "g_intmix_n1200_d128", "g_unitmix_n1000_d96_deg64")
EFS = (10, 20, 50, 100)
K = 10

# name: (process kind, N, D, deg, build EF, queries, generator kwargs, seed)
INDEXES = {
    "g_uniform_n2000_d64": ("uniform", 2000, 64, 32, 200, 200, {}, 11),
    "g_intmix_n1200_d128": ("int_mix", 1200, 128, 32, 200, 200, dict(centres=8, sigma=20.0), 12),
    "g_unitmix_n1000_d96_deg64": ("unit_mix", 1000, 96, 64, 200, 150,
                                  dict(centres=8, sigma=0.3), 13),
    # values in {0,1,2,3}: squared distances are small integers, so exact
    # ties are everywhere and the heap's tie order decides the output
    "g_ties_n1500_d64": ("ties", 1500, 64, 32, 100, 200, {}, 14),
}
FUNC_DIMS = (64, 96, 128, 768, 960, 80)


def run(*args):
    subprocess.run([DRIVER, *map(str, args)], check=True, stdout=subprocess.DEVNULL)


def make_index(name, kind, n, d, deg, ef_build, nq, kw, seed):
    if kind == "ties":
        rng = np.random.default_rng(seed)
        base = rng.integers(0, 4, size=(n, d)).astype(np.float32)
        queries = rng.integers(0, 4, size=(nq, d)).astype(np.float32)
    else:
        base, queries = datasets.make(kind, n, d, nq, seed=seed, **kw)
    with tempfile.TemporaryDirectory() as tmp:
        bpath = os.path.join(tmp, "base.f32")
        qpath = os.path.join(tmp, "q.f32")
        base.tofile(bpath)
        queries.tofile(qpath)
        ipath = os.path.join(HERE, name + ".index")
        run("build", bpath, n, d, deg, ef_build, 3, 8, ipath)
        out = {}
        for ef in EFS:
            opath = os.path.join(tmp, f"o{ef}.u32")
            run("search", ipath, n, d, deg, qpath, nq, ef, K, opath)
            out[f"ids_ef{ef}"] = np.fromfile(opath, np.uint32).reshape(nq, K)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), queries=queries,
                        shape=np.array([n, d, deg], np.int64), k=np.int64(K),
                        efs=np.array(EFS, np.int64), **out)


def padded(d):
    p = 1
    while p < d:
        p <<= 1
    return p


def make_h2(name, nq=100, nprep=8, seed=3):
    """Reference search_qg on the rebuilt row layout (oracle/ref_h2.cpp)."""
    z = np.load(os.path.join(HERE, name + ".npz"))
    n, d, deg = (int(x) for x in z["shape"])
    p = padded(d)
    rng = np.random.default_rng(seed + d)
    signs = ((rng.integers(0, 2, p) * 2 - 1).astype(np.float32) /
             np.sqrt(np.float32(p))).astype(np.float32)
    q = np.ascontiguousarray(z["queries"][:nq])
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        sp, qp = os.path.join(tmp, "s.f32"), os.path.join(tmp, "q.f32")
        signs.tofile(sp)
        q.tofile(qp)
        ipath = os.path.join(HERE, name + ".index")
        for ef in H2_EFS:
            o, sc, pr = (os.path.join(tmp, x) for x in ("o.u32", "side.bin", "prep.bin"))
            subprocess.run([DRIVER_H2, "search", ipath, str(n), str(d), str(deg), sp, qp,
                            str(nq), str(ef), str(K), o, sc, pr, str(nprep)], check=True,
                           stdout=subprocess.DEVNULL)
            out[f"ids_ef{ef}"] = np.fromfile(o, np.uint32).reshape(nq, K)
        side = np.fromfile(sc, np.uint8).reshape(n, deg * p // 8 + 12 * deg)
        prep = np.fromfile(pr, np.uint8).reshape(nprep, -1)
    codes = side[:, :deg * p // 8].copy()
    factors = side[:, deg * p // 8:].copy().view(np.float32).reshape(n, 3 * deg)
    o = 0
    rq = prep[:, o:o + 4 * p].copy().view(np.float32); o += 4 * p
    lo = prep[:, o:o + 4].copy().view(np.float32)[:, 0]; o += 4
    width = prep[:, o:o + 4].copy().view(np.float32)[:, 0]; o += 4
    sumq = prep[:, o:o + 4].copy().view(np.int32)[:, 0]; o += 4
    lut = prep[:, o:o + 4 * p].copy(); o += 4 * p
    sqr_y = prep[:, o:o + 4].copy().view(np.float32)[:, 0]; o += 4
    est = prep[:, o:o + 4 * deg].copy().view(np.float32)
    np.savez_compressed(os.path.join(HERE, "h2_" + name + ".npz"), signs=signs, queries=q,
                        codes=codes, factors=factors, prep_rq=rq, prep_lo=lo, prep_width=width,
                        prep_sumq=sumq, prep_lut=lut, prep_sqr_y=sqr_y, prep_est=est,
                        efs=np.array(H2_EFS, np.int64), k=np.int64(K), **out)


def make_h2_update(name, nq=100, seed=3):
    """search_qg at small ef / large k (update_results), same signs and
    queries as make_h2."""
    z = np.load(os.path.join(HERE, name + ".npz"))
    n, d, deg = (int(x) for x in z["shape"])
    h = np.load(os.path.join(HERE, "h2_" + name + ".npz"))
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        sp, qp = os.path.join(tmp, "s.f32"), os.path.join(tmp, "q.f32")
        h["signs"].tofile(sp)
        q = np.ascontiguousarray(h["queries"][:nq])
        q.tofile(qp)
        ipath = os.path.join(HERE, name + ".index")
        for ef, k in H2_UPDATE_CASES:
            o, sc, pr = (os.path.join(tmp, x) for x in ("o.u32", "side.bin", "prep.bin"))
            subprocess.run([DRIVER_H2, "search", ipath, str(n), str(d), str(deg), sp, qp,
                            str(nq), str(ef), str(k), o, sc, pr, "0"], check=True,
                           stdout=subprocess.DEVNULL)
            out[f"ids_ef{ef}_k{k}"] = np.fromfile(o, np.uint32).reshape(nq, k)
    np.savez_compressed(os.path.join(HERE, "h2u_" + name + ".npz"), queries=q,
                        cases=np.array(H2_UPDATE_CASES, np.int64), **out)


def make_buffers(cap=16, nops=4000, seed=9):
    """SearchBuffer / ResultBuffer operation traces (utils/buffer.hpp)."""
    rng = np.random.default_rng(seed)
    kind = rng.choice([0, 0, 0, 1, 2], size=nops).astype(np.uint32)
    ids = rng.integers(0, 1 << 20, size=nops).astype(np.uint32)
    dist = rng.integers(0, 40, size=nops).astype(np.float32)  # many equal distances
    ops = np.stack([kind, ids, dist.view(np.uint32)], axis=1).reshape(-1)
    with tempfile.TemporaryDirectory() as tmp:
        op, outp = os.path.join(tmp, "ops.u32"), os.path.join(tmp, "out.bin")
        ops.tofile(op)
        subprocess.run([DRIVER_H2, "buffers", op, str(nops), str(cap), outp], check=True)
        raw = np.fromfile(outp, np.uint32)
    npop = int(raw[0])
    np.savez_compressed(os.path.join(HERE, "buffers.npz"), ops=ops, cap=np.int64(cap),
                        popped=raw[1:1 + npop], results=raw[1 + npop:1 + npop + cap])


def make_funcs(d, m=None, seed=5):
    m = m or (64 if d <= 128 else 12)
    rng = np.random.default_rng(seed + d)
    codes = rng.integers(0, 256, size=(m, d), dtype=np.uint8)
    mn = rng.uniform(-2.0, 2.0, size=d).astype(np.float32)
    sc = rng.uniform(1e-4, 0.05, size=d).astype(np.float32)
    params = np.stack([mn, sc], axis=1).astype(np.float32)
    q = (mn + codes.astype(np.float32) * sc + rng.normal(0, 0.3, size=(m, d))).astype(np.float32)
    with tempfile.TemporaryDirectory() as tmp:
        p = [os.path.join(tmp, x) for x in ("c.u8", "p.f32", "q.f32", "dec.f32", "l2.f32")]
        codes.tofile(p[0])
        params.tofile(p[1])
        q.tofile(p[2])
        run("funcs", p[0], p[1], p[2], d, m, p[3], p[4])
        dec = np.fromfile(p[3], np.float32).reshape(m, d)
        l2 = np.fromfile(p[4], np.float32)
    np.savez_compressed(os.path.join(HERE, f"funcs_d{d}.npz"), codes=codes, params=params,
                        queries=q, decoded=dec, l2=l2)


def main():
    if not os.path.exists(DRIVER) or not os.path.exists(DRIVER_H2):
        sys.exit("oracle/_ref drivers missing: run `make -C oracle ref` first")
    if "--h2-only" not in sys.argv:
        for d in FUNC_DIMS:
            make_funcs(d)
    if "--h2-only" not in sys.argv:
        for name, spec in INDEXES.items():
            make_index(name, *spec)
            print("wrote", name)
    if "--h2u-only" not in sys.argv:
        make_buffers()
        for name in H2_INDEXES:
            make_h2(name)
            print("wrote h2", name)
    for name in H2_INDEXES:
        make_h2_update(name)
        print("wrote h2u", name)


if __name__ == "__main__":
    main()
