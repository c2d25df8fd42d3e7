"""The C ABI: symbols, argument validation, file handling -- no device calls."""
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

HEADER = os.path.join(ROOT, "include", "liteqg.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:lqg_status|const char\*)\s+(lqg_\w+)\s*\(", text, re.M)))


@pytest.fixture(scope="module")
def L():
    from liteqg_amd import _lib

    return _lib


def test_library_exports_every_declared_symbol(L):
    lib = L.lib()
    syms = declared_symbols()
    assert "lqg_search" in syms and "lqg_search_batch_device" in syms
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(L.EXPORTED)


def test_library_is_built_for_gfx950_only():
    data = open(os.path.join(ROOT, "liteqg_amd", "liblqg.so"), "rb").read()
    assert b"gfx950" in data
    for other in (b"gfx942", b"gfx90a", b"gfx1100"):
        assert b"amdgcn-amd-amdhsa--" + other not in data


def test_constructor_validation():
    from liteqg_amd.symphonyqg import Index

    Index("QG", "L2", 100, 16, 32)
    Index("QG", "L2", 100, 16, degree_bound=64)
    with pytest.raises(ValueError, match="L2"):
        Index("QG", "IP", 100, 16)
    with pytest.raises(ValueError, match="multiple of 32"):
        Index("QG", "L2", 100, 16, 48)
    with pytest.raises(ValueError, match="multiple of 32"):
        Index("QG", "L2", 100, 16, 0)
    with pytest.raises(ValueError, match="not supported"):
        Index("HNSW", "L2", 100, 16)
    with pytest.raises(ValueError):
        Index("QG", "L2", 0, 16)


def test_keyword_names_match_reference():
    """bindings.cpp:104-124 keyword names."""
    import inspect

    from liteqg_amd.symphonyqg import Index

    assert list(inspect.signature(Index.__init__).parameters)[1:] == [
        "index_type", "metric", "num_elements", "dimension", "degree_bound"]
    assert list(inspect.signature(Index.load).parameters)[1:] == ["filename"]
    assert list(inspect.signature(Index.save).parameters)[1:] == ["filename"]
    assert list(inspect.signature(Index.set_ef).parameters)[1:] == ["EF"]
    # the reference's four parameters, in order; additions are keyword-only
    bp = inspect.signature(Index.build_index).parameters
    assert list(bp)[1:5] == ["data", "EF", "num_iter", "num_thread"]
    assert all(v.kind is inspect.Parameter.KEYWORD_ONLY for v in list(bp.values())[5:])
    assert list(inspect.signature(Index.search).parameters)[1:] == ["query", "k"]
    import symphonyqg

    assert symphonyqg.Index is Index


def test_load_save_roundtrip_is_byte_identical(tmp_path):
    from liteqg_amd.symphonyqg import Index

    src = f"{GOLDEN}/g_intmix_n1200_d128.index"
    ix = Index("QG", "L2", 1200, 128, 32)
    ix.load(src)
    info = ix.info()
    assert info["entry_point"] == int(np.fromfile(src, np.uint32, 1)[0])
    out = tmp_path / "copy.index"
    ix.save(str(out))
    assert out.read_bytes() == open(src, "rb").read()


def test_load_rejects_wrong_shapes(tmp_path):
    from liteqg_amd.symphonyqg import Index, LqgError

    src = f"{GOLDEN}/g_intmix_n1200_d128.index"
    with pytest.raises(LqgError, match="expected"):
        Index("QG", "L2", 1201, 128, 32).load(src)
    with pytest.raises(LqgError, match="expected"):
        Index("QG", "L2", 1200, 128, 64).load(src)
    with pytest.raises(LqgError, match="cannot open"):
        Index("QG", "L2", 1200, 128, 32).load(str(tmp_path / "missing.index"))
    trunc = tmp_path / "trunc.index"
    trunc.write_bytes(open(src, "rb").read()[:-1])
    with pytest.raises(LqgError, match="expected"):
        Index("QG", "L2", 1200, 128, 32).load(str(trunc))


def test_load_rejects_out_of_range_ids(tmp_path):
    from liteqg_amd.symphonyqg import Index, LqgError

    src = open(f"{GOLDEN}/g_uniform_n2000_d64.index", "rb").read()
    bad = bytearray(src)
    off = 4 + 64 * 4  # first neighbour id of row 0
    bad[off:off + 4] = np.uint32(2000).tobytes()
    p = tmp_path / "bad.index"
    p.write_bytes(bytes(bad))
    with pytest.raises(LqgError, match="out of range"):
        Index("QG", "L2", 2000, 64, 32).load(str(p))


def test_call_order_errors():
    from liteqg_amd.symphonyqg import Index, LqgError

    ix = Index("QG", "L2", 1200, 128, 32)
    with pytest.raises(LqgError, match="set_ef"):
        ix.search(np.zeros(128, np.float32), 10)
    with pytest.raises(LqgError):
        ix.set_ef(0)
    with pytest.raises(LqgError, match="1024"):
        ix.set_ef(5000)
    ix.set_ef(10)
    with pytest.raises(LqgError, match="no data"):
        ix.search(np.zeros(128, np.float32), 10)
    with pytest.raises(LqgError, match="no data"):
        ix.save("/tmp/never_written.index")
    with pytest.raises(ValueError):
        ix.search(np.zeros(5, np.float32), 10)
    with pytest.raises(ValueError):
        ix.set_order("farthest")
    ix.load(f"{GOLDEN}/g_intmix_n1200_d128.index")
    with pytest.raises(LqgError, match="positive"):
        ix.search(np.zeros(128, np.float32), 0)
    q = np.zeros(128, np.float32)
    q[3] = np.nan
    with pytest.raises(LqgError, match="non-finite"):
        ix.search(q, 10)
    with pytest.raises(LqgError, match="kmax"):
        ix.search_batch(np.zeros((3, 128), np.float32), np.array([1, 2, 0]))


def test_build_index_shape_check():
    from liteqg_amd.symphonyqg import Index, LqgError

    ix = Index("QG", "L2", 100, 16, 32)
    with pytest.raises(LqgError, match="shape"):
        ix.build_index(np.zeros((99, 16), np.float32), 50)
    with pytest.raises(ValueError, match="shape"):
        ix.build_index(np.zeros((2, 3, 4), np.float32), 50)


def test_fastscan_graph_and_serving_validation():
    """Argument and call-order checks of the FastScan, graph and serving entry
    points fire before any device work."""
    import ctypes as C

    from liteqg_amd import _lib
    from liteqg_amd.symphonyqg import Index

    L = _lib.lib()
    ix = Index("QG", "L2", 100, 64, 32)
    # FastScan mode needs a sidecar
    assert L.lqg_set_mode(ix._h, 1) == 2
    assert L.lqg_set_mode(ix._h, 5) == 1
    with pytest.raises(ValueError):
        ix.set_mode("bogus")
    # sidecar entry points need a loaded index
    assert L.lqg_h2_build_sidecar(ix._h, None, 0) == 2
    assert L.lqg_h2_get_sidecar(ix._h, None, None, None) == 2
    assert L.lqg_h2_save_sidecar(ix._h, b"/nonexistent/x") == 2
    q = np.zeros(64, np.float32)
    lut = np.zeros(256, np.uint8)
    lw = np.zeros(2, np.float32)
    sq = np.zeros(1, np.int32)
    assert L.lqg_h2_probe(ix._h, _lib.ptr(q), 0, None, _lib.ptr(lut), _lib.ptr(lw),
                          _lib.ptr(sq), None, None) == 1
    # rotation method
    assert L.lqg_h2_set_rotation(ix._h, 2) == 1
    with pytest.raises(ValueError):
        ix.set_rotation("qr")
    ix.set_rotation("gemm")
    assert ix.get_rotation() == "gemm"
    ix.set_rotation("fht")
    assert L.lqg_h2_rotate(ix._h, _lib.ptr(q), 1, 0, _lib.ptr(q)) == 2  # no sidecar yet
    assert L.lqg_h2_rotate(ix._h, _lib.ptr(q), 1, 3, _lib.ptr(q)) == 1
    # register caps: SQ8 takes -1 (uncapped), 0 (default), 6, 8 (FastScan's 5
    # is not an SQ8 setting)
    for occ in (5, 7, -2):
        with pytest.raises(_lib.LqgError) as e:
            ix.set_launch(0, waves_per_group=2, cache_bits=10, occupancy=occ)
        assert e.value.status == 1
    for occ in (-1, 6):
        ix.set_launch(0, waves_per_group=2, cache_bits=10, occupancy=occ)
        assert ix.get_launch(64)["occupancy"] == occ
    ix.reset_launch(0)
    # diagnostic knobs: the integer SQ8 kernel, the FastScan overflow table
    assert L.lqg_set_sq8_integer(ix._h, 0) == 0
    assert L.lqg_set_sq8_integer(ix._h, 1) == 0
    assert L.lqg_set_sq8_integer(None, 1) == 1
    for bits in (1, 7, 17):
        assert L.lqg_h2_set_overflow_bits(ix._h, bits) == 1
    for bits in (0, 8, 16, 14):
        assert L.lqg_h2_set_overflow_bits(ix._h, bits) == 0
    # settings are per pool-capacity class; ef 129..192 is one class
    ix.set_launch(150, waves_per_group=3, cache_bits=7, occupancy=0)
    assert ix.get_launch(129)["cache_bits"] == 7 and ix.get_launch(192)["cache_bits"] == 7
    assert ix.get_launch(128)["cache_bits"] != 7 and ix.get_launch(193)["cache_bits"] != 7
    ix.reset_launch(0)
    assert ix.get_launch(150)["cache_bits"] != 7
    # graph capture: null output, zero queries, unknown order
    g = C.c_void_p()
    assert L.lqg_graph_create(ix._h, None, 1, None, None, 10, 10, 1, None, None, None,
                              C.byref(g)) == 1
    assert L.lqg_graph_create(ix._h, C.c_void_p(16), 0, None, None, 10, 10, 1, C.c_void_p(16),
                              None, None, C.byref(g)) == 1
    assert L.lqg_graph_launch(None, None) == 1
    # serving configuration bounds
    s = C.c_void_p()
    for cfg in [(0, 8, 8, 64, 10, 1), (17, 8, 8, 64, 10, 1), (2, 8, 4, 64, 10, 1),
                (2, 8, 8, 0, 10, 1), (2, 8, 8, 2048, 10, 1), (2, 8, 8, 64, 0, 1),
                (2, 8, 8, 64, 10, 7)]:
        assert L.lqg_serve_start(ix._h, C.byref(_lib.ServeConfig(*cfg)), C.byref(s)) == 1, cfg
    # a valid configuration on an index with no data fails on state, not args
    assert L.lqg_serve_start(ix._h, C.byref(_lib.ServeConfig(2, 8, 8, 64, 10, 1)),
                             C.byref(s)) == 2
    assert L.lqg_serve_stop(None) == 0


def test_header_is_plain_c(tmp_path):
    """include/liteqg.h compiles as C99 (pedantic, warnings as errors) and as C++,
    so C, cgo-style and ctypes hosts can use it as is."""
    import shutil
    import subprocess

    src = tmp_path / "use.c"
    src.write_text('#include "liteqg.h"\n'
                   "int f(lqg_index* h) {\n"
                   "    lqg_launch c; lqg_mode m; lqg_rotation r; lqg_serve_config s;\n"
                   "    (void)s; lqg_get_mode(h, &m); lqg_h2_get_rotation(h, &r);\n"
                   "    return lqg_get_launch(h, 50, &c) + (int)LQG_ORDER_NEAREST;\n"
                   "}\n")
    inc = os.path.join(ROOT, "include")
    for cc, std in (("gcc", "-std=c99"), ("g++", "-std=c++17")):
        if not shutil.which(cc):
            pytest.skip(f"{cc} not available")
        args = [cc, std, "-Wall", "-Wextra", "-pedantic", "-Werror", "-fsyntax-only", "-I", inc]
        if cc == "g++":
            args += ["-x", "c++"]
        r = subprocess.run(args + [str(src)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
