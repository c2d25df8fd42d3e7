"""The allocation-free batch entry point (lqg_search_batch_device_noalloc).

Bar: called through the C ABI on device arrays, two calls in a row (the
second with fewer queries) return the same ids, distances and per-query
counters as lqg_search_batch_device on the same inputs, in the SQ8 mode
(both output orders), the FastScan mode with the FHT rotation and with the
GEMM rotation; device free memory is the same before and after them; its
limits are refused with an error instead of an allocation.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import make_random_index

pytestmark = pytest.mark.gpu


def _index(tmp_path, n=3000, d=64, deg=32, seed=4):
    from liteqg_amd.symphonyqg import Index
    from oracle import oracle as O

    arr, _ = make_random_index(n, d, deg, seed=seed, knn=True)
    p = tmp_path / "na.index"
    O.write_index(str(p), arr)
    ix = Index("QG", "L2", n, d, deg)
    ix.load(str(p))
    return ix, arr


class _Bufs:
    """Device arrays of one batch, allocated once (so that later calls on them
    allocate nothing on the torch side either)."""

    def __init__(self, Q, ef, k, kmax):
        import torch

        nq = Q.shape[0]
        self.nq, self.kmax, self.ef_max = nq, kmax, int(ef.max())
        self.q = torch.from_numpy(Q).cuda()
        self.ef = torch.from_numpy(ef.view(np.int32)).cuda()
        self.k = torch.from_numpy(k.view(np.int32)).cuda()
        self.ids = torch.empty((nq, kmax), dtype=torch.int32, device="cuda")
        self.dis = torch.empty((nq, kmax), dtype=torch.float32, device="cuda")
        self.st = torch.empty((nq, 2), dtype=torch.int32, device="cuda")

    def run(self, fn, ix, order):
        """One call of `fn` (an entry point with the signature of
        lqg_search_batch_device); returns host copies of its outputs."""
        import torch

        from liteqg_amd import _lib

        self.ids.fill_(-7)
        self.dis.fill_(-7.0)
        self.st.fill_(-7)
        torch.cuda.synchronize()
        p = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
        _lib.check(fn(ix._h, p(self.q), self.nq, p(self.ef), p(self.k), self.ef_max, self.kmax,
                      order, p(self.ids), p(self.dis), p(self.st), None))
        torch.cuda.synchronize()
        return (self.ids.cpu().numpy().view(np.uint32).copy(), self.dis.cpu().numpy().copy(),
                self.st.cpu().numpy().view(np.uint32).copy())


@pytest.mark.parametrize("mode", ["sq8_reference", "sq8_nearest", "fastscan_fht",
                                  "fastscan_gemm"])
def test_noalloc_equals_allocating_entry_point(tmp_path, mode):
    import torch

    from liteqg_amd import _lib

    L = _lib.lib()
    ix, arr = _index(tmp_path)
    order = 0 if mode == "sq8_reference" else 1
    if mode.startswith("fastscan"):
        ix.build_sidecar(seed=2)
        ix.set_mode("fastscan")
        ix.set_rotation("gemm" if mode.endswith("gemm") else "fht")
    rng = np.random.default_rng(8)
    kmax = 12
    sizes = (1500, 333)  # the second call smaller than the first
    inputs = []
    for nq in sizes:
        Q = rng.random((nq, arr.dim), dtype=np.float32)
        ef = rng.integers(4, 65, nq).astype(np.uint32)
        k = rng.integers(1, kmax + 1, nq).astype(np.uint32)
        inputs.append((Q, ef, k))
    bufs = [_Bufs(Q, ef, k, kmax) for Q, ef, k in inputs]
    # the allocating entry point first: it also puts the index on the device
    want = [b.run(L.lqg_search_batch_device, ix, order) for b in bufs]
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    got = [b.run(L.lqg_search_batch_device_noalloc, ix, order) for b in bufs]
    free1 = torch.cuda.mem_get_info()[0]
    for (Q, ef, k), (wi, wd, ws), (gi, gd, gs) in zip(inputs, want, got):
        assert (gs[:, 0] != 0xFFFFFFFF).all()  # no query needed an overflow table
        for i in range(Q.shape[0]):
            assert np.array_equal(gi[i, :k[i]], wi[i, :k[i]]), i
            assert np.array_equal(gd[i, :k[i]].view(np.uint32), wd[i, :k[i]].view(np.uint32)), i
        assert np.array_equal(gs, ws)
    # nothing else allocates between the two readings, so a hipMalloc by the
    # library would show as less free device memory
    assert free1 == free0


def test_noalloc_limits(tmp_path):
    import torch

    from liteqg_amd import _lib

    L = _lib.lib()
    ix, arr = _index(tmp_path, n=2000)
    Q = np.random.default_rng(1).random((8, arr.dim), dtype=np.float32)
    ef = np.full(8, 32, np.uint32)
    k = np.full(8, 10, np.uint32)
    # not on the device yet: refused (uploading would allocate)
    dq = torch.from_numpy(Q).cuda()
    ids = torch.empty((8, 10), dtype=torch.int32, device="cuda")
    assert L.lqg_search_batch_device_noalloc(ix._h, C.c_void_p(dq.data_ptr()), 8, None, None,
                                             32, 10, 1, C.c_void_p(ids.data_ptr()), None, None,
                                             None) == 2
    b = _Bufs(Q, ef, k, 10)
    b.run(L.lqg_search_batch_device, ix, 1)
    # GEMM rotation: nq x P above 262,144 floats is refused
    ix.build_sidecar(seed=1)
    ix.set_mode("fastscan")
    ix.set_rotation("gemm")
    nq = 262144 // ix.padded_dim + 1
    big = torch.zeros((nq, arr.dim), dtype=torch.float32, device="cuda")
    bids = torch.empty((nq, 10), dtype=torch.int32, device="cuda")
    assert L.lqg_search_batch_device_noalloc(ix._h, C.c_void_p(big.data_ptr()), nq, None, None,
                                             32, 10, 1, C.c_void_p(bids.data_ptr()), None, None,
                                             None) == 1
    assert b"262,144" in L.lqg_last_error()
    # a batch within the limit runs, with the allocating entry point's results
    got = b.run(L.lqg_search_batch_device_noalloc, ix, 1)
    want = b.run(L.lqg_search_batch_device, ix, 1)
    assert np.array_equal(got[0], want[0])
    assert np.array_equal(got[2], want[2])
