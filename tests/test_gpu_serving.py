"""HIP-graph replays and the serving engine on the device.

Bar: every graph replay equals the eager batch launch (replayed several
times with eager work and synchronisations in between, and two graphs
replayed in turn on two streams so their launches overlap); every served
request's ids and distances equal those of the same query searched alone,
whether it was served alone or in a mix of thousands of requests with other
ef and k; the engine really batches (batches > 1 request) and overlaps
launches (more than one lane in flight).
"""
import numpy as np
import pytest

from conftest import make_random_index

pytestmark = pytest.mark.gpu


def _index(tmp_path, n=3000, d=64, deg=32, seed=1, fastscan=False):
    from liteqg_amd.symphonyqg import Index
    from oracle import oracle as O

    arr, _ = make_random_index(n, d, deg, seed=seed, knn=True)
    p = tmp_path / "s.index"
    O.write_index(str(p), arr)
    ix = Index("QG", "L2", n, d, deg)
    ix.load(str(p))
    if fastscan:
        ix.build_sidecar(seed=seed)
        ix.set_mode("fastscan")
    return ix, arr


@pytest.mark.parametrize("fastscan", [False, True], ids=["sq8", "fastscan"])
def test_graph_replays_equal_eager(tmp_path, fastscan):
    import torch

    from liteqg_amd.serving import BatchGraph

    ix, arr = _index(tmp_path, fastscan=fastscan)
    rng = np.random.default_rng(0)
    nq, k = 700, 10
    Q = rng.random((nq, arr.dim), dtype=np.float32)
    ef = rng.integers(8, 200, nq).astype(np.uint32)
    kk = rng.integers(1, k + 1, nq).astype(np.uint32)
    want, want_d = ix.search_batch(Q, kk, ef=ef, order="nearest", return_distances=True)
    dq = torch.from_numpy(Q).cuda()
    de = torch.from_numpy(ef.view(np.int32)).cuda()
    dk = torch.from_numpy(kk.view(np.int32)).cuda()
    ids = torch.full((nq, k), -1, dtype=torch.int32, device="cuda")
    dis = torch.zeros((nq, k), dtype=torch.float32, device="cuda")
    g = BatchGraph(ix, dq.data_ptr(), nq, ids.data_ptr(), k, 200, d_ef=de.data_ptr(),
                   d_k=dk.data_ptr(), d_dists=dis.data_ptr())
    s = torch.cuda.Stream()
    for rep in range(4):
        ids.fill_(-1)
        torch.cuda.synchronize()
        g.launch(s.cuda_stream)
        torch.cuda.synchronize()
        got = ids.cpu().numpy().view(np.uint32)
        got_d = dis.cpu().numpy()
        for i in range(nq):
            assert np.array_equal(got[i, :kk[i]], want[i, :kk[i]]), (rep, i)
            assert np.array_equal(got_d[i, :kk[i]].view(np.uint32),
                                  want_d[i, :kk[i]].view(np.uint32)), (rep, i)
        # eager work between replays (another search on the handle's stream)
        ix.search_batch(Q[:50], 5, ef=32)
    # a replay after the inputs change sees the new inputs
    dq.copy_(torch.from_numpy(Q[::-1].copy()))
    torch.cuda.synchronize()
    g.launch(s.cuda_stream)
    torch.cuda.synchronize()
    got = ids.cpu().numpy().view(np.uint32)
    want2 = ix.search_batch(Q[::-1].copy(), kk, ef=ef, order="nearest")
    for i in range(nq):
        assert np.array_equal(got[i, :kk[i]], want2[i, :kk[i]])
    g.close()


@pytest.mark.parametrize("fastscan", [False, True], ids=["sq8", "fastscan"])
def test_graphs_back_to_back_on_two_streams(tmp_path, fastscan):
    """Two graphs of the same batch, each with its own results and scratch,
    replayed in turn on two streams with no synchronisation between them, so
    their launches overlap on the device: every replay equals the eager
    launch."""
    import torch

    from liteqg_amd.serving import BatchGraph

    ix, arr = _index(tmp_path, fastscan=fastscan)
    rng = np.random.default_rng(3)
    nq, k, ef = 2000, 10, 64
    Q = rng.random((nq, arr.dim), dtype=np.float32)
    want = ix.search_batch(Q, k, ef=ef, order="nearest")
    dq = torch.from_numpy(Q).cuda()
    outs = [torch.full((nq, k), -1, dtype=torch.int32, device="cuda") for _ in range(2)]
    graphs = [BatchGraph(ix, dq.data_ptr(), nq, o.data_ptr(), k, ef) for o in outs]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    try:
        for rnd in range(3):
            for o in outs:
                o.fill_(-1)
            torch.cuda.synchronize()
            for i in range(6):
                graphs[i % 2].launch(streams[i % 2].cuda_stream)
            torch.cuda.synchronize()
            for j, o in enumerate(outs):
                assert np.array_equal(o.cpu().numpy().view(np.uint32), want), (rnd, j)
    finally:
        for g in graphs:
            g.close()


@pytest.mark.parametrize("fastscan", [False, True], ids=["sq8", "fastscan"])
def test_served_results_equal_alone_and_in_mix(tmp_path, fastscan):
    from liteqg_amd.serving import Server, poisson_arrivals

    ix, arr = _index(tmp_path, fastscan=fastscan)
    rng = np.random.default_rng(5)
    nq = 4000
    Q = (arr.raw[rng.integers(0, arr.n, nq)] + rng.normal(0, 0.02, (nq, arr.dim))).astype(np.float32)
    ef = rng.integers(10, 257, nq).astype(np.uint32)
    k = rng.integers(1, 33, nq).astype(np.uint32)
    want, want_d = ix.search_batch(Q, k, ef=ef, order="nearest", return_distances=True)
    with Server(ix, lanes=4, max_batch=256, slots=2048, ef_max=256, kmax=32) as srv:
        # alone: one request at a time
        for i in range(40):
            t = srv.submit(Q[i], int(ef[i]), int(k[i]))
            ids, dis = srv.wait(t, int(k[i]))
            assert np.array_equal(ids, want[i, :k[i]]), i
            assert np.array_equal(dis.view(np.uint32), want_d[i, :k[i]].view(np.uint32)), i
        # in the mix: all requests, Poisson arrivals at 2M/s (faster than one
        # lane serves them, so batches form and lanes overlap)
        lat, ids = srv.replay(Q, ef, k, poisson_arrivals(nq, 2.0e6, seed=1))
        for i in range(nq):
            assert np.array_equal(ids[i, :k[i]], want[i, :k[i]]), i
        st = srv.stats()
    assert st["completed"] == st["submitted"] == nq + 40
    assert st["max_batch_seen"] > 1
    assert st["max_lanes_busy"] > 1
    assert np.all(lat > 0)


@pytest.mark.parametrize("fastscan", [False, True], ids=["sq8", "fastscan"])
def test_served_768d_results_equal_alone_and_in_mix(oracle_mod, tmp_path, fastscan):
    """The 768-d serving shape: unit-norm mixture, degree 32, per-request ef in
    64..512 and k in {10, 100}.  Each served result equals the batch search of
    the same query (and, for a sample, the oracle's), alone and in the mix."""
    from liteqg_amd import datasets
    from liteqg_amd.serving import Server, poisson_arrivals
    from liteqg_amd.symphonyqg import Index

    O = oracle_mod
    n, d, nq = 3000, 768, 1500
    base, Q = datasets.make("unit_mix", n, d, nq, seed=2, centres=16, sigma=0.25,
                            centre_law="normal")
    ix = Index("QG", "L2", n, d, 32)
    ix.build_index(base, 200, num_iter=2)
    p = tmp_path / "s768.index"
    ix.save(str(p))
    arr = O.read_index(str(p), n, d, 32)
    H = None
    if fastscan:
        ix.build_sidecar(seed=6)
        ix.set_mode("fastscan")
        s, codes, fac = ix.get_sidecar()
        H = O.H2Arrays(arr, s, codes, fac, 1)
    rng = np.random.default_rng(8)
    ef = rng.integers(64, 513, nq).astype(np.uint32)
    k = rng.choice([10, 100], nq).astype(np.uint32)
    want, want_d = ix.search_batch(Q, k, ef=ef, order="nearest", return_distances=True)
    for i in range(0, nq, 50):
        ki = int(k[i])
        if H is None:
            e_ids, e_d, _ = O.search_canon_batch(arr, Q[i:i + 1], int(ef[i]), ki, "nearest")
        else:
            e_ids, e_d, _ = O.search_h2_batch(H, Q[i:i + 1], int(ef[i]), ki)
        assert np.array_equal(want[i, :ki], e_ids[0]), i
        assert np.array_equal(want_d[i, :ki].view(np.uint32), e_d[0].view(np.uint32)), i
    with Server(ix, lanes=4, max_batch=256, slots=2048, ef_max=512, kmax=100) as srv:
        for i in range(20):
            t = srv.submit(Q[i], int(ef[i]), int(k[i]))
            ids, dis = srv.wait(t, int(k[i]))
            assert np.array_equal(ids, want[i, :k[i]]), i
            assert np.array_equal(dis.view(np.uint32), want_d[i, :k[i]].view(np.uint32)), i
        lat, ids = srv.replay(Q, ef, k, poisson_arrivals(nq, 1.0e6, seed=3))
        for i in range(nq):
            assert np.array_equal(ids[i, :k[i]], want[i, :k[i]]), i
        st = srv.stats()
    assert st["completed"] == st["submitted"] == nq + 20
    assert st["max_batch_seen"] > 1


def test_server_admission_and_errors(tmp_path):
    from liteqg_amd._lib import LqgError
    from liteqg_amd.serving import Server

    ix, arr = _index(tmp_path, n=1500)
    q = np.random.default_rng(0).random(arr.dim, dtype=np.float32)
    with Server(ix, lanes=1, max_batch=4, slots=4, ef_max=64, kmax=10) as srv:
        with pytest.raises(LqgError):
            srv.submit(q, 65, 10)  # ef above ef_max
        with pytest.raises(LqgError):
            srv.submit(q, 32, 11)  # k above kmax
        tickets = [srv.submit(q, 32, 10) for _ in range(4)]
        assert all(t is not None for t in tickets)
        # every slot taken until a result is collected
        extra = srv.submit(q, 32, 10)
        got = [srv.wait(t, 10)[0] for t in tickets]
        assert all(np.array_equal(g, got[0]) for g in got)
        if extra is None:
            extra = srv.submit(q, 32, 10)
        assert np.array_equal(srv.wait(extra, 10)[0], got[0])
        with pytest.raises(LqgError):
            srv.poll(tickets[0], 10)  # already released


@pytest.mark.parametrize("fastscan", [False, True], ids=["sq8", "fastscan"])
def test_requests_complete_before_their_batch(tmp_path, fastscan):
    """Each request is handed back when its own walk ends (the kernels write
    its results and then its flag), not when its batch's slowest request does:
    with one request at ef 1024 among requests at ef 8, the fast ones finish
    well before it, and every result equals the same query searched alone."""
    from liteqg_amd.serving import Server

    ix, arr = _index(tmp_path, fastscan=fastscan)
    rng = np.random.default_rng(9)
    nq = 400
    Q = rng.random((nq, arr.dim), dtype=np.float32)
    ef = np.full(nq, 8, np.uint32)
    ef[0] = 1024
    k = np.full(nq, 10, np.uint32)
    want = ix.search_batch(Q, k, ef=ef, order="nearest")
    with Server(ix, lanes=2, max_batch=nq, slots=nq, ef_max=1024, kmax=10) as srv:
        srv.replay(Q[:8], ef[:8], k[:8], np.zeros(8, np.uint64))  # warm-up
        lat, ids = srv.replay(Q, ef, k, np.zeros(nq, np.uint64))
        st = srv.stats()
    assert np.array_equal(ids, want)
    assert st["completed"] == st["submitted"]
    assert np.median(lat[1:]) < 0.5 * lat[0], (np.median(lat[1:]), lat[0])


def test_poll_capacity_below_k_raises_and_keeps_ticket(tmp_path):
    """A poll whose result buffer is smaller than the request's k is refused
    (no write past it) and the ticket stays valid for a correct poll."""
    from liteqg_amd.serving import Server
    from liteqg_amd.symphonyqg import LqgError

    ix, arr = _index(tmp_path, n=2000)
    q = np.random.default_rng(3).random(arr.dim, dtype=np.float32)
    want = ix.search_batch(q[None], 12, ef=64, order="nearest")[0]
    with Server(ix, lanes=1, max_batch=8, ef_max=64, kmax=16) as srv:
        t = srv.submit(q, 64, 12)
        assert t is not None
        with pytest.raises(LqgError) as e:
            srv.wait(t, 11)
        assert e.value.status == 1
        ids, _ = srv.wait(t, 12)
        assert np.array_equal(ids, want)
        with pytest.raises(LqgError):  # released now
            srv.poll(t, 12)


def test_overflow_fails_only_its_own_request(tmp_path):
    """FastScan visited-set overflow is per request.  With no device overflow
    table and a 64-entry LDS table, a query fails once its visited set passes
    48 nodes; the device API marks exactly those queries (E = 0xFFFFFFFF, ids
    0xFFFFFFFF), the host batch call fails naming the overflow, and in serving
    only those requests fail while every other request of the same batches
    is handed back with the results it has searched alone.  The kernel exits
    normally (the limit is a host setting)."""
    import ctypes as C

    import torch

    from liteqg_amd import _lib
    from liteqg_amd.serving import Server
    from liteqg_amd.symphonyqg import LqgError

    ix, arr = _index(tmp_path, n=3000, fastscan=True)
    rng = np.random.default_rng(11)
    nq, k = 64, 10
    Q = rng.random((nq, arr.dim), dtype=np.float32)
    ef = np.where(np.arange(nq) % 4 == 0, 200, 4).astype(np.uint32)
    want = ix.search_batch(Q, k, ef=ef)
    ix.set_overflow_bits(0)
    ix.set_launch(0, waves_per_group=1, cache_bits=6, occupancy=0)
    try:
        L = _lib.lib()
        dq = torch.from_numpy(Q).cuda()
        de = torch.from_numpy(ef.view(np.int32)).cuda()
        dids = torch.empty((nq, k), dtype=torch.int32, device="cuda")
        dst = torch.empty((nq, 2), dtype=torch.int32, device="cuda")
        _lib.check(L.lqg_search_batch_device(ix._h, C.c_void_p(dq.data_ptr()), nq,
                                             C.c_void_p(de.data_ptr()), None, 200, k, 1,
                                             C.c_void_p(dids.data_ptr()), None,
                                             C.c_void_p(dst.data_ptr()), None))
        torch.cuda.synchronize()
        failed = dst.cpu().numpy().view(np.uint32)[:, 0] == 0xFFFFFFFF
        got = dids.cpu().numpy().view(np.uint32)
        assert failed.any() and (~failed).any()
        assert (got[failed] == 0xFFFFFFFF).all()
        assert np.array_equal(got[~failed], want[~failed])
        with pytest.raises(LqgError, match="overflowed"):
            ix.search_batch(Q, k, ef=ef)
        with Server(ix, lanes=2, max_batch=16, ef_max=200, kmax=k) as srv:
            tickets = [srv.submit(Q[i], int(ef[i]), k) for i in range(nq)]
            for i, t in enumerate(tickets):
                if failed[i]:
                    with pytest.raises(LqgError, match="overflowed"):
                        srv.wait(t, k)
                else:
                    ids, _ = srv.wait(t, k)
                    assert np.array_equal(ids, want[i]), i
    finally:
        ix.set_overflow_bits(14)
        ix.reset_launch(0)
    assert np.array_equal(ix.search_batch(Q, k, ef=ef), want)


def test_graph_refuses_replay_after_reload(tmp_path):
    """A graph holds the index's device addresses; after the index is reloaded
    (or its sidecar replaced) its replay is refused instead of reading freed
    memory."""
    import torch

    from liteqg_amd.serving import BatchGraph
    from liteqg_amd.symphonyqg import LqgError

    ix, arr = _index(tmp_path, n=2000, fastscan=True)
    Q = np.random.default_rng(5).random((100, arr.dim), dtype=np.float32)
    dq = torch.from_numpy(Q).cuda()
    ids = torch.empty((100, 10), dtype=torch.int32, device="cuda")
    g = BatchGraph(ix, dq.data_ptr(), 100, ids.data_ptr(), 10, 50)
    g.launch()
    torch.cuda.synchronize()
    ix.build_sidecar(seed=9)  # new records
    with pytest.raises(LqgError, match="captured"):
        g.launch()
    g.close()
    ix.set_mode("fastscan")
    g = BatchGraph(ix, dq.data_ptr(), 100, ids.data_ptr(), 10, 50)
    g.launch()
    torch.cuda.synchronize()
    ix.load(str(tmp_path / "s.index"))
    with pytest.raises(LqgError, match="captured"):
        g.launch()
    g.close()
