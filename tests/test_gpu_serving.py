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
