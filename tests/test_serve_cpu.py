"""The serving loop's host logic on the CPU: continuous batching across
requests and MFMA row tiles, results routed to the right requests, failure
propagation and the HBM budget.  The GPU pieces (streams, events, the split
prep and the cross-Gram) are replaced by CPU stand-ins whose results come
from the FP64 oracle, so only the scheduling code is under test here; the
GPU path is tested in test_gpu_serve.py."""
import contextlib

import numpy as np
import pytest
import torch

from mdtraj_amd import gen, serve
from oracle import fp64


class _Cuda:
    class Stream:
        waited = []

        def wait_event(self, ev):
            _Cuda.Stream.waited.append(ev)

    @staticmethod
    @contextlib.contextmanager
    def stream(s):
        yield

    @staticmethod
    def synchronize():
        pass

    class Event:
        def record(self, stream=None):
            pass

        def synchronize(self):
            pass

        def query(self):  # work on this stand-in is done when it is recorded
            return True


class _Torch:
    cuda = _Cuda
    float32 = torch.float32

    def __getattr__(self, name):
        return getattr(torch, name)

    @staticmethod
    def empty(*a, pin_memory=False, **k):
        return torch.empty(*a, **k)


class _Split:
    def __init__(self, xyz, mode="f16exact"):
        self.x = np.asarray(xyz, np.float32)
        self.n_frames = self.x.shape[0]
        self.nbytes = self.x.nbytes


def _cross(qs, ps):
    rows = [fp64.rmsd(ps.x, qs.x, i) for i in range(qs.n_frames)]
    return torch.from_numpy(np.asarray(rows, np.float32))


@pytest.fixture
def cpu_server(monkeypatch):
    monkeypatch.setattr(serve.dev, "require_torch_gpu", lambda: _Torch())
    monkeypatch.setattr(serve.dev, "device", lambda: "cpu")
    monkeypatch.setattr(serve.dev, "to_device_f32",
                        lambda x: torch.from_numpy(np.ascontiguousarray(x, np.float32)))
    monkeypatch.setattr(serve, "SplitSet", _Split)
    monkeypatch.setattr(serve, "rmsd_cross", _cross)
    return serve.RMSDServer


def test_batches_merge_and_rows_reach_their_requests(cpu_server):
    pools = [gen.walk(seed=20 + i, n_frames=f, n_atoms=n)
             for i, (f, n) in enumerate([(40, 12), (25, 30)])]
    rng = np.random.default_rng(1)
    srv = cpu_server(hbm_budget=1e9, max_batch_queries=128)
    ids = [srv.add_pool(p) for p in pools]
    reqs = []
    for _ in range(30):
        pid = int(rng.integers(0, 2))
        q = int(rng.integers(1, 65))
        queries = pools[pid][rng.integers(0, pools[pid].shape[0], q)]
        reqs.append((pid, queries, srv.submit(ids[pid], queries)))
    with srv:
        for pid, queries, r in reqs:
            got = r.wait(30)
            want = np.asarray([fp64.rmsd(pools[pid], queries, i) for i in range(len(queries))])
            assert got.shape == (len(queries), pools[pid].shape[0])
            assert np.abs(got - want).max() < 1e-6
            assert r.latency is not None and r.latency >= 0
    assert 64 < srv.largest_batch <= 128
    assert srv.batches < len(reqs)


def test_request_limits(cpu_server):
    srv = cpu_server(hbm_budget=1e9)
    pid = srv.add_pool(gen.walk(seed=1, n_frames=10, n_atoms=8))
    with pytest.raises(ValueError, match="1..64"):
        srv.submit(pid, np.zeros((65, 8, 3), np.float32))
    with pytest.raises(ValueError):
        srv.submit(pid, np.zeros((2, 9, 3), np.float32))
    with pytest.raises(ValueError):
        cpu_server(max_batch_queries=32, max_request_queries=64)


def test_budget_refuses_pools(cpu_server):
    srv = cpu_server(hbm_budget=10_000)
    with pytest.raises(MemoryError):
        srv.add_pool(gen.walk(seed=1, n_frames=500, n_atoms=30))


def test_without_copy_out_only_latency_is_recorded(cpu_server):
    srv = cpu_server(hbm_budget=1e9, keep_results=False)
    pool = gen.walk(seed=2, n_frames=20, n_atoms=10)
    pid = srv.add_pool(pool)
    with srv:
        r = srv.submit(pid, pool[:3])
        assert r.wait(30) is None
        assert r.latency is not None


def test_failure_reaches_every_waiter(cpu_server, monkeypatch):
    def boom(*a, **k):
        raise MemoryError("forced failure")

    srv = cpu_server(hbm_budget=1e9)
    pool = gen.walk(seed=3, n_frames=30, n_atoms=10)
    pid = srv.add_pool(pool)
    monkeypatch.setattr(serve, "rmsd_cross", boom)
    reqs = [srv.submit(pid, pool[:2]) for _ in range(6)]
    with srv:
        for r in reqs:
            with pytest.raises(RuntimeError):
                r.wait(30)
        with pytest.raises(RuntimeError):
            srv.submit(pid, pool[:1])


def test_batches_wait_for_the_pool_preparation(cpu_server):
    """Pools are prepared on the caller's stream and batches run on the
    server's: every batch's stream waits on its pool's ready event first."""
    _Cuda.Stream.waited.clear()
    srv = cpu_server(hbm_budget=1e9)
    pools = [gen.walk(seed=5 + i, n_frames=20, n_atoms=10) for i in range(2)]
    ids = [srv.add_pool(p) for p in pools]
    with srv:
        for pid in ids:
            srv.submit(pid, pools[pid][:2]).wait(30)
    readies = [srv.pools[pid].ready for pid in ids]
    assert len(_Cuda.Stream.waited) == srv.batches >= 2
    assert all(any(w is r for r in readies) for w in _Cuda.Stream.waited)
    assert any(w is readies[0] for w in _Cuda.Stream.waited)
    assert any(w is readies[1] for w in _Cuda.Stream.waited)


def test_pool_that_fits_but_not_with_a_full_batch(cpu_server):
    """The budget holds the pool and one 64-query batch but not a full
    256-query batch: requests still complete (batches are sized to what is
    free), in several batches."""
    pool = gen.walk(seed=6, n_frames=400, n_atoms=50)
    probe = cpu_server(hbm_budget=1e12)
    p = probe.add_pool(pool)
    one = probe._batch_bytes(probe.pools[p], 64)
    full = probe._batch_bytes(probe.pools[p], 256)
    budget = probe.pools[p].nbytes + one + 1000
    assert probe.pools[p].nbytes + full > budget
    srv = cpu_server(hbm_budget=budget)
    pid = srv.add_pool(pool)
    reqs = [srv.submit(pid, pool[i:i + 40]) for i in range(0, 320, 40)]
    with srv:
        for i, r in zip(range(0, 320, 40), reqs):
            got = r.wait(30)
            want = np.asarray([fp64.rmsd(pool, pool[i:i + 40], k) for k in range(40)])
            assert np.abs(got - want).max() < 1e-6
    assert srv.largest_batch <= 64
    assert srv.batches >= 5
    assert srv.inflight_bytes == 0


def test_budget_counts_one_request_batch(cpu_server):
    """add_pool refuses a pool that fits only without room for one request."""
    pool = gen.walk(seed=7, n_frames=300, n_atoms=40)
    probe = cpu_server(hbm_budget=1e12)
    p = probe.add_pool(pool)
    srv = cpu_server(hbm_budget=probe.pools[p].nbytes + 10)
    with pytest.raises(MemoryError, match="one 64-query batch"):
        srv.add_pool(pool)


def test_batch_work_cap_keeps_large_pool_batches_short(cpu_server):
    """A batch covers at most max_batch_work query x pool frame x atom
    products beyond its first request: on a large pool requests are batched
    few at a time, on a small pool up to max_batch_queries, and a request
    larger than the cap still runs alone."""
    big = gen.walk(seed=8, n_frames=200, n_atoms=30)     # 6,000 per query
    small = gen.walk(seed=9, n_frames=20, n_atoms=10)    # 200 per query
    srv = cpu_server(hbm_budget=1e9, max_batch_queries=128, max_batch_work=6000 * 20)
    pb, ps = srv.add_pool(big), srv.add_pool(small)
    reqs = [srv.submit(pb, big[i:i + 8]) for i in range(0, 80, 8)]      # 8 queries each
    reqs += [srv.submit(ps, small[:16]) for _ in range(6)]              # 16 queries each
    reqs.append(srv.submit(pb, big[:40]))                               # 40 > the cap alone
    sizes = []
    orig = srv._take_batch

    def spy(avail):
        pool, batch, nq = orig(avail)
        sizes.append((pool is srv.pools[pb], nq))
        return pool, batch, nq

    srv._take_batch = spy
    with srv:
        for r in reqs:
            r.wait(30)
    big_sizes = [nq for is_big, nq in sizes if is_big]
    small_sizes = [nq for is_big, nq in sizes if not is_big]
    assert max(nq for nq in big_sizes if nq != 40) <= 16   # 2 requests of 8 = 96,000 work
    assert 40 in big_sizes                                  # the oversized request, alone
    assert max(small_sizes) == 96                           # all six small requests at once
