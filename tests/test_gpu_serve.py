"""Serving: results equal the FP64 oracle, a request's answer is bit-identical
alone and in a mix of concurrent requests, and the budget refuses pools that
do not fit."""
import time

import numpy as np
import pytest

from mdtraj_amd import gen
from mdtraj_amd.serve import RMSDServer
from oracle import fp64

pytestmark = pytest.mark.gpu


def make_pools():
    return [gen.walk(seed=10 + i, n_frames=f, n_atoms=n)
            for i, (f, n) in enumerate([(300, 120), (1000, 40), (90, 333)])]


def test_serve_matches_fp64_and_is_batch_independent(gpu):
    pools = make_pools()
    rng = np.random.default_rng(0)
    with RMSDServer(hbm_budget=8e9) as srv:
        ids = [srv.add_pool(p) for p in pools]
        # alone: one request at a time
        probes = []
        for k in range(6):
            pid = int(rng.integers(0, len(pools)))
            q = int(rng.integers(1, 9))
            sel = rng.integers(0, pools[pid].shape[0], size=q)
            queries = pools[pid][sel] + rng.normal(0, 0.01, (q,) + pools[pid].shape[1:]).astype(np.float32)
            r = srv.submit(ids[pid], queries)
            probes.append((pid, queries, r.wait(60).copy()))
    # in a mix: the same requests among many others, all queued before a
    # second server starts, so its batches merge requests (and row tiles)
    srv = RMSDServer(hbm_budget=8e9)
    ids = [srv.add_pool(p) for p in pools]
    mixed = []
    for pid, queries, _ in probes:
        for _ in range(5):
            opid = int(rng.integers(0, len(pools)))
            oq = int(rng.integers(1, 33))
            srv.submit(ids[opid], pools[opid][rng.integers(0, pools[opid].shape[0], oq)])
        mixed.append(srv.submit(ids[pid], queries))
    with srv:
        for (pid, queries, alone), r in zip(probes, mixed):
            got = r.wait(60)
            assert np.array_equal(got, alone)
            want = fp64.rmsd_matrix(np.concatenate([queries, pools[pid]]))[:len(queries), len(queries):]
            assert np.abs(got - want).max() < 1e-5
        assert srv.batches >= 2
        assert srv.largest_batch > 64   # requests merged across MFMA row tiles


def test_serve_budget(gpu):
    srv = RMSDServer(hbm_budget=1e6)
    with pytest.raises(MemoryError):
        srv.add_pool(gen.walk(seed=1, n_frames=500, n_atoms=300))


def test_serve_failure_reaches_every_waiter(gpu, monkeypatch):
    """A failing batch (forced from the host: the cross-Gram call raises)
    must fail the waiting requests, not leave them hanging."""
    from mdtraj_amd import serve

    def boom(*a, **k):
        raise MemoryError("forced failure")

    pool = gen.walk(seed=3, n_frames=50, n_atoms=20)
    with RMSDServer(hbm_budget=1e9) as srv:
        pid = srv.add_pool(pool)
        monkeypatch.setattr(serve, "rmsd_cross", boom)
        reqs = [srv.submit(pid, pool[:2]) for _ in range(5)]
        for r in reqs:
            with pytest.raises(RuntimeError):
                r.wait(30)
        with pytest.raises(RuntimeError):
            srv.submit(pid, pool[:1])
