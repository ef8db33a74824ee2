"""Serving one-vs-pool RMSD queries against device-resident trajectory pools.

Pools are uploaded once and kept in HBM in prepared split form
(``allpairs.SplitSet``: centred, hi/lo planes, traces).  A request carries 1 to
``max_request_queries`` (64) query frames with the same atom count as its pool
and asks for the RMSD of every query frame to every pool frame.

A failure in the scheduler or completion thread (for example a failed
allocation) fails every outstanding request: ``Request.wait`` raises it.

Continuous batching: a scheduler thread takes the oldest pending request,
gathers every other pending request for the same pool up to
``max_batch_queries`` query frames (fewer on large pools, under
``max_batch_work``), and launches the batch (upload from a pinned staging
buffer, split prep, one cross-Gram pass over the pool, copy of the results
to pinned host memory) on the next stream of a small ring as soon as a slot
is free and the HBM budget admits its buffers.  ``start()`` allocates what
the largest batch needs before serving (pinned buffers, device memory on
each stream) and freezes the objects alive at that point out of Python's
garbage collection: either, left to happen while serving, held the
scheduler for tens of milliseconds.  Requests that arrive while earlier batches
are on the GPU join later batches; batches on different streams overlap on
the device.  A completion thread retires batches as their events complete,
in whatever order that is, and hands every request its rows.  Each (query, pool frame) result depends only on those two
frames (``hr_rmsd_cross``), so a request's answer is bit-identical whether
it was scored alone or batched with others.
"""
from __future__ import annotations

import collections
import gc
import threading
import time

import numpy as np

from . import _device as dev
from ._lib import load
from .allpairs import MODES, SplitSet, rmsd_cross


class Request:
    __slots__ = ("pool", "queries", "t_arrival", "t_done", "result", "error", "_done")

    def __init__(self, pool, queries, t_arrival):
        self.pool = pool
        self.queries = queries
        self.t_arrival = t_arrival
        self.t_done = None
        self.result = None
        self.error = None
        self._done = threading.Event()

    def wait(self, timeout=None):
        if not self._done.wait(timeout):
            raise TimeoutError("request not finished")
        if self.error is not None:
            raise RuntimeError("RMSD server failed") from self.error
        return self.result

    def _fail(self, exc):
        self.error = exc
        self._done.set()

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_arrival


class Pool:
    """A pool prepared on the caller's current stream.  ``ready`` is an event
    recorded after the preparation: batches run on the server's own streams,
    which are not ordered with the caller's, and wait on it before they read
    the prepared block."""

    def __init__(self, xyz, mode, torch_mod):
        self.n_frames, self.n_atoms = int(xyz.shape[0]), int(xyz.shape[1])
        self.split = SplitSet(xyz, mode=mode)
        self.nbytes = int(self.split.nbytes)
        self.ready = torch_mod.cuda.Event()
        self.ready.record()


class RMSDServer:
    """hbm_budget: bytes of HBM the server may hold (pools + batches in flight).
    max_request_queries: query frames one request may carry.
    max_batch_queries: query frames per batch; a batch reads its pool once
    whatever its size (one MFMA row tile per 64 query frames), so under load
    larger batches cost less per query.
    max_batch_work: query frames x pool frames x atoms one batch may cover
    beyond its first request (default 2.6e10, about 2 ms of split-FP16 MFMA
    work on MI355X): a batch's requests wait for all of its rows, so batches
    on the largest pools are kept short while those on small pools still
    fill up to max_batch_queries.
    n_streams: batches that may be in flight at once.
    keep_results: copy every request's rows out of the reused pinned buffer
    into its own array (False: only the arrival-to-rows-on-host latency is
    recorded, for clients that consume the rows as they land)."""

    def __init__(self, hbm_budget=200e9, mode="f16exact", max_batch_queries=256,
                 max_request_queries=64, n_streams=4, keep_results=True,
                 max_batch_work=2.6e10):
        self.t = dev.require_torch_gpu()
        self.mode = mode
        self.budget = int(hbm_budget)
        self.max_q = int(max_batch_queries)
        self.max_work = float(max_batch_work)
        self.max_req_q = int(max_request_queries)
        if not 1 <= self.max_req_q <= self.max_q:
            raise ValueError("need 1 <= max_request_queries <= max_batch_queries")
        self.keep_results = bool(keep_results)
        self.largest_batch = 0
        self.pools = []
        self.resident = 0
        self.inflight_bytes = 0
        self.streams = [self.t.cuda.Stream() for _ in range(n_streams)]
        self.free_streams = collections.deque(range(n_streams))
        # one pinned result buffer per stream slot, reused by every batch on
        # that slot (the slot is free again only after its rows were copied out)
        self.pinned = [None] * n_streams
        # and one pinned staging buffer per slot for the batch's query frames
        self.pinned_in = [None] * n_streams
        self.error = None
        self.pending = collections.deque()
        self.cv = threading.Condition()
        self.completions = collections.deque()
        self.done_cv = threading.Condition()
        self.running = False
        self.completed = []
        self.pairs_done = 0
        self.batches = 0

    # ---------------------------------------------------------------- pools
    def add_pool(self, xyz):
        """Prepare a pool (float32 (F, N, 3), host or device) in HBM; returns its id.
        Raises MemoryError when the HBM budget cannot hold it together with
        one batch of ``max_request_queries`` query frames against it (so that
        every request it may receive can run once nothing else is in flight)."""
        with dev.torch_buffers():
            pool = Pool(dev.to_device_f32(xyz), self.mode, self.t)
        with self.cv:
            one = self._batch_bytes(pool, self.max_req_q)
            if self.resident + pool.nbytes + one > self.budget:
                raise MemoryError(f"pool of {pool.nbytes} bytes plus one {self.max_req_q}-query "
                                  f"batch ({one} bytes) exceeds the HBM budget")
            self.resident += pool.nbytes
            self.pools.append(pool)
            return len(self.pools) - 1

    # -------------------------------------------------------------- requests
    def submit(self, pool_id, queries, t_arrival=None):
        if self.error is not None:
            raise RuntimeError("RMSD server failed") from self.error
        q = np.ascontiguousarray(queries, dtype=np.float32)
        if q.ndim == 2:
            q = q[None]
        p = self.pools[pool_id]
        if q.shape[1] != p.n_atoms or q.shape[2] != 3:
            raise ValueError(f"queries must be (q, {p.n_atoms}, 3), got {q.shape}")
        if not 1 <= q.shape[0] <= self.max_req_q:
            raise ValueError(f"a request carries 1..{self.max_req_q} query frames")
        r = Request(pool_id, q, time.perf_counter() if t_arrival is None else t_arrival)
        with self.cv:
            self.pending.append(r)
            self.cv.notify_all()
        return r

    def _batch_bytes(self, pool, nq):
        """HBM a batch of nq query frames holds while in flight: the uploaded
        queries, their prepared split block and the float32 result rows."""
        split = load().hr_split_bytes(nq, pool.n_atoms, MODES[self.mode])
        return 12 * nq * pool.n_atoms + int(split) + 4 * nq * pool.n_frames

    def _take_batch(self, avail):
        """Oldest pending request plus later ones for the same pool (FIFO
        within the pool), up to max_q query frames, max_work and `avail`
        bytes of HBM in flight.  The oldest request is always taken (it must
        fit `avail`)."""
        first = self.pending[0]
        pool = self.pools[first.pool]
        batch, nq = [], 0
        keep = collections.deque()
        while self.pending:
            r = self.pending.popleft()
            q = r.queries.shape[0]
            if (r.pool == first.pool and nq + q <= self.max_q
                    and (not batch or (nq + q <= self._max_batch_queries(pool)
                                       and self._batch_bytes(pool, nq + q) <= avail))):
                batch.append(r)
                nq += q
            else:
                keep.append(r)
        self.pending = keep
        return pool, batch, nq

    def _wait_for_launch(self):
        """Under self.cv: wait until the oldest pending request can be launched
        (a free stream slot and HBM for it); False when stopped with nothing
        pending.  Batches are sized by what they hold, so an admitted request
        never waits for memory that no completion will free."""
        while True:
            if not self.pending:
                if not self.running:
                    return False
                self.cv.wait()
                continue
            if not self.free_streams:
                self.cv.wait()
                continue
            first = self.pending[0]
            need = self._batch_bytes(self.pools[first.pool], first.queries.shape[0])
            if self.resident + self.inflight_bytes + need <= self.budget:
                return True
            if self.inflight_bytes > 0:
                self.cv.wait()     # a completion frees memory
                continue
            # cannot happen for pools admitted by add_pool; never wait forever
            self.pending.popleft()
            first._fail(MemoryError(f"a batch of {first.queries.shape[0]} query frames "
                                    f"({need} bytes) does not fit the HBM budget"))

    def _scheduler(self):
        try:
            while True:
                with self.cv:
                    if not self._wait_for_launch():
                        return
                    avail = self.budget - self.resident - self.inflight_bytes
                    pool, batch, nq = self._take_batch(avail)
                    need = self._batch_bytes(pool, nq)
                    self.largest_batch = max(self.largest_batch, nq)
                    si = self.free_streams.popleft()
                    self.inflight_bytes += need
                try:
                    self._launch(pool, batch, nq, si, need)
                except BaseException:
                    for r in batch:
                        r._fail(RuntimeError("batch launch failed"))
                    raise
        except BaseException as exc:  # noqa: BLE001 - handed to every waiter
            self._fail_all(exc)

    def _max_batch_queries(self, pool):
        """Query frames a batch on `pool` may grow to past its first request:
        max_q, or fewer under max_batch_work."""
        by_work = int(self.max_work // (float(pool.n_frames) * pool.n_atoms))
        return min(self.max_q, by_work)

    def _largest_batch(self, pool):
        """The most query frames a batch on `pool` can hold: its cap, or one
        request of max_request_queries, which is always admitted."""
        return max(self.max_req_q, self._max_batch_queries(pool))

    def _reserve(self):
        """Allocate, before serving, what the largest batch on any pool needs:
        each stream slot's pinned result buffer, and device memory for the
        batches in flight (kept by torch's caching allocator).  Allocated
        lazily, a pinned buffer of ~190 MB for a slot's first batch on a
        large pool held the scheduler for tens of ms, and every request that
        arrived meanwhile queued behind it."""
        if not self.pools:
            return
        t = self.t
        rows = max(self._largest_batch(p) * p.n_frames for p in self.pools)
        for si in range(len(self.pinned)):
            if self.pinned[si] is None or self.pinned[si].numel() < rows:
                self.pinned[si] = t.empty(rows, dtype=t.float32, pin_memory=True)
        # device memory: torch's caching allocator keeps a freed block for
        # the stream it was used on, so each slot reserves on its own stream
        one = max(self._batch_bytes(p, self._largest_batch(p)) for p in self.pools)
        room = max(0, self.budget - self.resident)
        qbytes = max(12 * self._largest_batch(p) * p.n_atoms for p in self.pools)
        for si, stream in enumerate(self.streams):
            if self.pinned_in[si] is None or self.pinned_in[si].numel() < qbytes // 4:
                self.pinned_in[si] = t.empty(qbytes // 4, dtype=t.float32, pin_memory=True)
            if one <= room:
                room -= one
                with t.cuda.stream(stream), dev.torch_buffers():
                    block = t.empty(one, dtype=t.uint8, device=dev.device())
                    del block
        t.cuda.synchronize()

    def _pinned_rows(self, si, nq, nf):
        """(nq, nf) view of stream slot si's pinned result buffer."""
        t = self.t
        need = nq * nf
        buf = self.pinned[si]
        if buf is None or buf.numel() < need:  # a pool added after start()
            buf = t.empty(max(need, self.max_q * nf), dtype=t.float32, pin_memory=True)
            self.pinned[si] = buf
        return buf[:need].view(nq, nf)

    def _launch(self, pool, batch, nq, si, need):
        t = self.t
        stream = self.streams[si]
        n_q = sum(r.queries.shape[0] for r in batch)
        staged = self.pinned_in[si]
        if staged is None or staged.numel() < n_q * pool.n_atoms * 3:  # a pool added after start()
            staged = t.empty(n_q * pool.n_atoms * 3, dtype=t.float32, pin_memory=True)
            self.pinned_in[si] = staged
        host_q = staged[: n_q * pool.n_atoms * 3].view(n_q, pool.n_atoms, 3)
        np.concatenate([r.queries for r in batch], axis=0, out=host_q.numpy())
        stream.wait_event(pool.ready)
        with t.cuda.stream(stream), dev.torch_buffers():
            # the slot's staging buffer is free: its previous batch completed
            dq = host_q.to(device=dev.device(), non_blocking=True)
            qs = SplitSet(dq, mode=self.mode)
            out = rmsd_cross(qs, pool.split)
            host_out = self._pinned_rows(si, nq, pool.n_frames)
            host_out.copy_(out, non_blocking=True)
            ev = t.cuda.Event()
            ev.record(stream)
        with self.done_cv:
            # keep device buffers alive until the event has passed
            self.completions.append((ev, batch, host_out, (dq, qs, out), si, need, nq,
                                     pool.n_frames))
            self.done_cv.notify_all()

    def _completer(self):
        # batches on different streams finish out of launch order: retire
        # whichever has finished, so a small batch never waits behind a
        # larger one launched before it
        try:
            while True:
                with self.done_cv:
                    while not self.completions and (self.running or self._sched.is_alive()):
                        self.done_cv.wait(timeout=0.05)
                    if not self.completions:
                        if not self.running and not self._sched.is_alive():
                            return
                        continue
                    item = None
                    for i, c in enumerate(self.completions):
                        if c[0].query():
                            item = c
                            del self.completions[i]
                            break
                if item is None:
                    time.sleep(2e-5)
                    continue
                ev, batch, host_out, keep, si, need, nq, nf = item
                now = time.perf_counter()
                res = host_out.numpy()
                row = 0
                for r in batch:
                    q = r.queries.shape[0]
                    if self.keep_results:
                        r.result = res[row:row + q].copy()   # the pinned buffer is reused
                    row += q
                    r.t_done = now
                    r._done.set()
                del keep
                with self.cv:
                    self.free_streams.append(si)
                    self.inflight_bytes -= need
                    self.pairs_done += nq * nf
                    self.batches += 1
                    self.completed.extend(batch)
                    self.cv.notify_all()
        except BaseException as exc:  # noqa: BLE001 - handed to every waiter
            self._fail_all(exc)

    def _fail_all(self, exc):
        """Stop serving and fail every request not yet answered."""
        with self.cv:
            self.error = exc
            self.running = False
            pending = list(self.pending)
            self.pending.clear()
            self.cv.notify_all()
        with self.done_cv:
            inflight = [r for c in self.completions for r in c[1]]
            self.completions.clear()
            self.done_cv.notify_all()
        for r in pending + inflight:
            if not r._done.is_set():
                r._fail(exc)

    # ------------------------------------------------------------- lifecycle
    def start(self):
        self._reserve()
        # Python's oldest-generation collection walks every live object (with
        # torch loaded, hundreds of thousands): 35-50 ms in which neither
        # thread launches or retires a batch.  Objects that exist now (the
        # process's modules, the pools) are moved out of later collections.
        gc.freeze()
        self.running = True
        self._sched = threading.Thread(target=self._scheduler, daemon=True)
        self._comp = threading.Thread(target=self._completer, daemon=True)
        self._sched.start()
        self._comp.start()
        return self

    def stop(self):
        """Finish everything pending, then stop the threads."""
        with self.cv:
            self.running = False
            self.cv.notify_all()
        self._sched.join()
        with self.done_cv:
            self.done_cv.notify_all()
        self._comp.join()

    def __enter__(self):
        return self.start()

    def __exit__(self, *exc):
        self.stop()

    def latency_percentiles(self, ps=(50, 99)):
        lat = np.array([r.latency for r in self.completed if r.latency is not None])
        return {p: float(np.percentile(lat, p)) if lat.size else None for p in ps}
