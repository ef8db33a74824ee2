"""Autotuner for the batched plan's work split and tile shape.

Candidates are (frames per work unit, KiB cap per unit, waves per SIMD of
the scoring kernel, schedule).  Schedules 3 and 4 (one cost-balanced range
per workgroup, the reference in LDS; 4: one workgroup per CU whose waves
share each system's frames through an LDS counter) have no unit shape.  Under schedule 1
(work units from a ticket) every unit is scored by one wave, whose
bandwidth is bounded by the loads it keeps in flight, so the batch ends with
the largest units still running: small units (and a byte cap, which
shortens the units of the largest systems) shorten that tail, large units
amortise the per-unit reference reads and QCP; 4 waves per SIMD keeps more
loads in flight at the cost of a few spilled registers.  Each candidate's
plan is built and timed on the actual batch (graph launches, HIP events);
the winner is cached per batch signature (system count and log2 bins of
frame and atom counts) in memory and in a JSON file
(``$MDTRAJ_AMD_TUNE_CACHE``, default ``~/.cache/mdtraj_amd/autotune.json``).
"""
from __future__ import annotations

import json
import math
import os

from . import _device as dev

CANDIDATES = ([(64, 0, 3, 3), (64, 0, 3, 4)]
              + [(fpu, kib, 3, 1) for fpu in (8, 16, 32, 64) for kib in (0, 64, 192)]
              + [(16, 0, 4, 1), (8, 64, 4, 1)])
_memory = {}


def _cache_path():
    return os.environ.get("MDTRAJ_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "mdtraj_amd",
                                       "autotune.json"))


def signature(xyz_list):
    bins = {}
    for x in xyz_list:
        k = (int(math.log2(max(1, int(x.shape[0])))), int(math.log2(int(x.shape[1]))))
        bins[k] = bins.get(k, 0) + 1
    return f"batch-v3:{len(xyz_list)}:" + ",".join(f"{a}.{b}x{c}" for (a, b), c in sorted(bins.items()))


def _load_disk():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except Exception:
        return {}


def _store_disk(key, value):
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        data = _load_disk()
        data[key] = value
        with open(path, "w") as f:
            json.dump(data, f, indent=1)
    except OSError:
        pass


def time_plan(batch, reps=5):
    """Mean milliseconds of one graph launch of `batch`'s plan."""
    t = dev.torch()
    for _ in range(2):
        batch.run()
    t.cuda.synchronize()
    s, e = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        batch.run()
    e.record()
    t.cuda.synchronize()
    return s.elapsed_time(e) / reps


def tune_batch(xyz_list, ref_frames=None, candidates=CANDIDATES, reps=5):
    """Time every candidate; returns (best (fpu, kib, waves, schedule),
    {candidate: ms})."""
    from .batch import BatchRMSD

    results = {}
    for fpu, kib, waves, schedule in candidates:
        b = BatchRMSD(xyz_list, ref_frames, frames_per_unit=fpu, unit_kib=kib, waves=waves,
                      schedule=schedule)
        results[(fpu, kib, waves, schedule)] = time_plan(b, reps)
        b.close()
    best = min(results, key=results.get)
    return best, results


def tuned_shape(xyz_list, ref_frames=None):
    key = signature(xyz_list)
    if key in _memory:
        return _memory[key]
    disk = _load_disk().get(key)
    if disk is not None:
        _memory[key] = tuple(disk)
        return _memory[key]
    best, _ = tune_batch(xyz_list, ref_frames)
    _memory[key] = best
    _store_disk(key, list(best))
    return best
