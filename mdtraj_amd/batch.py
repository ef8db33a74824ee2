"""Batched one-vs-all RMSD over many independent systems in one graph launch.

The reference scores one trajectory per ``md.rmsd`` call; a batch of uneven
systems (tens to thousands of atoms and frames each) is a loop of small calls
there.  Here a plan (``hr_batch_plan_create``) captures a hipGraph whose
scoring kernel covers every system's frames (by default one cost-balanced
range per workgroup, each system's reference staged in LDS), so a whole
batch is one graph launch with no per-system launch cost.
"""
from __future__ import annotations

import ctypes
import threading

from . import _device as dev
from ._lib import HR_TUNE_BATCH_SCHEDULE, BatchShape, SystemDesc, call, load

_SCHEDULE_LOCK = threading.Lock()


class BatchRMSD:
    """RMSD of every frame of each system to that system's ``ref_frames[s]``.

    ``systems``: list of float32 (F_s, N_s, 3) arrays or CUDA tensors (moved
    to the device once).  ``schedule``: the work split (``HR_TUNE_BATCH_SCHEDULE``:
    3 one cost-balanced range per workgroup with the reference in LDS, 4 the
    same with one workgroup per CU whose waves share each system's frames, 1
    work units from a ticket, 2 one range per wave).  Work-unit shape of schedule
    1: ``frames_per_unit`` (1..64) and ``unit_kib`` (if > 0, fewer frames per
    unit, a multiple of 4, where frames_per_unit frames would exceed that
    many KiB); ``waves``: 3 or 4 waves per SIMD of the scoring kernel.  None
    picks the autotuned value for this batch (``mdtraj_amd.autotune``, which
    times with torch events).
    ``run()`` is one graph launch and returns the list of float32 device
    arrays of RMSDs: CUDA tensors on torch's current stream for tensor
    systems, DeviceArrays on the default stream otherwise.
    """

    def __init__(self, systems, ref_frames=None, frames_per_unit=None, waves=None,
                 unit_kib=None, schedule=None):
        load()
        with dev.using(*systems):
            self._init(systems, ref_frames, frames_per_unit, waves, unit_kib, schedule)

    def _init(self, systems, ref_frames, frames_per_unit, waves, unit_kib, schedule):
        self.xyz = [dev.to_device_f32(x) for x in systems]
        for i, x in enumerate(self.xyz):
            if x.dim() != 3 or x.shape[2] != 3:
                raise ValueError(f"system {i}: xyz must be (n_frames, n_atoms, 3)")
        n = len(self.xyz)
        if n == 0:
            raise ValueError("no systems")
        self.ref_frames = [0] * n if ref_frames is None else [int(r) for r in ref_frames]
        self.out = [dev.empty_f32(max(int(x.shape[0]), 1)) for x in self.xyz]
        # per-system result views, made once: run() only launches the graph
        self._views = [o[: int(x.shape[0])] for o, x in zip(self.out, self.xyz)]
        if frames_per_unit is None or waves is None or unit_kib is None:
            from .autotune import tuned_shape

            t_fpu, t_kib, t_waves, t_sched = tuned_shape(self.xyz, self.ref_frames)
            frames_per_unit = t_fpu if frames_per_unit is None else frames_per_unit
            unit_kib = t_kib if unit_kib is None else unit_kib
            waves = t_waves if waves is None else waves
            schedule = t_sched if schedule is None else schedule
        self.frames_per_unit = int(frames_per_unit)
        self.unit_kib = int(unit_kib)
        self.waves = int(waves)
        # None: the process-wide HR_TUNE_BATCH_SCHEDULE as it stands
        self.schedule = None if schedule is None else int(schedule)
        self._plan = None
        self._build()

    def _descs(self):
        arr = (SystemDesc * len(self.xyz))()
        for i, (x, o) in enumerate(zip(self.xyz, self.out)):
            arr[i] = SystemDesc(x.data_ptr(), int(x.shape[0]), int(x.shape[1]),
                                self.ref_frames[i], o.data_ptr())
        return arr

    def _build(self):
        plan = ctypes.c_void_p()
        descs = self._descs()
        shape = BatchShape(self.frames_per_unit, self.unit_kib, self.waves)
        args = (ctypes.addressof(descs), len(self.xyz), ctypes.addressof(shape),
                ctypes.byref(plan))
        if self.schedule is None:
            call("hr_batch_plan_create", *args)
        else:
            with _SCHEDULE_LOCK:
                # the work split is a process-wide knob read at plan creation
                prev = int(load().hr_get_tuning(HR_TUNE_BATCH_SCHEDULE))
                call("hr_set_tuning", HR_TUNE_BATCH_SCHEDULE, self.schedule)
                try:
                    call("hr_batch_plan_create", *args)
                finally:
                    call("hr_set_tuning", HR_TUNE_BATCH_SCHEDULE, prev)
        self._plan = plan

    @property
    def n_units(self):
        return int(load().hr_batch_plan_units(self._plan))

    @property
    def total_frames(self):
        return sum(int(x.shape[0]) for x in self.xyz)

    @property
    def total_bytes(self):
        return sum(12 * int(x.shape[0]) * int(x.shape[1]) for x in self.xyz)

    def run(self):
        """One graph launch; returns the per-system float32 device views of
        the results (the same arrays every call)."""
        with dev.using(self.xyz[0]):
            call("hr_batch_plan_launch", self._plan, dev.stream())
        return list(self._views)

    def close(self):
        if self._plan is not None and self._plan.value:
            load().hr_batch_plan_destroy(self._plan)
        self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
