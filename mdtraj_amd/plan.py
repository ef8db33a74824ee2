"""Reusable device-resident RMSD plans (no per-call validation or allocation).

``OneVsAll`` holds a device trajectory, its atom selection and a workspace;
each ``run`` scores every frame against one reference conformation with two
kernel launches (reference prep + per-frame kernel): on torch's current
stream for a tensor trajectory, else on the default stream.  This is what repeated calls (clustering loops, serving, benchmarks)
use instead of ``md.rmsd``, whose per-call work is host validation and
transfers.
"""
from __future__ import annotations

import numpy as np

from . import _device as dev
from ._lib import call, load


class OneVsAll:
    def __init__(self, target_xyz, atom_indices=None):
        with dev.using(target_xyz):
            self._init(target_xyz, atom_indices)

    def _init(self, target_xyz, atom_indices):
        lib = load()
        self.xyz = dev.to_device_f32(target_xyz)
        if self.xyz.dim() != 3 or self.xyz.shape[2] != 3:
            raise ValueError("target must be (n_frames, n_atoms, 3)")
        self.n_frames, self.n_atoms = int(self.xyz.shape[0]), int(self.xyz.shape[1])
        if atom_indices is not None:
            a = np.asarray(atom_indices)
            if a.ndim != 1 or a.size == 0 or a.min() < 0 or a.max() >= self.n_atoms:
                raise ValueError("atom_indices must be valid positive indices")
        self.idx = dev.to_device_i32(atom_indices)
        self.n_sel = self.n_atoms if atom_indices is None else len(atom_indices)
        self.ws = dev.workspace(lib.hr_rmsd_workspace_bytes(self.n_sel))
        self.out = dev.empty_f32(self.n_frames)
        self._ref_idx = None

    def set_reference(self, ref_frame_xyz, ref_atom_indices=None):
        """Gather and sum the reference conformation (N_ref, 3) on the device."""
        with dev.using(self.xyz):
            return self._set_reference(ref_frame_xyz, ref_atom_indices)

    def _set_reference(self, ref_frame_xyz, ref_atom_indices):
        R = dev.to_device_f32(ref_frame_xyz)
        if ref_atom_indices is None and self.idx is not None:
            ref_atom_indices = self.idx
        self._ref_idx = dev.to_device_i32(ref_atom_indices)
        self._ref = R
        call("hr_rmsd_prepare_reference", R.data_ptr(), dev.ptr(self._ref_idx), self.n_sel,
             self.ws.data_ptr(), dev.stream())
        return self

    def set_reference_frame(self, frame):
        """Use frame `frame` of the plan's own trajectory (same selection)."""
        self._self_frame = int(frame)
        return self.set_reference(self.xyz[frame], self.idx)

    def run(self, out=None, self_frame=None):
        """Per-frame kernel only (the reference must be prepared).  Returns
        the float32 device tensor of RMSDs."""
        o = self.out if out is None else out
        sf = getattr(self, "_self_frame", -1) if self_frame is None else self_frame
        with dev.using(self.xyz):
            call("hr_rmsd_one_vs_all_prepared", self.xyz.data_ptr(), self.n_frames, self.n_atoms,
                 dev.ptr(self.idx), self.n_sel, sf, o.data_ptr(), self.ws.data_ptr(),
                 dev.stream())
        return o

    def __call__(self, frame):
        self.set_reference_frame(frame)
        return self.run()
