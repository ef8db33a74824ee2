"""The hot-path subset of ``md.Trajectory``.

Mirrors ``core/trajectory.py`` for what the RMSD path touches (paths relative
to the mdtraj source tree): the coordinate container ``xyz`` (``:998-1029``,
float32, C-contiguous, resets the trace cache on assignment), the
``_rmsd_traces`` cache (``:1376-1384``), ``superpose`` (``:1083-1173``),
``center_coordinates`` (``:2114-2137``) and frame slicing that carries the
cache (``:1328-1359``).  Topology, I/O and the other analyses are outside this
package; ``topology`` is stored as given.

``xyz`` may be a numpy array (host), a ``DeviceArray`` or a float32 CUDA
tensor (device resident: the RMSD calls then read it in place from HBM, with
no transfer).
"""
from __future__ import annotations

import numpy as np

from . import _device as dev
from ._lib import call, load


class Trajectory:
    def __init__(self, xyz, topology=None, time=None, unitcell_lengths=None,
                 unitcell_angles=None):
        self.topology = topology
        self.xyz = xyz
        self._rmsd_traces = None
        self.time = time
        self.unitcell_lengths = unitcell_lengths
        self.unitcell_angles = unitcell_angles

    # ------------------------------------------------------------ container
    @property
    def top(self):
        return self.topology

    @property
    def xyz(self):
        return self._xyz

    @xyz.setter
    def xyz(self, value):
        """float32 C-contiguous (n_frames, n_atoms, 3); a 2-D array gains a
        frame axis (``ensure_type(..., add_newaxis_on_deficient_ndim=True)``)."""
        if dev.is_device_array(value):
            if value.dim() == 2:
                value = value.reshape(1, *value.shape)
            if value.dim() != 3 or value.shape[2] != 3:
                raise ValueError(f"xyz must be shape (n_frames, n_atoms, 3), got {tuple(value.shape)}")
            value = dev.to_device_f32(value)
        else:
            value = np.asarray(value)
            if value.ndim == 2:
                value = value[np.newaxis]
            if value.ndim != 3 or value.shape[2] != 3:
                raise ValueError(f"xyz must be shape (n_frames, n_atoms, 3), got {value.shape}")
            if value.dtype != np.float32 or not value.flags.c_contiguous:
                value = np.ascontiguousarray(value, dtype=np.float32)
        self._xyz = value
        self._rmsd_traces = None

    @property
    def n_frames(self):
        return int(self._xyz.shape[0])

    @property
    def n_atoms(self):
        return int(self._xyz.shape[1])

    def __len__(self):
        return self.n_frames

    def __getitem__(self, key):
        return self.slice(key)

    def slice(self, key, copy=True):
        """Frame slice carrying the trace cache (``core/trajectory.py:1328-1359``)."""
        xyz = self._xyz[key]
        if dev.is_device_array(xyz):
            xyz = xyz.clone() if copy else xyz
        elif copy:
            xyz = np.array(xyz, copy=True)
        new = Trajectory(xyz, self.topology)
        if self._rmsd_traces is not None:
            new._rmsd_traces = np.array(self._rmsd_traces[key], ndmin=1, copy=True)
        return new

    @property
    def is_on_device(self):
        return dev.is_device_array(self._xyz)

    def to_device(self):
        """Move the coordinates to HBM (in place); the trace cache is kept."""
        traces = self._rmsd_traces
        self._xyz = dev.to_device_f32(self._xyz)
        self._rmsd_traces = traces
        return self

    def to_host(self):
        traces = self._rmsd_traces
        if self.is_on_device:
            self._xyz = dev.to_host(self._xyz)
        self._rmsd_traces = traces
        return self

    # --------------------------------------------------------------- methods
    def center_coordinates(self, mass_weighted=False):
        """Centre every frame at the origin in place and cache the traces.

        ``mass_weighted=True`` needs atom masses from a topology, which this
        package does not carry: it is refused unless ``topology`` is None, in
        which case the reference also takes the unweighted path
        (``core/trajectory.py:2132-2135``).
        """
        from ._rmsd import _center_inplace_atom_major

        if mass_weighted and self.topology is not None:
            raise NotImplementedError("mass-weighted centering needs topology masses")
        self._rmsd_traces = _center_inplace_atom_major(self._xyz)
        return self

    def superpose(self, reference, frame=0, atom_indices=None, ref_atom_indices=None,
                  parallel=True):
        """Superpose every frame onto ``reference[frame]`` in place.

        Follows ``core/trajectory.py:1083-1173``: the fitted subset's centroid
        is removed from every atom of the frame, the frame is rotated by the
        optimal rotation (QCP, FP64 on the GPU) and moved to the reference
        subset's centroid.  ``parallel`` is accepted and ignored.
        """
        if atom_indices is not None and len(atom_indices) == 0:
            raise ValueError("Number of atom indices must be greater than 0")
        if ref_atom_indices is None:
            ref_atom_indices = atom_indices
        if atom_indices is None and ref_atom_indices is not None:
            raise ValueError("Number of atoms must be consistent!")
        if ref_atom_indices is not None and len(ref_atom_indices) != len(atom_indices):
            raise ValueError("Number of atoms must be consistent!")

        with dev.using(self._xyz):
            return self._superpose(reference, frame, atom_indices, ref_atom_indices)

    def _superpose(self, reference, frame, atom_indices, ref_atom_indices):
        lib = load()
        rxyz = reference.xyz if hasattr(reference, "xyz") else reference
        on_device = self.is_on_device
        # In place, as the reference (core/trajectory.py:1128 aliases
        # self.xyz): a device-resident float32 trajectory is transformed in
        # its own memory; the reference frame is gathered into the workspace
        # before any frame is moved, so superposing onto itself is safe.
        X = dev.to_device_f32(self._xyz)
        R = dev.to_device_f32(rxyz[frame])
        n_frames, n_atoms = int(X.shape[0]), int(X.shape[1])
        n_sel = n_atoms if atom_indices is None else len(atom_indices)
        a = None if atom_indices is None else dev.indices_on_host(atom_indices)
        r = None if ref_atom_indices is None else dev.indices_on_host(ref_atom_indices)
        if a is not None and a.size and (a.min() < 0 or a.max() >= n_atoms):
            raise IndexError("atom_indices out of range")
        if r is not None and r.size and (r.min() < 0 or r.max() >= int(R.shape[0])):
            raise IndexError("ref_atom_indices out of range")
        if atom_indices is None and int(R.shape[0]) != n_atoms:
            raise ValueError("Number of atoms must be consistent!")
        idx = dev.to_device_i32(a)
        ridx = dev.to_device_i32(r)
        if n_frames:
            ws = dev.workspace(lib.hr_superpose_workspace_bytes(n_frames, n_sel))
            call("hr_superpose", X.data_ptr(), n_frames, n_atoms, None, n_atoms, dev.ptr(idx),
                 n_sel, R.data_ptr(), dev.ptr(ridx), 1, None, None, ws.data_ptr(), dev.stream())
            # core/trajectory.py:1162-1169: an all-zero first frame means the
            # rotation overflowed or underflowed.
            if not dev.any_nonzero(X[0]):
                raise OverflowError(
                    "Encounted a potential overflow/underflow error during superpose() due to "
                    "the magnitude of your `_xyz` coordinates. To circumvent this, (1) reload "
                    "your trajectory and (2) divide and/or multiply your `_xyz` dataset by "
                    "multiples of 10 before running superpose again. Then, (3) revert that "
                    "multiplication/division after the calculations.")
        if on_device:
            self.xyz = X
        else:
            out = dev.to_host(X)
            a = self._xyz
            if (isinstance(a, np.ndarray) and a.dtype == np.float32 and a.flags.c_contiguous
                    and a.flags.writeable):
                a[...] = out  # the caller's array, as the reference modifies it
                self._rmsd_traces = None
            else:
                self.xyz = out
        return self
