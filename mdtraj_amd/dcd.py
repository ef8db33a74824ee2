"""DCD trajectories decoded on the GPU.

``DCDTrajectoryFile`` keeps the reading half of the reference's class
(formats/dcd/dcd.pyx:174-551): ``read(n_frames, stride, atom_indices)``
returning ``(xyz, cell_lengths, cell_angles)`` in angstroms and degrees, with
both cell arrays None when every length is below 1e-10; ``read_as_traj``;
``seek``/``tell``/``len``; ``distance_unit``.  ``load_dcd`` builds a
Trajectory in nanometres (dcd.pyx:102-171).

The file is read into host memory once; ``hr_dcd_scan`` parses its header on
the host; the bytes are copied to HBM once, and every ``read`` turns the
requested frames' axis-major records into atom-major coordinates on the
device (``hr_dcd_decode``), gathering ``atom_indices`` on the way.  Results
are bit-identical to the reference's reader.  As in the reference, a frame
whose records are malformed ends the read like the end of the file (a
warning names it here).  ``read(device=True)`` returns the coordinates as a
device tensor.  Writing DCD files is not provided.
"""
from __future__ import annotations

import ctypes
import os
import warnings

import numpy as np

from . import _device as dev
from ._lib import DcdHeader, call, load

_STATUS = {1: "a record length marker does not match the header",
           2: "the frame lies outside the file"}
_M_PI_2 = 1.57079632679489661922
_ANGSTROM_TO_NM = np.float32(0.1)


def cell_lengths_angles(cell):
    """(lengths, angles) float32 from raw cells (F, 6) = (A, cos gamma, B,
    cos beta, cos alpha, C): angles from the cosines when all three lie in
    [-1, 1], otherwise taken as degrees (dcdplugin.c:990-1010)."""
    c = np.asarray(cell, np.float32).reshape(-1, 6)
    lengths = c[:, [0, 2, 5]].copy()
    d = c.astype(np.float64)
    cos = d[:, [4, 3, 1]]  # alpha, beta, gamma
    use = np.all((cos >= -1.0) & (cos <= 1.0), axis=1, keepdims=True)
    with np.errstate(invalid="ignore"):
        from_cos = 90.0 - np.arcsin(np.where(use, cos, 0.0)) * 90.0 / _M_PI_2
    return lengths, np.where(use, from_cos, cos).astype(np.float32)


class DCDTrajectoryFile:
    """Reader for DCD files (mode 'r')."""

    distance_unit = "angstroms"

    def __init__(self, filename, mode="r", force_overwrite=True):
        if mode != "r":
            raise ValueError("only mode='r' is supported: this reader decodes DCD files "
                             "on the GPU and does not write them")
        self.filename = os.fspath(filename)
        if not os.path.exists(self.filename):
            raise IOError(f"The file '{self.filename}' doesn't exist")
        self._host = np.fromfile(self.filename, dtype=np.uint8)
        self._dev = None
        self._dev_slot = None
        load()
        h = DcdHeader()
        buf = self._host.ctypes.data_as(ctypes.c_void_p)
        try:
            call("hr_dcd_scan", buf, self._host.size, ctypes.byref(h), None, 0)
        except Exception as e:
            raise IOError(f"Could not open file: {self.filename} ({e})") from None
        self._free = None
        if h.n_fixed:
            free = np.zeros(h.n_atoms - h.n_fixed, np.int32)
            call("hr_dcd_scan", buf, self._host.size, ctypes.byref(h),
                 free.ctypes.data_as(ctypes.c_void_p), free.size)
            self._free = free
        self._header = h
        self.n_atoms = int(h.n_atoms)
        self.frame_counter = 0
        self.is_open = True
        self.mode = "r"

    # -- reference API -------------------------------------------------------
    def __len__(self):
        if not self.is_open:
            raise RuntimeError("I/O operation on closed file")
        return int(self._header.n_frames)

    def tell(self):
        return int(self.frame_counter)

    def seek(self, offset, whence=0):
        """Move to frame ``offset`` from the start (0) or relative to the
        current frame (1), as dcd.pyx:302-347; positions past the end stop at
        the end; offsets from the end are not supported (as the reference)."""
        n, cur = len(self), self.tell()
        if whence == 0 and offset >= 0:
            pos = offset
        elif whence == 1 and offset >= 0:
            pos = cur + offset
        elif whence == 1 and offset < 0:
            pos = max(0, cur + offset)
        elif whence == 2 and offset <= 0:
            raise NotImplementedError("offsets from the end are not supported yet")
        else:
            raise IOError("Invalid argument")
        self.frame_counter = int(min(pos, n))

    def close(self):
        self._dev = None
        self._dev_slot = None
        self.is_open = False

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def read(self, n_frames=None, stride=None, atom_indices=None, device=False):
        """Read frames from the current position: every ``stride``-th, at
        most ``n_frames``; returns ``(xyz, cell_lengths, cell_angles)`` in
        angstroms and degrees."""
        xyz, lengths, angles = self._read(n_frames, stride, atom_indices, np.float32(1.0))
        return (xyz if device else dev.to_host(xyz)), lengths, angles

    def read_as_traj(self, topology=None, n_frames=None, stride=None, atom_indices=None):
        """A Trajectory of the next frames, coordinates and cell lengths in
        nanometres on the device (dcd.pyx:381-423)."""
        from .trajectory import Trajectory

        initial = self.tell()
        xyz, lengths, angles = self._read(n_frames, stride, atom_indices, _ANGSTROM_TO_NM)
        if lengths is not None:
            lengths = lengths * _ANGSTROM_TO_NM
        step = 1 if stride is None else stride
        time = step * np.arange(xyz.shape[0]) + initial
        return Trajectory(xyz, topology, time=time, unitcell_lengths=lengths,
                          unitcell_angles=angles)

    # -- device decode -------------------------------------------------------
    def _read(self, n_frames, stride, atom_indices, scale):
        if not self.is_open:
            raise IOError("file is not open")
        count = len(self) - self.tell() if n_frames is None else int(n_frames)
        if count < 0:
            raise ValueError("negative dimensions are not allowed")
        sel = None
        if atom_indices is None:
            n_sel = self.n_atoms
        elif isinstance(atom_indices, slice):
            sel = np.arange(self.n_atoms)[atom_indices]
            n_sel = len(sel)
        else:
            sel = np.asarray(atom_indices)
            if sel.size and sel.min() < 0:
                raise ValueError("atom_indices should be zero indexed. you gave an index less than zero")
            if sel.size and sel.max() >= self.n_atoms:
                raise ValueError("atom indices should be zero indexed. you gave an index bigger "
                                 "than the number of atoms")
            n_sel = sel.size
        step = 1 if stride is None else max(1, int(stride))
        frames = np.arange(self.tell(), len(self), step, dtype=np.int64)[:count]
        with dev.using_loaded():
            return self._decode(frames, sel, n_sel, step, scale)

    def _decode(self, frames, sel, n_sel, step, scale):
        F = int(frames.size)
        xyz = dev.empty((F, n_sel, 3), np.float32)
        cell = np.zeros((0, 6), np.float32)
        if F and n_sel:
            data, slot = self._device_bytes()
            idx = None if sel is None else dev.to_device_i32(sel.astype(np.int32))
            fr = dev.to_device(frames, np.int64)
            cell_d = dev.empty((F, 6), np.float32)
            status = dev.empty(F, np.int32)
            call("hr_dcd_decode", data.data_ptr(), data.numel(), ctypes.byref(self._header),
                 fr.data_ptr(), F, dev.ptr(idx), n_sel, dev.ptr(slot), float(scale),
                 xyz.data_ptr(), cell_d.data_ptr(), status.data_ptr(), dev.stream())
            st = dev.to_host(status)
            cell = dev.to_host(cell_d)
            bad = np.flatnonzero(st)
            if bad.size:
                b = int(bad[0])
                warnings.warn(f"DCD frame {int(frames[b])}: {_STATUS.get(int(st[b]), 'unknown error')}; "
                              f"the read ends before it", RuntimeWarning, stacklevel=3)
                xyz, cell = xyz[:b], cell[:b]
                self.frame_counter = int(frames[b]) + 1
            else:
                self.frame_counter = int(min(len(self), frames[-1] + step))
        elif F:
            self.frame_counter = int(min(len(self), frames[-1] + step))
        lengths, angles = cell_lengths_angles(cell)
        if np.all(lengths < 1e-10):
            lengths = angles = None
        return xyz, lengths, angles

    def _device_bytes(self):
        if self._dev is None or dev.is_torch_cuda(self._dev) != dev.is_torch_mode():
            self._dev = dev.to_device(self._host, np.uint8)
            if self._free is not None:
                # an atom listed twice takes its last slot, as the reference's copy loop
                rev = self._free[::-1].astype(np.int64) - 1
                atoms, first_in_rev = np.unique(rev, return_index=True)
                slot = np.full(self.n_atoms, -1, np.int32)
                slot[atoms] = (self._free.size - 1 - first_in_rev).astype(np.int32)
                self._dev_slot = dev.to_device_i32(slot)
        return self._dev, self._dev_slot


def load_dcd(filename, top=None, stride=None, atom_indices=None, frame=None):
    """Trajectory of a DCD file with device-resident coordinates in
    nanometres (dcd.pyx:102-171; ``top`` is kept as the trajectory's topology
    and is not required)."""
    with DCDTrajectoryFile(filename) as f:
        if frame is not None:
            f.seek(frame)
            n_frames = 1
        else:
            n_frames = None
        return f.read_as_traj(top, n_frames=n_frames, stride=stride, atom_indices=atom_indices)
