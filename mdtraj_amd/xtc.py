"""XTC trajectories decoded on the GPU.

``XTCTrajectoryFile`` keeps the reading half of the reference's class
(formats/xtc/xtc.pyx:190-760: ``read(n_frames, stride, atom_indices)``
returning ``(xyz, time, step, box)`` with ``box`` None when every entry is
below 1e-10 in magnitude, ``seek``/``tell``, ``len``, ``offsets``,
``n_atoms``) and ``load_xtc`` builds a Trajectory (xtc.pyx:111-187).

The file is read into host memory once; ``hr_xtc_scan`` indexes its frames
on the host; the bytes are copied to HBM once, and every ``read`` decodes
the requested frames in parallel on the device (``hr_xtc_decode``), with
results bit-identical to the reference's reader.  ``read(device=True)``
returns the coordinates as a device tensor (a serving pool or a one-vs-all
query reads them from HBM directly); the default returns numpy arrays like
the reference.  Writing XTC files is not provided.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _device as dev
from ._lib import call, load

_STATUS = {1: "not an XTC frame (bad magic)", 2: "atom count differs between frames",
           3: "frame ends early (truncated file)", 4: "malformed compressed coordinates"}


class XTCTrajectoryFile:
    """Reader for GROMACS XTC files (mode 'r')."""

    def __init__(self, filename, mode="r", force_overwrite=True, **kwargs):
        if mode != "r":
            raise ValueError("only mode='r' is supported: this reader decodes XTC files "
                             "on the GPU and does not write them")
        self.filename = os.fspath(filename)
        if not os.path.exists(self.filename):
            raise IOError(f"The file '{self.filename}' doesn't exist")
        self._host = np.fromfile(self.filename, dtype=np.uint8)
        self._dev = None
        load()
        n_frames, n_atoms = ctypes.c_int64(), ctypes.c_int32()
        buf = self._host.ctypes.data_as(ctypes.c_void_p)
        call("hr_xtc_scan", buf, self._host.size, None, 0, ctypes.byref(n_frames),
             ctypes.byref(n_atoms))
        offs = np.zeros(max(1, n_frames.value), np.int64)
        call("hr_xtc_scan", buf, self._host.size, offs.ctypes.data_as(ctypes.c_void_p),
             n_frames.value, ctypes.byref(n_frames), ctypes.byref(n_atoms))
        self._offsets = offs[: n_frames.value]
        self.n_atoms = int(n_atoms.value)
        self.frame_counter = 0
        self.is_open = True
        self.mode = "r"

    # -- reference API -------------------------------------------------------
    def __len__(self):
        return int(self._offsets.size)

    @property
    def offsets(self):
        """Byte offset of every frame (xtc.pyx:716-721)."""
        return self._offsets.copy()

    def tell(self):
        return int(self.frame_counter)

    def seek(self, offset, whence=0):
        """Move to frame ``offset`` relative to the start (0), the current
        frame (1) or the end (2), as xtc.pyx:602-640."""
        if whence == 0:
            pos = offset
        elif whence == 1:
            pos = self.frame_counter + offset
        elif whence == 2:
            pos = len(self) + offset
        else:
            raise IOError("Invalid argument")
        if not 0 <= pos <= len(self):
            raise IOError("Invalid argument")
        self.frame_counter = int(pos)

    def close(self):
        self._dev = None
        self.is_open = False

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def read(self, n_frames=None, stride=None, atom_indices=None, device=False):
        """Decode frames from the current position: every ``stride``-th, at
        most ``n_frames``; returns ``(xyz, time, step, box)``."""
        if not self.is_open:
            raise IOError("file must be open to read from it.")
        stride = int(stride) if stride is not None else 1
        if stride < 1:
            raise ValueError("stride must be positive")
        if n_frames is not None and not int(n_frames) == n_frames:
            raise ValueError(f'n_frames must be an int, you supplied "{n_frames}"')
        frames = np.arange(self.frame_counter, len(self), stride)
        if n_frames is not None:
            frames = frames[: max(0, int(n_frames))]
        if atom_indices is not None and not isinstance(atom_indices, slice):
            atom_indices = np.asarray(atom_indices)
            if atom_indices.size and atom_indices.min() < 0:
                raise ValueError("atom_indices should be zero indexed. you gave an index less than zero")
            if atom_indices.size and atom_indices.max() >= self.n_atoms:
                raise ValueError("atom indices should be zero indexed. you gave an index bigger "
                                 "than the number of atoms")
        with dev.using_loaded():
            xyz, time, step, box = self._decode(frames)
            if atom_indices is not None:
                sel = (np.arange(self.n_atoms)[atom_indices] if isinstance(atom_indices, slice)
                       else atom_indices)
                xyz = self._select_atoms(xyz, sel)
        if frames.size:
            self.frame_counter = int(min(len(self), frames[-1] + stride))
        if np.all(np.logical_and(box < 1e-10, box > -1e-10)):
            box = None
        return (xyz if device else dev.to_host(xyz)), time, step, box

    # -- device decode -------------------------------------------------------
    def _device_bytes(self):
        if self._dev is None or dev.is_torch_cuda(self._dev) != dev.is_torch_mode():
            self._dev = dev.to_device(self._host, np.uint8)
        return self._dev

    @staticmethod
    def _select_atoms(xyz, sel):
        """(F, n_sel, 3) device copy of the atoms `sel` of every frame."""
        sel = np.ascontiguousarray(sel, dtype=np.int32)
        F, N = int(xyz.shape[0]), int(xyz.shape[1])
        out = dev.empty((F, sel.size, 3), np.float32)
        if F and sel.size:
            idx = dev.to_device_i32(sel)
            call("hr_gather_atoms", xyz.data_ptr(), F, N, idx.data_ptr(), int(sel.size),
                 out.data_ptr(), dev.stream())
        return out

    def _decode(self, frames):
        F, N = int(frames.size), self.n_atoms
        xyz = dev.empty((F, N, 3), np.float32)
        time = dev.empty(F, np.float32)
        step = dev.empty(F, np.int32)
        box = dev.empty((F, 3, 3), np.float32)
        prec = dev.empty(F, np.float32)
        status = dev.empty(F, np.int32)
        if F:
            data = self._device_bytes()
            offs = dev.to_device(self._offsets[frames], np.int64)
            call("hr_xtc_decode", data.data_ptr(), data.numel(), offs.data_ptr(), F, N,
                 xyz.data_ptr(), time.data_ptr(), step.data_ptr(), box.data_ptr(),
                 prec.data_ptr(), status.data_ptr(), dev.stream())
            st = dev.to_host(status)
            bad = np.flatnonzero(st)
            if bad.size:
                f = int(frames[bad[0]])
                raise RuntimeError(f"XTC read error: {_STATUS.get(int(st[bad[0]]), 'unknown')} "
                                   f"(frame {f})")
        self.last_precision = dev.to_host(prec)
        return xyz, dev.to_host(time), dev.to_host(step), dev.to_host(box)


def load_xtc(filename, top=None, stride=None, atom_indices=None, frame=None):
    """Trajectory of an XTC file with device-resident coordinates
    (xtc.pyx:111-187; ``top`` is kept as the trajectory's topology and is
    not required)."""
    from .trajectory import Trajectory

    with XTCTrajectoryFile(filename) as f:
        if frame is not None:
            f.seek(frame)
            xyz, time, step, box = f.read(n_frames=1, atom_indices=atom_indices, device=True)
        else:
            xyz, time, step, box = f.read(stride=stride, atom_indices=atom_indices, device=True)
    traj = Trajectory(xyz, top, time=time)
    traj.unitcell_vectors = box  # (F, 3, 3) or None, as the reference's box
    traj.step = step
    return traj
