"""LP-RMSD on the GPU: drop-in for ``md.lprmsd`` (rmsd/_lprmsd.pyx:72-228).

The RMSD of every target frame to ``reference[frame]``, minimised over the
rotation and over relabellings of the atoms inside each permute group.  The
steps, the argument validation and the error messages are the reference's
(_lprmsd.pyx:232-271).  Per frame, ``hr_lprmsd`` runs one wave:
  (1) centre the selection;
  (2) fit the atoms in no group and rotate by that fit;
  (3) solve each group's assignment exactly (Hungarian method);
  (4) QCP on the relabelled selection.
``superpose=True`` centres and rotates every frame of ``target`` in place,
as the reference does.  Distances come back as float32 numpy, or as a device
tensor when ``target.xyz`` is one.
"""
from __future__ import annotations

import numpy as np

from . import _device as dev
from ._lib import call, load
from ._rmsd import _xyz_of


def _validate_shapes(txyz, rxyz, frame):
    assert len(txyz.shape) == 3 and len(rxyz.shape) == 3 and txyz.shape[2] == 3 and rxyz.shape[2] == 3
    if txyz.shape[1] != rxyz.shape[1]:
        raise ValueError("Input trajectories must have same number of atoms. "
                         "found %d and %d." % (txyz.shape[1], rxyz.shape[1]))
    if frame >= rxyz.shape[0]:
        raise ValueError("Cannot calculate RMSD of frame %d: reference has "
                         "only %d frames." % (frame, rxyz.shape[0]))


def _validate_atom_indices(atom_indices, n_atoms):
    if atom_indices is None:
        return np.arange(n_atoms, dtype=int)
    atom_indices = np.unique(np.asarray(atom_indices)).astype(int)
    if atom_indices.ndim != 1:
        raise ValueError("atom_indices must be one-dimensional")
    if not np.all((atom_indices >= 0) & (atom_indices < n_atoms)):
        raise ValueError("atom_indices must be valid positive indices")
    return atom_indices


def _validate_permute_groups(permute_groups, atom_indices):
    if permute_groups is None:
        return [atom_indices]
    groups = [np.unique(np.asarray(g)).astype(int) for g in permute_groups]
    for g in groups:
        if len(np.setdiff1d(g, atom_indices)) > 0:
            raise ValueError("The elements in each permute group must be a subset of atom_indices")
    allp = np.concatenate(groups) if groups else np.zeros(0, int)
    if len(np.unique(allp)) != len(allp):  # an atom in two groups
        raise ValueError("permute_groups must be mutually disjoint sets")
    return groups


@dev.follows(0, "target")
def lprmsd(target, reference, frame=0, atom_indices=None, permute_groups=None, parallel=True,
           superpose=False):
    """LP-RMSD of every conformation in ``target`` to ``reference[frame]``
    (nm, float32); see the module docstring."""
    txyz = _xyz_of(target)
    rxyz = _xyz_of(reference)
    _validate_shapes(txyz, rxyz, frame)
    n_atoms = int(txyz.shape[1])
    atom_indices = _validate_atom_indices(atom_indices, n_atoms)
    groups = _validate_permute_groups(permute_groups, atom_indices)
    # empty groups relabel nothing (permute_groups=[[]] is the plain RMSD)
    rel = [np.searchsorted(atom_indices, g).astype(np.int32) for g in groups if len(g)]
    flat = np.concatenate(rel) if rel else np.zeros(0, np.int32)
    dis = np.setdiff1d(np.arange(len(atom_indices)), flat).astype(np.int32)
    offsets = np.cumsum([0] + [len(g) for g in rel]).astype(np.int32)
    max_group = max([len(g) for g in rel], default=0)

    lib = load()
    on_device = dev.is_device_array(txyz)
    # the reference frame is copied first (the reference copies it too), so a
    # target that shares its buffer is read unmoved
    R = dev.to_device_f32(rxyz[frame]).clone()
    if superpose and on_device:
        f32 = dev.torch().float32 if dev.is_torch_cuda(txyz) else np.dtype(np.float32)
        if not (txyz.is_contiguous() and txyz.dtype == f32):
            raise ValueError("superpose needs target.xyz as a C-contiguous float32 tensor")
        T = txyz
    else:
        T = dev.to_device_f32(txyz)
        if superpose and T.data_ptr() == getattr(txyz, "data_ptr", lambda: None)():
            T = T.clone()
    n_frames = int(T.shape[0])
    out = dev.empty_f32(n_frames)
    if n_frames:
        d_idx = dev.to_device_i32(atom_indices)
        d_dis = dev.to_device_i32(dis) if dis.size else None
        d_grp = dev.to_device_i32(flat) if flat.size else None
        d_off = dev.to_device_i32(offsets)
        ws = dev.workspace(lib.hr_lprmsd_workspace_bytes(len(atom_indices), len(dis)))
        call("hr_lprmsd", T.data_ptr(), n_frames, n_atoms, R.data_ptr(), dev.ptr(d_idx),
             len(atom_indices), dev.ptr(d_dis), len(dis), dev.ptr(d_grp), dev.ptr(d_off),
             len(rel), max_group, int(bool(superpose)), out.data_ptr(), ws.data_ptr(),
             dev.stream())
    if superpose and not on_device:
        host = dev.to_host(T)
        if (isinstance(txyz, np.ndarray) and txyz.dtype == np.float32 and txyz.flags.c_contiguous
                and txyz.flags.writeable):
            txyz[...] = host
        elif hasattr(target, "xyz"):
            target.xyz = host
    return out if on_device else dev.to_host(out)
