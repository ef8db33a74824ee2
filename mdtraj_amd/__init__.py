"""mdtraj_amd — MDTraj's RMSD hot path, MI355X-native (gfx950 HIP kernels).

Drop-in surface (names and semantics of ``mdtraj``):
    rmsd(target, reference, frame=0, atom_indices=None, ref_atom_indices=None,
         parallel=True, precentered=False, superpose=True)
    rmsf(target, reference, frame=0, atom_indices=None, ref_atom_indices=None,
         parallel=True, precentered=False, mode='atom')
    lprmsd(target, reference, frame=0, atom_indices=None, permute_groups=None,
           parallel=True, superpose=False)
    Trajectory(xyz).superpose(...), .center_coordinates(), ._rmsd_traces
    _rmsd: the ``mdtraj._rmsd`` module functions
    capi(): paths of the C header and shared library (as ``md.capi()``)
Readers (the reference's reading API, decoded on the GPU):
    xtc.XTCTrajectoryFile, xtc.load_xtc; dcd.DCDTrajectoryFile, dcd.load_dcd
Beyond the reference:
    rmsd_matrix(traj, ...)   all-pairs RMSD (FP32 or split-precision MFMA)
    allpairs.assign_to_leaders / nearest, allpairs.centroid_scores / centroid
                             consumers of the matrix (clustering notebooks)
    batch.BatchRMSD, serve.RMSDServer   uneven batches and continuous batching
"""
from __future__ import annotations

import os

from . import _rmsd
from ._lprmsd import lprmsd
from ._rmsd import rmsd, rmsf, set_inplace_centering
from .trajectory import Trajectory

__version__ = "0.1.0"


def capi():
    """Locations of the C ABI (as ``mdtraj.capi()``, mdtraj/__init__.py:217-225)."""
    from ._lib import HEADER_PATH, LIB_PATH

    return {"lib_dir": os.path.dirname(LIB_PATH), "include_dir": os.path.dirname(HEADER_PATH),
            "lib": LIB_PATH, "header": HEADER_PATH}


def rmsd_matrix(*args, **kwargs):
    from .allpairs import rmsd_matrix as _rm

    return _rm(*args, **kwargs)


__all__ = ["rmsd", "rmsf", "lprmsd", "Trajectory", "capi", "rmsd_matrix", "_rmsd",
           "set_inplace_centering"]
