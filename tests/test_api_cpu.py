"""Host-side behaviour of the drop-in API: validation errors of the reference
(rmsd/_rmsd.pyx:154-184, core/trajectory.py:1117-1124) and no CPU fallback."""
import numpy as np
import pytest

import mdtraj_amd as md
from mdtraj_amd import _rmsd


def traj(F=4, N=10, seed=0):
    return md.Trajectory(np.random.default_rng(seed).standard_normal((F, N, 3)).astype(np.float32))


def test_invalid_atom_indices():
    t = traj()
    with pytest.raises(ValueError, match="atom_indices must be valid positive indices"):
        md.rmsd(t, t, 0, atom_indices=[0, 10])
    with pytest.raises(ValueError, match="atom_indices must be valid positive indices"):
        md.rmsd(t, t, 0, atom_indices=[-1, 2])


def test_mismatched_ref_indices():
    t = traj()
    with pytest.raises(ValueError, match="same number of atom indices"):
        md.rmsd(t, t, 0, atom_indices=[0, 1, 2], ref_atom_indices=[0, 1])
    with pytest.raises(ValueError, match="ref_atom_indices must be valid positive indices"):
        md.rmsd(t, t, 0, atom_indices=[0, 1], ref_atom_indices=[0, 99])


def test_frame_out_of_range():
    t = traj()
    with pytest.raises(ValueError, match="Cannot calculate RMSD of frame 4"):
        md.rmsd(t, t, 4)


def test_atom_count_mismatch():
    with pytest.raises(ValueError, match="same number of atoms"):
        md.rmsd(traj(N=10), traj(N=11), 0)


def test_bad_shape_asserts():
    t = traj()
    bad = md.Trajectory(np.zeros((2, 10, 3), np.float32))
    bad._xyz = np.zeros((2, 10, 4), np.float32)
    with pytest.raises(AssertionError):
        md.rmsd(bad, t, 0)


def test_superpose_empty_indices():
    t = traj()
    with pytest.raises(ValueError, match="Number of atom indices must be greater than 0"):
        t.superpose(t, 0, atom_indices=[])
    with pytest.raises(ValueError, match="Number of atoms must be consistent"):
        t.superpose(t, 0, atom_indices=[0, 1], ref_atom_indices=[1])


def test_xyz_setter_resets_traces_and_casts():
    t = traj()
    t._rmsd_traces = np.ones(4, np.float32)
    t.xyz = np.zeros((3, 5, 3), np.float64)
    assert t._rmsd_traces is None and t.xyz.dtype == np.float32 and t.xyz.flags.c_contiguous
    t.xyz = np.zeros((5, 3))
    assert t.xyz.shape == (1, 5, 3)


def test_no_cpu_fallback_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    t = traj()
    with pytest.raises(RuntimeError, match="GPU"):
        md.rmsd(t, t, 0)
    with pytest.raises(RuntimeError, match="GPU"):
        _rmsd._center_inplace_atom_major(t.xyz.copy())


def test_default_ref_indices_checked_against_reference():
    # _rmsd.pyx:163-174: ref_atom_indices defaults to atom_indices and is
    # range-checked against the reference's atoms
    big, small = traj(N=12), traj(N=5)
    with pytest.raises(ValueError, match="ref_atom_indices must be valid positive indices"):
        md.rmsd(big, small, 0, atom_indices=[0, 8])


def test_lprmsd_validation_on_the_host():
    """md.lprmsd checks its arguments before any device work, with the
    reference's messages (_lprmsd.pyx:232-271)."""
    import mdtraj_amd as md

    x = np.zeros((3, 6, 3), np.float32)
    with pytest.raises(ValueError, match="only 3 frames"):
        md.lprmsd(x, x, 7)
    with pytest.raises(ValueError, match="same number of atoms"):
        md.lprmsd(x, np.zeros((1, 5, 3), np.float32))
    with pytest.raises(ValueError, match="valid positive indices"):
        md.lprmsd(x, x, 0, atom_indices=[0, 9])
    with pytest.raises(ValueError, match="subset of atom_indices"):
        md.lprmsd(x, x, 0, atom_indices=[0, 1, 2], permute_groups=[[4]])
    with pytest.raises(ValueError, match="mutually disjoint"):
        md.lprmsd(x, x, 0, permute_groups=[[0, 1], [1, 2]])
