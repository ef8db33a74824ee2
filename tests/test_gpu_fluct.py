"""GPU parity of md.rmsf (atom and residue modes, fitted and prealigned,
with and without a selection) and of getMultipleAlignDisplaceRMSDs_atom_major
against the FP64 oracle (1e-5 nm)."""
import numpy as np
import pytest

import mdtraj_amd as md
from mdtraj_amd import _rmsd, gen
from oracle import fp64

pytestmark = pytest.mark.gpu

TOL = 1e-5


class _Atom:
    def __init__(self, res, mass):
        self.residue = type("R", (), {"index": res})()
        self.element = type("E", (), {"mass": mass})()


class _Top:
    def __init__(self, res, masses, n_res):
        self._atoms = [_Atom(int(r), float(m)) for r, m in zip(res, masses)]
        self.n_residues = n_res

    def atom(self, i):
        return self._atoms[i]


@pytest.mark.parametrize("n_frames,n_atoms", [(1, 5), (65, 40), (300, 700), (2000, 64)])
def test_rmsf_fitted_vs_fp64(gpu, n_frames, n_atoms):
    x = gen.walk(seed=n_frames + n_atoms, n_frames=n_frames, n_atoms=n_atoms)
    t = md.Trajectory(x)
    got = md.rmsf(t, t, n_frames // 2)
    want = fp64.rmsf(x, x, n_frames // 2)
    assert got.shape == (n_atoms,) and got.dtype == np.float32
    assert np.abs(got - want).max() < TOL


def test_rmsf_selection_and_ref_selection(gpu):
    x = gen.walk(seed=3, n_frames=120, n_atoms=200)
    sel = gen.subset(4, 200, 150)
    rsel = np.sort(np.random.default_rng(1).choice(200, 150, replace=False))
    t = md.Trajectory(x)
    got = md.rmsf(t, t, 7, atom_indices=sel)
    assert np.abs(got - fp64.rmsf(x, x, 7, atom_indices=sel)).max() < TOL
    got = md.rmsf(t, t, 7, atom_indices=sel, ref_atom_indices=rsel)
    want = fp64.rmsf(x, x, 7, atom_indices=sel, ref_atom_indices=rsel)
    assert np.abs(got - want).max() < TOL


def test_rmsf_prealigned(gpu):
    # tests/test_rmsd.py:255-267 on the GPU path
    x = gen.walk(seed=11, n_frames=90, n_atoms=60)
    t = md.Trajectory(x.copy())
    t.superpose(t, 9, atom_indices=np.arange(30))
    other = np.arange(30, 60)
    got = md.rmsf(t, None, atom_indices=other)
    avg = t.xyz.astype(np.float64).mean(axis=0)
    want = np.sqrt(3 * np.mean((t.xyz - avg) ** 2, axis=(0, 2)))[other]
    assert np.abs(got - want).max() < TOL
    got_all = md.rmsf(t, None)
    assert np.abs(got_all - fp64.rmsf(t.xyz, None)).max() < TOL


def test_rmsf_residue_mode(gpu):
    x = gen.walk(seed=12, n_frames=50, n_atoms=24)
    res = np.repeat(np.arange(6), 4)
    masses = np.tile([12.0, 1.0, 14.0, 16.0], 6)
    t = md.Trajectory(x)
    t.topology = _Top(res, masses, 7)
    got = md.rmsf(t, t, 0, mode="residue")
    want = fp64.rmsf_by_residue(fp64.rmsf(x, x, 0), res, masses, 7)
    assert got[-1] == -1.0
    assert np.abs(got - want).max() < TOL
    sel = np.arange(0, 24, 2)
    got = md.rmsf(t, t, 0, atom_indices=sel, mode="residue")
    want = fp64.rmsf_by_residue(fp64.rmsf(x, x, 0, atom_indices=sel), res[sel], masses[sel], 7)
    assert np.abs(got - want).max() < TOL


def test_rmsf_validation(gpu):
    t = md.Trajectory(gen.walk(seed=1, n_frames=4, n_atoms=10))
    with pytest.raises(ValueError, match="Mode must be one of"):
        md.rmsf(t, t, 0, mode="chain")
    with pytest.raises(ValueError, match="Cannot calculate RMSF of frame 9"):
        md.rmsf(t, t, 9)
    with pytest.raises(ValueError, match="atom_indices must be valid positive indices"):
        md.rmsf(t, t, 0, atom_indices=[0, 10])


@pytest.mark.parametrize("n_frames,n_align,n_displ", [(1, 4, 4), (77, 37, 50), (400, 300, 1000)])
def test_align_displace_vs_fp64(gpu, n_frames, n_align, n_displ):
    rng = np.random.default_rng(n_frames)
    x = gen.walk(seed=n_frames, n_frames=n_frames, n_atoms=n_align + n_displ)
    pad = lambda k: (k + 3) // 4 * 4
    align = np.zeros((n_frames, pad(n_align), 3), np.float32)
    displ = np.zeros((n_frames, pad(n_displ), 3), np.float32)
    a = x[:, :n_align]
    align[:, :n_align] = a - a.mean(axis=1, keepdims=True)
    displ[:, :n_displ] = x[:, n_align:]
    frame = int(rng.integers(0, n_frames))
    g = np.zeros(n_frames, np.float32)
    got, rot = _rmsd.getMultipleAlignDisplaceRMSDs_atom_major(
        align, align, g, g, displ, displ, n_align, n_displ, frame)
    want, R = fp64.align_displace_rmsd(align[:, :n_align], align[:, :n_align],
                                       displ[:, :n_displ], displ[:, :n_displ], frame)
    assert got.shape == (n_frames,) and rot.shape == (n_frames, 3, 3)
    assert np.abs(got - want).max() < TOL
    assert np.abs(rot - R).max() < 1e-5


def test_align_displace_validation(gpu):
    z = np.zeros((2, 5, 3), np.float32)
    g = np.zeros(2, np.float32)
    with pytest.raises(ValueError, match="middle dimension of 4"):
        _rmsd.getMultipleAlignDisplaceRMSDs_atom_major(z, z, g, g, z, z, 5, 5, 0)
