"""Pins the oracle's rmsf and align/displace functions.

rmsf follows rmsd/_rmsd.pyx:246-484; the checks are the reference's own test
formulas (tests/test_rmsd.py:244-286: superpose, then the fluctuation about
the mean computed directly), rigid copies (zero), a statistical law
(isotropic noise of sigma per axis gives sqrt(3) sigma), and the residue
mode's mass-weighted formula.  align/displace follows rmsd/_rmsd.pyx:677-759;
with the displacement set equal to the alignment set it must reproduce the
QCP eigenvalue RMSD, and exact rigid copies give zero."""
import numpy as np
import pytest

from mdtraj_amd import gen
from oracle import fp64


def _rigid(seed, n_frames, n_atoms, noise=0.0):
    rng = np.random.default_rng(seed)
    base = rng.uniform(0, 3, size=(n_atoms, 3))
    R = gen.rotation_matrix_from_quaternion(gen.uniform_quaternion(rng, n_frames))
    x = np.einsum("na,fba->fnb", base, R) + rng.uniform(-2, 2, size=(n_frames, 1, 3))
    return x + noise * rng.normal(size=x.shape)


def test_rmsf_of_rigid_copies_is_zero():
    x = _rigid(0, 40, 30)
    assert np.abs(fp64.rmsf(x, x, 3)).max() < 1e-12


def test_rmsf_matches_superpose_then_fluctuation():
    # tests/test_rmsd.py:244-252
    x = gen.walk(seed=5, n_frames=60, n_atoms=50).astype(np.float64)
    got = fp64.rmsf(x, x, 0)
    s = fp64.superpose(x, x, 0)
    avg = s.mean(axis=0)
    want = np.sqrt(3 * np.mean((s - avg) ** 2, axis=(0, 2)))
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


def test_rmsf_prealigned_uses_raw_coordinates_for_a_selection():
    # tests/test_rmsd.py:255-267: superpose on one half, rmsf of the other
    # half with reference None equals the fluctuation of the raw coordinates
    x = gen.walk(seed=6, n_frames=50, n_atoms=40).astype(np.float64)
    half = np.arange(20)
    s = fp64.superpose(x, x, 9, atom_indices=half)
    other = np.arange(20, 40)
    got = fp64.rmsf(s, None, atom_indices=other)
    avg = s.mean(axis=0)
    want = np.sqrt(3 * np.mean((s - avg) ** 2, axis=(0, 2)))[other]
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


def test_rmsf_selection_is_not_recentred():
    """Rigid copies with translations: without a selection the centred
    copies coincide after the fit (zero); with an explicit all-atom selection
    the rotation acts about the origin on uncentred coordinates, as the
    reference does, and the translations leave a fluctuation."""
    x = _rigid(1, 30, 12)
    assert np.abs(fp64.rmsf(x, x, 0)).max() < 1e-12
    assert np.abs(fp64.rmsf(x, x, 0, atom_indices=np.arange(12))).min() > 0.1


def test_rmsf_noise_law():
    rng = np.random.default_rng(7)
    base = rng.uniform(0, 3, size=(16, 3))
    sigma = np.linspace(0.01, 0.1, 16)
    x = base + sigma[None, :, None] * rng.normal(size=(20000, 16, 3))
    got = fp64.rmsf(x, None)   # prealigned: per-frame centroid removed, no fit
    # centring on the 16-atom centroid removes a little of each atom's variance
    want = np.sqrt(3 * (sigma ** 2 * (1 - 2 / 16) + (sigma ** 2).sum() / 16 ** 2))
    np.testing.assert_allclose(got, want, rtol=0.02)


def test_rmsf_by_residue_formula():
    # tests/test_rmsd.py:270-288
    x = gen.walk(seed=8, n_frames=30, n_atoms=24).astype(np.float64)
    res = np.repeat(np.arange(6), 4)
    masses = np.tile([12.0, 1.0, 14.0, 16.0], 6)
    f = fp64.rmsf(x, x, 0)
    got = fp64.rmsf_by_residue(f, res, masses, 7)
    s = fp64.superpose(x, x, 0)
    avg = s.mean(axis=0)
    want = np.full(7, -1.0)
    for r in range(6):
        ind = np.where(res == r)[0]
        msf = 3 * np.mean((s[:, ind] - avg[ind]) ** 2, axis=(0, 2))
        want[r] = np.sqrt(np.sum(msf * masses[ind]) / np.sum(masses[ind]))
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


def test_align_displace_on_the_alignment_set_is_qcp():
    x = gen.walk(seed=9, n_frames=40, n_atoms=30).astype(np.float64)
    c = x - x.mean(axis=1, keepdims=True)
    r, R = fp64.align_displace_rmsd(c, c, c, c, 4)
    # atol: the eigenvalue route leaves sqrt(eps G / N) on the self pair
    np.testing.assert_allclose(r, fp64.rmsd(x, x, 4), rtol=0, atol=1e-7)
    assert np.allclose(np.linalg.det(R), 1.0)


def test_align_displace_rigid_is_zero_and_rotation_maps_reference():
    rng = np.random.default_rng(10)
    a_ref = rng.normal(size=(10, 3))
    a_ref -= a_ref.mean(axis=0)
    d_ref = rng.normal(size=(7, 3)) + 3.0
    Q = gen.rotation_matrix_from_quaternion(gen.uniform_quaternion(rng, 25))
    Qt = np.transpose(Q, (0, 2, 1))          # x' = x Q^T = (Q x^T)^T
    align = np.einsum("nk,fkj->fnj", a_ref, Qt)
    displ = np.einsum("nk,fkj->fnj", d_ref, Qt)
    r, R = fp64.align_displace_rmsd(np.repeat(a_ref[None], 2, 0), align,
                                    np.repeat(d_ref[None], 2, 0), displ, 1)
    assert r.max() < 1e-12
    np.testing.assert_allclose(R, Qt, atol=1e-12)


def test_planted_transposed_rotation_is_caught(monkeypatch):
    real = fp64.kabsch_rotation
    monkeypatch.setattr(fp64, "kabsch_rotation",
                        lambda m, t: np.transpose(real(m, t), (0, 2, 1)))
    x = _rigid(2, 10, 12)
    assert np.abs(fp64.rmsf(x, x, 0)).max() > 1e-3
