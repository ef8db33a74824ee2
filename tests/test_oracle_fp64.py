"""Pins the FP64 oracle (oracle/fp64.py) by checks that do not restate it:
known answers, a second algorithm (Kabsch/SVD), brute force over rotations,
invariants, and the compiled reference's golden vectors within the
reference's own float32 error."""
import numpy as np
import pytest

from mdtraj_amd import gen
from oracle import fp64

XYZ1 = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [1, 0, 1], [0, 1, 1.0]])
ROT90 = np.array([[1, 0, 0], [0, 0, -1], [0, 1, 0]])


def test_known_answer_rotated_copy_is_zero():
    # tests/test_alignment.py:29-43: a rotated + translated copy has RMSD 0
    xyz2 = ROT90.dot(XYZ1.T).T + np.array([0.7, -0.2, 1.9])
    X = np.stack([XYZ1, xyz2])
    assert fp64.rmsd(X, X, 0)[1] < 1e-7
    assert fp64.rmsd_kabsch(X, X, 0)[1] < 1e-7


def test_known_answer_extra_points_transform():
    # tests/test_alignment.py:34,59-65: fit on the first 6 points, move all 11
    extra = np.array([[1, 2, 3], [4, 5, 6], [7, 8, 9], [10, 11, 12], [13, 14, 15.0]])
    allp = np.concatenate([XYZ1, extra])
    xyz4 = ROT90.dot(allp.T).T + np.array([0.5, 0.5, -3.0])
    moved = fp64.superpose(xyz4[None], allp[None], 0, atom_indices=np.arange(6))
    assert np.abs(moved[0, :6] - XYZ1).max() < 1e-12


@pytest.mark.parametrize("seed", range(5))
def test_eigenvalue_route_equals_kabsch(seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((40, 17, 3))
    a = fp64.rmsd(x, x, seed)
    b = fp64.rmsd_kabsch(x, x, seed)
    # compare MSDs: the self pair's RMSD is sqrt of an FP64 rounding residual
    assert np.abs(a ** 2 - b ** 2).max() < 1e-13


@pytest.mark.parametrize("seed", range(3))
def test_brute_force_rotation_search(seed):
    rng = np.random.default_rng(100 + seed)
    a = rng.standard_normal((5, 3))
    b = rng.standard_normal((5, 3))
    want = fp64.brute_force_rmsd(a, b, n_grid=14, refine=300, seed=seed)
    got = fp64.rmsd(a[None], b[None], 0)[0]
    # brute force finds an upper bound that converges onto the optimum
    assert got <= want + 1e-9
    assert want - got < 1e-6


def test_invariants():
    rng = np.random.default_rng(7)
    x = rng.standard_normal((30, 25, 3))
    d = fp64.rmsd(x, x, 4)
    moved = gen.random_rigid(rng, x, translate=10.0)
    assert np.abs(fp64.rmsd(moved, x, 4) - d).max() < 1e-10  # rigid motion of targets
    assert np.abs(fp64.rmsd(x, moved, 4) - d).max() < 1e-10  # and of the reference
    # symmetry: rmsd(i -> j) == rmsd(j -> i)
    D = fp64.rmsd_matrix(x)
    assert np.abs(D - D.T).max() < 1e-10
    assert np.abs(np.diag(D)).max() < 1e-6
    # superposition can only lower the deviation of centred frames
    c, _ = fp64.center_and_trace(x)
    assert np.all(d <= fp64.rmsd_nosuperpose(c, c, 4) + 1e-12)
    # triangle inequality of the RMSD metric
    for i, j, k in [(0, 1, 2), (3, 9, 17), (5, 5, 20)]:
        assert D[i, k] <= D[i, j] + D[j, k] + 1e-9


def test_degenerate_systems_are_zero():
    one = np.random.default_rng(1).standard_normal((3, 1, 3))
    assert np.all(fp64.rmsd(one, one, 0) == 0.0)
    zeros = np.zeros((2, 5, 3))
    assert np.all(fp64.rmsd(zeros, zeros, 0) == 0.0)


def test_centering_closed_form():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((4, 9, 3)) + 5
    c, g = fp64.center_and_trace(x)
    assert np.abs(c.mean(axis=1)).max() < 1e-14
    assert np.allclose(g, [np.sum((f - f.mean(0)) ** 2) for f in x])


def test_all_pairs_row_equals_one_vs_all():
    x = gen.walk(seed=5, n_frames=40, n_atoms=30)
    D = fp64.rmsd_matrix(x)
    for i in (0, 13, 39):
        assert np.abs(D[i] ** 2 - fp64.rmsd(x, x, i) ** 2).max() < 1e-13
    assert fp64.condensed(D).shape == (40 * 39 // 2,)


def test_superpose_moves_onto_reference():
    rng = np.random.default_rng(11)
    ref = rng.standard_normal((20, 3)) + 3
    frames = gen.random_rigid(rng, np.repeat(ref[None], 5, 0))
    moved = fp64.superpose(frames, ref[None], 0)
    assert np.abs(moved - ref[None]).max() < 1e-10


def test_reference_golden_within_float32_error(golden):
    """The compiled reference's float32 outputs sit within its own float32
    error model around the FP64 oracle (sqrt(eps32 (G_a + G_b)/N) near 0,
    eps32 (G_a + G_b) / (2 N r) elsewhere, with a safety factor)."""
    eps = np.finfo(np.float32).eps
    x, r = golden["cube_xyz"], golden["cube_ref"]
    o = fp64.rmsd(x, r[None], 0)
    assert np.abs(golden["cube_rmsd"] - o).max() < 5e-6
    w, idx = golden["walk_xyz"], golden["walk_idx"]
    o = fp64.rmsd(w, w, 5, idx)
    _, g = fp64.center_and_trace(w[:, idx])
    s = (g + g[5]) / len(idx)
    bound = 8 * np.minimum(np.sqrt(eps * s), eps * s / (2 * np.maximum(o, 1e-12))) + 1e-6
    assert np.all(np.abs(golden["walk_rmsd_sel_f5"] - o) <= bound)
    assert np.abs(golden["align_rmsd"] - fp64.rmsd(golden["align_xyz"], golden["align_xyz"], 0)).max() < 1e-3


def test_reference_superpose_golden(golden):
    x, sel = golden["sup_xyz"], golden["sup_sel"]
    o = fp64.superpose(x, x, 4, atom_indices=sel)
    assert np.abs(o - golden["sup_moved_sel_f4"]).max() < 2e-5
    o = fp64.superpose(x, x, 0)
    assert np.abs(o - golden["sup_moved_all_f0"]).max() < 2e-5


def test_planted_mistake_is_caught(monkeypatch):
    """A sign error in one key-matrix entry must be caught by the Kabsch check."""
    real = fp64.key_matrix

    def broken(M):
        K = real(M)
        K[..., 0, 1] = K[..., 1, 0] = -K[..., 0, 1]
        return K

    monkeypatch.setattr(fp64, "key_matrix", broken)
    rng = np.random.default_rng(0)
    x = rng.standard_normal((20, 12, 3))
    assert np.abs(fp64.rmsd(x, x, 0) - fp64.rmsd_kabsch(x, x, 0)).max() > 1e-3
