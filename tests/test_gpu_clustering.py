"""Consumers of the RMSD matrix on the GPU.

Nearest leader (hr_rmsd_cross_argmin): the fused row minimum equals the
argmin of the materialised cross matrix from the same kernel bit for bit
(ties to the lowest index, as np.argmin), and the oracle's leader
assignment wherever the best two leaders are more than 1e-4 nm apart.
Centroid scores (hr_centroid_scores): within 1e-9 relative of the
notebook's formula on the same condensed matrix, the two-frame closed form,
the centroid of a two-cluster trajectory, and the oracle's centroid."""
import numpy as np
import pytest

from mdtraj_amd import allpairs, gen
from oracle import fp64

pytestmark = pytest.mark.gpu


def square(cond, n):
    D = np.zeros((n, n))
    iu = np.triu_indices(n, 1)
    D[iu] = cond
    return D + D.T


@pytest.mark.parametrize("mode", ["f16exact", "f16x3", "bf16exact"])
def test_nearest_equals_cross_argmin(gpu, mode):
    x = gen.walk(seed=13, n_frames=300, n_atoms=90)
    leaders = np.concatenate([x[[5, 77, 150, 222]], x[[77]]])  # a duplicate: lowest index wins
    a = allpairs.SplitSet(x, mode=mode)
    b = allpairs.SplitSet(leaders, mode=mode)
    idx, val = allpairs.nearest(a, b)
    full = allpairs.rmsd_cross(a, b).cpu().numpy()
    assert np.array_equal(idx.cpu().numpy(), np.argmin(full, axis=1))
    assert np.array_equal(val.cpu().numpy(), full.min(axis=1))
    assert 4 not in set(idx.cpu().numpy().tolist())


def test_nearest_against_oracle_and_edges(gpu):
    rng = np.random.default_rng(3)
    x = gen.walk(seed=17, n_frames=200, n_atoms=60)
    leaders = x[rng.choice(200, 7, replace=False)] + rng.normal(0, 0.01, (7, 60, 3)).astype(np.float32)
    sel = np.sort(rng.choice(60, 45, replace=False))
    idx, val = allpairs.assign_to_leaders(x, leaders, atom_indices=sel)
    want_i, want_v = fp64.nearest_leader(x, leaders, sel)
    r = np.array([fp64.rmsd(leaders, x[f:f + 1], 0, sel) for f in range(200)])
    r.sort(axis=1)
    clear = r[:, 1] - r[:, 0] > 1e-4
    assert clear.sum() > 150
    assert np.array_equal(idx[clear], want_i[clear])
    assert np.abs(val - want_v).max() < 1e-5
    # more than one column tile of leaders, rows not a multiple of 64
    big = gen.walk(seed=19, n_frames=130, n_atoms=40)
    i2, v2 = allpairs.assign_to_leaders(big[:97], big)
    # a frame against itself: no exact zero from the eigenvalue (about 1e-5 nm),
    # far below the 0.005 nm per-coordinate step to the next frame
    assert np.array_equal(i2, np.arange(97)) and v2.max() < 1e-3


def test_centroid_scores(gpu):
    x = gen.walk(seed=23, n_frames=257, n_atoms=50)
    cond = allpairs.rmsd_matrix(x)
    got = allpairs.centroid_scores(cond, 257, beta=1.0)
    want = fp64.centroid_scores(square(cond.astype(np.float64), 257), 1.0)
    assert np.allclose(got, want, rtol=1e-9, atol=0)
    # against the oracle's own distances (FP64 RMSD rows): same centroid
    D = np.array([fp64.rmsd(x, x, i) for i in range(257)])
    assert int(np.argmax(got)) == int(np.argmax(fp64.centroid_scores(D)))
    assert allpairs.centroid(x) == int(np.argmax(got))
    # two frames: 1 + exp(-2 beta) whatever the distance
    y = gen.walk(seed=29, n_frames=2, n_atoms=20)
    two = allpairs.centroid_scores(allpairs.rmsd_matrix(y), 2, beta=0.7)
    assert np.allclose(two, 1 + np.exp(-1.4), rtol=1e-12)
    # one frame: no entries, the score is the diagonal's exp(0)
    one = allpairs.centroid_scores(np.zeros(0, np.float32), 1)
    assert np.array_equal(one, [1.0])


def test_centroid_two_clusters_and_large_n(gpu):
    rng = np.random.default_rng(31)
    a = rng.normal(0, 1, (1, 40, 3)) + rng.normal(0, 0.01, (2500, 40, 3))
    c = rng.normal(0, 1, (1, 40, 3)) + rng.normal(0, 0.01, (1700, 40, 3))
    y = np.concatenate([a, c]).astype(np.float32)  # 4,200 frames: several column chunks
    cond = allpairs.rmsd_matrix(y)
    got = allpairs.centroid_scores(cond, len(y))
    assert int(np.argmax(got)) < 2500
    sub = np.r_[0:50, 2500:2530, 4190:4200]
    D = square(cond.astype(np.float64), len(y))
    want = np.exp(-D / D.std()).sum(axis=1)
    assert np.allclose(got[sub], want[sub], rtol=1e-9)


def test_centroid_scores_of_a_misaligned_view(gpu):
    import torch

    x = gen.walk(seed=37, n_frames=300, n_atoms=30)
    cond = allpairs.rmsd_matrix(x)
    buf = torch.zeros(cond.size + 1, dtype=torch.float32, device="cuda")
    buf[1:] = torch.from_numpy(cond).cuda()
    view = buf[1:]  # 4 bytes past a 16-byte boundary: the wrapper copies it
    assert view.data_ptr() % 16 == 4
    got = allpairs.centroid_scores(view, 300)
    assert np.array_equal(got, allpairs.centroid_scores(cond, 300))
