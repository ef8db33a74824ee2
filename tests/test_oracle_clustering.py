"""The clustering consumers of the oracle (fp64.centroid_scores,
fp64.nearest_leader), pinned by closed forms and invariances rather than by
their own formulas."""
import numpy as np

from mdtraj_amd import gen
from oracle import fp64


def test_centroid_two_frames_closed_form():
    # D = [[0, a], [a, 0]]: mean a/2, mean square a^2/2, std a/2, so every
    # score is 1 + exp(-2 beta), whatever a
    for a, beta in ((0.3, 1.0), (7.0, 0.5)):
        D = np.array([[0.0, a], [a, 0.0]])
        assert np.allclose(fp64.centroid_scores(D, beta), 1 + np.exp(-2 * beta), rtol=1e-14)


def test_centroid_scale_and_permutation_invariance():
    rng = np.random.default_rng(4)
    x = gen.walk(seed=6, n_frames=40, n_atoms=30)
    D = np.array([fp64.rmsd(x, x, i) for i in range(len(x))])
    s = fp64.centroid_scores(D)
    assert np.allclose(fp64.centroid_scores(3.0 * D), s, rtol=1e-12)  # d / std is scale-free
    p = rng.permutation(len(x))
    assert np.allclose(fp64.centroid_scores(D[np.ix_(p, p)]), s[p], rtol=1e-12)
    # 30 frames of one shape and 10 of another (RMSD ignores translation, so
    # the clusters differ in shape): the centroid is in the larger cluster
    a = rng.normal(0, 1, (1, 30, 3)) + rng.normal(0, 0.01, (30, 30, 3))
    c = rng.normal(0, 1, (1, 30, 3)) + rng.normal(0, 0.01, (10, 30, 3))
    y = np.concatenate([a, c]).astype(np.float32)
    Dy = np.array([fp64.rmsd(y, y, i) for i in range(len(y))])
    assert int(np.argmax(fp64.centroid_scores(Dy))) < 30


def test_nearest_leader_brute_force():
    x = gen.walk(seed=8, n_frames=25, n_atoms=40)
    leaders = x[[3, 11, 19]]
    idx, val = fp64.nearest_leader(x, leaders)
    # each leader is nearest to itself at RMSD 0
    assert list(idx[[3, 11, 19]]) == [0, 1, 2] and np.all(val[[3, 11, 19]] < 1e-6)
    for f in range(len(x)):
        r = [fp64.rmsd(x[f:f + 1], leaders, k)[0] for k in range(3)]  # RMSD is symmetric
        assert idx[f] == int(np.argmin(r)) and abs(val[f] - min(r)) < 1e-9
