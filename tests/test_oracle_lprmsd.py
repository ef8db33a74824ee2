"""The LP-RMSD oracle, pinned.

Against the reference's own code: tests/golden/lprmsd_golden.npz holds what
its compiled Munkres / euclidean_permutation / QCP loop returned
(tests/golden/make_lprmsd_golden.py), and where oracle/_ref is built, more
random cases are compared live.  Against known answers: the reference's
tests/test_lprmsd.py cases (a copy, a relabelled copy, a rotated copy with no
permutable atoms, a rotated and partly relabelled copy: all 0) and
permute_groups=[[]] reducing to the plain RMSD.  Against brute force: the
assignment step's cost equals the minimum over every relabelling of small
groups.  A planted mistake (assigning target to reference, the transposed
cost) is caught."""
import ctypes
import os

import numpy as np
import pytest

from oracle import fp64
from oracle import lprmsd as olp

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "lprmsd_golden.npz"))
REF = os.path.join(HERE, "..", "oracle", "_ref", "libref_rmsd.so")


def golden_cases():
    import importlib.util

    spec = importlib.util.spec_from_file_location(
        "mk", os.path.join(HERE, "golden", "make_lprmsd_golden.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    return mk


def rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def test_matches_reference_golden():
    mk = golden_cases()
    for name, target, ref, frame, sel, groups, sup in mk.cases():
        got = olp.lprmsd(target, ref, frame, sel, groups, superpose=sup)
        want = GOLD[f"{name}/lprmsd"]
        if sup:
            got, moved = got
            assert np.abs(moved - GOLD[f"{name}/moved"]).max() < 1e-4
        # float32 reference arithmetic against float64: ~1e-6 nm, plus the
        # reference's float32 self-RMSD noise (up to ~5e-4 nm at 0)
        assert np.abs(got - want).max() < 6e-4, name
        assert np.abs(got - want)[want > 1e-2].max(initial=0) < 2e-5, name


def test_reference_known_answers():
    rng = np.random.default_rng(5)
    ref = rng.normal(size=(1, 10, 3)).astype(np.float32)
    assert olp.lprmsd(ref, ref)[0] < 1e-6  # test_lprmsd_null
    assert olp.lprmsd(ref[:, rng.permutation(10)], ref)[0] < 1e-6  # test_lprmsd_0
    ref = rng.normal(size=(1, 50, 3)).astype(np.float32)
    R = rotation(rng)
    assert olp.lprmsd(ref @ R, ref, permute_groups=[[]])[0] < 1e-6  # test_lprmsd_1
    mp = np.concatenate((rng.permutation(10), 10 + np.arange(40)))
    assert olp.lprmsd(ref[:, mp] @ R, ref, permute_groups=[np.arange(10)])[0] < 1e-6  # _2
    # permute_groups=[[]]: no relabelling, the plain RMSD (test_lprmsd_4/_5)
    x = rng.normal(size=(7, 30, 3))
    sel = rng.permutation(30)[:25]
    assert np.allclose(olp.lprmsd(x, x, 2, sel, [[]]), fp64.rmsd(x, x, 2, np.sort(sel)), atol=1e-7)


def test_assignment_is_the_brute_force_optimum():
    rng = np.random.default_rng(9)
    for trial in range(20):
        k = 2 + trial % 6
        ref = rng.normal(size=(k + 3, 3))
        tgt = ref[np.r_[rng.permutation(k), k + np.arange(3)]] + rng.normal(0, 0.4, (k + 3, 3))
        g = np.arange(k)
        m = olp.assignment(ref, tgt, [g])
        cost = float(((ref[g] - tgt[m[g]]) ** 2).sum())
        assert abs(cost - olp.brute_force_assignment_cost(ref, tgt, g)) < 1e-12
        assert np.array_equal(m[k:], np.arange(k, k + 3))  # atoms in no group keep their labels
        # the transposed problem (target rows, reference columns) relabels the
        # other side: applied as if it were the mapping, it is not the optimum
        # once the optimal mapping is not its own inverse
        mt = olp.assignment(tgt, ref, [g])
        if not np.array_equal(m[mt[g]], g):
            assert float(((ref[g] - tgt[mt[g]]) ** 2).sum()) > cost + 1e-12 or np.array_equal(mt, m)


def test_validation_messages():
    with pytest.raises(ValueError, match="valid positive indices"):
        olp.normalize(5, [0, 7])
    with pytest.raises(ValueError, match="subset of atom_indices"):
        olp.normalize(5, [0, 1, 2], [[3]])
    with pytest.raises(ValueError, match="mutually disjoint"):
        olp.normalize(5, None, [[0, 1], [1, 2]])


@pytest.mark.skipif(not os.path.exists(REF), reason="reference library not built here")
def test_live_against_reference_code():
    mk = golden_cases()
    rng = np.random.default_rng(13)
    for trial in range(4):
        F, N = 5, 30 + 10 * trial
        x = (rng.normal(0, 0.7, (1, N, 3)) + rng.normal(0, 0.08, (F, N, 3))).astype(np.float32)
        groups = [list(range(0, 8 + trial)), list(range(15, 22))]
        sel = np.r_[0:N - 3]
        want, _ = mk.ref_lprmsd(x, x, 1, sel, groups, False)
        got = olp.lprmsd(x, x, 1, sel, groups)
        ok = want > 1e-2
        assert np.abs(got - want)[ok].max() < 2e-5
