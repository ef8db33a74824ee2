"""LP-RMSD on the GPU (hr_lprmsd) against the reference's own code and the
oracle.

tests/golden/lprmsd_golden.npz holds the reference's values (its Munkres /
euclidean_permutation / QCP loop compiled from source).  The GPU computes in
FP64 from float32 input where the reference computes in float32, so values
agree to ~1e-5 nm, and the reference's own float32 noise at RMSD 0 (up to
~5e-4 nm) is allowed there.  The reference's test_lprmsd.py cases, the
superposed coordinates, groups wider than a wave, device-tensor input and
the host-side limits are covered too."""
import os

import numpy as np
import pytest

import mdtraj_amd as md
from oracle import fp64
from oracle import lprmsd as olp

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "lprmsd_golden.npz"))


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
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]],
                    np.float32)


def test_against_reference_golden(gpu):
    mk = golden_cases()
    for name, target, ref, frame, sel, groups, sup in mk.cases():
        t = np.array(target, np.float32, copy=True)
        got = md.lprmsd(t, np.asarray(ref, np.float32), frame, sel, groups, superpose=sup)
        want = GOLD[f"{name}/lprmsd"]
        assert got.dtype == np.float32 and got.shape == want.shape
        assert np.abs(got - want).max() < 6e-4, name
        assert np.abs(got - want)[want > 1e-2].max(initial=0) < 2e-5, name
        if sup:  # target centred and rotated in place, as the reference
            assert np.abs(t - GOLD[f"{name}/moved"]).max() < 1e-4


def test_reference_test_cases(gpu):
    rng = np.random.default_rng(21)
    ref = rng.normal(size=(1, 10, 3)).astype(np.float32)
    assert md.lprmsd(ref.copy(), ref)[0] < 1e-3  # test_lprmsd_null
    assert md.lprmsd(ref[:, rng.permutation(10)].copy(), ref)[0] < 1e-3  # test_lprmsd_0
    ref = rng.normal(size=(1, 50, 3)).astype(np.float32)
    R = rotation(rng)
    assert md.lprmsd((ref @ R).astype(np.float32), ref, permute_groups=[[]])[0] < 1e-2  # _1
    mp = np.concatenate((rng.permutation(10), 10 + np.arange(40)))
    new = (ref[:, mp] @ R).astype(np.float32)
    assert md.lprmsd(new, ref, permute_groups=[np.arange(10)])[0] < 1e-2  # test_lprmsd_2
    # permute_groups=[[]] is the plain RMSD (test_lprmsd_4, _5)
    x = (rng.normal(0, 1, (1, 60, 3)) + rng.normal(0, 0.05, (9, 60, 3))).astype(np.float32)
    sel = rng.permutation(60)[:55]
    got = md.lprmsd(x, x, 0, sel, [[]])
    assert np.abs(got - md.rmsd(x, x, 0, np.sort(sel))).max() < 1e-5
    y = x.copy()
    got = md.lprmsd(y, x, 0, permute_groups=[[]], superpose=True)
    assert np.abs(got - md.rmsd(x, x, 0)).max() < 1e-5


def test_against_oracle_with_wide_groups(gpu):
    rng = np.random.default_rng(31)
    F, N = 24, 400
    base = rng.normal(0, 1.5, (N, 3))
    x = (base[None] + rng.normal(0, 0.12, (F, N, 3))).astype(np.float32)
    for f in range(F):  # relabel inside the groups and rotate every frame
        x[f, :150] = x[f, rng.permutation(150)]
        x[f, 200:270] = x[f, 200 + rng.permutation(70)]
        x[f] = x[f] @ rotation(rng)
    groups = [np.arange(150), np.arange(200, 270)]
    sel = np.r_[0:380]
    want = olp.lprmsd(x, x, 4, sel, groups)
    got = md.lprmsd(x, x, 4, sel, groups)
    ok = want > 1e-2
    assert np.abs(got - want)[ok].max() < 2e-5
    assert got[4] < 1e-3
    # the default (one group of every atom, no first fit)
    want = olp.lprmsd(x[:6, :200], x[:6, :200], 1)
    got = md.lprmsd(x[:6, :200].copy(), x[:6, :200].copy(), 1)
    assert np.abs(got - want)[want > 1e-2].max() < 2e-5


def test_device_input_and_superpose_in_place(gpu):
    import torch

    rng = np.random.default_rng(41)
    x = (rng.normal(0, 1, (1, 40, 3)) + rng.normal(0, 0.1, (5, 40, 3))).astype(np.float32)
    groups = [np.arange(10), np.arange(20, 30)]
    want, moved = olp.lprmsd(x, x, 0, None, groups, superpose=True)
    t = torch.from_numpy(x.copy()).cuda()
    got = md.lprmsd(t, x, 0, None, groups, superpose=True)
    assert got.is_cuda
    assert np.abs(got.cpu().numpy() - want)[want > 1e-2].max() < 2e-5
    assert np.abs(t.cpu().numpy() - moved).max() < 1e-4
    # without superpose the device input is left alone
    t2 = torch.from_numpy(x.copy()).cuda()
    md.lprmsd(t2, x, 0, None, groups)
    assert np.array_equal(t2.cpu().numpy(), x)


def test_limits_and_validation(gpu):
    x = np.zeros((2, 3000, 3), np.float32)
    with pytest.raises(RuntimeError, match="LDS"):
        md.lprmsd(x, x, 0)  # one group of 3,000 atoms does not fit one wave's LDS
    with pytest.raises(ValueError, match="mutually disjoint"):
        md.lprmsd(x, x, 0, None, [[0, 1], [1, 2]])
    with pytest.raises(ValueError, match="only 2 frames"):
        md.lprmsd(x, x, 5)
