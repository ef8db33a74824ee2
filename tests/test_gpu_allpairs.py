"""GPU all-pairs RMSD: rows equal one-vs-all, condensed layout, accuracy vs
the FP64 oracle per mode (tolerances stated per mode)."""
import numpy as np
import pytest

import mdtraj_amd as md
from mdtraj_amd import _device as dev
from mdtraj_amd import gen
from oracle import fp64

pytestmark = pytest.mark.gpu

# Plain FP32 kernel: M accumulated in FP32 over the atoms, error ~ eps32 K |x|^2
# / (N r); measured 3.5e-4 nm on the 300 x 500 walk (near-identical frames).
# f16x3 / bf16x3: one FP32 accumulator of |M| ~ N sigma^2 -> ~2e-4 nm for
# near-identical frames (CPU emulation 1.7e-4 / 2.2e-4 on a 200 x 500 walk).
# f16exact: exact hi.hi accumulation -> emulated 7.4e-7 nm, bar 1e-5 nm.
# bf16exact: BF16 lo parts carry 8 bits -> coordinates to ~2^-16 of max |x|.
TOL = {"fp32": 1e-3, "f16x3": 1e-3, "bf16x3": 2e-3, "f16exact": 1e-5, "bf16exact": 2e-4}


@pytest.mark.parametrize("mode", ["fp32", "f16x3", "bf16x3", "f16exact", "bf16exact"])
@pytest.mark.parametrize("F,N,n_sel", [(2, 5, None), (65, 30, None), (200, 100, 77), (300, 500, None)])
def test_all_pairs_vs_fp64(gpu, mode, F, N, n_sel):
    x = gen.walk(seed=F + N, n_frames=F, n_atoms=N)
    idx = None if n_sel is None else gen.subset(1, N, n_sel)
    got = md.rmsd_matrix(md.Trajectory(x), atom_indices=idx, mode=mode)
    want = fp64.condensed(fp64.rmsd_matrix(x, atom_indices=idx))
    assert got.shape == want.shape
    err = np.abs(got - want).max()
    print(f"all-pairs {mode} F={F} N={N}: max err {err:.3e} nm")
    assert err < TOL[mode]


def test_square_and_rows(gpu):
    x = gen.walk(seed=9, n_frames=50, n_atoms=40)
    D = md.rmsd_matrix(md.Trajectory(x), mode="fp32", square=True)
    assert D.shape == (50, 50) and np.all(np.diag(D) == 0) and np.array_equal(D, D.T)
    row = md.rmsd(md.Trajectory(x), md.Trajectory(x), 17)
    assert np.abs(D[17] - row).max() < TOL["fp32"]


@pytest.mark.parametrize("mode", ["f16exact", "f16x3"])
@pytest.mark.parametrize("QA,FB,N", [(1, 100, 40), (5, 300, 200), (64, 130, 500), (70, 65, 33)])
def test_cross_vs_fp64(gpu, mode, QA, FB, N):
    from mdtraj_amd.allpairs import SplitSet, rmsd_cross

    x = gen.walk(seed=QA + FB, n_frames=QA + FB, n_atoms=N)
    qa, pool = x[:QA], x[QA:]
    got = rmsd_cross(SplitSet(qa, mode=mode), SplitSet(pool, mode=mode)).cpu().numpy()
    D = fp64.rmsd_matrix(x)[:QA, QA:]
    err = np.abs(got - D).max()
    print(f"cross {mode} {QA}x{FB} N={N}: max err {err:.3e} nm")
    assert err < TOL[mode]


def test_cross_batch_independence(gpu):
    """A query's row is bit-identical whether it is scored alone or with
    other queries in the same call."""
    from mdtraj_amd.allpairs import SplitSet, rmsd_cross

    x = gen.walk(seed=77, n_frames=300, n_atoms=120)
    pool = SplitSet(x[40:])
    alone = rmsd_cross(SplitSet(x[7:8]), pool)
    mixed = rmsd_cross(SplitSet(x[:40]), pool)
    assert np.array_equal(dev.to_host(alone)[0], dev.to_host(mixed)[7])


@pytest.mark.parametrize("mode", ["f16exact", "f16x3", "bf16exact", "bf16x3"])
def test_half_tiles_equal_full_tiles(gpu, mode):
    """The 64 x 32 half-tile kernel (two workgroups per CU) and its variants
    (26-slot rows, register epilogue) run the same MFMA sequence per pair
    block and the same QCP as the 64 x 64 kernel: the triangle, the
    rectangle and the nearest column agree bit for bit, over sizes that leave
    ragged tiles and a 32-column half past the end."""
    from mdtraj_amd import _lib
    from mdtraj_amd.allpairs import SplitSet, nearest, rmsd_cross

    x = gen.walk(seed=61, n_frames=333, n_atoms=150)
    got = {}
    try:
        for tile in (1, 2, 3, 4, 5):
            _lib.call("hr_set_tuning", _lib.HR_TUNE_PAIRS_TILE, tile)
            tri = md.rmsd_matrix(x, mode=mode)
            a, b = SplitSet(x[:97], mode=mode), SplitSet(x[97:], mode=mode)
            rect = dev.to_host(rmsd_cross(a, b))
            idx, val = nearest(a, b)
            got[tile] = (tri, rect, dev.to_host(idx), dev.to_host(val))
    finally:
        _lib.call("hr_set_tuning", _lib.HR_TUNE_PAIRS_TILE, 0)
    for tile in (2, 3, 4, 5):
        for u, v in zip(got[1], got[tile]):
            assert np.array_equal(u, v), tile
    want = fp64.condensed(fp64.rmsd_matrix(x))
    assert np.abs(got[2][0] - want).max() < TOL[mode]
