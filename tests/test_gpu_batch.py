"""Batched plan: each system's RMSDs equal the FP64 oracle's one-vs-all for
that system (1e-5 nm), the reference frame reads exactly 0, and a captured
graph gives the same result on every replay (eager work in between)."""
import numpy as np
import pytest

from mdtraj_amd import gen
from mdtraj_amd import _device as dev
from mdtraj_amd.batch import BatchRMSD
from oracle import fp64

pytestmark = pytest.mark.gpu


def systems(seed=0, n=24):
    rng = np.random.default_rng(seed)
    out = []
    for s in range(n):
        F = int(rng.integers(1, 140))
        N = int(rng.choice([1, 2, 3, 17, 50, 64, 65, 200, 700, 1300]))
        x, _ = gen.small_cube(seed=1000 + s, n_frames=F, n_atoms=N, sigma=0.03)
        out.append(x)
    return out


@pytest.mark.parametrize("schedule", [1, 2, 3, 4])
@pytest.mark.parametrize("fpu,kib,waves", [(64, 0, 3), (16, 0, 3), (7, 0, 4), (64, 0, 4),
                                           (64, 8, 3), (8, 64, 3)])
def test_batch_vs_fp64(gpu, fpu, kib, waves, schedule):
    """Every work split (ticket units, static cost split per wave, per
    workgroup with the reference in LDS) over systems of 1 - 1,300 atoms and
    1 - 139 frames."""
    from mdtraj_amd import _lib

    xs = systems()
    refs = [int(np.random.default_rng(i).integers(0, x.shape[0])) for i, x in enumerate(xs)]
    _lib.call("hr_set_tuning", _lib.HR_TUNE_BATCH_SCHEDULE, schedule)
    try:
        b = BatchRMSD(xs, refs, frames_per_unit=fpu, unit_kib=kib, waves=waves)
    finally:
        _lib.call("hr_set_tuning", _lib.HR_TUNE_BATCH_SCHEDULE, 0)
    res = [r.cpu().numpy() for r in b.run()]
    for x, r, got in zip(xs, refs, res):
        want = fp64.rmsd(x, x, r)
        assert got.shape == want.shape
        assert np.abs(got - want).max() < 1e-5
        assert got[r] == 0.0
    b.close()


@pytest.mark.parametrize("schedule", [3, 4])
@pytest.mark.parametrize("waves", [3, 4])
def test_block_split_reference_beyond_lds(gpu, waves, schedule):
    """The per-workgroup split stages references of up to ~4,400 atoms (3
    waves/SIMD) or ~3,400 (4) in LDS, the per-CU split up to ~13,600; the
    rest of a larger reference is read from global memory.  Mixed in one
    plan, both paths match the FP64 oracle."""
    xs = []
    for s, (F, N) in enumerate([(9, 3500), (70, 40), (5, 5000), (33, 4400), (3, 9000), (65, 1),
                                (2, 15000), (130, 700)]):
        x, _ = gen.small_cube(seed=2000 + s, n_frames=F, n_atoms=N, sigma=0.03)
        xs.append(x)
    refs = [x.shape[0] // 2 for x in xs]
    b = BatchRMSD(xs, refs, frames_per_unit=64, unit_kib=0, waves=waves, schedule=schedule)
    res = [dev.to_host(r) for r in b.run()]
    for x, r, got in zip(xs, refs, res):
        want = fp64.rmsd(x, x, r)
        assert np.abs(got - want).max() < 1e-5
        assert got[r] == 0.0
    b.close()


def test_batch_graph_replays_are_identical(gpu):
    torch = gpu
    xs = systems(seed=3, n=16)
    b = BatchRMSD(xs, frames_per_unit=32, unit_kib=0, waves=3)
    first = [r.clone() for r in b.run()]
    torch.cuda.synchronize()
    for k in range(4):
        # eager work between replays, then the replay must match exactly
        junk = torch.randn(1 << 20, device="cuda").sum()
        for o in b.out:
            o.fill_(-1.0)
        again = b.run()
        torch.cuda.synchronize()
        for a, f in zip(again, first):
            assert np.array_equal(dev.to_host(a), dev.to_host(f))
        del junk
    b.close()


def test_autotuner_candidates_agree(gpu):
    from mdtraj_amd import autotune

    xs = systems(seed=5, n=12)
    best, results = autotune.tune_batch(xs, candidates=[(16, 0, 3, 1), (64, 32, 4, 1),
                                                        (64, 0, 3, 3)], reps=2)
    assert best in results and all(v > 0 for v in results.values())
    a = BatchRMSD(xs, frames_per_unit=16, unit_kib=0, waves=3, schedule=1).run()
    b = BatchRMSD(xs, frames_per_unit=64, unit_kib=32, waves=4, schedule=1).run()
    for u, v in zip(a, b):
        assert np.array_equal(dev.to_host(u), dev.to_host(v))
