"""Parity at the benchmark workloads' own sizes.

* all-pairs: the 2,000-frame x 500-atom FP64 subset in every mode (bounds per
  mode below), and one full 20,000 x 500 split-precision launch checked by
  properties: sampled rows equal one-vs-all ``md.rmsd`` within 1e-5 nm,
  every entry finite, the last condensed entry written;
* batched: systems up to 5,000 atoms x 2,000 frames under the autotuned work
  unit (64 frames, 192 KiB, 3 waves/SIMD), every system against the FP64
  oracle on sampled frames;
* one-vs-all: 100,000 x 2,000 with the seeded 1,500-atom selection, against
  the FP64 oracle on a seeded 2,000-frame sample.
"""
import numpy as np
import pytest

import mdtraj_amd as md
from mdtraj_amd import _device as dev
from mdtraj_amd import gen
from oracle import fp64

pytestmark = pytest.mark.gpu

# Max |GPU - FP64| (nm) on the 2,000-frame random-walk subset, per mode:
#   f16exact   exact hi.hi accumulation (~1e-6 measured): the 1e-5 bar
#   bf16exact  BF16 lo parts keep 8 bits (~1e-5 measured on small walks)
#   f16x3 / bf16x3 / fp32  one FP32 accumulator of |M| ~ N sigma^2 per entry:
#              ~2e-4 .. 4e-4 nm between adjacent walk frames
SUBSET_TOL = {"f16exact": 1e-5, "bf16exact": 5e-5, "f16x3": 1e-3, "bf16x3": 2e-3, "fp32": 1e-3}


@pytest.fixture(scope="module")
def subset_2000():
    x = gen.walk(seed=2, n_frames=2000, n_atoms=500)
    want = fp64.condensed(fp64.rmsd_matrix(x))
    return x, want


@pytest.mark.parametrize("mode", ["f16exact", "bf16exact", "f16x3", "bf16x3", "fp32"])
def test_all_pairs_2000_subset_every_mode(gpu, subset_2000, mode):
    x, want = subset_2000
    print(f"all-pairs 2,000 x 500, mode {mode}", flush=True)
    got = md.rmsd_matrix(x, mode=mode)
    assert got.shape == want.shape == (2000 * 1999 // 2,)
    err = float(np.abs(got - want).max())
    print(f"all-pairs 2,000 x 500 {mode}: max |gpu - fp64| = {err:.3e} nm")
    assert err < SUBSET_TOL[mode]


def test_all_pairs_20000_full_launch(gpu):
    torch = gpu
    F, N = 20_000, 500
    print("all-pairs 20,000 x 500 f16exact, one launch", flush=True)
    X = gen.walk_device(seed=2, n_frames=F, n_atoms=N)
    n_pairs = F * (F - 1) // 2
    out = torch.full((n_pairs,), float("nan"), dtype=torch.float32, device="cuda")
    md.rmsd_matrix(X, mode="f16exact", out=out)
    torch.cuda.synchronize()
    assert bool(torch.isfinite(out).all())
    assert not bool(torch.isnan(out[-1]))
    assert float(out.min()) > 0.0
    rows = [0, 1, 313, 9_999, 12_345, F - 2, F - 1]
    t = md.Trajectory(X)
    for i in rows:
        # row i of the condensed upper triangle: (j, i) for j < i, (i, j) for j > i
        j = torch.arange(F, device="cuda", dtype=torch.int64)
        lo, hi = torch.minimum(j, torch.tensor(i, device="cuda")), torch.maximum(
            j, torch.tensor(i, device="cuda"))
        c = lo * F - lo * (lo + 1) // 2 + (hi - lo - 1)
        keep = j != i
        got = out[c[keep]].cpu().numpy()
        want = md.rmsd(t, t, i).cpu().numpy()[keep.cpu().numpy()]
        err = float(np.abs(got - want).max())
        assert err < 1e-5, (i, err)
    del out, X


def test_batched_config_scale_autotuned_shape(gpu):
    from mdtraj_amd.batch import BatchRMSD

    sizes = [(2000, 5000), (1999, 4999), (10, 50), (1500, 3000), (64, 5000), (300, 777),
             (2000, 51), (11, 4096)]
    xs = [gen.small_cube(seed=500 + s, n_frames=f, n_atoms=n, sigma=0.03)[0]
          for s, (f, n) in enumerate(sizes)]
    refs = [0, 1998, 9, 700, 63, 0, 1, 5]
    print("batched: 8 systems up to 5,000 atoms x 2,000 frames, shape (64, 192, 3)",
          flush=True)
    b = BatchRMSD(xs, refs, frames_per_unit=64, unit_kib=192, waves=3)
    res = [dev.to_host(r) for r in b.run()]
    rng = np.random.default_rng(0)
    for x, r, got in zip(xs, refs, res):
        F = x.shape[0]
        pick = np.unique(np.concatenate([[0, r, F - 1], rng.integers(0, F, size=min(F, 200))]))
        want = fp64.rmsd(x[pick], x[r:r + 1], 0)
        assert got.shape == (F,)
        assert np.abs(got[pick] - want).max() < 1e-5
        assert got[r] == 0.0
    b.close()


def test_one_vs_all_config_scale(gpu):
    torch = gpu
    F, N, NSEL = 100_000, 2000, 1500
    print("one-vs-all 100,000 x 2,000, 1,500 selected", flush=True)
    X = gen.walk_device(seed=1, n_frames=F, n_atoms=N)
    idx = gen.subset(2, N, NSEL)
    t = md.Trajectory(X)
    got = md.rmsd(t, t, 0, atom_indices=idx).cpu().numpy()
    pick = np.sort(np.random.default_rng(7).choice(F, size=2000, replace=False))
    sample = X[torch.as_tensor(pick, device="cuda")].cpu().numpy()
    want = fp64.rmsd(sample, X[0:1].cpu().numpy(), 0, idx)
    err = float(np.abs(got[pick] - want).max())
    print(f"one-vs-all 100,000 x 2,000: max |gpu - fp64| on 2,000 frames = {err:.3e} nm")
    assert err < 1e-5
    assert got[0] == 0.0 and np.all(np.isfinite(got))
    del X
