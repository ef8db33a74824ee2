"""Memory-mapped coordinates (the reference's tests/test_rmsd_memmap.py, on a
seeded walk in place of frame0.h5, which the reference ships only inside its
test-data archive).

By default inputs are never modified, so a read-only map works.  With
``set_inplace_centering(True)`` the reference's behaviour returns: a read-only
map raises ``ValueError`` (the reference centres ``target.xyz`` in place,
rmsd/_rmsd.pyx:197,212), a copy-on-write map works and ends centred exactly as
the reference's float32 centring leaves it."""
import numpy as np
import pytest

import mdtraj_amd as md
from mdtraj_amd import gen
from oracle import fp64

pytestmark = pytest.mark.gpu


@pytest.fixture
def saved(tmp_path):
    x = gen.walk(seed=71, n_frames=120, n_atoms=60)
    fn = tmp_path / "temp.npy"
    np.save(fn, x)
    return x, fn


def test_read_only_map_works_by_default(gpu, saved):
    x, fn = saved
    t = md.Trajectory(np.load(fn, mmap_mode="r"))
    got = md.rmsd(t, t, 0)
    assert np.abs(got - fp64.rmsd(x, x, 0)).max() < 1e-5
    assert np.array_equal(np.asarray(t.xyz), x)


def test_reference_semantics_opt_in(gpu, saved, golden):
    x, fn = saved
    prev = md.set_inplace_centering(True)
    try:
        t = md.Trajectory(np.load(fn, mmap_mode="r"))
        with pytest.raises(ValueError):
            md.rmsd(t, t, 0)
        t = md.Trajectory(np.load(fn, mmap_mode="c"))
        got = md.rmsd(t, t, 0)
        assert np.abs(got - fp64.rmsd(x, x, 0)).max() < 1e-5
        # centred in place like the reference's inplace_center_and_trace
        assert np.abs(np.asarray(t.xyz).mean(axis=1)).max() < 1e-5
        # with a selection the reference copies first: nothing is modified
        w = golden["walk_xyz"].copy()
        tw = md.Trajectory(w)
        md.rmsd(tw, tw, 5, atom_indices=golden["walk_idx"])
        assert np.array_equal(w, golden["walk_xyz"])
        # without one, target.xyz ends as the compiled reference centred it
        # (the reference frame, a view of frame 5, is then centred again)
        md.rmsd(tw, tw, 5)
        keep = np.arange(w.shape[0]) != 5
        assert np.array_equal(w[keep], golden["walk_centered"][keep])
        assert np.abs(w[5] - golden["walk_centered"][5]).max() < 1e-6
    finally:
        md.set_inplace_centering(prev)
    assert md.set_inplace_centering(prev) is prev
