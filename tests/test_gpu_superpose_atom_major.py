"""GPU parity of the low-level ``_rmsd.superpose_atom_major`` route.

The reference's ``Trajectory.superpose`` (core/trajectory.py:1114-1172) does
its own numpy preparation and then calls ``_rmsd.superpose_atom_major``
(rmsd/_rmsd.pyx:620-674): it centres a fancy-index copy of the fitted atoms
and the whole displaced trajectory on the fitted atoms' FP64 centroid, centres
a (1, n, 3) copy of the reference selection, forms float32 traces with
``einsum``, calls the kernel with ``target_frame=0``, and adds the reference
centroid back.  A maintainer who swaps ``mdtraj._rmsd`` for this package's
module reaches ``hr_superpose(align_xyz != NULL, translate=0)`` through exactly
that sequence, so it is driven here the same way.

Bars: moved coordinates within 2e-5 nm of the FP64 oracle and within 5e-5 nm
of the compiled reference's float32 output (the golden ``sup_moved_*``
arrays), with the rotation pass split into 1 MiB chunks and in one pass.
"""
import numpy as np
import pytest

from mdtraj_amd import _lib, _rmsd, gen
from oracle import fp64

pytestmark = pytest.mark.gpu


def reference_superpose_body(xyz, ref_xyz, frame, atom_indices=None, ref_atom_indices=None,
                             rmsd_module=_rmsd):
    """The numpy steps of core/trajectory.py:1114-1172 around the module's
    ``superpose_atom_major``; returns the moved (F, N, 3) float32 array."""
    if atom_indices is None:
        atom_indices = slice(None)
    elif len(atom_indices) == 0:
        raise ValueError("Number of atom indices must be greater than 0")
    if ref_atom_indices is None:
        ref_atom_indices = atom_indices
    n_frames = xyz.shape[0]
    # fitted atoms: a copy under fancy indexing, the same memory under a slice
    align = np.asarray(xyz[:, atom_indices, :], order="c")
    displace = np.asarray(xyz, order="c")
    ref_align = np.array(ref_xyz[frame, ref_atom_indices, :], copy=True, order="c").reshape(1, -1, 3)
    offset = np.mean(align, axis=1, dtype=np.float64).reshape(n_frames, 1, 3)
    align -= offset
    if align.ctypes.data != displace.ctypes.data:
        displace -= offset
    ref_offset = ref_align[0].astype("float64").mean(0)
    ref_align[0] -= ref_offset
    g_mobile = np.einsum("ijk,ijk->i", align, align)
    g_ref = np.einsum("ijk,ijk->i", ref_align, ref_align)
    rmsd_module.superpose_atom_major(ref_align, align, g_ref, g_mobile, displace, 0)
    if not np.any(displace[0]):
        raise OverflowError("potential overflow/underflow during superpose()")
    displace += ref_offset
    return displace


@pytest.fixture(params=[0, 1], ids=["one-pass", "1MiB-chunks"])
def chunk_mib(request, gpu):
    mib = request.param
    _lib.call("hr_set_tuning", _lib.HR_TUNE_SUPERPOSE_CHUNK_MIB, mib)
    yield mib
    _lib.call("hr_set_tuning", _lib.HR_TUNE_SUPERPOSE_CHUNK_MIB, 0)


def test_swap_body_vs_golden_and_fp64(chunk_mib, golden):
    x, sel = golden["sup_xyz"], golden["sup_sel"]
    moved = reference_superpose_body(x.copy(), x, 4, sel)
    assert moved.dtype == np.float32 and moved.shape == x.shape
    assert np.abs(moved - fp64.superpose(x, x, 4, sel)).max() < 2e-5
    assert np.abs(moved - golden["sup_moved_sel_f4"]).max() < 5e-5
    moved = reference_superpose_body(x.copy(), x, 0)
    assert np.abs(moved - fp64.superpose(x, x, 0)).max() < 2e-5
    assert np.abs(moved - golden["sup_moved_all_f0"]).max() < 5e-5


def test_direct_call_displace_wider_than_align(chunk_mib):
    """The call as the reference body makes it, with 1,200 frames of 1,000
    atoms displaced and a 300-atom fitted subset (about 4.6 MB of moved
    coordinates, so 1 MiB chunks split the rotation pass five ways)."""
    F, N = 1200, 1000
    w = gen.walk(seed=21, n_frames=F, n_atoms=N)
    sel = np.sort(np.random.default_rng(21).choice(N, size=300, replace=False))
    frame = 17
    align = np.ascontiguousarray(w[:, sel, :])
    displace = w.copy()
    ref_align = np.array(w[frame, sel, :], copy=True).reshape(1, -1, 3)
    offset = align.mean(axis=1, dtype=np.float64).reshape(F, 1, 3)
    align -= offset
    displace -= offset
    ref_offset = ref_align[0].astype(np.float64).mean(0)
    ref_align[0] -= ref_offset
    g_m = np.einsum("ijk,ijk->i", align, align)
    g_r = np.einsum("ijk,ijk->i", ref_align, ref_align)
    align_before = align.copy()
    ref_before = ref_align.copy()
    _rmsd.superpose_atom_major(ref_align, align, g_r, g_m, displace, 0)
    # only the displaced array moves
    assert np.array_equal(align, align_before) and np.array_equal(ref_align, ref_before)
    want = fp64.superpose(w, w, frame, sel)
    assert np.abs((displace + ref_offset) - want).max() < 2e-5
    # the fitted atoms land on the reference: their RMSD equals the QCP RMSD
    moved_sel = (displace + ref_offset)[:, sel, :].astype(np.float64)
    disp = np.sqrt(((moved_sel - w[frame, sel].astype(np.float64)) ** 2).sum(2).mean(1))
    assert np.abs(disp - fp64.rmsd(w, w, frame, sel)).max() < 2e-5


def test_target_frame_other_than_zero(gpu):
    """superpose_atom_major fits onto xyz_align_target[target_frame]."""
    w = gen.walk(seed=22, n_frames=50, n_atoms=64)
    c = (w - w.mean(axis=1, keepdims=True, dtype=np.float64)).astype(np.float32)
    g = np.einsum("ijk,ijk->i", c, c)
    d = c.copy()
    _rmsd.superpose_atom_major(c, c.copy(), g, g, d, 9)
    want = fp64.superpose(c, c, 9) - c[9].astype(np.float64).mean(0)
    assert np.abs(d - want).max() < 2e-5


def test_validation_messages(gpu):
    a = np.zeros((3, 10, 3), np.float32)
    with pytest.raises(ValueError, match="same number of atoms"):
        _rmsd.superpose_atom_major(a[:1, :9], a, None, None, a.copy(), 0)
    with pytest.raises(ValueError, match="same number of frames"):
        _rmsd.superpose_atom_major(a[:1], a, None, None, a[:2].copy(), 0)
