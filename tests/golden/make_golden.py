"""Write the golden vectors of the RMSD path from the compiled reference core.

    make -C oracle && python tests/golden/make_golden.py

Runs only where oracle/_ref/libref_rmsd.so was built from the reference's
C++ sources (rmsd/src/{theobald_rmsd,center,rotation}.cpp).  Each fixture
holds seeded inputs and the reference's float32 outputs; the numpy prep of
Trajectory.superpose (core/trajectory.py:1127-1150) is restated here around
the reference's superpose_atom_major loop (rmsd/_rmsd.pyx:620-674).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from mdtraj_amd import gen  # noqa: E402
from oracle import native  # noqa: E402


def superpose_reference(ref, xyz, ref_xyz, frame, atom_indices=None, ref_atom_indices=None):
    """core/trajectory.py:1114-1172 around the compiled superpose loop."""
    if atom_indices is None:
        atom_indices = slice(None)
    if ref_atom_indices is None:
        ref_atom_indices = atom_indices
    xyz = np.array(xyz, dtype=np.float32, copy=True)
    n_frames = xyz.shape[0]
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
    rot = ref.superpose_atom_major(np.ascontiguousarray(ref_align, np.float32),
                                   np.ascontiguousarray(align, np.float32),
                                   np.ascontiguousarray(g_ref, np.float32),
                                   np.ascontiguousarray(g_mobile, np.float32), displace, 0)
    displace += ref_offset
    return displace, rot


def main():
    ref = native.reference()
    out = {}

    # 1. small cube: reference conformation vs 200 noisy rigidly moved frames
    x, r = gen.small_cube(seed=0, n_frames=200, n_atoms=22)
    out["cube_xyz"] = x
    out["cube_ref"] = r
    out["cube_rmsd"] = ref.rmsd(x, r)

    # 2. random walk, 64 frames x 300 atoms, 200-atom subset, vs frame 5
    w = gen.walk(seed=11, n_frames=64, n_atoms=300)
    idx = gen.subset(12, 300, 200)
    out["walk_xyz"] = w
    out["walk_idx"] = idx
    out["walk_rmsd_sel_f5"] = ref.rmsd(w[:, idx, :], w[5, idx, :])
    out["walk_rmsd_all_f0"] = ref.rmsd(w, w[0])
    # md.rmsd(t, t, 3) with atom_indices=None: aliased reference frame
    out["walk_rmsd_self_f3"] = ref.rmsd_self(w, 3)
    cx, cg = ref.center(w)
    out["walk_centered"] = cx
    out["walk_traces"] = cg

    # 3. tests/test_alignment.py:29-36 points, 90 degree rotation, fixed offset
    xyz1 = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [1, 0, 1], [0, 1, 1.0]])
    rotation = np.array([[1, 0, 0], [0, 0, -1], [0, 1, 0]])
    offset = np.array([0.3, -1.1, 2.4])
    xyz2 = rotation.dot(xyz1.T).T + offset
    pair = np.stack([xyz1, xyz2]).astype(np.float32)
    out["align_xyz"] = pair
    out["align_rmsd"] = ref.rmsd(pair, pair[0])

    # 4. superpose: 12 frames x 40 atoms far from the origin, fitted on 25 atoms
    rng = np.random.default_rng(21)
    base = rng.standard_normal((40, 3)) + 7.0
    frames = gen.random_rigid(rng, base[None] + 0.05 * rng.standard_normal((12, 40, 3)))
    frames = frames.astype(np.float32)
    sel = np.sort(rng.choice(40, 25, replace=False))
    out["sup_xyz"] = frames
    out["sup_sel"] = sel
    disp, rot = superpose_reference(ref, frames, frames, 4, sel)
    out["sup_moved_sel_f4"] = disp
    out["sup_rot_sel_f4"] = rot
    disp, rot = superpose_reference(ref, frames, frames, 0)
    out["sup_moved_all_f0"] = disp
    out["sup_rot_all_f0"] = rot

    # 5. degenerate systems: one atom; all zeros
    one = rng.standard_normal((3, 1, 3)).astype(np.float32)
    out["one_atom_xyz"] = one
    out["one_atom_rmsd"] = ref.rmsd(one, one[0])
    zeros = np.zeros((2, 5, 3), np.float32)
    out["zeros_rmsd"] = ref.rmsd(zeros, zeros[0])

    path = os.path.join(HERE, "rmsd_reference_golden.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
