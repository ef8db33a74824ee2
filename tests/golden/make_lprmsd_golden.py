"""Writes tests/golden/lprmsd_golden.npz: LP-RMSD values computed by the
reference's own code (oracle/_ref/libref_rmsd.so, compiled from
rmsd/src/{theobald_rmsd,center,rotation,Munkres,euclidean_permutation}.cpp
and driven by oracle/ref_driver.cpp::ref_lprmsd, the loop of
rmsd/_lprmsd.pyx:137-228).

Cases (seeded): a random-walk trajectory with three permute groups and
distinguishable atoms, with and without superpose; the default single group
of every atom; a scrambled-and-rotated copy of a frame (LP-RMSD near 0).
Run from the repository root after `make -C oracle ref`:
python tests/golden/make_lprmsd_golden.py
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
P = ctypes.c_void_p


def cases():
    rng = np.random.default_rng(77)
    base = rng.normal(0, 0.6, (48, 3))
    steps = rng.normal(0, 0.04, (12, 48, 3))
    walk = (base[None] + np.cumsum(steps, axis=0)).astype(np.float32)
    groups = [list(range(0, 9)), list(range(14, 26)), [30, 31, 32, 33]]
    sel = list(range(0, 40))
    scr = walk[3].copy()
    perm = np.r_[rng.permutation(12), 12 + np.arange(36)]
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    scrambled = (scr[perm] @ R).astype(np.float32)[None]
    return [
        ("groups", walk, walk, 2, sel, groups, False),
        ("groups_superpose", walk, walk, 2, sel, groups, True),
        ("all_permutable", walk[:6], walk, 0, None, None, False),
        ("scrambled", scrambled, walk, 3, None, [list(range(12))], False),
    ]


def normalize(n_atoms, atom_indices, permute_groups):
    ai = np.arange(n_atoms) if atom_indices is None else np.unique(atom_indices)
    groups = [ai] if permute_groups is None else [np.unique(np.asarray(g, int)) for g in permute_groups]
    return ai, [np.searchsorted(ai, g) for g in groups]


def ref_lprmsd(target, ref, frame, atom_indices, permute_groups, superpose):
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_rmsd.so"))
    target = np.ascontiguousarray(target, np.float32).copy()
    ai, rel = normalize(target.shape[1], atom_indices, permute_groups)
    flat = np.concatenate(rel).astype(np.int32) if rel else np.zeros(0, np.int32)
    off = np.cumsum([0] + [len(g) for g in rel]).astype(np.int32)
    rf = np.ascontiguousarray(ref[frame], np.float32)
    ai32 = ai.astype(np.int32)
    out = np.zeros(len(target), np.float32)
    lib.ref_lprmsd(target.ctypes.data_as(P), len(target), target.shape[1], rf.ctypes.data_as(P),
                   ai32.ctypes.data_as(P), len(ai32), flat.ctypes.data_as(P) if flat.size else None,
                   off.ctypes.data_as(P), len(rel), int(superpose), out.ctypes.data_as(P))
    return out, target


def main():
    store = {}
    for name, target, ref, frame, sel, groups, sup in cases():
        out, moved = ref_lprmsd(target, ref, frame, sel, groups, sup)
        store[f"{name}/target"] = np.asarray(target, np.float32)
        store[f"{name}/ref"] = np.asarray(ref, np.float32)
        store[f"{name}/lprmsd"] = out
        if sup:
            store[f"{name}/moved"] = moved
        print(name, out[:4])
    np.savez_compressed(os.path.join(HERE, "lprmsd_golden.npz"), **store)


if __name__ == "__main__":
    main()
