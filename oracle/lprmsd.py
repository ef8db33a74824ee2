"""LP-RMSD restated in float64 numpy (test infrastructure only: the checker
for the GPU kernel, never called by the product).

Follows rmsd/_lprmsd.pyx:72-228 step by step.
  selection   atom_indices sorted and unique, permute_groups sorted and
              unique, None meaning one group of every selected atom
              (_lprmsd.pyx:232-271); the distinguishable atoms are the
              selected atoms in no group.
  per frame   (1) gather the selection and centre it;
              (2) when distinguishable atoms exist, fit them alone (each set
                  centred on its own centroid) and rotate the whole centred
                  selection by that rotation;
              (3) for each permute group, assign reference atom i to target
                  atom j minimising the summed squared distances
                  (euclidean_permutation.cpp:17-71 fills the cost matrix
                  A[i, j] = |ref_i - target_j|^2 and solves it with the
                  Hungarian method, Munkres.cpp); atoms in no group keep
                  their labels;
              (4) the QCP RMSD of the relabelled selection against the
                  reference; with superpose, every frame (all atoms, centred)
                  is rotated by rot1 rot2.
The assignment step uses scipy.optimize.linear_sum_assignment, an exact
solver of the same problem: with a unique optimum both give the same labels.
Rotations and RMSDs come from oracle.fp64 (QCP eigenvalue, Kabsch).

Pinned by tests/test_oracle_lprmsd.py against the reference's own code
(oracle/_ref/libref_rmsd.so: its Munkres, euclidean_permutation, centring,
QCP and rotation compiled from source, driven by ref_driver.cpp::ref_lprmsd),
the known answers of the reference's tests/test_lprmsd.py, and a brute-force
search over every relabelling of small groups.
"""
from __future__ import annotations

import numpy as np
from scipy.optimize import linear_sum_assignment

from . import fp64


def normalize(n_atoms, atom_indices=None, permute_groups=None):
    """(atom_indices, groups as indices into the selection, distinguishable
    indices into the selection), validated as _lprmsd.pyx:232-271."""
    if atom_indices is None:
        atom_indices = np.arange(n_atoms)
    else:
        atom_indices = np.unique(np.asarray(atom_indices, dtype=int))
        if not np.all((atom_indices >= 0) & (atom_indices < n_atoms)):
            raise ValueError("atom_indices must be valid positive indices")
    if permute_groups is None:
        groups = [atom_indices]
    else:
        groups = [np.unique(np.asarray(g, dtype=np.int64)) for g in permute_groups]
        for g in groups:
            if len(np.setdiff1d(g, atom_indices)) > 0:
                raise ValueError("The elements in each permute group must be a subset of atom_indices")
        allp = np.concatenate(groups) if groups else np.zeros(0, int)
        if len(np.unique(allp)) != len(allp):  # an atom in two groups
            raise ValueError("permute_groups must be mutually disjoint sets")
    rel = [np.searchsorted(atom_indices, g) for g in groups]
    used = np.concatenate(rel) if rel else np.zeros(0, int)
    dis = np.setdiff1d(np.arange(len(atom_indices)), used)
    return atom_indices, rel, dis


def assignment(ref, tgt, groups):
    """mapping with tgt[mapping] the relabelled target: mapping[i] = j
    assigns reference atom i to target atom j within each group."""
    mapping = np.arange(len(ref))
    for g in groups:
        if len(g) == 0:
            continue
        d = ref[g][:, None, :] - tgt[g][None, :, :]
        cost = (d * d).sum(-1)
        rows, cols = linear_sum_assignment(cost)
        mapping[g[rows]] = g[cols]
    return mapping


def lprmsd(target_xyz, ref_xyz, frame=0, atom_indices=None, permute_groups=None,
           superpose=False):
    """LP-RMSD of every frame of target_xyz (F, N, 3) to ref_xyz[frame];
    with superpose also returns the target rotated as the reference does."""
    target = np.asarray(target_xyz, np.float64)
    atom_indices, groups, dis = normalize(target.shape[1], atom_indices, permute_groups)
    ref = np.asarray(ref_xyz, np.float64)[frame, atom_indices]
    ref = ref - ref.mean(0)
    ref_dis = ref[dis] - ref[dis].mean(0) if len(dis) else None
    out = np.empty(len(target))
    moved = target.copy()
    for f in range(len(target)):
        t = target[f, atom_indices]
        t = t - t.mean(0)
        rot1 = np.eye(3)
        if len(dis):
            td = t[dis] - t[dis].mean(0)
            rot1 = fp64.kabsch_rotation(td[None], ref_dis)[0]
            t = t @ rot1
        t2 = t[assignment(ref, t, groups)]
        out[f] = fp64.rmsd(t2[None], ref[None], 0)[0]
        if superpose:
            rot2 = fp64.kabsch_rotation(t2[None], ref)[0]
            x = moved[f] - moved[f].mean(0)
            moved[f] = x @ (rot1 @ rot2)
    return (out, moved) if superpose else out


def brute_force_assignment_cost(ref, tgt, group):
    """Smallest summed squared distance over every relabelling of one small
    group (exhaustive, for groups of up to 7 atoms)."""
    import itertools

    best = np.inf
    for perm in itertools.permutations(range(len(group))):
        d = ref[group] - tgt[group[list(perm)]]
        best = min(best, float((d * d).sum()))
    return best
