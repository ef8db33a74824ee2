"""FP64 numpy oracle for the RMSD hot path (test infrastructure only).

Every function works in float64 on float32 (or float64) inputs and restates
the mathematics the reference implements; the reference citation is given per
function (paths relative to the mdtraj source tree).

Pinned by (tests/test_oracle_fp64.py):
* the known-answer data of ``tests/test_alignment.py:29-43`` (six points, 90°
  rotation → RMSD 0);
* Kabsch/SVD (``geometry/alignment.py:117-181``) agreeing with the QCP
  eigenvalue route on random inputs;
* brute-force minimisation over rotations on tiny inputs;
* invariants: rigid-motion invariance, symmetry, ``rmsd <= no-superpose rmsd``,
  1-atom and all-zero systems give 0;
* golden vectors written by the compiled reference C++ core
  (``tests/golden/``): within the reference's own float32 error.
"""
from __future__ import annotations

import numpy as np


def center_and_trace(xyz):
    """Centre each frame and return (centred coordinates, traces) in float64.

    Follows ``rmsd/src/center_generic.h:3-44`` (mean subtraction per frame,
    trace G = sum of squared centred coordinates), without its float casts.
    ``xyz``: (F, N, 3).  Returns ((F, N, 3) float64, (F,) float64).
    """
    x = np.asarray(xyz, dtype=np.float64)
    mu = x.mean(axis=1, keepdims=True)
    c = x - mu
    g = np.einsum("fnk,fnk->f", c, c)
    return c, g


def correlation_matrix(a, b):
    """M[f, r, c] = sum_i a[f, i, r] * b[i, c] for centred a (F,N,3), b (N,3).

    Follows the M layout of ``rmsd/src/theobald_rmsd_generic.h:137-171``
    (``M[r*3+c] = sum a_r b_c``; a is the first argument of ``msd_atom_major``).
    """
    return np.einsum("fnr,nc->frc", np.asarray(a, np.float64), np.asarray(b, np.float64))


def key_matrix(M):
    """Symmetric 4x4 key matrix K of Horn/Theobald from M (..., 3, 3).

    Same entries as ``rmsd/src/theobald_rmsd.cpp:227-236``, which indexes the
    row-major M as ``M[i + 3 j]``, i.e. builds K from M^T (its top
    eigenvector is then the quaternion rotating the first frame onto the
    second; the eigenvalues do not depend on the transpose).
    """
    M = np.swapaxes(np.asarray(M, dtype=np.float64), -1, -2)
    xx, xy, xz = M[..., 0, 0], M[..., 0, 1], M[..., 0, 2]
    yx, yy, yz = M[..., 1, 0], M[..., 1, 1], M[..., 1, 2]
    zx, zy, zz = M[..., 2, 0], M[..., 2, 1], M[..., 2, 2]
    K = np.empty(M.shape[:-2] + (4, 4), dtype=np.float64)
    K[..., 0, 0] = xx + yy + zz
    K[..., 0, 1] = K[..., 1, 0] = yz - zy
    K[..., 0, 2] = K[..., 2, 0] = zx - xz
    K[..., 0, 3] = K[..., 3, 0] = xy - yx
    K[..., 1, 1] = xx - yy - zz
    K[..., 1, 2] = K[..., 2, 1] = xy + yx
    K[..., 1, 3] = K[..., 3, 1] = zx + xz
    K[..., 2, 2] = -xx + yy - zz
    K[..., 2, 3] = K[..., 3, 2] = yz + zy
    K[..., 3, 3] = -xx - yy + zz
    return K


def lambda_max(M):
    """Largest eigenvalue of K(M) by a dense symmetric eigensolver (FP64).

    The reference solves the same eigenvalue via a closed-form quartic
    (``rmsd/src/theobald_rmsd.cpp:81-193``); the eigenvalue is the same number.
    """
    return np.linalg.eigvalsh(key_matrix(M))[..., -1]


def msd_from_M_G(M, Ga, Gb, n_atoms):
    """MSD = max(0, (G_a + G_b - 2 lambda_max) / N) in FP64.

    Follows ``rmsd/src/theobald_rmsd.cpp:273-277``.
    """
    lam = lambda_max(M)
    msd = (np.asarray(Ga, np.float64) + np.asarray(Gb, np.float64) - 2.0 * lam) / n_atoms
    return np.maximum(msd, 0.0)


def _select(xyz, indices):
    x = np.asarray(xyz)
    if indices is None:
        return x
    return x[..., np.asarray(indices), :]


def rmsd(target_xyz, ref_xyz, frame=0, atom_indices=None, ref_atom_indices=None):
    """One-vs-all RMSD (superpose=True) in FP64.

    Mirrors ``rmsd/_rmsd.pyx:154-224``: gather ``atom_indices`` from every
    target frame and ``ref_atom_indices`` (default: ``atom_indices``) from
    ``ref_xyz[frame]``, centre both, QCP per frame.  Returns (F,) float64.
    """
    if ref_atom_indices is None:
        ref_atom_indices = atom_indices
    t = _select(target_xyz, atom_indices)
    r = _select(np.asarray(ref_xyz)[frame][None], ref_atom_indices)
    tc, tg = center_and_trace(t)
    rc, rg = center_and_trace(r)
    M = correlation_matrix(tc, rc[0])
    n = t.shape[1]
    return np.sqrt(msd_from_M_G(M, tg, rg[0], n))


def rmsd_kabsch(target_xyz, ref_xyz, frame=0, atom_indices=None, ref_atom_indices=None):
    """One-vs-all RMSD via Kabsch SVD, the second FP64 route.

    Restates ``geometry/alignment.py:117-181``: correlation matrix, SVD,
    reflection fix, RMSD = sqrt((G_a + G_b - 2 (s1 + s2 + d s3)) / N).
    """
    if ref_atom_indices is None:
        ref_atom_indices = atom_indices
    t = _select(target_xyz, atom_indices)
    r = _select(np.asarray(ref_xyz)[frame][None], ref_atom_indices)
    tc, tg = center_and_trace(t)
    rc, rg = center_and_trace(r)
    M = correlation_matrix(tc, rc[0])
    U, S, Vt = np.linalg.svd(M)
    d = np.sign(np.linalg.det(U) * np.linalg.det(Vt))
    d = np.where(d == 0, 1.0, d)
    lam = S[:, 0] + S[:, 1] + d * S[:, 2]
    n = t.shape[1]
    return np.sqrt(np.maximum((tg + rg[0] - 2 * lam) / n, 0.0))


def rmsd_nosuperpose(target_xyz, ref_xyz, frame=0, atom_indices=None, ref_atom_indices=None):
    """superpose=False: sqrt(mean over atoms of |a - b|^2), no centering.

    Follows ``rmsd/_rmsd.pyx:228-241,765-793``.
    """
    if ref_atom_indices is None:
        ref_atom_indices = atom_indices
    t = np.asarray(_select(target_xyz, atom_indices), np.float64)
    r = np.asarray(_select(np.asarray(ref_xyz)[frame][None], ref_atom_indices), np.float64)
    d = t - r
    return np.sqrt(np.einsum("fnk,fnk->f", d, d) / t.shape[1])


def kabsch_rotation(mobile_centred, target_centred):
    """Rotation R (row-vector convention, x' = x R) aligning mobile to target.

    Restates ``geometry/alignment.py:151-156`` (SVD with reflection fix).
    Works on stacks: mobile (F, N, 3), target (N, 3) → (F, 3, 3).
    """
    M = correlation_matrix(mobile_centred, target_centred)
    U, S, Vt = np.linalg.svd(M)
    det = np.linalg.det(U) * np.linalg.det(Vt)
    U = U.copy()
    U[det < 0, :, -1] *= -1.0
    return U @ Vt


def superpose(xyz, ref_xyz, frame=0, atom_indices=None, ref_atom_indices=None):
    """FP64 ``Trajectory.superpose``: returns the moved coordinates (F, N, 3).

    Follows ``core/trajectory.py:1114-1172``: centre the fitted subset of
    every frame (displacing all atoms by that centroid), rotate by the optimal
    rotation onto the centred reference subset, add the reference centroid.
    """
    if ref_atom_indices is None:
        ref_atom_indices = atom_indices
    x = np.asarray(xyz, np.float64)
    align = _select(x, atom_indices)
    ref = np.asarray(_select(np.asarray(ref_xyz)[frame][None], ref_atom_indices), np.float64)[0]
    mu = align.mean(axis=1, keepdims=True)
    ref_mu = ref.mean(axis=0)
    R = kabsch_rotation(align - mu, ref - ref_mu)
    return np.einsum("fnk,fkj->fnj", x - mu, R) + ref_mu


def rmsf(target_xyz, ref_xyz=None, frame=0, atom_indices=None, ref_atom_indices=None):
    """FP64 ``md.rmsf`` in atom mode: per-atom RMS fluctuation (n_sel,).

    Follows ``rmsd/_rmsd.pyx:246-443``.  Unless ``ref_xyz`` is None (frames
    already aligned), every frame's selection is fitted onto the centred
    reference selection and the fitting rotation is applied to the selection
    (``:405-414``).  The rotated coordinates are the selection centred on its
    centroid when ``atom_indices`` is None, and the raw selection otherwise:
    there the reference centres a view of ``target.xyz`` in place before it
    copies the coordinates it rotates (``:389,395-401,404``), while a fancy
    index takes a copy first.  Fluctuation = mean over frames of |x - <x>|^2
    (``:416-440``), then sqrt.
    """
    x = _select(np.asarray(target_xyz, np.float64), atom_indices)
    mu = x.mean(axis=1, keepdims=True)
    disp = x - mu if atom_indices is None else x.copy()
    if ref_xyz is not None:
        if ref_atom_indices is None:
            ref_atom_indices = atom_indices
        ref = np.asarray(_select(np.asarray(ref_xyz)[frame][None], ref_atom_indices),
                         np.float64)[0]
        R = kabsch_rotation(x - mu, ref - ref.mean(axis=0))
        disp = np.einsum("fnk,fkj->fnj", disp, R)
    avg = disp.mean(axis=0)
    return np.sqrt(((disp - avg) ** 2).sum(axis=2).mean(axis=0))


def rmsf_by_residue(rmsf_atoms, residue_of_atom, masses, n_residues):
    """Residue mode of ``md.rmsf`` (``rmsd/_rmsd.pyx:445-484``): per residue,
    sqrt(sum m_i f_i^2 / sum m_i) over its selected atoms, f_i the per-atom
    RMSF (the reference combines the mean-square fluctuations, before any
    square root); -1 for residues without selected atoms."""
    num = np.zeros(n_residues)
    den = np.zeros(n_residues)
    msf = np.asarray(rmsf_atoms, np.float64) ** 2
    np.add.at(num, residue_of_atom, msf * masses)
    np.add.at(den, residue_of_atom, masses)
    out = np.full(n_residues, -1.0)
    ok = den > 0
    out[ok] = np.sqrt(num[ok] / den[ok])
    return out


def align_displace_rmsd(align1, align2, displ1, displ2, frame):
    """FP64 ``getMultipleAlignDisplaceRMSDs_atom_major`` (``rmsd/_rmsd.pyx:677-759``):
    the rotation R_i that best fits ``align1[frame]`` onto ``align2[i]``
    (``msd_atom_major(align1[frame], align2[i])``, applied as x R), then the
    RMS displacement of ``displ1[frame] R_i`` from ``displ2[i]``
    (``rot_msd_atom_major``, ``rmsd/src/rotation_generic.h:47-102``), with no
    centering of the displacement sets.  The alignment sets are centred here
    (the reference requires the caller to have centred them).  Returns
    (rmsd (F,), R (F, 3, 3))."""
    a1 = np.asarray(align1, np.float64)[frame]
    a2 = np.asarray(align2, np.float64)
    ref = a1 - a1.mean(axis=0)
    mob = a2 - a2.mean(axis=1, keepdims=True)
    # optimal rotation of mobile frame i onto the reference, transposed:
    # the rotation taking the reference onto frame i
    R = np.transpose(kabsch_rotation(mob, ref), (0, 2, 1))
    d1 = np.asarray(displ1, np.float64)[frame]
    d2 = np.asarray(displ2, np.float64)
    moved = np.einsum("nk,fkj->fnj", d1, R)
    return np.sqrt(((moved - d2) ** 2).sum(axis=2).mean(axis=1)), R


def rmsd_matrix(xyz, atom_indices=None):
    """All-pairs RMSD matrix D[i, j] in FP64 (F, F).

    The reference has no all-pairs routine; its users loop
    ``md.rmsd(traj, traj, i)`` (examples/clustering.ipynb:80-81), so row i is
    exactly ``rmsd(xyz, xyz, i)``.  Computed blockwise with one batched
    eigensolve per block.
    """
    x = _select(xyz, atom_indices)
    c, g = center_and_trace(x)
    F, n = c.shape[0], c.shape[1]
    D = np.zeros((F, F), dtype=np.float64)
    planes = [np.ascontiguousarray(c[:, :, r]) for r in range(3)]
    # M[i, j, r, s] = sum_k c[i,k,r] c[j,k,s]: nine plane products per row block
    blk = max(1, min(F, 1_000_000 // max(1, F)))
    for i0 in range(0, F, blk):
        i1 = min(F, i0 + blk)
        M = np.empty((i1 - i0, F, 3, 3), dtype=np.float64)
        for r in range(3):
            for s in range(3):
                M[:, :, r, s] = planes[r][i0:i1] @ planes[s].T
        msd = msd_from_M_G(M, g[i0:i1, None], g[None, :], n)
        D[i0:i1] = np.sqrt(msd)
    return D


def condensed(D):
    """Upper triangle (i < j) of a square matrix, row-major, as scipy's
    ``squareform`` lays it out."""
    F = D.shape[0]
    iu = np.triu_indices(F, k=1)
    return D[iu]


def brute_force_rmsd(a, b, n_grid=24, refine=200, seed=0):
    """RMSD of two tiny frames by direct minimisation over rotations.

    Independent of M/K/SVD: coarse grid over unit quaternions, then a random
    local refinement.  Used only to pin the eigenvalue route on tiny inputs.
    """
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    a = a - a.mean(0)
    b = b - b.mean(0)

    def rot(q):
        q = q / np.linalg.norm(q)
        w, x, y, z = q
        return np.array([
            [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
            [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
            [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
        ])

    def cost(q):
        d = a @ rot(q) - b
        return np.mean(np.sum(d * d, axis=1))

    g = np.linspace(-1, 1, n_grid)
    best_q, best = None, np.inf
    for w in np.linspace(0, 1, n_grid // 2 + 1):
        for x in g:
            for y in g:
                for z in g:
                    q = np.array([w, x, y, z])
                    if np.linalg.norm(q) < 1e-9:
                        continue
                    c = cost(q)
                    if c < best:
                        best, best_q = c, q
    rng = np.random.default_rng(seed)
    step = 0.1
    q = best_q / np.linalg.norm(best_q)
    for _ in range(refine):
        improved = False
        for _ in range(20):
            cand = q + step * rng.standard_normal(4)
            c = cost(cand)
            if c < best:
                best, q, improved = c, cand / np.linalg.norm(cand), True
        if not improved:
            step *= 0.6
        if step < 1e-10:
            break
    return np.sqrt(max(best, 0.0))


def centroid_scores(D, beta=1.0):
    """examples/centroids.ipynb cells 3-4 on a full square distance matrix:
    sum_j exp(-beta D_ij / D.std()) in float64 (the notebook fills a float64
    matrix with md.rmsd rows and takes numpy's population std)."""
    D = np.asarray(D, np.float64)
    return np.exp(-beta * D / D.std()).sum(axis=1)


def nearest_leader(frames, leaders, atom_indices=None):
    """examples/two-pass-clustering.ipynb cell 6: for every frame,
    argmin over leaders of md.rmsd(leaders, frame, 0), here with this
    oracle's rmsd; returns (index, rmsd) arrays."""
    frames = np.asarray(frames)
    idx = np.empty(len(frames), np.int64)
    val = np.empty(len(frames), np.float64)
    for f in range(len(frames)):
        r = rmsd(leaders, frames[f:f + 1], 0, atom_indices)
        idx[f] = int(np.argmin(r))
        val[f] = r[idx[f]]
    return idx, val
