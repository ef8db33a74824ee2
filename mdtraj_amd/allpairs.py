"""All-pairs RMSD matrix.

The reference has no all-pairs routine; its users loop ``md.rmsd(t, t, i)``
(examples/clustering.ipynb:80-81, examples/centroids.ipynb:79-82), so row i
of this matrix is one-vs-all with reference frame i.  The output is the
condensed upper triangle (i < j, row-major), the layout
``scipy.spatial.distance.squareform`` and ``scipy.cluster.hierarchy.linkage``
take.
"""
from __future__ import annotations

import numpy as np

from . import _device as dev
from ._lib import (HR_PAIRS_BF16_EXACT, HR_PAIRS_BF16X3, HR_PAIRS_F16_EXACT, HR_PAIRS_F16X3,
                   HR_PAIRS_FP32, call, load)

MODES = {"fp32": HR_PAIRS_FP32, "f16x3": HR_PAIRS_F16X3, "bf16x3": HR_PAIRS_BF16X3,
         "f16exact": HR_PAIRS_F16_EXACT, "bf16exact": HR_PAIRS_BF16_EXACT}


@dev.follows(0, "traj")
def rmsd_matrix(traj, atom_indices=None, mode="f16exact", square=False, out=None):
    """RMSD between every pair of frames of ``traj``.

    mode: "f16exact" (default; FP16 split with an exactly accumulated
    fixed-point hi part and a small cross-term accumulator, four MFMA
    products, <= 1e-5 nm vs FP64), "fp32" (FP32 FMA Gram), "f16x3" / "bf16x3"
    (hi+lo split, three MFMA products in one FP32 accumulator: fastest, about
    1e-4 nm on near-identical frames) or "bf16exact".  Returns the
    condensed float32 vector (a CUDA tensor when the input is on the device,
    numpy otherwise), or the square matrix with ``square=True``.
    """
    if mode not in MODES:
        raise ValueError(f"mode must be one of {sorted(MODES)}")
    lib = load()
    xyz = traj.xyz if hasattr(traj, "xyz") else traj
    on_device = dev.is_device_array(xyz)
    X = dev.to_device_f32(xyz)
    F, N = int(X.shape[0]), int(X.shape[1])
    if atom_indices is not None:
        a = np.asarray(atom_indices)
        if a.ndim != 1 or a.size == 0 or a.min() < 0 or a.max() >= N:
            raise ValueError("atom_indices must be valid positive indices")
    idx = dev.to_device_i32(atom_indices)
    n_sel = N if atom_indices is None else len(atom_indices)
    n_pairs = F * (F - 1) // 2
    if out is None:
        out = dev.empty_f32(max(n_pairs, 1))
    nbytes = lib.hr_rmsd_all_pairs_workspace_bytes(F, n_sel, MODES[mode])
    ws = dev.workspace(nbytes)
    call("hr_rmsd_all_pairs", X.data_ptr(), F, N, dev.ptr(idx), n_sel, MODES[mode],
         out.data_ptr(), ws.data_ptr(), nbytes, dev.stream())
    out = out[:n_pairs]
    if square:
        if dev.is_torch_cuda(out):
            t = dev.torch()
            D = t.zeros((F, F), dtype=t.float32, device=out.device)
            iu = t.triu_indices(F, F, offset=1, device=out.device)
            D[iu[0], iu[1]] = out
            D = D + D.T
            return D if on_device else dev.to_host(D)
        D = np.zeros((F, F), np.float32)
        D[np.triu_indices(F, k=1)] = dev.to_host(out)
        D = D + D.T
        return dev.to_device_f32(D) if on_device else D
    return out if on_device else dev.to_host(out)


class SplitSet:
    """A frame set prepared for the split-precision Gram (centred, split into
    hi/lo planes, traces and scales), resident in HBM."""

    def __init__(self, traj, atom_indices=None, mode="f16exact"):
        if mode not in MODES or mode == "fp32":
            raise ValueError("SplitSet needs a split mode")
        xyz = traj.xyz if hasattr(traj, "xyz") else traj
        with dev.using(xyz):
            self._prepare(xyz, atom_indices, mode)

    def _prepare(self, xyz, atom_indices, mode):
        lib = load()
        X = dev.to_device_f32(xyz)
        self.n_frames, n_atoms = int(X.shape[0]), int(X.shape[1])
        if atom_indices is not None:
            a = np.asarray(atom_indices)
            if a.ndim != 1 or a.size == 0 or a.min() < 0 or a.max() >= n_atoms:
                raise ValueError("atom_indices must be valid positive indices")
        idx = dev.to_device_i32(atom_indices)
        self.n_sel = n_atoms if atom_indices is None else len(atom_indices)
        self.mode = mode
        self.nbytes = lib.hr_split_bytes(self.n_frames, self.n_sel, MODES[mode])
        self.block = dev.workspace(self.nbytes)
        call("hr_split_prepare", X.data_ptr(), self.n_frames, n_atoms, dev.ptr(idx), self.n_sel,
             MODES[mode], self.block.data_ptr(), self.nbytes, dev.stream())


def rmsd_cross(a, b, out=None):
    """RMSD between every frame of prepared set ``a`` (rows) and of ``b``
    (columns): float32 CUDA tensor (a.n_frames, b.n_frames)."""
    if a.mode != b.mode or a.n_sel != b.n_sel:
        raise ValueError("sets must share the mode and the number of selected atoms")
    with dev.using(a.block):
        if out is None:
            out = dev.empty((a.n_frames, b.n_frames), np.float32)
        call("hr_rmsd_cross", a.block.data_ptr(), a.n_frames, b.block.data_ptr(), b.n_frames,
             a.n_sel, MODES[a.mode], out.data_ptr(), out.stride(0), dev.stream())
    return out


def nearest(a, b):
    """For every frame of prepared set ``a``, the nearest frame of set ``b``:
    (index int32 tensor, RMSD float32 tensor), the minimum taken inside the
    cross kernel (``hr_rmsd_cross_argmin``; ties go to the lowest index, as
    ``np.argmin``).  The leader-assignment step of
    examples/two-pass-clustering.ipynb (cell 6) with ``b`` the leaders."""
    if a.mode != b.mode or a.n_sel != b.n_sel:
        raise ValueError("sets must share the mode and the number of selected atoms")
    if b.n_frames == 0 and a.n_frames:
        raise ValueError("no frames to take the nearest of")
    lib = load()
    with dev.using(a.block):
        index = dev.empty(a.n_frames, np.int32)
        value = dev.empty(a.n_frames, np.float32)
        ws = dev.workspace(lib.hr_rmsd_cross_argmin_workspace_bytes(a.n_frames))
        call("hr_rmsd_cross_argmin", a.block.data_ptr(), a.n_frames, b.block.data_ptr(),
             b.n_frames, a.n_sel, MODES[a.mode], index.data_ptr(), value.data_ptr(),
             ws.data_ptr(), dev.stream())
    return index, value


def topk_rows(matrix, k):
    """The k smallest entries of every row of a float32 device matrix (an
    ``rmsd_cross`` output), ascending, ties to the lower column, NaN last
    (``np.argsort(row, kind="stable")[:k]``): (index int32 (rows, k), value
    float32 (rows, k)) device arrays (``hr_topk_rows``)."""
    k = int(k)
    with dev.using(matrix):
        n_rows, n_cols = int(matrix.shape[0]), int(matrix.shape[1])
        if not 1 <= k <= n_cols:
            raise ValueError(f"k must be in [1, {n_cols}]")
        index = dev.empty((n_rows, k), np.int32)
        value = dev.empty((n_rows, k), np.float32)
        call("hr_topk_rows", matrix.data_ptr(), n_rows, n_cols, matrix.stride(0), k,
             index.data_ptr(), value.data_ptr(), dev.stream())
    return index, value


def nearest_k(a, b, k):
    """For every frame of prepared set ``a``, its k nearest frames of set
    ``b``, nearest first (ties to the lower index): (index, RMSD) device
    arrays (a.n_frames, k).  The k-leader generalisation of ``nearest``:
    the cross matrix (4 a.n_frames b.n_frames bytes of HBM, freed on
    return), then ``topk_rows``."""
    if a.mode != b.mode or a.n_sel != b.n_sel:
        raise ValueError("sets must share the mode and the number of selected atoms")
    if not 1 <= int(k) <= b.n_frames:
        raise ValueError(f"k must be in [1, {b.n_frames}]")
    d = rmsd_cross(a, b)
    return topk_rows(d, k)


def assign_to_leaders(frames, leaders, atom_indices=None, mode="f16exact"):
    """Index of the nearest leader for every frame, and that RMSD (numpy):
    ``argmin(md.rmsd(leaders, frame, 0))`` for all frames in one pass."""
    idx, val = nearest(SplitSet(frames, atom_indices, mode), SplitSet(leaders, atom_indices, mode))
    return dev.to_host(idx), dev.to_host(val)


@dev.follows(0, "condensed")
def centroid_scores(condensed, n_frames, beta=1.0):
    """``sum_j exp(-beta d_ij / d.std())`` for every frame i of the full
    symmetric matrix whose condensed upper triangle is ``condensed``
    (examples/centroids.ipynb, cells 3-4); float64 numpy array."""
    lib = load()
    n = int(n_frames)
    if n < 0:
        raise ValueError("n_frames must be >= 0")
    D = dev.to_device_f32(condensed).reshape(-1)
    if D.data_ptr() % 16:  # the kernel reads 16-byte groups
        D = D.clone()
    if D.numel() != n * (n - 1) // 2:
        raise ValueError(f"condensed length {D.numel()} does not match {n} frames")
    scores = dev.empty(max(n, 1), np.float64)
    ws = dev.workspace(lib.hr_centroid_workspace_bytes(n))
    call("hr_centroid_scores", D.data_ptr() if D.numel() else None, n, float(beta),
         scores.data_ptr(), None, ws.data_ptr(), dev.stream())
    return dev.to_host(scores[:n])


@dev.follows(0, "traj")
def centroid(traj, atom_indices=None, beta=1.0, mode="f16exact"):
    """Index of the frame with the largest summed similarity to all others
    (examples/centroids.ipynb): all-pairs RMSD, then centroid_scores."""
    xyz = traj.xyz if hasattr(traj, "xyz") else traj
    n = int(xyz.shape[0])
    d = rmsd_matrix(dev.to_device_f32(xyz), atom_indices=atom_indices, mode=mode)
    return int(np.argmax(centroid_scores(d, n, beta)))
