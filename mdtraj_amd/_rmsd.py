"""The ``mdtraj._rmsd`` surface on the MI355X kernels.

Mirrors ``mdtraj/rmsd/_rmsd.pyx`` (paths relative to the mdtraj source tree):
same names, argument meanings, defaults, validation order, exception types
and messages.  Every computation runs in ``libhiprmsd.so`` on the GPU.

Documented deviations from the reference:
* inputs are never modified: the reference centres ``target.xyz`` in place
  when ``atom_indices`` is None (``_rmsd.pyx:197,212``); here centering is
  fused into the RMSD kernel and nothing is written back (so read-only and
  memory-mapped arrays work, where the reference raises ``ValueError``);
* ``precentered=True`` is accepted with the reference's warnings, but the
  cached float32 traces are not used for arithmetic: centroids and traces are
  recomputed in FP64 inside the same single pass, which costs nothing extra
  and keeps the 1e-5 nm accuracy;
* the reference frame's RMSD to itself is exactly 0 whenever the target IS
  the reference and the selections agree (the reference gets 0 only through a
  pointer-equality early-out, ``rmsd/src/theobald_rmsd_sse.h:256-262``);
* ``parallel`` is accepted and ignored (the GPU path is always parallel);
* when the two trajectories have different atom counts and no selection is
  given, ``ValueError`` is raised (the reference's check there calls
  ``len(slice)`` and raises ``TypeError``).
* device-resident inputs (torch CUDA tensors as ``.xyz``) return a CUDA tensor.

``set_inplace_centering(True)`` restores the reference's side effect for
callers that depend on it (tests/test_rmsd_memmap.py): with ``superpose`` and
no ``atom_indices``, ``rmsd`` centres ``target.xyz`` and ``reference.xyz[frame]``
in place (bit-exact to the reference's float32 centring) unless cached traces
are used, and a read-only ``xyz`` raises ``ValueError`` as the reference's
memoryview does.
"""
from __future__ import annotations

import ctypes
import warnings

import numpy as np

from . import _device as dev
from ._lib import call, load


_INPLACE = False


def set_inplace_centering(enabled):
    """Opt in (True) or out (False, the default) of the reference's in-place
    centring of ``target.xyz`` and ``reference.xyz[frame]`` by ``rmsd``
    (rmsd/_rmsd.pyx:197,199,212-213, reached when ``atom_indices`` is None).
    Returns the previous setting."""
    global _INPLACE
    prev, _INPLACE = _INPLACE, bool(enabled)
    return prev


def _xyz_of(obj):
    return obj.xyz if hasattr(obj, "xyz") else obj


def _check_indices(idx, n_atoms, name):
    """ensure_type(np.asarray(idx), dtype=int, ndim=1) plus the range check of
    ``_rmsd.pyx:158-161,171-174``."""
    if dev.is_torch_cuda(idx):
        a = idx
        if a.dim() != 1:
            raise ValueError(f"{name} must be ndim 1. You supplied {a.dim()}")
        if a.numel() and (int(a.min()) < 0 or int(a.max()) >= n_atoms):
            raise ValueError(f"{name} must be valid positive indices")
        return a
    if isinstance(idx, dev.DeviceArray):
        idx = idx.numpy()
    a = np.asarray(idx)
    if a.ndim != 1:
        raise ValueError(f"{name} must be ndim 1. You supplied {a.ndim}")
    a = a.astype(np.int64)
    if not np.all((a >= 0) * (a < n_atoms)):
        raise ValueError(f"{name} must be valid positive indices")
    return a


def _same_buffer(a, b) -> bool:
    if a is b:
        return True
    if dev.is_device_array(a) and dev.is_device_array(b):
        return (a.data_ptr() == b.data_ptr() and tuple(a.shape) == tuple(b.shape)
                and type(a) is type(b))
    if isinstance(a, np.ndarray) and isinstance(b, np.ndarray):
        return (a.__array_interface__["data"][0] == b.__array_interface__["data"][0]
                and a.shape == b.shape and a.strides == b.strides)
    return False


def _validate(target_xyz, reference_xyz, frame, atom_indices, ref_atom_indices):
    """Validation of ``_rmsd.pyx:154-184``; returns (idx, ref_idx, n_sel, frame)."""
    if atom_indices is not None:
        atom_indices = _check_indices(atom_indices, target_xyz.shape[1], "atom_indices")
    if ref_atom_indices is None:
        ref_atom_indices = atom_indices
    else:
        if atom_indices is None:
            raise TypeError("ref_atom_indices requires atom_indices "
                            "(the reference takes len() of the default selection)")
        if len(ref_atom_indices) != len(atom_indices):
            raise ValueError(
                "atom_indices and ref_atom_indices must have same number of atom indices. "
                "found %d and %d." % (len(atom_indices), len(ref_atom_indices)))
    if ref_atom_indices is not None:
        # also when it defaults to atom_indices: those index the reference too
        ref_atom_indices = _check_indices(ref_atom_indices, reference_xyz.shape[1],
                                          "ref_atom_indices")
    assert (len(target_xyz.shape) == 3 and len(reference_xyz.shape) == 3
            and target_xyz.shape[2] == 3 and reference_xyz.shape[2] == 3)
    if atom_indices is None and target_xyz.shape[1] != reference_xyz.shape[1]:
        raise ValueError("Input trajectories must have same number of atoms. "
                         "found %d and %d." % (target_xyz.shape[1], reference_xyz.shape[1]))
    n_ref_frames = reference_xyz.shape[0]
    if frame >= n_ref_frames:
        raise ValueError("Cannot calculate RMSD of frame %d: reference has "
                         "only %d frames." % (frame, n_ref_frames))
    if frame < -n_ref_frames:
        raise IndexError("index %d is out of bounds for axis 0 with size %d"
                         % (frame, n_ref_frames))
    if frame < 0:
        frame += n_ref_frames
    n_sel = target_xyz.shape[1] if atom_indices is None else len(atom_indices)
    return atom_indices, ref_atom_indices, n_sel, frame


def _result(out, like_device):
    return out if like_device else dev.to_host(out)


def _centre_like_reference(txyz, rxyz, frame):
    """The reference's side effect on float32 C-ordered host arrays: every
    target frame centred, then the reference frame (a view, so already
    centred when target is reference) centred again (rmsd/_rmsd.pyx:212-213)."""
    if txyz.dtype == np.float32 and txyz.flags.c_contiguous and txyz.shape[0]:
        _center_inplace_atom_major(txyz)
    r = rxyz[frame:frame + 1] if isinstance(rxyz, np.ndarray) else None
    if r is not None and r.dtype == np.float32 and r.flags.c_contiguous:
        _center_inplace_atom_major(r)


def _i32_host(idx):
    return None if idx is None else np.ascontiguousarray(dev.indices_on_host(idx), dtype=np.int32)


def _rmsd_host(txyz, rxyz, frame, idx, ridx, n_sel, same):
    """Host arrays in, host array out: one synchronous call of
    hr_rmsd_one_vs_all_host, which gathers the selection into pinned staging
    chunks and scores each chunk on the GPU as it arrives."""
    dev.require_gpu()
    T = np.ascontiguousarray(txyz, dtype=np.float32)
    n_frames, n_atoms = T.shape[0], T.shape[1]
    if n_frames == 0:
        return np.empty(0, np.float32)
    t_ptr = T.ctypes.data
    if same:
        # the reference frame is frame `frame` of the same (converted) array
        r_ptr, n_ref_atoms = t_ptr + frame * 12 * n_atoms, n_atoms
    else:
        R = np.ascontiguousarray(np.asarray(rxyz)[frame], dtype=np.float32)
        r_ptr, n_ref_atoms = R.ctypes.data, R.shape[0]
    # the result's buffer comes from ctypes, whose address is at hand (an
    # ndarray's .ctypes.data costs about a microsecond per call)
    buf = (ctypes.c_float * n_frames)()
    out = np.frombuffer(buf, dtype=np.float32)
    a, r = _i32_host(idx), _i32_host(ridx)
    call("hr_rmsd_one_vs_all_host", t_ptr, n_frames, n_atoms,
         None if a is None else a.ctypes.data, n_sel, r_ptr, n_ref_atoms,
         None if r is None else r.ctypes.data, frame if same else -1, ctypes.addressof(buf))
    return out


@dev.follows(0, "target")
def rmsd(target, reference, frame=0, atom_indices=None, ref_atom_indices=None,
         parallel=True, precentered=False, superpose=True):
    """RMSD of every conformation in ``target`` to ``reference[frame]``.

    Drop-in for ``md.rmsd`` (``rmsd/_rmsd.pyx:65-243``); see the module
    docstring for the deviations.  Returns float32 (n_frames,) in nm — a numpy
    array, or a CUDA tensor when ``target.xyz`` is one.
    """
    txyz = _xyz_of(target)
    rxyz = _xyz_of(reference)
    idx, ridx, n_sel, frame = _validate(txyz, rxyz, frame, atom_indices, ref_atom_indices)
    usable = False
    if superpose:
        # _rmsd.pyx:203-211: the cached traces are usable only with both
        # caches present and no selection; otherwise the reference warns.
        usable = (precentered and getattr(reference, "_rmsd_traces", None) is not None
                  and getattr(target, "_rmsd_traces", None) is not None
                  and atom_indices is None)
        if precentered and not usable:
            warnings.warn("in rmsd(), precentered is ignored when atom_indices != None",
                          RuntimeWarning)
    elif precentered:
        warnings.warn("in rmsd(), precentered is ignored when superpose=False", RuntimeWarning)

    same = _same_buffer(txyz, rxyz) and ref_atom_indices is None
    centre_after = (_INPLACE and superpose and atom_indices is None and not usable
                    and isinstance(txyz, np.ndarray))
    if centre_after:
        # the reference centres views of both arrays in place
        for a in (txyz, rxyz):
            if isinstance(a, np.ndarray) and not a.flags.writeable:
                raise ValueError("buffer source array is read-only")
    if superpose and isinstance(txyz, np.ndarray) and not dev.is_device_array(rxyz):
        out = _rmsd_host(txyz, rxyz, frame, idx, ridx, n_sel, same)
        if centre_after:
            _centre_like_reference(txyz, rxyz, frame)
        return out
    lib = load()
    on_device = dev.is_device_array(txyz)
    T = dev.to_device_f32(txyz)
    if same:
        R = T[frame]
    else:
        R = dev.to_device_f32(rxyz[frame])
    n_frames, n_atoms = int(T.shape[0]), int(T.shape[1])
    out = dev.empty_f32(n_frames)
    if n_frames == 0:
        return _result(out, on_device)
    d_idx = dev.to_device_i32(idx)
    d_ridx = dev.to_device_i32(ridx)
    ws = dev.workspace(lib.hr_rmsd_workspace_bytes(n_sel))
    if superpose:
        call("hr_rmsd_one_vs_all", T.data_ptr(), n_frames, n_atoms, dev.ptr(d_idx), n_sel,
             R.data_ptr(), dev.ptr(d_ridx), frame if same else -1, out.data_ptr(),
             ws.data_ptr(), dev.stream())
    else:
        call("hr_rmsd_nosuperpose", T.data_ptr(), n_frames, n_atoms, dev.ptr(d_idx), n_sel,
             R.data_ptr(), dev.ptr(d_ridx), out.data_ptr(), ws.data_ptr(), dev.stream())
    return _result(out, on_device)


@dev.follows(0, "xyz")
def _center_inplace_atom_major(xyz):
    """Centre (F, N, 3) float32 coordinates in place and return the traces.

    Mirrors ``rmsd/_rmsd.pyx:487-491``; rounding as the reference (double
    sums, float mean and subtraction, float squares summed in double).
    """
    if dev.is_device_array(xyz):
        assert xyz.dim() == 3 and xyz.shape[2] == 3
        f32 = dev.torch().float32 if dev.is_torch_cuda(xyz) else np.dtype(np.float32)
        if not xyz.is_contiguous() or xyz.dtype != f32:
            raise ValueError("xyz must be a C-contiguous float32 tensor")
        traces = dev.empty_f32(xyz.shape[0])
        call("hr_center_and_trace", xyz.data_ptr(), traces.data_ptr(), int(xyz.shape[0]),
             int(xyz.shape[1]), dev.stream())
        return dev.to_host(traces)
    if not isinstance(xyz, np.ndarray) or xyz.dtype != np.float32 or not xyz.flags.c_contiguous:
        raise ValueError("xyz must be a C-contiguous float32 array")
    if not xyz.flags.writeable:
        raise ValueError("buffer source array is read-only")
    assert xyz.shape[2] == 3
    X = dev.to_device_f32(xyz)
    traces = dev.empty_f32(xyz.shape[0])
    if xyz.shape[0]:
        call("hr_center_and_trace", X.data_ptr(), traces.data_ptr(), int(xyz.shape[0]),
             int(xyz.shape[1]), dev.stream())
    xyz[...] = dev.to_host(X)
    return dev.to_host(traces)


@dev.follows(1, "xyz2")
def getMultipleRMSDs_atom_major(xyz1, xyz2, g1, g2, frame, parallel=True):
    """RMSDs between xyz1[frame] and every frame of xyz2 (atom-major).

    Mirrors ``rmsd/_rmsd.pyx:562-615``.  The caller's traces g1/g2 are
    accepted for signature compatibility; centroids and traces are recomputed
    in FP64 on the GPU.
    """
    assert xyz1.ndim == 3 and xyz2.ndim == 3 and xyz1.shape[2] == 3 and xyz2.shape[2] == 3
    if not (xyz1.shape[1] == xyz2.shape[1]):
        raise ValueError("Input arrays must have same number of atoms. "
                         "found %d and %d." % (xyz1.shape[1], xyz2.shape[1]))
    if frame >= xyz1.shape[0]:
        raise ValueError("Cannot calculate RMSD of frame %d: xyz1 has "
                         "only %d frames." % (frame, xyz1.shape[0]))
    lib = load()
    T = dev.to_device_f32(xyz2)
    R = dev.to_device_f32(xyz1[frame])
    n_frames, n_atoms = int(T.shape[0]), int(T.shape[1])
    out = dev.empty_f32(n_frames)
    if n_frames:
        ws = dev.workspace(lib.hr_rmsd_workspace_bytes(n_atoms))
        call("hr_rmsd_one_vs_all", T.data_ptr(), n_frames, n_atoms, None, n_atoms,
             R.data_ptr(), None, -1, out.data_ptr(), ws.data_ptr(), dev.stream())
    return _result(out, dev.is_device_array(xyz2))


@dev.follows(1, "xyz2")
def getMultipleRMSDs_axis_major(xyz1, xyz2, g1, g2, frame, parallel=True):
    """RMSDs between xyz1[frame] and every frame of xyz2, axis-major
    (n_frames, 3, n_atoms).  Mirrors ``rmsd/_rmsd.pyx:502-557``."""
    assert xyz1.ndim == 3 and xyz2.ndim == 3 and xyz1.shape[1] == 3 and xyz2.shape[1] == 3
    if not (xyz1.shape[2] == xyz2.shape[2]):
        raise ValueError("Input arrays must have same number of atoms. "
                         "found %d and %d." % (xyz1.shape[2], xyz2.shape[2]))
    if frame >= xyz1.shape[0]:
        raise ValueError("Cannot calculate RMSD of frame %d: xyz1 has "
                         "only %d frames." % (frame, xyz1.shape[0]))
    lib = load()
    T = dev.to_device_f32(xyz2)
    R = dev.to_device_f32(xyz1[frame])
    n_frames, n_atoms = int(T.shape[0]), int(T.shape[2])
    out = dev.empty_f32(n_frames)
    if n_frames:
        ws = dev.workspace(lib.hr_rmsd_workspace_bytes(n_atoms))
        call("hr_rmsd_one_vs_all_axis_major", T.data_ptr(), n_frames, n_atoms, n_atoms,
             R.data_ptr(), n_atoms, -1, out.data_ptr(), ws.data_ptr(), dev.stream())
    return _result(out, dev.is_device_array(xyz2))


@dev.follows(4, "xyz_displace_mobile")
def superpose_atom_major(xyz_align_target, xyz_align_mobile, g_target, g_mobile,
                         xyz_displace_mobile, target_frame, parallel=True):
    """Rotate every frame of ``xyz_displace_mobile`` in place by the rotation
    that best fits ``xyz_align_mobile[i]`` onto ``xyz_align_target[target_frame]``
    (inputs centred by the caller).  Mirrors ``rmsd/_rmsd.pyx:620-674``."""
    if not (xyz_align_target.shape[1] == xyz_align_mobile.shape[1]):
        raise ValueError("Input arrays must have same number of atoms. "
                         "found %d and %d." % (xyz_align_target.shape[1],
                                               xyz_align_mobile.shape[1]))
    if not xyz_align_mobile.shape[0] == xyz_displace_mobile.shape[0]:
        raise ValueError("xyz_align_mobile and xyz_displace_mobile must contain the same "
                         "number of frames")
    lib = load()
    on_device = dev.is_device_array(xyz_displace_mobile)
    D = dev.to_device_f32(xyz_displace_mobile)
    A = dev.to_device_f32(xyz_align_mobile)
    R = dev.to_device_f32(xyz_align_target[target_frame])
    n_frames, n_disp, n_align = int(D.shape[0]), int(D.shape[1]), int(A.shape[1])
    if n_frames:
        ws = dev.workspace(lib.hr_superpose_workspace_bytes(n_frames, n_align))
        call("hr_superpose", D.data_ptr(), n_frames, n_disp, A.data_ptr(), n_align, None,
             n_align, R.data_ptr(), None, 0, None, None, ws.data_ptr(), dev.stream())
    if not on_device:
        xyz_displace_mobile[...] = dev.to_host(D)
    elif D.data_ptr() != xyz_displace_mobile.data_ptr():
        xyz_displace_mobile.copy_(D)


@dev.follows(0, "target")
def rmsf(target, reference, frame=0, atom_indices=None, ref_atom_indices=None,
         parallel=True, precentered=False, mode="atom"):
    """Root-mean-square fluctuation of the selected atoms of ``target``.

    Drop-in for ``md.rmsf`` (``rmsd/_rmsd.pyx:246-484``): unless ``reference``
    is None (frames already aligned), each frame's selection is fitted onto
    ``reference[frame]``'s and rotated by the fit; the fluctuation is taken
    about the mean structure.  As in the reference, the fluctuating
    coordinates are each frame's selection centred on its centroid when
    ``atom_indices`` is None, and the selection as stored (rotated about the
    origin) otherwise.  ``mode='residue'`` combines the atoms' mean-square
    fluctuations per residue with mass weights (needs ``target.topology``
    with ``atom(i).residue.index``, ``atom(i).element.mass`` and
    ``n_residues``); residues without a selected atom read -1.  Deviations:
    inputs are never modified (the reference centres ``target.xyz`` in place
    when ``atom_indices`` is None); sums are FP64.
    """
    prealigned = reference is None
    if reference is None:
        reference = target
    txyz = _xyz_of(target)
    rxyz = _xyz_of(reference)
    n_common = min(txyz.shape[1], rxyz.shape[1])
    if atom_indices is not None:
        atom_indices = _check_indices(atom_indices, n_common, "atom_indices")
    if ref_atom_indices is None:
        ref_atom_indices = atom_indices
    else:
        if atom_indices is None:
            raise TypeError("ref_atom_indices requires atom_indices "
                            "(the reference takes len() of the default selection)")
        if len(ref_atom_indices) != len(atom_indices):
            raise ValueError(
                "atom_indices and ref_atom_indices must have same number of atom indices. "
                "found %d and %d." % (len(atom_indices), len(ref_atom_indices)))
        ref_atom_indices = _check_indices(ref_atom_indices, n_common, "ref_atom_indices")
    if mode not in {"atom", "residue"}:
        raise ValueError('Mode must be one of "atom" or "residue". "%s" supplied.' % mode)
    assert (len(txyz.shape) == 3 and len(rxyz.shape) == 3
            and txyz.shape[2] == 3 and rxyz.shape[2] == 3)
    if atom_indices is None and txyz.shape[1] != rxyz.shape[1]:
        raise ValueError("Input trajectories must have same number of atoms. "
                         "found %d and %d." % (txyz.shape[1], rxyz.shape[1]))
    if frame >= rxyz.shape[0]:
        raise ValueError("Cannot calculate RMSF of frame %d: reference has "
                         "only %d frames." % (frame, rxyz.shape[0]))
    if frame < 0:
        frame += rxyz.shape[0]
    usable = (getattr(reference, "_rmsd_traces", None) is not None
              and getattr(target, "_rmsd_traces", None) is not None and atom_indices is None)
    if precentered and not usable:
        warnings.warn("in rmsd(), precentered is ignored when atom_indices != None",
                      RuntimeWarning)

    lib = load()
    on_device = dev.is_device_array(txyz)
    T = dev.to_device_f32(txyz)
    n_frames, n_atoms = int(T.shape[0]), int(T.shape[1])
    n_sel = n_atoms if atom_indices is None else len(atom_indices)
    out = dev.empty_f32(n_sel)
    if n_frames == 0:
        out.zero_()
    else:
        d_idx = dev.to_device_i32(atom_indices)
        d_ridx = dev.to_device_i32(ref_atom_indices)
        R = None if prealigned else dev.to_device_f32(rxyz[frame])
        ws = dev.workspace(lib.hr_rmsf_workspace_bytes(n_frames, n_sel))
        call("hr_rmsf", T.data_ptr(), n_frames, n_atoms, dev.ptr(d_idx), n_sel,
             None if R is None else R.data_ptr(), dev.ptr(d_ridx), int(atom_indices is None),
             out.data_ptr(), ws.data_ptr(), dev.stream())
    if mode == "atom":
        return _result(out, on_device)
    top = target.topology
    sel = np.arange(n_atoms) if atom_indices is None else np.asarray(
        dev.to_host(atom_indices) if dev.is_device_array(atom_indices) else atom_indices)
    residue_of = np.array([top.atom(int(i)).residue.index for i in sel], dtype=np.int64)
    masses = np.array([float(top.atom(int(i)).element.mass) for i in sel], dtype=np.float64)
    msf = dev.to_host(out).astype(np.float64) ** 2
    num = np.zeros(top.n_residues)
    den = np.zeros(top.n_residues)
    np.add.at(num, residue_of, msf * masses)
    np.add.at(den, residue_of, masses)
    res = np.full(top.n_residues, -1.0, dtype=np.float64)
    ok = den > 0
    res[ok] = np.sqrt(num[ok] / den[ok])
    return res.astype(np.float32)


@dev.follows(1, "xyz_align2")
def getMultipleAlignDisplaceRMSDs_atom_major(xyz_align1, xyz_align2, g_align1, g_align2,
                                             xyz_displ1, xyz_displ2, n_atoms_align,
                                             n_atoms_displ, frame, parallel=True):
    """RMSDs of ``xyz_displ1[frame]``, rotated by the rotation that best fits
    ``xyz_align1[frame]`` onto each ``xyz_align2[i]``, from ``xyz_displ2[i]``.

    Mirrors ``rmsd/_rmsd.pyx:677-759`` (arrays padded to a multiple of 4
    atoms, the first ``n_atoms_*`` used).  Returns (rmsds float32 (F,),
    rotations float32 (F, 3, 3), applied as x' = x R).  The alignment sets'
    centroids are removed on the GPU (the reference expects them centred and
    takes the traces from the caller; g_align1/g_align2 are accepted for
    signature compatibility).
    """
    assert (xyz_align1.shape[2] == 3) and (xyz_align2.shape[2] == 3)
    assert (xyz_displ1.shape[2] == 3) and (xyz_displ2.shape[2] == 3)
    if not ((xyz_align1.shape[1] % 4 == 0) & (xyz_align2.shape[1] % 4 == 0)):
        raise ValueError("Input arrays must have middle dimension of 4*n, "
                         "found %d and %d." % (xyz_align1.shape[1], xyz_align2.shape[1]))
    if not ((xyz_displ1.shape[1] % 4 == 0) & (xyz_displ2.shape[1] % 4 == 0)):
        raise ValueError("Input arrays must have middle dimension of 4*n, "
                         "found %d and %d." % (xyz_displ1.shape[1], xyz_displ2.shape[1]))
    if not xyz_align1.shape[0] == xyz_displ1.shape[0]:
        raise ValueError("xyz_align1 and xyz_displ1 must contain the same number of frames")
    if not xyz_align2.shape[0] == xyz_displ2.shape[0]:
        raise ValueError("xyz_align2 and xyz_displ2 must contain the same number of frames")
    if frame >= xyz_align1.shape[0]:
        raise ValueError("Cannot calculate RMSD of frame %d: xyz1 has "
                         "only %d frames." % (frame, xyz_align1.shape[0]))
    n_atoms_align, n_atoms_displ = int(n_atoms_align), int(n_atoms_displ)
    if not (0 < n_atoms_align <= min(xyz_align1.shape[1], xyz_align2.shape[1])
            and 0 < n_atoms_displ <= min(xyz_displ1.shape[1], xyz_displ2.shape[1])):
        raise ValueError("n_atoms_align / n_atoms_displ must be within the arrays' atoms")
    lib = load()
    on_device = dev.is_device_array(xyz_align2)
    A = dev.to_device_f32(xyz_align2)
    D = dev.to_device_f32(xyz_displ2)
    A_ref = dev.to_device_f32(xyz_align1[frame][:n_atoms_align]).contiguous()
    D_ref = dev.to_device_f32(xyz_displ1[frame][:n_atoms_displ]).contiguous()
    n_frames = int(A.shape[0])
    out = dev.empty_f32(n_frames)
    rot = dev.empty_f32(n_frames * 9)
    if n_frames:
        ws = dev.workspace(lib.hr_align_displace_workspace_bytes(n_frames, n_atoms_align))
        call("hr_rmsd_align_displace", A.data_ptr(), n_frames, int(A.shape[1]), n_atoms_align,
             A_ref.data_ptr(), D.data_ptr(), int(D.shape[1]), n_atoms_displ, D_ref.data_ptr(),
             out.data_ptr(), rot.data_ptr(), ws.data_ptr(), dev.stream())
    rot = rot.reshape(n_frames, 3, 3)
    return _result(out, on_device), _result(rot, on_device)
