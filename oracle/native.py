"""ctypes wrappers for the oracle's native libraries (test infrastructure).

``port``       -> oracle/librmsd32.so, the float32 restatement (rmsd32.c)
``reference``  -> oracle/_ref/libref_rmsd.so, the reference's own C++ core
                  compiled from its sources (present only where it was built)

Both are used only as checkers or as the timed CPU baseline.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PORT_PATH = os.path.join(_HERE, "librmsd32.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libref_rmsd.so")

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_c_int = ctypes.c_int
_c_float = ctypes.c_float


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class _Lib:
    def __init__(self, path, prefix):
        self.path = path
        self.lib = ctypes.CDLL(path)
        self.prefix = prefix

    def fn(self, name, argtypes, restype=None):
        f = getattr(self.lib, self.prefix + name)
        f.argtypes = argtypes
        f.restype = restype
        return f


_port = None
_ref = None


def port():
    """The float32 restatement (always buildable: ``make -C oracle``)."""
    global _port
    if _port is None:
        if not os.path.exists(_PORT_PATH):
            raise RuntimeError(f"{_PORT_PATH} missing: run `make -C oracle`")
        _port = _Port(_PORT_PATH)
    return _port


def reference_available():
    return os.path.exists(_REF_PATH)


def reference():
    """The compiled reference core, or RuntimeError when it was not built."""
    global _ref
    if _ref is None:
        if not reference_available():
            raise RuntimeError(f"{_REF_PATH} missing (reference tree absent at build time)")
        _ref = _Ref(_REF_PATH)
    return _ref


class _Port(_Lib):
    def __init__(self, path):
        super().__init__(path, "o32_")
        self._rmsd = self.fn("rmsd", [_f32p, _c_int, _c_int, _f32p, _c_int, _f32p])
        self._center = self.fn("center_and_trace", [_f32p, _f32p, _c_int, _c_int])
        self._msd = self.fn("msd_atom_major",
                            [_c_int, _f32p, _f32p, _c_float, _c_float, _c_int, _f32p], _c_float)
        self._sup = self.fn("superpose_atom_major",
                            [_f32p, _f32p, _f32p, _f32p, _f32p, _c_int, _c_int, _c_int, _c_int, _f32p])
        self._nosup = self.fn("msd_nosuperpose", [_c_int, _f32p, _f32p], _c_float)

    def rmsd(self, target_sel, ref_sel, parallel=True):
        """Centre copies of the gathered target (F,n,3) and reference (n,3),
        then float32 QCP per frame.  Returns float32 (F,)."""
        t = np.array(target_sel, dtype=np.float32, order="C", copy=True)
        r = np.array(ref_sel, dtype=np.float32, order="C", copy=True)
        out = np.zeros(t.shape[0], np.float32)
        self._rmsd(t, t.shape[0], t.shape[1], r, int(parallel), out)
        return out

    def rmsd_inplace(self, target_sel, ref_sel, out, parallel=True):
        """As ``rmsd`` but centres the given C-contiguous float32 buffers in place."""
        self._rmsd(target_sel, target_sel.shape[0], target_sel.shape[1], ref_sel, int(parallel), out)

    def center(self, xyz):
        x = np.array(xyz, dtype=np.float32, order="C", copy=True)
        g = np.zeros(x.shape[0], np.float32)
        self._center(x, g, x.shape[0], x.shape[1])
        return x, g

    def msd_pair(self, a, b, Ga, Gb, rot=False):
        r = np.zeros(9, np.float32)
        m = self._msd(a.shape[0], _f32(a), _f32(b), Ga, Gb, int(rot), r)
        return (m, r.reshape(3, 3)) if rot else m

    def nosuperpose(self, target_sel, ref_sel):
        t = _f32(target_sel)
        r = _f32(ref_sel)
        return np.sqrt(np.array([self._nosup(t.shape[1], t[i], r) for i in range(t.shape[0])],
                                np.float32))

    def superpose_atom_major(self, align_target, align_mobile, g_target, g_mobile, displace,
                             target_frame=0):
        rot = np.zeros((align_mobile.shape[0], 9), np.float32)
        self._sup(align_target, align_mobile, g_target, g_mobile, displace,
                  align_mobile.shape[0], align_mobile.shape[1], displace.shape[1],
                  target_frame, rot)
        return rot.reshape(-1, 3, 3)


class _Ref(_Lib):
    def __init__(self, path):
        super().__init__(path, "ref_")
        self._rmsd = self.fn("rmsd", [_f32p, _c_int, _c_int, _f32p, _c_int, _f32p])
        self._self = self.fn("rmsd_self", [_f32p, _c_int, _c_int, _c_int, _c_int, _f32p])
        self._center = self.fn("center", [_f32p, _c_int, _c_int, _f32p])
        self._sup = self.fn("superpose_atom_major",
                            [_f32p, _f32p, _f32p, _f32p, _f32p, _c_int, _c_int, _c_int, _c_int, _f32p])
        self._pair = self.fn("msd_from_pair",
                             [_c_int, _f32p, _f32p, _c_float, _c_float, _c_int, _f32p], _c_float)

    def rmsd(self, target_sel, ref_sel, parallel=True):
        t = np.array(target_sel, dtype=np.float32, order="C", copy=True)
        r = np.array(ref_sel, dtype=np.float32, order="C", copy=True)
        out = np.zeros(t.shape[0], np.float32)
        self._rmsd(t, t.shape[0], t.shape[1], r, int(parallel), out)
        return out

    def rmsd_inplace(self, target_sel, ref_sel, out, parallel=True):
        self._rmsd(target_sel, target_sel.shape[0], target_sel.shape[1], ref_sel, int(parallel), out)

    def rmsd_self(self, xyz, frame, parallel=True):
        """md.rmsd(t, t, frame) with atom_indices=None (aliased reference)."""
        x = np.array(xyz, dtype=np.float32, order="C", copy=True)
        out = np.zeros(x.shape[0], np.float32)
        self._self(x, x.shape[0], x.shape[1], frame, int(parallel), out)
        return out

    def center(self, xyz):
        x = np.array(xyz, dtype=np.float32, order="C", copy=True)
        g = np.zeros(x.shape[0], np.float32)
        self._center(x, x.shape[0], x.shape[1], g)
        return x, g

    def msd_pair(self, a, b, Ga, Gb, rot=False):
        r = np.zeros(9, np.float32)
        m = self._pair(a.shape[0], _f32(a), _f32(b), Ga, Gb, int(rot), r)
        return (m, r.reshape(3, 3)) if rot else m

    def superpose_atom_major(self, align_target, align_mobile, g_target, g_mobile, displace,
                             target_frame=0):
        rot = np.zeros((align_mobile.shape[0], 9), np.float32)
        self._sup(align_target, align_mobile, g_target, g_mobile, displace,
                  align_mobile.shape[0], align_mobile.shape[1], displace.shape[1],
                  target_frame, rot)
        return rot.reshape(-1, 3, 3)
