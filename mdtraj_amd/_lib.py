"""ctypes binding of ``libhiprmsd.so`` (the C ABI declared in include/hiprmsd.h).

The library is the only compute path: if it is missing or cannot be loaded,
every call raises ``RuntimeError`` — there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhiprmsd.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "hiprmsd.h")

HR_OK = 0
HR_ERR_INVALID = 1
HR_ERR_HIP = 2
HR_ERR_UNSUPPORTED = 3
HR_TUNE_FRAMES_WAVES_PER_SIMD = 0
HR_TUNE_FRAMES_GRID = 1
HR_TUNE_FRAMES_ROW_LANES = 2
HR_TUNE_BATCH_SCHEDULE = 3
HR_TUNE_SUPERPOSE_CHUNK_MIB = 4
HR_TUNE_HOST_DIRECT_KIB = 5
HR_TUNE_PAIRS_TILE = 6
HR_TUNE_FRAMES_STATIC_PCT = 7
HR_TUNE_FRAMES_CHUNK_KIB = 8

HR_PAIRS_FP32 = 0
HR_PAIRS_F16X3 = 1
HR_PAIRS_BF16X3 = 2
HR_PAIRS_F16_EXACT = 3
HR_PAIRS_BF16_EXACT = 4

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_sz = ctypes.c_size_t

# name -> (restype, argtypes)
_SIGNATURES = {
    "hr_version": (ctypes.c_int, []),
    "hr_last_error": (ctypes.c_char_p, []),
    "hr_init": (ctypes.c_int, [ctypes.c_int]),
    "hr_device_cu_count": (ctypes.c_int, []),
    "hr_set_tuning": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "hr_get_tuning": (ctypes.c_int, [ctypes.c_int]),
    "hr_center_and_trace": (ctypes.c_int, [_vp, _vp, _i64, _i32, _vp]),
    "hr_rmsd_workspace_bytes": (_sz, [_i32]),
    "hr_rmsd_one_vs_all": (ctypes.c_int, [_vp, _i64, _i32, _vp, _i32, _vp, _vp, _i64, _vp, _vp, _vp]),
    "hr_rmsd_prepare_reference": (ctypes.c_int, [_vp, _vp, _i32, _vp, _vp]),
    "hr_rmsd_one_vs_all_prepared": (ctypes.c_int, [_vp, _i64, _i32, _vp, _i32, _i64, _vp, _vp, _vp]),
    "hr_rmsd_one_vs_all_noalloc_max_sel": (_sz, []),
    "hr_rmsd_one_vs_all_noalloc": (ctypes.c_int, [_vp, _i64, _i32, _vp, _i32, _vp, _vp, _i64, _vp,
                                                  _vp]),
    "hr_rmsd_one_vs_all_axis_major": (ctypes.c_int, [_vp, _i64, _i32, _i32, _vp, _i32, _i64, _vp, _vp, _vp]),
    "hr_rmsd_nosuperpose": (ctypes.c_int, [_vp, _i64, _i32, _vp, _i32, _vp, _vp, _vp, _vp, _vp]),
    "hr_superpose_workspace_bytes": (_sz, [_i64, _i32]),
    "hr_superpose": (ctypes.c_int, [_vp, _i64, _i32, _vp, _i32, _vp, _i32, _vp, _vp, _i32, _vp, _vp, _vp, _vp]),
    "hr_align_displace_workspace_bytes": (_sz, [_i64, _i32]),
    "hr_rmsd_align_displace": (ctypes.c_int, [_vp, _i64, _i32, _i32, _vp, _vp, _i32, _i32, _vp, _vp, _vp, _vp, _vp]),
    "hr_rmsf_workspace_bytes": (_sz, [_i64, _i32]),
    "hr_rmsf": (ctypes.c_int, [_vp, _i64, _i32, _vp, _i32, _vp, _vp, _i32, _vp, _vp, _vp]),
    "hr_rmsd_all_pairs_workspace_bytes": (_sz, [_i64, _i32, _i32]),
    "hr_rmsd_all_pairs": (ctypes.c_int, [_vp, _i64, _i32, _vp, _i32, _i32, _vp, _vp, _sz, _vp]),
    "hr_split_bytes": (_sz, [_i64, _i32, _i32]),
    "hr_split_prepare": (ctypes.c_int, [_vp, _i64, _i32, _vp, _i32, _i32, _vp, _sz, _vp]),
    "hr_rmsd_cross": (ctypes.c_int, [_vp, _i64, _vp, _i64, _i32, _i32, _vp, _i64, _vp]),
    "hr_batch_plan_create": (ctypes.c_int, [_vp, _i32, _vp, _vp]),
    "hr_batch_plan_launch": (ctypes.c_int, [_vp, _vp]),
    "hr_batch_plan_destroy": (ctypes.c_int, [_vp]),
    "hr_batch_plan_units": (_i64, [_vp]),
    "hr_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "hr_alloc": (ctypes.c_int, [ctypes.POINTER(_vp), _sz]),
    "hr_free": (ctypes.c_int, [_vp]),
    "hr_alloc_host": (ctypes.c_int, [ctypes.POINTER(_vp), _sz]),
    "hr_free_host": (ctypes.c_int, [_vp]),
    "hr_memcpy_h2d": (ctypes.c_int, [_vp, _vp, _sz, _vp]),
    "hr_memcpy_d2h": (ctypes.c_int, [_vp, _vp, _sz, _vp]),
    "hr_memcpy_d2d": (ctypes.c_int, [_vp, _vp, _sz, _vp]),
    "hr_memset": (ctypes.c_int, [_vp, ctypes.c_int, _sz, _vp]),
    "hr_stream_create": (ctypes.c_int, [ctypes.POINTER(_vp)]),
    "hr_stream_destroy": (ctypes.c_int, [_vp]),
    "hr_stream_synchronize": (ctypes.c_int, [_vp]),
    "hr_rmsd_one_vs_all_host": (ctypes.c_int, [_vp, _i64, _i32, _vp, _i32, _vp, _i32, _vp, _i64,
                                               _vp]),
    "hr_gather_atoms": (ctypes.c_int, [_vp, _i64, _i32, _vp, _i32, _vp, _vp]),
    "hr_hbm_read_probe": (ctypes.c_int, [_vp, _sz, _vp, ctypes.c_int, _vp]),
    "hr_hbm_read_probe_nt": (ctypes.c_int, [_vp, _sz, _vp, ctypes.c_int, _vp]),
    "hr_xtc_scan": (ctypes.c_int, [_vp, _i64, _vp, _i64, _vp, _vp]),
    "hr_rmsd_cross_argmin_workspace_bytes": (_sz, [_i64]),
    "hr_rmsd_cross_argmin": (ctypes.c_int, [_vp, _i64, _vp, _i64, _i32, _i32, _vp, _vp, _vp, _vp]),
    "hr_centroid_workspace_bytes": (_sz, [_i64]),
    "hr_centroid_scores": (ctypes.c_int, [_vp, _i64, ctypes.c_double, _vp, _vp, _vp, _vp]),
    "hr_topk_rows": (ctypes.c_int, [_vp, _i64, _i64, _i64, ctypes.c_int32, _vp, _vp, _vp]),
    "hr_xtc_decode": (ctypes.c_int, [_vp, _i64, _vp, _i64, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "hr_dcd_scan": (ctypes.c_int, [_vp, _i64, _vp, _vp, _i32]),
    "hr_lprmsd_workspace_bytes": (_sz, [_i32, _i32]),
    "hr_lprmsd": (ctypes.c_int, [_vp, _i64, _i32, _vp, _vp, _i32, _vp, _i32, _vp, _vp, _i32,
                                 _i32, _i32, _vp, _vp, _vp]),
    "hr_dcd_decode": (ctypes.c_int, [_vp, _i64, _vp, _vp, _i64, _vp, _i32, _vp, ctypes.c_float,
                                     _vp, _vp, _vp, _vp]),
}

_lock = threading.Lock()
_lib = None


class SystemDesc(ctypes.Structure):
    """hr_system_desc of include/hiprmsd.h."""
    _fields_ = [("xyz", ctypes.c_void_p), ("n_frames", ctypes.c_int64),
                ("n_atoms", ctypes.c_int32), ("ref_frame", ctypes.c_int32),
                ("out", ctypes.c_void_p)]


class DcdHeader(ctypes.Structure):
    """hr_dcd_header of include/hiprmsd.h."""
    _fields_ = [("n_atoms", ctypes.c_int32), ("n_fixed", ctypes.c_int32),
                ("n_frames", ctypes.c_int64), ("first_frame", ctypes.c_int64),
                ("first_frame_bytes", ctypes.c_int64), ("frame_bytes", ctypes.c_int64),
                ("flags", ctypes.c_int32), ("reverse_endian", ctypes.c_int32),
                ("n_set", ctypes.c_int32), ("istart", ctypes.c_int32),
                ("nsavc", ctypes.c_int32), ("pad", ctypes.c_int32), ("delta", ctypes.c_double)]


class BatchShape(ctypes.Structure):
    """hr_batch_shape of include/hiprmsd.h."""
    _fields_ = [("frames_per_unit", ctypes.c_int32), ("unit_kib", ctypes.c_int32),
                ("waves_per_simd", ctypes.c_int32)]


class HipRmsdError(RuntimeError):
    """A hiprmsd entry point returned an error status."""


def exported_symbols():
    """Names the C ABI declares (used by the ABI test)."""
    return list(_SIGNATURES)


def load():
    """Load the library once; RuntimeError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is not built: run `python -m mdtraj_amd._build` "
                    "(hipcc, gfx950). There is no CPU fallback.")
            lib = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in _SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def check(status: int, what: str) -> None:
    if status != HR_OK:
        msg = load().hr_last_error()
        msg = msg.decode() if msg else ""
        raise HipRmsdError(f"{what} failed (status {status}): {msg}")


def call(name: str, *args):
    """Call entry point `name` and raise on a non-zero status."""
    fn = getattr(load(), name)
    check(fn(*args), name)
