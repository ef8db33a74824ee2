"""The C-ABI library loads and exports every symbol include/hiprmsd.h declares
(no compute call: this runs without a GPU)."""
import ctypes
import os
import re

from mdtraj_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "hiprmsd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hr_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "hr_rmsd_one_vs_all" in names and "hr_superpose" in names
    assert set(names) == set(_lib.exported_symbols())


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name


def test_host_only_entry_points():
    lib = _lib.load()
    assert lib.hr_version() >= 100
    assert lib.hr_rmsd_workspace_bytes(1500) >= 1500 * 12
    assert lib.hr_superpose_workspace_bytes(1000, 100) >= 1000 * (24 + 36)


def _status(lib, name, *args):
    rc = getattr(lib, name)(*args)
    return rc, (lib.hr_last_error() or b"").decode()


def test_invalid_arguments_are_rejected_on_the_host():
    """Argument checks run before any device work, so they answer here:
    a positive HR_ERR_* status and a message naming the problem."""
    lib = _lib.load()
    HR_ERR_INVALID = 1
    ws = ctypes.create_string_buffer(64)
    cases = [
        ("hr_rmsd_one_vs_all", (None, -1, 10, None, 10, None, None, -1, None, None, None),
         "invalid sizes"),
        ("hr_rmsd_one_vs_all", (None, 5, 10, None, 7, None, None, -1, None, None, None),
         "n_sel"),
        ("hr_rmsd_one_vs_all", (None, 5, 10, None, 10, None, None, -1, None, None, None),
         "NULL"),
        ("hr_rmsd_all_pairs", (None, 10, 5, None, 5, 9, None, None, 0, None), "unknown mode"),
        ("hr_rmsd_cross", (None, 2, None, 2, 5, 0, None, 2, None), "not a split mode"),
        ("hr_rmsf", (None, 0, 5, None, 5, None, None, 1, None, None, None), "invalid sizes"),
        ("hr_rmsd_align_displace", (None, 3, 2, 4, None, None, 4, 4, None, None, None, None,
                                    None), "invalid sizes"),
        ("hr_batch_plan_create", (None, 0, None, ctypes.addressof(ws)), "no systems"),
        ("hr_set_tuning", (9, 1), "unknown tuning knob"),
        ("hr_set_tuning", (_lib.HR_TUNE_FRAMES_STATIC_PCT, 101), "must be 0 .. 100"),
        ("hr_set_tuning", (_lib.HR_TUNE_FRAMES_CHUNK_KIB, -1), "must be 0 or positive"),
        ("hr_set_tuning", (_lib.HR_TUNE_FRAMES_WAVES_PER_SIMD, 5), "must be 0, 3 or 4"),
        ("hr_set_tuning", (_lib.HR_TUNE_FRAMES_GRID, 3), "must be 0, 1 or 2"),
        ("hr_set_tuning", (_lib.HR_TUNE_FRAMES_ROW_LANES, 12), "must be 0, 4, 8, 16, 32 or 64"),
        ("hr_set_tuning", (_lib.HR_TUNE_BATCH_SCHEDULE, 5), "must be 0 .. 4"),
        ("hr_set_tuning", (_lib.HR_TUNE_SUPERPOSE_CHUNK_MIB, -1), "must be 0 or positive"),
        ("hr_centroid_scores", (ctypes.addressof(ws) + 4, 10, 1.0, ctypes.addressof(ws), None,
                                ctypes.addressof(ws), None), "16-byte aligned"),
        ("hr_centroid_scores", (None, -1, 1.0, None, None, None, None), "invalid n_frames"),
        ("hr_rmsd_cross_argmin", (None, 2, None, 2, 5, 0, None, None, None, None),
         "not a split mode"),
        ("hr_dcd_decode", (None, 0, None, None, 0, None, 0, None, 1.0, None, None, None, None),
         "NULL pointer"),
        ("hr_lprmsd", (None, 1, 10, None, None, 12, None, 0, None, None, 0, 0, 0, None, None, None),
         "invalid sizes"),
        ("hr_hbm_read_probe", (ctypes.addressof(ws), 24, ctypes.addressof(ws), 1, None),
         "multiple of 16"),
        ("hr_hbm_read_probe_nt", (ctypes.addressof(ws) + 4, 16, ctypes.addressof(ws), 1, None),
         "16-byte aligned"),
        ("hr_hbm_read_probe_nt", (ctypes.addressof(ws), 16, ctypes.addressof(ws), 0, None),
         "invalid probe arguments"),
    ]
    for name, args, text in cases:
        rc, msg = _status(lib, name, *args)
        assert rc == HR_ERR_INVALID, (name, rc, msg)
        assert text in msg, (name, msg)


def test_batch_shape_is_validated():
    lib = _lib.load()
    desc = (_lib.SystemDesc * 1)(_lib.SystemDesc(None, 10, 5, 0, None))
    plan = ctypes.c_void_p()
    for shape, text in (((65, 0, 3), "frames_per_unit"), ((8, -1, 3), "unit_kib"),
                        ((8, 0, 5), "waves_per_simd")):
        s = _lib.BatchShape(*shape)
        rc, msg = _status(lib, "hr_batch_plan_create", ctypes.addressof(desc), 1,
                          ctypes.addressof(s), ctypes.byref(plan))
        assert rc == 1 and text in msg, (shape, rc, msg)


def test_lprmsd_lds_limit_is_a_host_error():
    """A group too large for one wave's LDS is refused before any launch."""
    lib = _lib.load()
    rc, msg = _status(lib, "hr_lprmsd", None, 1, 3000, None, None, 3000, None, 0, None, None, 1,
                      3000, 0, None, None, None)
    assert rc == _lib.HR_ERR_UNSUPPORTED and "LDS" in msg
