"""hr_rmsd_one_vs_all_noalloc: the one-vs-all entry point with its workspace
in a static per-device scratch array (no allocation, no caller workspace).
Its output must equal hr_rmsd_one_vs_all's with a caller workspace bit for
bit (the same kernels read the same sums), on two calls in a row, the second
smaller than the first, and match the FP64 oracle within 1e-5 nm."""
import ctypes

import numpy as np
import pytest

from mdtraj_amd import _lib, gen
from mdtraj_amd._device import DeviceArray
from oracle import fp64

pytestmark = pytest.mark.gpu

_vp, _i32, _i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64


def _noalloc():
    lib = _lib.load()
    fn = lib.hr_rmsd_one_vs_all_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [_vp, _i64, _i32, _vp, _i32, _vp, _vp, _i64, _vp, _vp]
    lib.hr_rmsd_one_vs_all_noalloc_max_sel.restype = ctypes.c_size_t
    return fn, int(lib.hr_rmsd_one_vs_all_noalloc_max_sel())


def _both(x, idx, frame):
    """(no-alloc output, workspace output) for x (F, N, 3) against x[frame]."""
    fn, _ = _noalloc()
    F, N = x.shape[0], x.shape[1]
    n_sel = N if idx is None else len(idx)
    d_x = DeviceArray.from_host(x)
    d_idx = None if idx is None else DeviceArray.from_host(np.asarray(idx, np.int32))
    ip = None if d_idx is None else d_idx.data_ptr()
    out_a = DeviceArray.empty(F, np.float32)
    out_b = DeviceArray.empty(F, np.float32)
    ws = DeviceArray.empty(_lib.load().hr_rmsd_workspace_bytes(n_sel), np.uint8)
    ref = d_x[frame]
    _lib.check(fn(d_x.data_ptr(), F, N, ip, n_sel, ref.data_ptr(), ip, frame, out_a.data_ptr(),
                  None), "hr_rmsd_one_vs_all_noalloc")
    _lib.call("hr_rmsd_one_vs_all", d_x.data_ptr(), F, N, ip, n_sel, ref.data_ptr(), ip, frame,
              out_b.data_ptr(), ws.data_ptr(), None)
    return out_a.numpy(), out_b.numpy()


def test_noalloc_equals_workspace_call_twice(gpu):
    # first call: 3,000 frames x 500 atoms, a 400-atom selection
    x1 = gen.walk(seed=41, n_frames=3000, n_atoms=500)
    i1 = np.sort(np.random.default_rng(41).choice(500, size=400, replace=False))
    print("no-alloc call 1: 3000 x 500, 400 selected, frame 11", flush=True)
    a1, b1 = _both(x1, i1, 11)
    assert np.array_equal(a1, b1)
    assert np.abs(a1 - fp64.rmsd(x1, x1, 11, i1)).max() < 1e-5
    assert a1[11] == 0.0
    # second call, smaller: 700 frames x 120 atoms, all atoms; the scratch
    # array still holds the first call's larger reference
    x2 = gen.walk(seed=42, n_frames=700, n_atoms=120)
    print("no-alloc call 2: 700 x 120, all atoms, frame 3", flush=True)
    a2, b2 = _both(x2, None, 3)
    assert np.array_equal(a2, b2)
    assert np.abs(a2 - fp64.rmsd(x2, x2, 3)).max() < 1e-5
    assert a2[3] == 0.0


def test_noalloc_refuses_selections_beyond_its_scratch(gpu):
    fn, max_sel = _noalloc()
    assert max_sel == 65536
    x = DeviceArray.empty((1, max_sel + 1, 3), np.float32)
    out = DeviceArray.empty(1, np.float32)
    rc = fn(x.data_ptr(), 1, max_sel + 1, None, max_sel + 1, x.data_ptr(), None, -1,
            out.data_ptr(), None)
    assert rc == _lib.HR_ERR_INVALID
    assert b"65536" in _lib.load().hr_last_error()
