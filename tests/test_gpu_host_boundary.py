"""The host-memory boundary: hr_rmsd_one_vs_all_host (host pointers in and
out, as the reference's C ABI, rmsd/include/theobald_rmsd.h:16-18), the
library's own device buffers (hr_alloc / hr_memcpy_*), a Python process that
never imports torch, and a plain C caller.

Bars: within 1e-5 nm of the FP64 oracle; the reference frame scored against
itself is exactly 0; host and device paths agree to 1e-6 nm (both accumulate
in FP64; the host path gathers the selection before the copy, so the kernel's
row width can differ)."""
import ctypes
import os
import struct
import subprocess
import sys
import textwrap

import numpy as np
import pytest

import mdtraj_amd as md
from mdtraj_amd import _lib, gen
from oracle import fp64

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOL_FP64 = 1e-5


@pytest.mark.parametrize("F,N,n_sel,frame", [
    (1000, 22, None, 0),        # the reference's CPU-sized case, one chunk
    (3000, 300, 220, 17),       # gathered selection
    (50_000, 500, 400, 20_000),  # 240 MB selected: four 64 MiB chunks, slot 0 reused
])
def test_host_entry_vs_fp64(gpu, F, N, n_sel, frame):
    x = gen.walk(seed=F % 97, n_frames=F, n_atoms=N)
    idx = None if n_sel is None else np.sort(
        np.random.default_rng(F).choice(N, size=n_sel, replace=False))
    t = md.Trajectory(x)
    got = md.rmsd(t, t, frame, atom_indices=idx)
    assert isinstance(got, np.ndarray) and got.dtype == np.float32 and got.shape == (F,)
    want = fp64.rmsd(x, x, frame, idx)
    assert np.abs(got - want).max() < TOL_FP64
    assert got[frame] == 0.0
    # the device-resident path scores the same frames
    d = md.rmsd(md.Trajectory(gpu.from_numpy(x).cuda()), t, frame, atom_indices=idx)
    assert np.abs(d.cpu().numpy() - got).max() < 1e-6


@pytest.mark.parametrize("F,N,n_sel", [(1000, 22, None), (5000, 300, 220), (200, 9000, 7000)])
def test_direct_and_copied_staging_agree(gpu, F, N, n_sel):
    """Small inputs are read by the kernels straight from pinned memory;
    forcing the copied path (HR_TUNE_HOST_DIRECT_KIB = -1) gives the same
    bits, since the same kernels read the same gathered values."""
    x = gen.walk(seed=N, n_frames=F, n_atoms=N)
    idx = None if n_sel is None else np.sort(
        np.random.default_rng(N).choice(N, size=n_sel, replace=False))
    t = md.Trajectory(x)
    got = {}
    try:
        for kib in (0, -1, 1 << 20):
            _lib.call("hr_set_tuning", _lib.HR_TUNE_HOST_DIRECT_KIB, kib)
            got[kib] = md.rmsd(t, t, 3, atom_indices=idx)
    finally:
        _lib.call("hr_set_tuning", _lib.HR_TUNE_HOST_DIRECT_KIB, 0)
    assert np.array_equal(got[0], got[-1]) and np.array_equal(got[0], got[1 << 20])
    assert np.abs(got[0] - fp64.rmsd(x, x, 3, idx)).max() < TOL_FP64
    assert got[0][3] == 0.0


@pytest.mark.parametrize("F", [1, 2, 15, 16, 17, 63, 64, 65, 191, 193, 1000])
@pytest.mark.parametrize("N", [3, 22, 150])
def test_single_launch_small_calls(gpu, F, N):
    """Small host calls run one kernel that prepares the reference itself and
    assigns every frame set statically when there is at most one per wave;
    frame counts around the row and wave boundaries give the same bits as the
    copied path (reference prep kernel, then the per-frame kernel)."""
    x = gen.walk(seed=F * 7 + N, n_frames=F, n_atoms=N)
    t = md.Trajectory(x)
    frame = F // 2
    got = {}
    try:
        for kib in (0, -1):
            _lib.call("hr_set_tuning", _lib.HR_TUNE_HOST_DIRECT_KIB, kib)
            got[kib] = md.rmsd(t, t, frame)
    finally:
        _lib.call("hr_set_tuning", _lib.HR_TUNE_HOST_DIRECT_KIB, 0)
    assert np.array_equal(got[0], got[-1])
    assert got[0][frame] == 0.0
    assert np.abs(got[0] - fp64.rmsd(x, x, frame)).max() < TOL_FP64


def test_host_entry_ref_indices_and_other_reference(gpu):
    x = gen.walk(seed=5, n_frames=400, n_atoms=120)
    r = gen.walk(seed=6, n_frames=3, n_atoms=150)
    a = np.arange(10, 100)
    b = np.arange(50, 140)
    got = md.rmsd(md.Trajectory(x), md.Trajectory(r), 2, atom_indices=a, ref_atom_indices=b)
    want = fp64.rmsd(x, r, 2, a, b)
    assert np.abs(got - want).max() < TOL_FP64
    assert got.min() > 0   # a different reference: no exact zero


def test_host_entry_refuses_bad_indices(gpu):
    x = np.zeros((4, 10, 3), np.float32)
    out = np.zeros(4, np.float32)
    bad = np.array([0, 10], np.int32)
    with pytest.raises(_lib.HipRmsdError, match="atom_idx"):
        _lib.call("hr_rmsd_one_vs_all_host", x.ctypes.data, 4, 10, bad.ctypes.data, 2,
                  x.ctypes.data, 10, None, -1, out.ctypes.data)
    good = np.array([0, 9], np.int32)
    with pytest.raises(_lib.HipRmsdError, match="ref_idx"):
        _lib.call("hr_rmsd_one_vs_all_host", x.ctypes.data, 4, 10, good.ctypes.data, 2,
                  x.ctypes.data, 5, good.ctypes.data, -1, out.ctypes.data)


def test_device_array_trajectory(gpu):
    """Library-owned device arrays carry a trajectory without torch types:
    md.rmsd reads it in place and returns a DeviceArray; superpose moves it in
    its own memory."""
    from mdtraj_amd._device import DeviceArray

    x = gen.walk(seed=8, n_frames=300, n_atoms=90)
    d = DeviceArray.from_host(x)
    t = md.Trajectory(d)
    assert t.is_on_device and t.xyz is d
    got = md.rmsd(t, t, 4)
    assert isinstance(got, DeviceArray)
    assert np.abs(got.numpy() - fp64.rmsd(x, x, 4)).max() < TOL_FP64
    ptr = d.data_ptr()
    t.superpose(t, 4, atom_indices=np.arange(0, 90, 2))
    assert t.xyz.data_ptr() == ptr
    assert np.abs(t.xyz.numpy() - fp64.superpose(x, x, 4, np.arange(0, 90, 2))).max() < 2e-5


_NO_TORCH_SCRIPT = textwrap.dedent("""
    import sys
    sys.path.insert(0, {root!r})
    import numpy as np
    import mdtraj_amd as md
    from mdtraj_amd import _rmsd, dcd, gen, xtc
    from mdtraj_amd._device import DeviceArray

    x = gen.walk(seed=31, n_frames=500, n_atoms=200)
    idx = np.arange(0, 200, 3)
    t = md.Trajectory(x)
    out = {{}}
    out["rmsd_host"] = md.rmsd(t, t, 7, atom_indices=idx)
    out["rmsd_nosup"] = md.rmsd(t, t, 7, atom_indices=idx, superpose=False)
    out["rmsd_device"] = md.rmsd(md.Trajectory(DeviceArray.from_host(x)), t, 7).numpy()
    s = md.Trajectory(x.copy())
    s.superpose(t, 7, atom_indices=idx)
    out["superpose"] = s.xyz
    c = md.Trajectory(x.copy())
    c.center_coordinates()
    out["centered"] = c.xyz
    out["traces"] = c._rmsd_traces
    out["rmsf"] = md.rmsf(t, t, 7, atom_indices=idx)
    out["matrix"] = md.rmsd_matrix(x[:120])
    out["lprmsd"] = md.lprmsd(t, t, 7, atom_indices=idx, permute_groups=[idx[:20]])
    with xtc.XTCTrajectoryFile({xtc_path!r}) as f:
        xyz, *_ = f.read(atom_indices=[0, 2, 4])
    out["xtc_sel"] = xyz
    with xtc.XTCTrajectoryFile({xtc_path!r}) as f:
        dxyz, *_ = f.read(atom_indices=[0, 2, 4], device=True)
    assert isinstance(dxyz, DeviceArray)
    out["xtc_sel_device"] = dxyz.numpy()
    with dcd.DCDTrajectoryFile({dcd_path!r}) as f:
        out["dcd_sel"], out["dcd_lengths"], _ = f.read(atom_indices=[1, 3, 5])
    out["dcd_nm"] = dcd.load_dcd({dcd_path!r}).xyz.numpy()
    assert "torch" not in sys.modules, "torch was imported"
    np.savez({out_path!r}, **out)
    print("no-torch process done")
""")


def test_process_without_torch(gpu, tmp_path, golden):
    """md.rmsd, superpose, centring, rmsf, the all-pairs matrix, LP-RMSD and
    the XTC and DCD readers in a Python process that never imports torch."""
    xtc_path = os.path.join(ROOT, "tests", "golden", "xtc_small10.xtc")
    out_path = str(tmp_path / "no_torch.npz")
    dcd_path = os.path.join(ROOT, "tests", "golden", "dcd_charmm_cell.dcd")
    script = _NO_TORCH_SCRIPT.format(root=ROOT, xtc_path=xtc_path, dcd_path=dcd_path,
                                     out_path=out_path)
    print("running the no-torch child process", flush=True)
    r = subprocess.run([sys.executable, "-c", script], capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.load(out_path)
    x = gen.walk(seed=31, n_frames=500, n_atoms=200)
    idx = np.arange(0, 200, 3)
    assert np.abs(got["rmsd_host"] - fp64.rmsd(x, x, 7, idx)).max() < TOL_FP64
    assert got["rmsd_host"][7] == 0.0
    assert np.abs(got["rmsd_nosup"] - fp64.rmsd_nosuperpose(x, x, 7, idx)).max() < TOL_FP64
    assert np.abs(got["rmsd_device"] - fp64.rmsd(x, x, 7)).max() < TOL_FP64
    assert np.abs(got["superpose"] - fp64.superpose(x, x, 7, idx)).max() < 2e-5
    c, g = fp64.center_and_trace(x)
    assert np.abs(got["centered"] - c).max() < 2e-6
    assert np.abs(got["traces"] - g).max() < 1e-5 * np.abs(g).max()
    assert np.abs(got["rmsf"] - fp64.rmsf(x, x, 7, idx)).max() < TOL_FP64
    assert np.abs(got["matrix"] - fp64.condensed(fp64.rmsd_matrix(x[:120]))).max() < TOL_FP64
    lp = got["lprmsd"]
    assert lp.shape == (500,) and np.all(np.isfinite(lp))
    assert np.all(lp <= fp64.rmsd(x, x, 7, idx) + 1e-5)   # relabelling only lowers it
    ref = np.load(os.path.join(ROOT, "tests", "golden", "xtc_golden.npz"))
    full = ref["small10_xyz"]
    assert np.array_equal(got["xtc_sel"], full[:, [0, 2, 4]])
    assert np.array_equal(got["xtc_sel_device"], full[:, [0, 2, 4]])
    d = np.load(os.path.join(ROOT, "tests", "golden", "dcd_golden.npz"))
    assert np.array_equal(got["dcd_sel"], d["charmm_cell/xyz"][:, [1, 3, 5]])
    assert np.array_equal(got["dcd_lengths"], d["charmm_cell/cell"][:, :3].astype(np.float32))
    assert np.allclose(got["dcd_nm"], d["charmm_cell/xyz"] * np.float32(0.1), rtol=1e-6)


def test_c_caller_without_hip_headers(gpu, tmp_path):
    """tests/c_host_example.c: the host-pointer entry point and the
    hr_alloc / hr_memcpy route, from plain C linked with gcc."""
    exe = str(tmp_path / "c_host_example")
    lib_dir = os.path.join(ROOT, "mdtraj_amd")
    subprocess.run(["gcc", "-O2", "-Wall", f"-I{os.path.join(ROOT, 'include')}",
                    os.path.join(ROOT, "tests", "c_host_example.c"), f"-L{lib_dir}", "-lhiprmsd",
                    f"-Wl,-rpath,{lib_dir}", "-o", exe], check=True)
    F, N = 700, 150
    x = gen.walk(seed=12, n_frames=F, n_atoms=N)
    idx = np.arange(3, N, 2).astype(np.int32)
    inp, outp = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    with open(inp, "wb") as f:
        f.write(struct.pack("<qii", F, N, idx.size) + x.tobytes() + idx.tobytes())
    print("running the C caller", flush=True)
    r = subprocess.run([exe, inp, outp], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    res = np.fromfile(outp, np.float32)
    host, devp = res[:F], res[F:]
    want = fp64.rmsd(x, x, 0, idx)
    assert np.abs(host - want).max() < TOL_FP64
    assert np.abs(devp - want).max() < TOL_FP64
    assert host[0] == 0.0 and devp[0] == 0.0


def test_host_entry_small_case_latency(gpu):
    """The reference's CPU-sized case through md.rmsd with host arrays: one
    host-pointer call per query.  Reported, not asserted beyond a sanity
    bound (the bench line carries the comparison with the CPU core)."""
    import time

    x, _ = gen.small_cube(seed=0, n_frames=1000, n_atoms=22, sigma=0.02)
    t = md.Trajectory(x)
    for _ in range(20):
        md.rmsd(t, t, 0)
    ts = []
    for _ in range(200):
        t0 = time.perf_counter()
        md.rmsd(t, t, 0)
        ts.append(time.perf_counter() - t0)
    med = float(np.median(ts))
    print(f"md.rmsd 1,000 x 22 host arrays: median {med * 1e6:.1f} us per call")
    assert med < 5e-3
