"""DCD decoding on the GPU against the reference's reader.

Every fixture's coordinates and cells are compared bit for bit with what the
reference's reader returned (tests/golden/dcd_golden.npz), for full,
strided and streaming reads and atom subsets; load_dcd's nanometres against
the reference's float32 conversion; a larger file with fixed atoms in the
opposite byte order against the oracle; out-of-range frame numbers through
the C ABI end with a status word, not an access outside the bytes."""
import ctypes
import importlib.util
import os
import tempfile
import warnings

import numpy as np
import pytest

import mdtraj_amd as md
from mdtraj_amd import _lib
from mdtraj_amd.dcd import DCDTrajectoryFile, load_dcd
from oracle import dcd as odcd
from oracle import fp64

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
G = np.load(os.path.join(HERE, "dcd_golden.npz"))
CASES = sorted({k.split("/")[0] for k in G.files})


def path(name):
    return os.path.join(HERE, f"dcd_{name}.dcd")


def check(got, xyz, cell):
    x, lengths, angles = got
    assert np.array_equal(x, xyz)
    if lengths is None:
        assert np.all(cell[:, :3] < 1e-10)
    else:
        assert np.array_equal(lengths, cell[:, :3]) and np.array_equal(angles, cell[:, 3:])


@pytest.mark.parametrize("name", CASES)
def test_read_matches_reference(gpu, name):
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)  # the bad-record file warns
        with DCDTrajectoryFile(path(name)) as f:
            check(f.read(), G[f"{name}/xyz"], G[f"{name}/cell"])
        with DCDTrajectoryFile(path(name)) as f:
            f.seek(2)
            check(f.read(stride=3), G[f"{name}/xyz_2_3"], G[f"{name}/cell_2_3"])


def test_streaming_subsets_and_positions(gpu):
    want, cell = G["fixed_bigendian/xyz"], G["fixed_bigendian/cell"]
    with DCDTrajectoryFile(path("fixed_bigendian")) as f:
        for i in range(4):  # one frame at a time (the reference's test_read_3)
            x, L, A = f.read(1)
            assert np.array_equal(x, want[i:i + 1]) and np.array_equal(L, cell[i:i + 1, :3])
        x, _, _ = f.read()  # the rest (test_read_4)
        assert np.array_equal(x, want[4:]) and f.tell() == len(want)
        assert f.read()[0].shape == (0, 26, 3)
        f.seek(0)
        x, _, _ = f.read(n_frames=1000, stride=2, atom_indices=[1, 2, 5])  # test_read_5
        assert np.array_equal(x, want[::2][:, [1, 2, 5]])
        f.seek(1)
        e = f.read(atom_indices=slice(3, 20, 4), device=True)[0]
        assert e.is_cuda and np.array_equal(e.cpu().numpy(), want[1:, 3:20:4])
    with DCDTrajectoryFile(path("badrecord")) as f:
        with pytest.warns(RuntimeWarning, match="frame 3"):
            x, _, _ = f.read()
        assert len(x) == 3 and f.tell() == 4


def test_load_dcd_in_nanometres(gpu):
    t = load_dcd(path("charmm_cell"), stride=2)
    assert t.is_on_device
    want = G["charmm_cell/xyz"][::2] * np.float32(0.1)  # the reference's in-place float32 scaling
    assert np.array_equal(t.xyz.cpu().numpy(), want)
    assert np.array_equal(t.unitcell_lengths, G["charmm_cell/cell"][::2, :3] * np.float32(0.1))
    assert np.array_equal(t.unitcell_angles, G["charmm_cell/cell"][::2, 3:])
    assert np.array_equal(t.time, 2 * np.arange(6))
    one = load_dcd(path("xplor"), frame=5, atom_indices=[0, 3])
    assert one.n_frames == 1 and np.array_equal(one.time, [5])
    assert np.array_equal(one.xyz.cpu().numpy(), G["xplor/xyz"][5:6, [0, 3]] * np.float32(0.1))
    assert one.unitcell_lengths is None
    # file -> HBM -> md.rmsd against the oracle on the reference-decoded coordinates
    r = md.rmsd(t, t, 1).cpu().numpy()
    assert np.abs(r - fp64.rmsd(want, want, 1)).max() < 1e-5


def _writer():
    spec = importlib.util.spec_from_file_location("make_dcd_golden",
                                                  os.path.join(HERE, "make_dcd_golden.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    return mk


@pytest.mark.parametrize("endian,fourdim", [(">", False), ("<", True)])
def test_large_fixed_atoms(gpu, endian, fourdim):
    # (the reference tests the fourth-dimension flag without byte swapping,
    # so it never sees one in an opposite-order file: that pair is pinned on
    # the CPU against its reader)
    mk = _writer()
    rng = np.random.default_rng(7)
    F, N = 1500, 700
    x = mk.walk(rng, F, N)
    free = np.sort(rng.choice(N, 500, replace=False)) + 1
    with tempfile.NamedTemporaryFile(suffix=".dcd", delete=False) as fh:
        name = fh.name
    try:
        mk.write_dcd(name, x, endian, cell=mk.cosine_cell(rng, F), free=free, fourdim=fourdim)
        data = open(name, "rb").read()
        want, L, A, _ = odcd.read(data)
        with DCDTrajectoryFile(name) as f:
            got, gl, ga = f.read()
        assert np.array_equal(got, want) and np.array_equal(gl, L) and np.array_equal(ga, A)
        fixed = np.setdiff1d(np.arange(N), free - 1)
        assert np.array_equal(got[:, fixed], np.broadcast_to(x[0, fixed], got[:, fixed].shape))
        assert np.array_equal(got[:, free - 1], x[:, free - 1])  # float32 stored as written
    finally:
        os.unlink(name)


def test_frames_outside_the_file_get_a_status(gpu):
    import torch

    d = open(path("charmm_cell"), "rb").read()
    lib = _lib.load()
    h = _lib.DcdHeader()
    host = np.frombuffer(d, np.uint8)
    assert lib.hr_dcd_scan(host.ctypes.data_as(ctypes.c_void_p), host.size, ctypes.byref(h), None, 0) == 0
    dev = torch.from_numpy(host.copy()).cuda()
    frames = torch.tensor([0, 11, 12, -1, 5], dtype=torch.int64, device="cuda")
    xyz = torch.zeros((5, 23, 3), dtype=torch.float32, device="cuda")
    status = torch.full((5,), 99, dtype=torch.int32, device="cuda")
    rc = lib.hr_dcd_decode(ctypes.c_void_p(dev.data_ptr()), dev.numel(), ctypes.byref(h),
                           ctypes.c_void_p(frames.data_ptr()), 5, None, 23, None,
                           ctypes.c_float(1.0), ctypes.c_void_p(xyz.data_ptr()), None,
                           ctypes.c_void_p(status.data_ptr()), None)
    torch.cuda.synchronize()
    assert rc == 0
    assert status.cpu().tolist() == [0, 0, 2, 2, 0]
    out = xyz.cpu().numpy()
    assert np.array_equal(out[[0, 1, 4]], G["charmm_cell/xyz"][[0, 11, 5]])
    assert np.isnan(out[[2, 3]]).all()
