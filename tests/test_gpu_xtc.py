"""XTC decoding on the GPU: every fixture decodes bit-identically to the
reference's reader (xyz, time, step, box, precision), the reader's
n_frames / stride / atom_indices / seek semantics match slices of the
whole-file decode, a corrupted stream is refused or changes the frame
(never a silent pass), and a trajectory loaded from a file feeds the RMSD
kernels directly from HBM."""
import os

import numpy as np
import pytest

import mdtraj_amd as md
from mdtraj_amd.xtc import XTCTrajectoryFile, load_xtc
from oracle import fp64

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
G = np.load(os.path.join(HERE, "xtc_golden.npz"))
CASES = sorted({k.split("_")[0] for k in G.files})


def path(name):
    return os.path.join(HERE, f"xtc_{name}.xtc")


@pytest.mark.parametrize("name", CASES)
def test_decode_equals_reference_reader(gpu, name):
    with XTCTrajectoryFile(path(name)) as f:
        xyz, time, step, box = f.read()
        prec = f.last_precision
    assert xyz.dtype == np.float32 and np.array_equal(xyz, G[f"{name}_xyz"])
    assert np.array_equal(time, G[f"{name}_time"])
    assert np.array_equal(step, G[f"{name}_step"])
    want_box = G[f"{name}_box"]
    if np.all(np.abs(want_box) < 1e-10):
        assert box is None
    else:
        assert np.array_equal(box, want_box)
    if xyz.shape[1] > 9:
        assert np.array_equal(prec, G[f"{name}_prec"])


def test_read_semantics(gpu):
    want = G["frame0_xyz"]
    with XTCTrajectoryFile(path("frame0")) as f:
        a = f.read(n_frames=10)[0]
        assert np.array_equal(a, want[:10]) and f.tell() == 10
        b = f.read(n_frames=7, stride=3)[0]
        assert np.array_equal(b, want[10:31:3]) and f.tell() == 31
        f.seek(100)
        c = f.read(stride=50, atom_indices=[3, 1, 20])[0]
        assert np.array_equal(c, want[100::50][:, [3, 1, 20]])
        f.seek(-2, whence=2)
        d, t, s, _ = f.read()
        assert np.array_equal(d, want[-2:]) and np.array_equal(t, G["frame0_time"][-2:])
        assert f.read()[0].shape == (0, 22, 3)
        f.seek(0)
        e = f.read(atom_indices=slice(5, 9), device=True)[0]
        assert e.is_cuda and np.array_equal(e.cpu().numpy(), want[:, 5:9])


def test_corrupted_stream_is_not_silently_accepted(gpu):
    import tempfile

    buf = bytearray(open(path("walk"), "rb").read())
    buf[52 + 40 + 200] ^= 0x24  # inside frame 0's compressed bytes
    with tempfile.NamedTemporaryFile(suffix=".xtc", delete=False) as fh:
        fh.write(bytes(buf))
    try:
        with XTCTrajectoryFile(fh.name) as f:
            try:
                xyz = f.read()[0]
            except RuntimeError as e:
                assert "XTC read error" in str(e)
                return
        assert not np.array_equal(xyz[0], G["walk_xyz"][0])
        assert np.array_equal(xyz[1:], G["walk_xyz"][1:])  # frames are independent
    finally:
        os.unlink(fh.name)


def test_file_to_rmsd_in_hbm(gpu):
    """load_xtc -> device Trajectory -> md.rmsd, against the oracle on the
    reference-decoded coordinates."""
    t = load_xtc(path("water"), stride=2)
    assert t.is_on_device
    want_xyz = G["water_xyz"][::2]
    got = md.rmsd(t, t, 3)
    assert got.is_cuda  # device in, device out
    got = got.cpu().numpy()
    assert np.abs(got - fp64.rmsd(want_xyz, want_xyz, 3)).max() < 1e-5
    assert np.array_equal(t.time, G["water_time"][::2])
