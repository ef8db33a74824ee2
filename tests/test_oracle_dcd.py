"""The DCD oracle against the reference's reader.

tests/golden/dcd_golden.npz holds what the reference's own reader
(dcdplugin.c, as DCDTrajectoryFile.read drives it) returned for each fixture
file; every branch of the reader is covered by one of them (byte order,
X-PLOR and CHARMM headers, fixed atoms, a fourth dimension, cells as cosines
and in degrees, titles, trailing bytes, a bad record).  Where the reference
library is built (oracle/_ref), reads at other positions and strides are
compared live."""
import ctypes
import os

import numpy as np
import pytest

from oracle import dcd

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
NAMES = ["charmm_cell", "charmm_nocell", "bigendian_cell", "xplor", "fixed_degrees",
         "fixed_bigendian", "fourdim", "titles", "trailing", "badrecord"]
REF = os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref", "libref_dcd.so")


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(GOLDEN, "dcd_golden.npz"))


def data(name):
    with open(os.path.join(GOLDEN, f"dcd_{name}.dcd"), "rb") as fh:
        return fh.read()


def check(got, xyz, cell):
    x, lengths, angles, _ = got
    assert np.array_equal(x, xyz)
    if lengths is None:
        assert np.all(cell[:, :3] < 1e-10)
    else:
        assert np.array_equal(lengths, cell[:, :3]) and np.array_equal(angles, cell[:, 3:])


@pytest.mark.parametrize("name", NAMES)
def test_matches_reference_reader(golden, name):
    d = data(name)
    assert dcd.parse_header(d)["n_frames"] == int(golden[f"{name}/nsets"])
    check(dcd.read(d), golden[f"{name}/xyz"], golden[f"{name}/cell"])
    check(dcd.read(d, start=2, stride=3), golden[f"{name}/xyz_2_3"], golden[f"{name}/cell_2_3"])


def test_branches_are_the_ones_intended():
    h = {n: dcd.parse_header(data(n)) for n in NAMES}
    assert h["bigendian_cell"]["endian"] == ">" and h["fixed_bigendian"]["endian"] == ">"
    assert h["xplor"]["flags"] == 0 and abs(h["xplor"]["delta"] - 0.002) < 1e-15  # a double
    assert h["fourdim"]["flags"] & dcd.FOUR_DIMS
    assert h["fixed_degrees"]["namnf"] == 10 and h["fixed_bigendian"]["namnf"] == 17
    assert h["charmm_nocell"]["flags"] == dcd.CHARMM
    assert h["charmm_cell"]["flags"] == dcd.CHARMM | dcd.EXTRA_BLOCK
    # trailing bytes shorter than a frame do not make a frame
    assert h["trailing"]["n_frames"] == 8
    assert len(data("trailing")) - h["trailing"]["first"] == 8 * h["trailing"]["frame_bytes"] + 37


def test_closed_forms():
    # fixed atoms keep frame 0's coordinates in every frame; free ones move
    x, _, _, _ = dcd.read(data("fixed_degrees"))
    h = dcd.parse_header(data("fixed_degrees"))
    fixed = np.setdiff1d(np.arange(30), h["free"] - 1)
    assert np.array_equal(x[:, fixed], np.broadcast_to(x[0, fixed], x[:, fixed].shape))
    assert not np.array_equal(x[1, h["free"] - 1], x[0, h["free"] - 1])
    # cells stored in degrees come back as stored
    _, L, A, _ = dcd.read(data("fixed_degrees"))
    assert np.all(L == 50.0) and np.array_equal(A[0], np.float32([85.0, 95.0, 100.0]))
    # orthogonal cosines (0) give exactly 90 degrees
    assert dcd.cell_lengths_angles(np.float32([[10, 0, 11, 0, 0, 12]]))[1].tolist() == [[90, 90, 90]]
    # the reference writer's cosines of (90, 80, 70) come back within float rounding
    _, L, A, _ = dcd.read(data("charmm_cell"))
    assert np.allclose(A, [90.0, 80.0, 70.0], atol=1e-4)
    # a bad record ends the read at the frame before it, as end of file does
    x, _, _, pos = dcd.read(data("badrecord"))
    assert len(x) == 3 and pos == 4


def test_refuses_malformed_headers():
    d = bytearray(data("charmm_cell"))
    with pytest.raises(dcd.DcdError):
        dcd.parse_header(bytes(d[:60]))
    bad = bytearray(d)
    bad[4:8] = b"XXXX"
    with pytest.raises(dcd.DcdError):
        dcd.parse_header(bytes(bad))
    h = dcd.parse_header(bytes(d))
    with pytest.raises(dcd.DcdError):  # no room for the first frame
        dcd.parse_header(bytes(d[:h["first"] + h["first_bytes"] - 1]))


@pytest.mark.skipif(not os.path.exists(REF), reason="reference DCD library not built here")
@pytest.mark.parametrize("name", ["charmm_cell", "fixed_degrees", "fixed_bigendian", "fourdim"])
def test_live_positions_and_strides(name):
    lib = ctypes.CDLL(REF)
    P = ctypes.c_void_p
    path = os.path.join(GOLDEN, f"dcd_{name}.dcd").encode()
    d = data(name)
    h = dcd.parse_header(d)
    n = h["n_frames"]
    for start, stride in [(0, 2), (1, 1), (3, 4), (n - 1, 1), (5, 7)]:
        xyz = np.zeros((n, h["natoms"], 3), np.float32)
        cell = np.zeros((n, 6), np.float32)
        got = lib.refdcd_read(path, start, stride, n, xyz.ctypes.data_as(P), cell.ctypes.data_as(P))
        check(dcd.read(d, start=start, stride=stride), xyz[:got], cell[:got])


@pytest.mark.skipif(not os.path.exists(REF), reason="reference DCD library not built here")
def test_fourth_dimension_flag_in_opposite_byte_order(tmp_path):
    """The reference tests the flag word without byte swapping
    (dcdplugin.c:175-176), so an opposite-order file's fourth-dimension
    record is not expected and the read ends after frame 0; the oracle (and
    the GPU decoder, which shares its layout) return the same frames."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("mk", os.path.join(GOLDEN, "make_dcd_golden.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    rng = np.random.default_rng(1)
    x = mk.walk(rng, 6, 13)
    p = str(tmp_path / "be4d.dcd")
    mk.write_dcd(p, x, ">", cell=mk.cosine_cell(rng, 6), fourdim=True)
    lib = ctypes.CDLL(REF)
    P = ctypes.c_void_p
    xyz = np.zeros((64, 13, 3), np.float32)
    cell = np.zeros((64, 6), np.float32)
    got = lib.refdcd_read(p.encode(), 0, 1, 64, xyz.ctypes.data_as(P), cell.ctypes.data_as(P))
    d = open(p, "rb").read()
    assert not dcd.parse_header(d)["flags"] & dcd.FOUR_DIMS
    assert got == 1
    check(dcd.read(d), xyz[:got], cell[:got])
