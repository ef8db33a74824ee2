"""XTC reader, host side: the frame index (hr_xtc_scan) against the
oracle's header walk, host validation of non-XTC and truncated bytes, the
reader's position API, and a loud failure (no CPU decode) without a GPU."""
import ctypes
import os
import struct

import numpy as np
import pytest

from mdtraj_amd import _lib
from mdtraj_amd.xtc import XTCTrajectoryFile
from oracle import xtc as oxtc

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
G = np.load(os.path.join(HERE, "xtc_golden.npz"))
CASES = sorted({k.split("_")[0] for k in G.files})


def path(name):
    return os.path.join(HERE, f"xtc_{name}.xtc")


def scan(buf):
    lib = _lib.load()
    a = np.frombuffer(buf, np.uint8)
    n, na = ctypes.c_int64(), ctypes.c_int32()
    offs = np.zeros(4096, np.int64)
    rc = lib.hr_xtc_scan(a.ctypes.data_as(ctypes.c_void_p), a.size,
                         offs.ctypes.data_as(ctypes.c_void_p), offs.size,
                         ctypes.byref(n), ctypes.byref(na))
    msg = lib.hr_last_error().decode()
    return rc, offs[: min(n.value, offs.size)], n.value, na.value, msg


@pytest.mark.parametrize("name", CASES)
def test_scan_matches_oracle_offsets(name):
    buf = open(path(name), "rb").read()
    rc, offs, n, na, _ = scan(buf)
    assert rc == 0
    assert list(offs) == oxtc.frame_offsets(buf)
    assert n == G[f"{name}_xyz"].shape[0] and na == G[f"{name}_xyz"].shape[1]


def test_scan_rejects_bad_bytes():
    buf = open(path("walk"), "rb").read()
    rc, _, _, _, msg = scan(buf[:-10])
    assert rc == _lib.HR_ERR_INVALID and "truncated" in msg
    rc, _, _, _, msg = scan(b"\x00\x00\x07\xcc" + buf[4:])  # magic 1996
    assert rc == _lib.HR_ERR_INVALID and "not an XTC frame" in msg
    rc, _, n, na, _ = scan(b"")
    assert rc == 0 and n == 0 and na == 0


def test_reader_position_api():
    with XTCTrajectoryFile(path("frame0")) as f:
        assert len(f) == 501 and f.n_atoms == 22 and f.tell() == 0
        assert list(f.offsets) == oxtc.frame_offsets(open(path("frame0"), "rb").read())
        f.seek(10)
        f.seek(5, whence=1)
        assert f.tell() == 15
        f.seek(-1, whence=2)
        assert f.tell() == 500
        with pytest.raises(IOError):
            f.seek(502)
    with pytest.raises(ValueError):
        XTCTrajectoryFile(path("walk"), mode="w")
    with pytest.raises(IOError):
        XTCTrajectoryFile(path("missing"))


def test_no_cpu_decode():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with XTCTrajectoryFile(path("walk")) as f:
        with pytest.raises(RuntimeError):
            f.read()


def test_scan_and_oracle_agree_on_mutated_headers():
    """Random byte changes and truncations near frame headers: the C frame
    index and the oracle's header walk accept and refuse the same inputs
    (the C index also requires one atom count throughout), with the same
    offsets when both accept."""
    rng = np.random.default_rng(13)
    for name in CASES:
        base = open(path(name), "rb").read()
        offs = oxtc.frame_offsets(base)
        for trial in range(40):
            buf = bytearray(base)
            if trial % 4 == 0:
                buf = buf[: int(rng.integers(0, len(buf)))]
            else:
                o = offs[int(rng.integers(0, len(offs)))]
                for _ in range(int(rng.integers(1, 3))):
                    pos = o + int(rng.integers(0, 96))
                    if pos < len(buf):
                        buf[pos] = int(rng.integers(0, 256))
            data = bytes(buf)
            rc, got, n, na, _ = scan(data)
            try:
                want = oxtc.frame_offsets(data)
                natoms = {struct.unpack_from(">i", data, p + 4)[0] for p in want}
                ok = len(natoms) <= 1
            except oxtc.XtcError:
                want, ok = None, False
            assert (rc == 0) == ok, (name, trial, rc)
            if ok:
                assert list(got) == want[:4096] and n == len(want), (name, trial)
