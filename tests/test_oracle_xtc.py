"""The XTC oracle (oracle/xtc.py) against the reference's decoder.

The golden arrays were decoded by the reference's own xdrfile.c
(tests/golden/make_xtc_golden.py); the files are the reference's
examples/data/frame0.xtc and trajectories its encoder wrote.  Beyond
equality with the reference: the decoded coordinates of the synthetic
cases lie within half a quantum (1 / precision) of what was written, a
corrupted stream is never silently accepted, and every decoder branch is
exercised (runs, the three-field path, uncompressed small frames)."""
import os

import numpy as np
import pytest

from oracle import xtc

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
G = np.load(os.path.join(HERE, "xtc_golden.npz"))
CASES = sorted({k.split("_")[0] for k in G.files})


def data(name):
    with open(os.path.join(HERE, f"xtc_{name}.xtc"), "rb") as f:
        return f.read()


@pytest.mark.parametrize("name", CASES)
def test_oracle_equals_reference_decoder(name):
    xyz, time, step, box, prec = xtc.read(data(name))
    assert np.array_equal(xyz, G[f"{name}_xyz"])
    assert np.array_equal(time, G[f"{name}_time"])
    assert np.array_equal(step, G[f"{name}_step"])
    assert np.array_equal(box, G[f"{name}_box"])
    if xyz.shape[1] > 9:  # the reference leaves the precision unset for raw frames
        assert np.array_equal(prec, G[f"{name}_prec"])


def test_cases_cover_every_branch():
    """Every decoder branch runs on some fixture: the three-field path
    (a coordinate range above 2^24 quanta), uncompressed frames of <= 9
    atoms, and runs of small deltas (whose first two atoms are swapped)."""
    st = {}
    xtc.read(data("wide"), st)
    assert st["large"] == G["wide_xyz"].shape[0]
    st = {}
    xtc.read(data("tiny9"), st)
    assert st["raw"] == G["tiny9_xyz"].shape[0]
    for name in ("water", "frame0", "walk"):
        st = {}
        xtc.read(data(name), st)
        assert st["runs"] > 0 and st["small_atoms"] >= st["runs"]
    st = {}
    xtc.read(data("small10"), st)
    assert st["raw"] == 0 and st["large"] == 0


def test_round_trip_within_half_a_quantum():
    """What the reference's encoder wrote decodes to within half a quantum
    (0.5 / precision) of the written coordinates, plus the float32 rounding
    of integer x (1 / precision); uncompressed frames decode exactly."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("make_xtc_golden", os.path.join(HERE, "make_xtc_golden.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    written, _ = mk.cases()
    for name, (xyz, prec, _) in written.items():
        dec = xtc.read(data(name))[0]
        if xyz.shape[1] <= 9:
            assert np.array_equal(dec, xyz)
            continue
        err = np.abs(dec.astype(np.float64) - xyz.astype(np.float64))
        bound = 0.5 / prec + 2 * np.finfo(np.float32).eps * np.abs(xyz.astype(np.float64)) + 1e-12
        assert np.all(err <= bound), (name, float((err - bound).max()))


def test_frame_offsets_walk_the_file():
    for name in CASES:
        d = data(name)
        offs = xtc.frame_offsets(d)
        assert len(offs) == G[f"{name}_xyz"].shape[0]
        assert offs[0] == 0 and all(b > a for a, b in zip(offs, offs[1:]))


def test_truncated_and_corrupt_streams_are_refused():
    d = data("walk")
    with pytest.raises(xtc.XtcError):
        xtc.read(d[:-7])
    bad = bytearray(d)
    bad[0:4] = (1996).to_bytes(4, "big")
    with pytest.raises(xtc.XtcError):
        xtc.read(bytes(bad))
    # a flipped bit inside the compressed bytes changes the frame or is refused
    bad = bytearray(d)
    bad[52 + 4 + 36 + 40] ^= 0x10
    try:
        xyz = xtc.read(bytes(bad))[0]
    except xtc.XtcError:
        return
    assert not np.array_equal(xyz, G["walk_xyz"])
