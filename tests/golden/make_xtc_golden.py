"""Write the XTC golden fixtures (run in the build container only, where
oracle/_ref/libref_xtc.so was compiled from the reference's xdrfile.c and
xdrfile_xtc.c by oracle/Makefile).

Each case is a small trajectory written by the reference's encoder
(write_xtc) and read back by the reference's decoder (read_xtc); the .xtc
files and the decoded arrays (xtc_golden.npz) are committed.  The
reference's own examples/data/frame0.xtc is copied as a fixture with its
decoded arrays.  Cases cover every branch of the decoder: the compressed
path with water-like runs (the first two atoms of a run are swapped),
coordinates spanning more than 2^24 precision units (the three-field path),
frames of <= 9 atoms (stored uncompressed), coarse and fine precisions,
negative coordinates, and an all-zero box.
"""
import ctypes
import os
import shutil

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_FRAME0 = "/root/reference/examples/data/frame0.xtc"
P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
LIB = None  # oracle/_ref/libref_xtc.so, loaded by main()


def write(path, xyz, time, step, box, prec):
    xyz = np.ascontiguousarray(xyz, np.float32)
    rc = LIB.refxtc_write(path.encode(), xyz.shape[1], xyz.shape[0], P(xyz),
                          P(np.ascontiguousarray(time, np.float32)),
                          P(np.ascontiguousarray(step, np.int32)),
                          P(np.ascontiguousarray(box, np.float32).reshape(-1, 9)),
                          ctypes.c_float(prec))
    assert rc == 0, rc


def read(path, n_frames):
    n = LIB.refxtc_natoms(path.encode())
    xyz = np.zeros((n_frames + 1, n, 3), np.float32)
    time = np.zeros(n_frames + 1, np.float32)
    step = np.zeros(n_frames + 1, np.int32)
    box = np.zeros((n_frames + 1, 9), np.float32)
    prec = np.zeros(n_frames + 1, np.float32)
    got = LIB.refxtc_read(path.encode(), n, n_frames + 1, P(xyz), P(time), P(step), P(box), P(prec))
    assert got == n_frames, got
    return xyz[:got], time[:got], step[:got], box[:got].reshape(-1, 3, 3), prec[:got]


def cases():
    """The synthetic trajectories (coordinates, precision, box edge), seeded;
    importable without the reference library."""
    rng = np.random.default_rng(11)
    out = {}

    def walk(F, N, lo, hi, sigma):
        x0 = rng.uniform(lo, hi, (N, 3))
        steps = rng.normal(0, sigma, (F, N, 3))
        return (x0 + np.cumsum(steps, 0)).astype(np.float32)

    def frame_meta(F, edge):
        box = np.zeros((F, 3, 3), np.float32)
        if edge:
            box[:, 0, 0], box[:, 1, 1], box[:, 2, 2] = edge, edge * 1.1, edge * 0.9
        return np.arange(F, dtype=np.float32) * 2.5, np.arange(F, dtype=np.int32) * 500, box

    out["walk"] = (walk(40, 300, 0, 5, 0.01), 1000.0, 5.0)
    # water-like: O at random, two H within 0.1 nm (runs of small deltas)
    F, W = 30, 200
    o = rng.uniform(0, 4, (W, 3))
    mol = np.stack([o, o + rng.normal(0, 0.06, (W, 3)), o + rng.normal(0, 0.06, (W, 3))], 1)
    water = mol.reshape(1, 3 * W, 3) + rng.normal(0, 0.002, (F, 3 * W, 3)) + \
        np.cumsum(rng.normal(0, 0.01, (F, 1, 3)), 0)
    out["water"] = (water.astype(np.float32), 1000.0, 4.0)
    out["wide"] = (walk(8, 200, 0, 30000, 1.0), 1000.0, 0.0)
    out["coarse"] = (walk(20, 500, -3, 3, 0.05), 10.0, 6.0)
    out["fine"] = (walk(10, 300, 0, 2, 0.001), 100000.0, 2.0)
    out["centered"] = (walk(12, 150, -2.5, 2.5, 0.02), 1000.0, 0.0)
    out["tiny5"] = (walk(12, 5, 0, 3, 0.05), 1000.0, 3.0)
    out["tiny9"] = (walk(10, 9, 0, 3, 0.05), 1000.0, 3.0)
    out["small10"] = (walk(10, 10, 0, 3, 0.05), 1000.0, 3.0)
    meta = {k: frame_meta(v[0].shape[0], v[2]) for k, v in out.items()}
    return out, meta


def main():
    global LIB
    LIB = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_xtc.so"))
    arrays = {}
    data, meta = cases()
    for name, (xyz, prec, _) in data.items():
        path = os.path.join(HERE, f"xtc_{name}.xtc")
        t, s, b = meta[name]
        write(path, xyz, t, s, b, prec)
        got = read(path, xyz.shape[0])
        for key, a in zip(("xyz", "time", "step", "box", "prec"), got):
            arrays[f"{name}_{key}"] = a
    dst = os.path.join(HERE, "xtc_frame0.xtc")
    shutil.copyfile(REF_FRAME0, dst)
    got = read(dst, 501)
    for key, a in zip(("xyz", "time", "step", "box", "prec"), got):
        arrays[f"frame0_{key}"] = a
    np.savez_compressed(os.path.join(HERE, "xtc_golden.npz"), **arrays)
    print("wrote", sorted({k.split("_")[0] for k in arrays}))


if __name__ == "__main__":
    main()
