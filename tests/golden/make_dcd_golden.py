"""Writes the DCD fixtures and what the reference's reader returns for them.

Two files come from the reference's own writer (dcdplugin.c write_dcdheader /
write_dcdstep, through oracle/_ref/libref_dcd.so): CHARMM format with and
without a unit cell.  The others are written here, byte by byte, to reach the
reader's other branches (dcdplugin.c:98-623): opposite byte order, an X-PLOR
header (double DELTA, no CHARMM version), fixed atoms (later frames hold only
the free atoms), a fourth-dimension record, a cell stored in degrees, more
title lines, trailing bytes shorter than a frame, and a bad record marker.

Expected outputs are the reference reader's (libref_dcd.so, read_next_timestep
as DCDTrajectoryFile.read calls it): coordinates, the (A, B, C, alpha, beta,
gamma) cell and the number of frames read, for a full read and for a read
from frame 2 with stride 3.  Run from the repository root after
`make -C oracle ref`:  python tests/golden/make_dcd_golden.py
"""
from __future__ import annotations

import ctypes
import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
P = ctypes.c_void_p


def write_dcd(path, xyz, endian="<", charmm=True, cell=None, fourdim=False, free=None,
              titles=(b"written for the dcd fixtures", b"second title line"), delta=0.002,
              marker_override=None):
    """A DCD file of float32 coordinates (F, N, 3).  cell: (F, 6) raw doubles
    (A, cos gamma, B, cos beta, cos alpha, C) or None; free: 1-based free-atom
    indices (the others are fixed at frame 0's coordinates);
    marker_override: (frame, value) replaces that frame's first X marker."""
    xyz = np.asarray(xyz, np.float32)
    F, N = xyz.shape[:2]
    e = endian
    out = bytearray()
    i32 = lambda v: out.extend(struct.pack(e + "i", v))
    hdr = [0] * 20
    hdr[0], hdr[1], hdr[2] = F, 100, 10
    namnf = 0 if free is None else N - len(free)
    hdr[8] = namnf
    body = bytearray()
    for k, v in enumerate(hdr):
        body += struct.pack(e + "i", v)
    if charmm:
        body[36:40] = struct.pack(e + "f", delta)
        body[40:44] = struct.pack(e + "i", 1 if cell is not None else 0)
        body[44:48] = struct.pack(e + "i", 1 if fourdim else 0)
        body[76:80] = struct.pack(e + "i", 24)
    else:
        body[36:44] = struct.pack(e + "d", delta)
    i32(84)
    out += b"CORD"
    out += body
    i32(84)
    i32(4 + 80 * len(titles))
    i32(len(titles))
    for t in titles:
        out += t.ljust(80)[:80]
    i32(4 + 80 * len(titles))
    i32(4)
    i32(N)
    i32(4)
    if free is not None:
        i32(4 * len(free))
        for v in free:
            i32(int(v))
        i32(4 * len(free))
    for f in range(F):
        if cell is not None:
            i32(48)
            out += struct.pack(e + "6d", *[float(v) for v in cell[f]])
            i32(48)
        sel = np.arange(N) if (free is None or f == 0) else np.asarray(free) - 1
        recs = [xyz[f, sel, c] for c in range(3)]
        if fourdim:
            recs.append(np.full(len(sel), 7.5, np.float32))
        for r, vals in enumerate(recs):
            m = 4 * len(sel)
            lead = marker_override[1] if (marker_override and marker_override[0] == f and r == 0) else m
            i32(lead)
            out += vals.astype(e + "f4").tobytes()
            i32(m)
    with open(path, "wb") as fh:
        fh.write(bytes(out))


def walk(rng, F, N, scale=20.0, step=0.3):
    x0 = rng.uniform(-scale, scale, (N, 3))
    return (x0 + np.cumsum(rng.normal(0, step, (F, N, 3)), axis=0)).astype(np.float32)


def cosine_cell(rng, F):
    """Raw cells with angle cosines, as CHARMM writes them."""
    c = np.zeros((F, 6))
    c[:, [0, 2, 5]] = rng.uniform(25, 40, (F, 3))
    ang = rng.uniform(60, 120, (F, 3))  # alpha, beta, gamma in degrees
    c[:, 4] = np.cos(np.radians(ang[:, 0]))
    c[:, 3] = np.cos(np.radians(ang[:, 1]))
    c[:, 1] = np.cos(np.radians(ang[:, 2]))
    return c


def cases():
    """(name, writer) pairs; writer(path) writes the file."""
    rng = np.random.default_rng(2024)
    out = []
    x = walk(rng, 12, 23)
    L = np.tile(np.float32([31.5, 32.25, 40.0]), (12, 1)) + rng.uniform(0, 1, (12, 3)).astype(np.float32)
    A = np.tile(np.float32([90.0, 80.0, 70.0]), (12, 1))
    out.append(("charmm_cell", lambda p, x=x, L=L, A=A: ref_write(p, x, L, A)))
    x = walk(rng, 9, 40)
    out.append(("charmm_nocell", lambda p, x=x: ref_write(p, x, None, None)))
    x, c = walk(rng, 10, 17), cosine_cell(rng, 10)
    out.append(("bigendian_cell", lambda p, x=x, c=c: write_dcd(p, x, ">", cell=c)))
    x = walk(rng, 8, 12)
    out.append(("xplor", lambda p, x=x: write_dcd(p, x, charmm=False)))
    x = walk(rng, 11, 30)
    free = np.sort(rng.choice(30, 20, replace=False)) + 1
    deg = np.zeros((11, 6))
    deg[:, [0, 2, 5]] = 50.0
    deg[:, [1, 3, 4]] = [100.0, 95.0, 85.0]  # degrees, not cosines (NAMD 2.5)
    out.append(("fixed_degrees", lambda p, x=x, f=free, c=deg: write_dcd(p, x, cell=c, free=f)))
    x = walk(rng, 9, 26)
    free = np.sort(rng.choice(26, 9, replace=False)) + 1
    c = cosine_cell(rng, 9)
    out.append(("fixed_bigendian", lambda p, x=x, f=free, c=c: write_dcd(p, x, ">", cell=c, free=f)))
    x, c = walk(rng, 7, 19), cosine_cell(rng, 7)
    out.append(("fourdim", lambda p, x=x, c=c: write_dcd(p, x, cell=c, fourdim=True)))
    x = walk(rng, 6, 15)
    out.append(("titles", lambda p, x=x: write_dcd(p, x, titles=(b"one", b"two", b"three"))))
    x = walk(rng, 8, 21)

    def trailing(p, x=x):
        write_dcd(p, x)
        with open(p, "ab") as fh:
            fh.write(b"\0" * 37)

    out.append(("trailing", trailing))
    x = walk(rng, 9, 14)
    out.append(("badrecord", lambda p, x=x: write_dcd(p, x, marker_override=(3, 4 * 14 + 4))))
    return out


_LIB = None


def ref_lib():
    global _LIB
    if _LIB is None:
        _LIB = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_dcd.so"))
    return _LIB


def ref_write(path, xyz, lengths, angles):
    xyz = np.ascontiguousarray(xyz, np.float32)
    F, N = xyz.shape[:2]
    Lp = None if lengths is None else np.ascontiguousarray(lengths, np.float32).ctypes.data_as(P)
    Ap = None if angles is None else np.ascontiguousarray(angles, np.float32).ctypes.data_as(P)
    if ref_lib().refdcd_write(path.encode(), N, F, xyz.ctypes.data_as(P), Lp, Ap) != 0:
        raise RuntimeError("reference writer failed")


def ref_read(path, start=0, stride=1):
    lib = ref_lib()
    na, ns = ctypes.c_int(), ctypes.c_int()
    if lib.refdcd_info(path.encode(), ctypes.byref(na), ctypes.byref(ns)) != 0:
        raise RuntimeError("reference reader refused " + path)
    n = max(ns.value, 1)
    xyz = np.zeros((n, na.value, 3), np.float32)
    cell = np.zeros((n, 6), np.float32)
    got = lib.refdcd_read(path.encode(), start, stride, n, xyz.ctypes.data_as(P), cell.ctypes.data_as(P))
    return xyz[:got], cell[:got], ns.value


def main():
    store = {}
    for name, writer in cases():
        path = os.path.join(HERE, f"dcd_{name}.dcd")
        writer(path)
        xyz, cell, nsets = ref_read(path)
        store[f"{name}/xyz"], store[f"{name}/cell"] = xyz, cell
        store[f"{name}/nsets"] = np.int64(nsets)
        xs, cs, _ = ref_read(path, 2, 3)
        store[f"{name}/xyz_2_3"], store[f"{name}/cell_2_3"] = xs, cs
        print(f"{name}: {nsets} frames in the file, {len(xyz)} read, {xyz.shape[1]} atoms",
              file=sys.stderr)
    np.savez_compressed(os.path.join(HERE, "dcd_golden.npz"), **store)


if __name__ == "__main__":
    main()
