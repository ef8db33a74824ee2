"""DCD reading restated in numpy (test infrastructure only: the checker for
the GPU decoder, never called by the product).

Follows the reference's reader, formats/dcd/src/dcdplugin.c:
  header           read_dcdheader (:98-374): the 84-byte control record
                   (NSET, ISTART, NSAVC, NAMNF, DELTA, the CHARMM version and
                   its extra-block / 4th-dimension flags), the title record,
                   the atom count and, with fixed atoms, the free-atom list;
                   byte order from the first record marker;
  frame count      open_dcd_read (:870-905): from the file size, with every
                   frame after the first holding only the free atoms;
  frame            read_dcdstep (:481-623) with read_charmm_extrablock
                   (:376-406, six doubles cast to float), read_fixed_atoms
                   (:409-443, free atoms over the first frame's coordinates)
                   and read_charmm_4dim (:445-466, skipped);
  unit cell        read_next_timestep (:975-1013): lengths A, B, C; angles
                   from cosines when all three lie in [-1, 1], else as stored;
and the loop of formats/dcd/dcd.pyx:463-551 (n_frames, stride, atom_indices,
a failed frame ends the read like end of file, no cell when every length is
below 1e-10).

Pinned by tests/test_oracle_dcd.py against the reference's own reader
(oracle/_ref/libref_dcd.so, compiled from dcdplugin.c) on files its writer
wrote and on files that exercise the other branches (opposite byte order,
X-PLOR header, fixed atoms, a fourth dimension, cells in degrees).
"""
from __future__ import annotations

import struct

import numpy as np

CHARMM, FOUR_DIMS, EXTRA_BLOCK = 0x01, 0x02, 0x04
M_PI_2 = 1.57079632679489661922


class DcdError(ValueError):
    pass


def parse_header(data):
    """Header fields and the frame layout of a DCD byte string."""
    data = bytes(data)
    if len(data) < 8:
        raise DcdError("file too short for a DCD header")
    w0, w1 = struct.unpack_from("<II", data, 0)
    cord = struct.unpack_from("<I", b"CORD")[0]
    if (w0 + w1) & 0xFFFFFFFF == 84:
        raise DcdError("CHARMM -i8 files (64-bit record markers) are not supported")
    if w0 == 84 and w1 == cord:
        e = "<"
    else:
        s0, s1 = struct.unpack_from(">II", data, 0)
        if (s0 + s1) & 0xFFFFFFFF == 84:
            raise DcdError("CHARMM -i8 files (64-bit record markers) are not supported")
        if s0 == 84 and w1 == cord:
            e = ">"
        else:
            raise DcdError("unrecognized DCD header")
    if len(data) < 92:
        raise DcdError("truncated header")
    hdr = data[8:88]
    raw = lambda off: struct.unpack_from("<i", hdr, off)[0]  # flags read unswapped, as the reference
    flags = 0
    if raw(76) != 0:
        flags = CHARMM
        if raw(40) != 0:
            flags |= EXTRA_BLOCK
        if raw(44) == 1:
            flags |= FOUR_DIMS
    i32 = lambda off: struct.unpack_from(e + "i", hdr, off)[0]
    nset, istart, nsavc, namnf = i32(0), i32(4), i32(8), i32(32)
    if flags & CHARMM:
        delta = float(struct.unpack_from(e + "f", hdr, 36)[0])
    else:
        delta = struct.unpack_from(e + "d", hdr, 36)[0]
    u32 = lambda off: struct.unpack_from(e + "I", data, off)[0]

    def need(n):
        if len(data) < n:
            raise DcdError("truncated header")

    if u32(88) != 84:
        raise DcdError("bad end of the control record")
    need(100)
    if ((u32(92) - 4) & 0xFFFFFFFF) % 80 != 0:
        raise DcdError("bad title record")
    ntitle = struct.unpack_from(e + "i", data, 96)[0]
    if ntitle < 0:
        raise DcdError("negative title count")
    if ntitle > 1000:
        ntitle = 2 if ntitle == 1095062083 else 0
    pos = 100 + 80 * ntitle + 4
    need(pos + 12)
    if u32(pos) != 4:
        raise DcdError("bad atom-count record")
    natoms = struct.unpack_from(e + "i", data, pos + 4)[0]
    if u32(pos + 8) != 4:
        raise DcdError("bad atom-count record")
    pos += 12
    if natoms <= 0 or natoms > (1 << 30):
        raise DcdError(f"unsupported atom count {natoms}")
    free = None
    if namnf != 0:
        nfree = natoms - namnf
        if nfree <= 0:
            raise DcdError("bad fixed-atom count")
        need(pos + 8 + 4 * nfree)
        if u32(pos) != 4 * nfree:
            raise DcdError("bad free-atom record")
        free = np.frombuffer(data, e + "i4", nfree, pos + 4).astype(np.int64)
        if u32(pos + 4 + 4 * nfree) != 4 * nfree:
            raise DcdError("bad free-atom record")
        if free.min() < 1 or free.max() > natoms:
            raise DcdError("free-atom index out of range")
        pos += 8 + 4 * nfree
    ndims = 4 if flags & FOUR_DIMS else 3
    extra = 56 if flags & EXTRA_BLOCK else 0
    first_bytes = (natoms + 2) * ndims * 4 + extra
    frame_bytes = (natoms - namnf + 2) * ndims * 4 + extra
    trj = len(data) - pos - first_bytes
    if trj < 0:
        raise DcdError("file contains no frames")
    return dict(endian=e, flags=flags, nset=nset, istart=istart, nsavc=nsavc, delta=delta,
                namnf=namnf, free=free, natoms=natoms, first=pos, first_bytes=first_bytes,
                frame_bytes=frame_bytes, n_frames=trj // frame_bytes + 1)


def frame_offset(h, k):
    return h["first"] + (0 if k == 0 else h["first_bytes"] + (k - 1) * h["frame_bytes"])


def read_frame(data, h, k, frame0=None):
    """Coordinates (N, 3) float32 and the raw unit cell (6 float32) of frame
    k; None when a record marker is wrong.  frame0 is frame 0's (N, 3)
    coordinates, needed for frames after the first when atoms are fixed."""
    e, N = h["endian"], h["natoms"]
    pos = frame_offset(h, k)
    u32 = lambda off: struct.unpack_from(e + "I", data, off)[0]
    cell = np.array([0, 90, 0, 90, 90, 0], np.float32)  # the reference's defaults
    if h["flags"] & EXTRA_BLOCK:
        if u32(pos) != 48:
            return None
        cell = np.frombuffer(data, e + "f8", 6, pos + 4).astype(np.float32)
        pos += 56
    full = h["namnf"] == 0 or k == 0
    n = N if full else N - h["namnf"]
    xyz = np.empty((N, 3), np.float32)
    for c in range(3):
        if u32(pos) != 4 * n or u32(pos + 4 + 4 * n) != 4 * n:
            return None
        v = np.frombuffer(data, e + "f4", n, pos + 4).astype(np.float32)
        if full:
            xyz[:, c] = v
        else:
            xyz[:, c] = frame0[:, c]
            xyz[h["free"] - 1, c] = v  # later entries win, as the reference's loop
        pos += 8 + 4 * n
    if h["flags"] & FOUR_DIMS and u32(pos) != 4 * n:
        return None
    return xyz, cell


def cell_lengths_angles(cell):
    """(lengths, angles) float32 from the raw cell (A, cos gamma, B, cos beta,
    cos alpha, C) as read_next_timestep converts it."""
    cell = np.asarray(cell, np.float32).reshape(-1, 6)
    lengths = cell[:, [0, 2, 5]].copy()
    c = cell.astype(np.float64)
    cosines = np.all((c[:, [1, 3, 4]] >= -1.0) & (c[:, [1, 3, 4]] <= 1.0), axis=1)
    ang = np.empty((len(cell), 3), np.float64)
    ang[:, 0] = np.where(cosines, 90.0 - np.arcsin(np.clip(c[:, 4], -1, 1)) * 90.0 / M_PI_2, c[:, 4])
    ang[:, 1] = np.where(cosines, 90.0 - np.arcsin(np.clip(c[:, 3], -1, 1)) * 90.0 / M_PI_2, c[:, 3])
    ang[:, 2] = np.where(cosines, 90.0 - np.arcsin(np.clip(c[:, 1], -1, 1)) * 90.0 / M_PI_2, c[:, 1])
    return lengths, ang.astype(np.float32)


def read(data, n_frames=None, stride=None, atom_indices=None, start=0):
    """DCDTrajectoryFile.read from frame `start`: (xyz, cell_lengths,
    cell_angles, position after the read), lengths and angles None when every
    length is below 1e-10."""
    data = bytes(data)
    h = parse_header(data)
    stride = 1 if stride is None else int(stride)
    count = h["n_frames"] - start if n_frames is None else int(n_frames)
    frame0 = None
    if h["namnf"]:
        r0 = read_frame(data, h, 0)
        frame0 = None if r0 is None else r0[0]
    xs, cells = [], []
    pos = start
    for _ in range(count):
        if pos >= h["n_frames"]:
            break
        r = None if (h["namnf"] and frame0 is None) else read_frame(data, h, pos, frame0)
        pos += 1
        if r is None:
            break
        xyz, cell = r
        xs.append(xyz if atom_indices is None else xyz[atom_indices])
        cells.append(cell)
        pos = min(h["n_frames"], pos + stride - 1)
    n_sel = h["natoms"] if atom_indices is None else len(np.arange(h["natoms"])[atom_indices])
    xyz = np.array(xs, np.float32).reshape(len(xs), n_sel, 3)
    lengths, angles = cell_lengths_angles(np.array(cells, np.float32).reshape(-1, 6))
    if np.all(lengths < 1e-10):
        lengths = angles = None
    return xyz, lengths, angles, pos
