"""XTC decoding restated in plain Python (test infrastructure only: the
checker for the GPU decoder, never called by the product).

Follows the reference's reader: a frame is the header of
formats/xtc/src/xdrfile_xtc.c:41-62 (magic 1995, natoms, step, time), the
box (:64-90, nine floats), then xdrfile_decompress_coord_float
(formats/xtc/src/xdrfile.c:778-1000).  All words are big-endian XDR; the
compressed bytes are padded to a multiple of four (xdr_opaque).  Bit fields
are read most significant bit first (decodebits, xdrfile.c:677-723); a
group of three "small" integers is one mixed-radix number (decodeints,
xdrfile.c:725-758), evaluated here with Python's exact integers.

Pinned by tests/test_oracle_xtc.py against the reference's own decoder
(oracle/_ref/libref_xtc.so, built from xdrfile.c) on the reference's
examples/data/frame0.xtc and on files its encoder wrote (tests/golden).
"""
from __future__ import annotations

import struct

import numpy as np

MAGIC = 1995

# xdrfile.c:760-776
MAGICINTS = (
    0, 0, 0, 0, 0, 0, 0, 0, 0, 8, 10, 12, 16, 20, 25, 32, 40, 50, 64,
    80, 101, 128, 161, 203, 256, 322, 406, 512, 645, 812, 1024, 1290,
    1625, 2048, 2580, 3250, 4096, 5060, 6501, 8192, 10321, 13003,
    16384, 20642, 26007, 32768, 41285, 52015, 65536, 82570, 104031,
    131072, 165140, 208063, 262144, 330280, 416127, 524287, 660561,
    832255, 1048576, 1321122, 1664510, 2097152, 2642245, 3329021,
    4194304, 5284491, 6658042, 8388607, 10568983, 13316085, 16777216)
FIRSTIDX = 9
LASTIDX = len(MAGICINTS)


class XtcError(ValueError):
    pass


def _bits_for(n):
    """Bits of the smallest field holding 0 .. n (sizeofint, xdrfile.c:490-501)."""
    b = 0
    while n >= (1 << b) and b < 32:
        b += 1
    return b


def _bits_for_product(sizes):
    """Bits of the mixed-radix number with these digit sizes (sizeofints,
    xdrfile.c:515-548): the bit length of prod(sizes)."""
    p = 1
    for s in sizes:
        p *= s
    return p.bit_length()


class _Bits:
    """MSB-first bit reader over one frame's compressed bytes."""

    def __init__(self, data):
        self.v = int.from_bytes(data, "big")
        self.left = 8 * len(data)

    def take(self, n):
        if n == 0:
            return 0
        if n > self.left:
            raise XtcError("compressed stream ends inside a field")
        self.left -= n
        return (self.v >> self.left) & ((1 << n) - 1)

    def ints3(self, nbits, sizes):
        """decodeints: byte chunks least significant first, then digits."""
        v, shift = 0, 0
        while nbits > 8:
            v |= self.take(8) << shift
            shift += 8
            nbits -= 8
        if nbits > 0:
            v |= self.take(nbits) << shift
        z = v % sizes[2]
        v //= sizes[2]
        y = v % sizes[1]
        x = v // sizes[1]
        return [x & 0xFFFFFFFF, y, z]


def _i32(u):
    u &= 0xFFFFFFFF
    return u - (1 << 32) if u >= 1 << 31 else u


def decode_coords(data, pos, natoms, stats=None):
    """Coordinates of one frame starting at the coordinate block (the
    repeated atom count); returns (float32 (natoms, 3), precision, next pos).
    `stats` (a dict) counts the branches taken: raw frames, three-field
    frames, runs and small atoms."""
    (lsize,) = struct.unpack_from(">i", data, pos)
    pos += 4
    if lsize != natoms:
        raise XtcError(f"frame holds {lsize} atoms, header says {natoms}")
    if stats is not None:
        for key in ("raw", "large", "runs", "small_atoms"):
            stats.setdefault(key, 0)
    if natoms <= 9:  # stored uncompressed (xdrfile.c:836-841)
        if stats is not None:
            stats["raw"] += 1
        xyz = np.frombuffer(data, ">f4", 3 * natoms, pos).astype(np.float32).reshape(natoms, 3)
        return xyz, np.float32(0.0), pos + 12 * natoms
    (prec,) = struct.unpack_from(">f", data, pos)
    minint = struct.unpack_from(">3i", data, pos + 4)
    maxint = struct.unpack_from(">3i", data, pos + 16)
    (smallidx,) = struct.unpack_from(">i", data, pos + 28)
    (nbytes,) = struct.unpack_from(">i", data, pos + 32)
    pos += 36
    blob = data[pos:pos + nbytes]
    if len(blob) != nbytes:
        raise XtcError("truncated frame")
    pos += (nbytes + 3) // 4 * 4
    sizeint = [(maxint[k] - minint[k] + 1) & 0xFFFFFFFF for k in range(3)]
    large = (sizeint[0] | sizeint[1] | sizeint[2]) > 0xFFFFFF
    bitsizeint = [_bits_for(s) for s in sizeint] if large else None
    bitsize = 0 if large else _bits_for_product(sizeint)
    if stats is not None and large:
        stats["large"] += 1
    if not FIRSTIDX <= smallidx < LASTIDX:
        raise XtcError(f"smallidx {smallidx} out of range")
    smaller = MAGICINTS[max(FIRSTIDX, smallidx - 1)] // 2
    smallnum = MAGICINTS[smallidx] // 2
    sizesmall = MAGICINTS[smallidx]
    inv = np.float32(np.float64(1.0) / np.float64(np.float32(prec)))
    bits = _Bits(blob)
    out = np.empty((natoms, 3), np.int64)
    n_out = 0  # atoms emitted
    run = 0  # a clear flag bit keeps the previous run length (xdrfile.c:924-931)
    while n_out < natoms:
        if large:
            big = [bits.take(bitsizeint[k]) for k in range(3)]
        else:
            big = bits.ints3(bitsize, sizeint)
        big = [_i32(big[k] + minint[k]) for k in range(3)]
        is_smaller = 0
        if bits.take(1):
            run = bits.take(5)
            is_smaller = run % 3
            run -= is_smaller
            is_smaller -= 1
        if n_out + 1 + run // 3 > natoms:
            raise XtcError("run past the frame's atom count")
        if run == 0:
            out[n_out] = big
            n_out += 1
        else:
            # each small atom is a delta from the one before it; the first
            # is stored before the large atom (swapped, for water), so the
            # order out is small_1, large, small_2, small_3, ...
            if stats is not None:
                stats["runs"] += 1
                stats["small_atoms"] += run // 3
            prev = big
            for k in range(0, run, 3):
                d = bits.ints3(smallidx, [sizesmall] * 3)
                t = [_i32(d[c] + prev[c] - smallnum) for c in range(3)]
                out[n_out] = t
                n_out += 1
                if k == 0:
                    out[n_out] = big
                    n_out += 1
                prev = t
        smallidx += is_smaller
        if is_smaller < 0:
            smallnum = smaller
            smaller = MAGICINTS[smallidx - 1] // 2 if smallidx > FIRSTIDX else 0
        elif is_smaller > 0:
            smaller = smallnum
            smallnum = MAGICINTS[smallidx] // 2
        if not FIRSTIDX <= smallidx < LASTIDX:
            raise XtcError("smallidx left the table")
        sizesmall = MAGICINTS[smallidx]
    xyz = out.astype(np.float32) * inv
    return xyz.astype(np.float32), np.float32(prec), pos


def frame_offsets(data):
    """Byte offset of every frame (header walk; xdrfile_xtc.c:41-90)."""
    offs = []
    pos = 0
    n = len(data)
    while pos + 4 <= n:
        if pos + 56 > n:
            raise XtcError(f"truncated frame header at byte {pos}")
        (magic, natoms) = struct.unpack_from(">ii", data, pos)
        if magic != MAGIC:
            raise XtcError(f"bad magic {magic} at byte {pos}")
        if natoms <= 0:
            raise XtcError(f"frame at byte {pos} holds {natoms} atoms")
        offs.append(pos)
        c = pos + 16 + 36 + 4  # header, box, atom count
        if natoms <= 9:
            pos = c + 12 * natoms
        else:
            if c + 36 > n:
                raise XtcError(f"truncated frame at byte {pos}")
            (nbytes,) = struct.unpack_from(">i", data, c + 32)
            if nbytes < 0:
                raise XtcError(f"negative byte count in frame at byte {pos}")
            pos = c + 36 + (nbytes + 3) // 4 * 4
        if pos > n:
            raise XtcError("truncated last frame")
    return offs


def read(data, stats=None):
    """Every frame of an XTC byte string: xyz (F, N, 3) float32, time (F,)
    float32, step (F,) int32, box (F, 3, 3) float32, precision (F,) float32
    (the reference's read_xtc output per frame, xtc.pyx:476-480)."""
    offs = frame_offsets(data)
    if not offs:
        return (np.zeros((0, 0, 3), np.float32), np.zeros(0, np.float32),
                np.zeros(0, np.int32), np.zeros((0, 3, 3), np.float32), np.zeros(0, np.float32))
    (natoms,) = struct.unpack_from(">i", data, 4)
    F = len(offs)
    xyz = np.empty((F, natoms, 3), np.float32)
    time = np.empty(F, np.float32)
    step = np.empty(F, np.int32)
    box = np.empty((F, 3, 3), np.float32)
    prec = np.zeros(F, np.float32)
    for f, pos in enumerate(offs):
        _, n, st, t = struct.unpack_from(">iiif", data, pos)
        if n != natoms:
            raise XtcError("atom count changes between frames")
        step[f], time[f] = st, t
        box[f] = np.frombuffer(data, ">f4", 9, pos + 16).reshape(3, 3)
        xyz[f], prec[f], _ = decode_coords(data, pos + 52, natoms, stats)
    return xyz, time, step, box, prec
