"""Seeded synthetic trajectories for the RMSD workloads.

Random rigid rotations are uniform quaternions by Shoemake's method (Graphics
Gems III, 1992), as ``utils/rotation.py:105-157`` generates them, turned into
matrices as ``utils/rotation.py:38-102`` does; frames are rotated as row
vectors (x' = x R) and then translated.

Workloads (sizes, nm, float32):
* ``small_cube``       one reference in a 1.0 nm cube (22 atoms), 1,000 frames =
                       reference + N(0, 0.02^2), random rotation, translation U(-2, 2)
* ``walk``             reference uniform in a 5 nm cube; frame k = reference +
                       cumulative N(0, 0.005^2) per-atom steps, then a random
                       rotation and translation per frame (not accumulated)
* ``uneven_systems``   independent systems with log-uniform atom and frame counts,
                       each generated like ``small_cube``
* ``pools``            serving pools with log-uniform frame and atom counts
"""
from __future__ import annotations

import numpy as np


def uniform_quaternion(rng, n):
    """(n, 4) unit quaternions (w, x, y, z), uniform on SO(3)."""
    u = rng.random((3, n))
    s1 = np.sqrt(1.0 - u[0])
    s2 = np.sqrt(u[0])
    a = 2.0 * np.pi * u[1]
    b = 2.0 * np.pi * u[2]
    return np.stack([s1 * np.sin(a), s1 * np.cos(a), s2 * np.sin(b), s2 * np.cos(b)], axis=-1)


def rotation_matrix_from_quaternion(q):
    """(..., 3, 3) rotation matrices of (..., 4) quaternions (need not be unit)."""
    q = np.asarray(q, dtype=np.float64)
    w, x, y, z = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    nq = np.sum(q * q, axis=-1)
    s = np.where(nq > 0, 2.0 / np.where(nq > 0, nq, 1.0), 0.0)
    R = np.empty(q.shape[:-1] + (3, 3))
    R[..., 0, 0] = 1.0 - s * (y * y + z * z)
    R[..., 0, 1] = s * (x * y - w * z)
    R[..., 0, 2] = s * (x * z + w * y)
    R[..., 1, 0] = s * (x * y + w * z)
    R[..., 1, 1] = 1.0 - s * (x * x + z * z)
    R[..., 1, 2] = s * (y * z - w * x)
    R[..., 2, 0] = s * (x * z - w * y)
    R[..., 2, 1] = s * (y * z + w * x)
    R[..., 2, 2] = 1.0 - s * (x * x + y * y)
    return R


def random_rigid(rng, xyz, translate=2.0):
    """Apply an independent uniform rotation and a U(-translate, translate)^3
    translation to every frame of xyz (F, N, 3); float64 in, float64 out."""
    F = xyz.shape[0]
    R = rotation_matrix_from_quaternion(uniform_quaternion(rng, F))
    t = rng.uniform(-translate, translate, size=(F, 1, 3))
    return np.einsum("fnk,fkj->fnj", xyz, R) + t


def small_cube(seed=0, n_frames=1000, n_atoms=22, sigma=0.02, box=1.0):
    """Reference + Gaussian noise, random rigid motion per frame.
    Returns (frames float32 (F, N, 3), reference float32 (N, 3))."""
    rng = np.random.default_rng(seed)
    ref = rng.uniform(0.0, box, size=(n_atoms, 3))
    frames = ref[None] + rng.normal(0.0, sigma, size=(n_frames, n_atoms, 3))
    frames = random_rigid(rng, frames)
    return frames.astype(np.float32), ref.astype(np.float32)


def walk(seed=1, n_frames=1000, n_atoms=500, sigma=0.005, box=5.0, chunk=4096):
    """Random-walk trajectory (host, numpy), float32 (F, N, 3).  Frame 0 is the
    reference itself (before its rigid motion)."""
    rng = np.random.default_rng(seed)
    ref = rng.uniform(0.0, box, size=(n_atoms, 3))
    out = np.empty((n_frames, n_atoms, 3), dtype=np.float32)
    cur = ref
    for f0 in range(0, n_frames, chunk):
        f1 = min(n_frames, f0 + chunk)
        steps = rng.normal(0.0, sigma, size=(f1 - f0, n_atoms, 3))
        if f0 == 0:
            steps[0] = 0.0
        pos = cur[None] + np.cumsum(steps, axis=0)
        cur = pos[-1]
        out[f0:f1] = random_rigid(rng, pos).astype(np.float32)
    return out


def walk_device(seed=1, n_frames=100_000, n_atoms=2000, sigma=0.005, box=5.0, chunk=8192):
    """Random-walk trajectory generated directly in HBM (torch, float32).

    Same construction as ``walk``; the draws come from torch's seeded CUDA
    generator, so the values differ from the numpy version.  Returns a CUDA
    tensor (F, N, 3)."""
    import torch

    dev = torch.device("cuda", torch.cuda.current_device())
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    ref = torch.rand((n_atoms, 3), generator=g, device=dev, dtype=torch.float64) * box
    out = torch.empty((n_frames, n_atoms, 3), device=dev, dtype=torch.float32)
    cur = ref
    for f0 in range(0, n_frames, chunk):
        f1 = min(n_frames, f0 + chunk)
        steps = torch.randn((f1 - f0, n_atoms, 3), generator=g, device=dev,
                            dtype=torch.float64) * sigma
        if f0 == 0:
            steps[0] = 0.0
        pos = cur[None] + torch.cumsum(steps, dim=0)
        cur = pos[-1]
        u = torch.rand((3, f1 - f0), generator=g, device=dev, dtype=torch.float64)
        s1, s2 = torch.sqrt(1 - u[0]), torch.sqrt(u[0])
        a, b = 2 * np.pi * u[1], 2 * np.pi * u[2]
        w, x, y, z = s1 * torch.sin(a), s1 * torch.cos(a), s2 * torch.sin(b), s2 * torch.cos(b)
        R = torch.stack([
            torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
            torch.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
            torch.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1),
        ], -2)
        t = (torch.rand((f1 - f0, 1, 3), generator=g, device=dev, dtype=torch.float64) * 4 - 2)
        out[f0:f1] = (torch.bmm(pos, R) + t).to(torch.float32)
        del steps, pos
    return out


def cube_device(gen_, n_frames, n_atoms, sigma=0.03, box=1.0, out=None):
    """``small_cube`` generated in HBM with torch generator ``gen_``: a
    reference uniform in a box, Gaussian noise per frame, a uniform random
    rotation and a U(-2, 2) translation per frame.  float32 CUDA tensor."""
    import torch

    dev = gen_.device
    ref = torch.rand((n_atoms, 3), generator=gen_, device=dev, dtype=torch.float64) * box
    x = ref[None] + torch.randn((n_frames, n_atoms, 3), generator=gen_, device=dev,
                                dtype=torch.float64) * sigma
    u = torch.rand((3, n_frames), generator=gen_, device=dev, dtype=torch.float64)
    s1, s2 = torch.sqrt(1 - u[0]), torch.sqrt(u[0])
    a, b = 2 * np.pi * u[1], 2 * np.pi * u[2]
    w, qx, qy, qz = s1 * torch.sin(a), s1 * torch.cos(a), s2 * torch.sin(b), s2 * torch.cos(b)
    R = torch.stack([
        torch.stack([1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - w * qz), 2 * (qx * qz + w * qy)], -1),
        torch.stack([2 * (qx * qy + w * qz), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - w * qx)], -1),
        torch.stack([2 * (qx * qz - w * qy), 2 * (qy * qz + w * qx), 1 - 2 * (qx * qx + qy * qy)], -1),
    ], -2)
    t = torch.rand((n_frames, 1, 3), generator=gen_, device=dev, dtype=torch.float64) * 4 - 2
    return (torch.bmm(x, R) + t).to(torch.float32)


def uneven_systems_device(seed=3, n_systems=512, atoms=(50, 5000), frames=(10, 2000),
                          sigma=0.03):
    """The batched workload in HBM: sizes from ``uneven_sizes`` (log-uniform),
    each system like ``small_cube`` with noise sigma.  List of CUDA tensors."""
    import torch

    nf, na = uneven_sizes(seed, n_systems, atoms, frames)
    g = torch.Generator(device=torch.device("cuda", torch.cuda.current_device()))
    g.manual_seed(seed)
    return [cube_device(g, int(f), int(n), sigma) for f, n in zip(nf, na)]


def subset(seed, n_atoms, n_sel):
    """Seeded sorted random atom subset (int64)."""
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(n_atoms, size=n_sel, replace=False)).astype(np.int64)


def log_uniform_int(rng, lo, hi, size):
    return np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi + 1), size=size))).astype(np.int64).clip(lo, hi)


def uneven_systems(seed=3, n_systems=512, atoms=(50, 5000), frames=(10, 2000), sigma=0.03):
    """List of (frames float32 (F_s, N_s, 3)) systems, each like small_cube."""
    rng = np.random.default_rng(seed)
    na = log_uniform_int(rng, atoms[0], atoms[1], n_systems)
    nf = log_uniform_int(rng, frames[0], frames[1], n_systems)
    out = []
    for s in range(n_systems):
        x, _ = small_cube(seed=int(rng.integers(1 << 31)), n_frames=int(nf[s]),
                          n_atoms=int(na[s]), sigma=sigma)
        out.append(x)
    return out


def uneven_sizes(seed=3, n_systems=512, atoms=(50, 5000), frames=(10, 2000)):
    rng = np.random.default_rng(seed)
    na = log_uniform_int(rng, atoms[0], atoms[1], n_systems)
    nf = log_uniform_int(rng, frames[0], frames[1], n_systems)
    return nf, na


def dcd_image(xyz, lengths=None):
    """Bytes of a CHARMM-format DCD file holding float32 coordinates xyz
    (F, N, 3), in angstroms, with an orthorhombic unit cell of ``lengths``
    (3 values) in every frame when given: the layout the reference's writer
    produces (formats/dcd/src/dcdplugin.c:677-790: an 84-byte control
    record, two title lines, the atom count, then per frame the cell record
    and the X, Y and Z records), built with array operations so that large
    synthetic files are cheap.  Returns a uint8 numpy array."""
    import struct

    xyz = np.ascontiguousarray(xyz, np.float32)
    F, N = xyz.shape[:2]
    ctl = bytearray(80)
    struct.pack_into("<iii", ctl, 0, F, 0, 1)  # NSET, ISTART, NSAVC
    struct.pack_into("<f", ctl, 36, 1.0)  # DELTA
    struct.pack_into("<i", ctl, 40, 0 if lengths is None else 1)  # unit cell in every frame
    struct.pack_into("<i", ctl, 76, 24)  # CHARMM version
    titles = b"synthetic trajectory".ljust(80) + b"mdtraj_amd.gen.dcd_image".ljust(80)
    head = (struct.pack("<i", 84) + b"CORD" + bytes(ctl) + struct.pack("<iii", 84, 164, 2)
            + titles + struct.pack("<iiii", 164, 4, N, 4))
    cw = 0 if lengths is None else 14
    rec = np.empty((F, cw + 3 * (N + 2)), np.uint32)
    if cw:
        a, b, c = (float(v) for v in lengths)
        rec[:, 0] = rec[:, 13] = 48
        rec[:, 1:13] = np.array([a, 0.0, b, 0.0, 0.0, c], np.float64).view(np.uint32)
    for k in range(3):
        o = cw + k * (N + 2)
        rec[:, o] = rec[:, o + N + 1] = 4 * N
        rec[:, o + 1:o + 1 + N] = xyz[:, :, k].view(np.uint32)
    return np.concatenate([np.frombuffer(head, np.uint8), rec.reshape(-1).view(np.uint8)])
