"""GPU parity of the one-vs-all, no-superpose, centering and superpose paths.

Bars: within 1e-5 nm of the FP64 oracle (the kernels accumulate in FP64,
so the observed error is ~1e-7); against the reference's float32 output,
within 1e-4 nm wherever the reference's own float32 error model is below
5e-5 nm, and within that model elsewhere (self pairs, r -> 0)."""
import numpy as np
import pytest

import mdtraj_amd as md
from mdtraj_amd import _rmsd, gen
from oracle import fp64

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-5
TOL_REF32 = 1e-4
EPS32 = np.finfo(np.float32).eps


def ref32_bound(o, g_over_n):
    return 8 * np.minimum(np.sqrt(EPS32 * g_over_n), EPS32 * g_over_n / (2 * np.maximum(o, 1e-12)))


def T(x):
    return md.Trajectory(x)


def test_cube_vs_fp64_and_reference(gpu, golden):
    x, r = golden["cube_xyz"], golden["cube_ref"]
    got = md.rmsd(T(x), T(r[None]), 0)
    assert got.dtype == np.float32 and got.shape == (200,)
    o = fp64.rmsd(x, r[None], 0)
    assert np.abs(got - o).max() < TOL_FP64
    assert np.abs(got - golden["cube_rmsd"]).max() < TOL_REF32


def test_walk_selection_vs_fp64_and_reference(gpu, golden):
    w, idx = golden["walk_xyz"], golden["walk_idx"]
    t = T(w)
    got = md.rmsd(t, t, 5, atom_indices=idx)
    o = fp64.rmsd(w, w, 5, idx)
    assert np.abs(got - o).max() < TOL_FP64
    _, g = fp64.center_and_trace(w[:, idx])
    b = ref32_bound(o, (g + g[5]) / len(idx))
    ref = golden["walk_rmsd_sel_f5"]
    well = b < 5e-5
    assert np.all(np.abs(got - ref)[well] < TOL_REF32)
    assert np.all(np.abs(got - ref)[~well] <= b[~well] + TOL_REF32)


def test_self_pair_is_exactly_zero(gpu, golden):
    w = golden["walk_xyz"]
    t = T(w)
    got = md.rmsd(t, t, 3)
    assert got[3] == 0.0
    o = fp64.rmsd(w, w, 3)
    assert np.abs(got - o).max() < TOL_FP64
    # the reference's aliased call: exact 0 there too
    assert golden["walk_rmsd_self_f3"][3] == 0.0


def test_all_atoms_vs_reference(gpu, golden):
    w = golden["walk_xyz"]
    got = md.rmsd(T(w), T(w.copy()), 0)
    assert np.abs(got - fp64.rmsd(w, w, 0)).max() < TOL_FP64


def test_alignment_known_answer(gpu, golden):
    # tests/test_alignment.py:29-43: rotated copy -> 0 (the reference's float32
    # gives 4e-4 here; FP64 accumulation gives ~1e-7)
    p = golden["align_xyz"]
    got = md.rmsd(T(p), T(p), 0)
    assert got[1] < 1e-6


def test_degenerate_and_tiny(gpu, golden):
    one = golden["one_atom_xyz"]
    assert np.all(md.rmsd(T(one), T(one), 0) == 0.0)
    z = np.zeros((2, 5, 3), np.float32)
    assert np.all(md.rmsd(T(z), T(z), 0) == 0.0)
    # tests/test_rmsd.py:291-316: 1-atom selections against other trajectories
    rng = np.random.default_rng(52)
    t1 = T(rng.standard_normal((1, 1, 3)).astype(np.float32))
    t2 = T(rng.standard_normal((1, 2, 3)).astype(np.float32))
    assert np.isfinite(md.rmsd(t1, t2, atom_indices=[0], ref_atom_indices=[1])).all()
    assert np.isfinite(md.rmsd(t2, t1, atom_indices=[1], ref_atom_indices=[0])).all()


def _line_and_plane_frames(seed, n_frames, n_atoms, kind):
    """Rigid copies (plus noise) of one collinear or planar structure."""
    rng = np.random.default_rng(seed)
    if kind == "line":
        base = np.outer(rng.normal(size=n_atoms), rng.normal(size=3))
    else:
        base = rng.normal(size=(n_atoms, 3))
        base[:, 2] = 0.0
    base *= 2.0
    q = gen.uniform_quaternion(rng, n_frames)
    R = gen.rotation_matrix_from_quaternion(q)
    noise = rng.choice([0.0, 1e-3, 3e-2], size=(n_frames, 1, 1))
    x = np.einsum("na,fba->fnb", base, R) + noise * rng.normal(size=(n_frames, n_atoms, 3))
    return (x + rng.normal(size=(n_frames, 1, 3))).astype(np.float32)


@pytest.mark.parametrize("kind,n_atoms", [("line", 2), ("line", 9), ("plane", 3), ("plane", 40)])
def test_collinear_and_planar_vs_fp64(gpu, kind, n_atoms):
    """Collinear structures give the key matrix a double top eigenvalue,
    where Newton on its characteristic polynomial alone is only
    sqrt(eps)-accurate; planar ones give det M = 0."""
    x = _line_and_plane_frames(n_atoms, 150, n_atoms, kind)
    t = T(x)
    got = md.rmsd(t, t, 5)
    o = fp64.rmsd(x, x, 5)
    assert np.abs(got - o).max() < TOL_FP64
    assert got[5] == 0.0
    # superpose reaches the optimal RMSD even where the rotation is not unique
    s = T(x.copy())
    s.superpose(T(x), 5)
    disp = np.sqrt(((s.xyz.astype(np.float64) - s.xyz[5]) ** 2).sum(2).mean(1))
    assert np.abs(disp - o).max() < 2e-5


def test_empty_target(gpu):
    t = T(np.zeros((0, 7, 3), np.float32))
    r = T(np.ones((1, 7, 3), np.float32))
    assert md.rmsd(t, r, 0).shape == (0,)


@pytest.mark.parametrize("n_frames,n_atoms,n_sel", [
    (1, 3, None), (65, 64, None), (130, 129, 77), (300, 5000, 4500), (50, 20000, 18000),
    (1000, 22, None), (3, 70000, None)])
def test_shapes_vs_fp64(gpu, n_frames, n_atoms, n_sel):
    """Frame counts around the 64-frame group, atom counts around the unroll
    width, selections that fit LDS and that do not (global path)."""
    x = gen.walk(seed=n_frames + n_atoms, n_frames=n_frames, n_atoms=n_atoms)
    idx = None if n_sel is None else gen.subset(n_atoms, n_atoms, n_sel)
    ref_frame = n_frames // 2
    t = T(x)
    got = md.rmsd(t, t, ref_frame, atom_indices=idx)
    o = fp64.rmsd(x, x, ref_frame, idx)
    assert np.abs(got - o).max() < TOL_FP64


@pytest.mark.parametrize("row_lanes", [4, 8, 16, 32, 64])
def test_row_lanes_layouts_vs_fp64(gpu, row_lanes):
    """Every lanes-per-frame layout of the per-frame kernel (16, 8, 4, 2 or
    1 frames per wave): frame counts not a multiple of the frames per wave,
    selections shorter than one row's atom block, gathered and contiguous
    atoms, the no-superpose and fit (superpose) kinds."""
    from mdtraj_amd import _lib

    _lib.call("hr_set_tuning", _lib.HR_TUNE_FRAMES_ROW_LANES, row_lanes)
    try:
        for n_frames, n_atoms, n_sel in [(131, 300, 220), (67, 2000, None), (5, 40, 17),
                                         (70, 1001, None)]:
            x = gen.walk(seed=n_frames * 7 + n_atoms, n_frames=n_frames, n_atoms=n_atoms)
            idx = None if n_sel is None else gen.subset(n_atoms, n_atoms, n_sel)
            f = n_frames // 3
            t = T(x)
            got = md.rmsd(t, t, f, atom_indices=idx)
            assert np.abs(got - fp64.rmsd(x, x, f, idx)).max() < TOL_FP64
            assert got[f] == 0.0
            ns = md.rmsd(t, t, f, atom_indices=idx, superpose=False)
            assert np.abs(ns - fp64.rmsd_nosuperpose(x, x, f, idx)).max() < TOL_FP64
            sup = T(x.copy()).superpose(t, f, atom_indices=idx)
            assert np.abs(sup.xyz - fp64.superpose(x, x, f, idx)).max() < 2e-5
    finally:
        _lib.call("hr_set_tuning", _lib.HR_TUNE_FRAMES_ROW_LANES, 0)


def test_ref_atom_indices_mapping(gpu):
    # tests/test_rmsd.py:143-171: swapping halves with ref_atom_indices
    rng = np.random.default_rng(52)
    base = rng.standard_normal((1, 100, 3))
    normal = gen.random_rigid(rng, base + 0.05 * rng.standard_normal((6, 100, 3))).astype(np.float32)
    flipped = np.concatenate([normal[:, 50:], normal[:, :50]], axis=1)
    got = md.rmsd(T(normal), T(flipped), 2, atom_indices=np.arange(50),
                  ref_atom_indices=np.arange(50, 100))
    o = fp64.rmsd(normal, flipped, 2, np.arange(50), np.arange(50, 100))
    assert np.abs(got - o).max() < TOL_FP64
    assert got[2] < 1e-6
    # wrong reference indices give strictly larger RMSDs (tests/test_rmsd.py:204-228)
    bad = md.rmsd(T(normal), T(flipped), 2, atom_indices=np.arange(50),
                  ref_atom_indices=np.arange(0, 50))
    assert np.all(bad > got)


def test_nosuperpose(gpu, golden):
    w, idx = golden["walk_xyz"], golden["walk_idx"]
    got = md.rmsd(T(w), T(w), 7, atom_indices=idx, superpose=False)
    o = fp64.rmsd_nosuperpose(w, w, 7, idx)
    assert np.abs(got - o).max() < TOL_FP64
    assert got[7] == 0.0


def test_precentered_equals_plain(gpu, golden):
    # tests/test_rmsd.py:51-70: precentered vs not, decimal 6
    w = golden["walk_xyz"]
    t1 = T(w.copy())
    t2 = T(w.copy())
    t2.center_coordinates()
    for i in (0, 9, 63):
        v1 = md.rmsd(t1, t1, i, precentered=False)
        v2 = md.rmsd(t2, t2, i, precentered=True)
        assert np.abs(v1 - v2).max() < 1.5e-6
    with pytest.warns(RuntimeWarning):
        md.rmsd(t2, t2, 0, atom_indices=[0, 1, 2], precentered=True)


def test_inputs_not_modified(gpu, golden):
    w = golden["walk_xyz"].copy()
    t = T(w)
    md.rmsd(t, t, 0)
    assert np.array_equal(t.xyz, golden["walk_xyz"])


def test_center_coordinates_matches_reference(gpu, golden):
    t = T(golden["walk_xyz"].copy())
    t.center_coordinates()
    assert np.array_equal(t.xyz, golden["walk_centered"])
    assert np.abs(t._rmsd_traces - golden["walk_traces"]).max() <= 2 * np.spacing(
        golden["walk_traces"]).max()
    # tests/test_trajectory.py:316-346: zero centroid
    assert np.abs(t.xyz.mean(axis=1)).max() < 1e-5


def test_center_inplace_device_tensor(gpu, golden):
    x = gpu.from_numpy(golden["walk_xyz"].copy()).cuda()
    g = _rmsd._center_inplace_atom_major(x)
    assert np.array_equal(x.cpu().numpy(), golden["walk_centered"])
    assert g.shape == (64,)


def test_superpose_vs_reference_and_fp64(gpu, golden):
    x, sel = golden["sup_xyz"], golden["sup_sel"]
    t = T(x.copy())
    t.superpose(T(x), 4, atom_indices=sel)
    assert np.abs(t.xyz - fp64.superpose(x, x, 4, sel)).max() < 2e-5
    assert np.abs(t.xyz - golden["sup_moved_sel_f4"]).max() < 5e-5
    t = T(x.copy())
    t.superpose(t, 0)
    assert np.abs(t.xyz - golden["sup_moved_all_f0"]).max() < 5e-5


def test_superpose_semantics(gpu):
    rng = np.random.default_rng(52)
    # tests/test_rmsd.py:111-128: far from origin, reference unmodified,
    # translation restored
    reference = T(rng.standard_normal((1, 100, 3)).astype(np.float32) + 100)
    ref_copy = reference.xyz.copy()
    for indices in [None, np.arange(90)]:
        query = T(reference.xyz + 0.01 * rng.standard_normal(reference.xyz.shape).astype(np.float32))
        query.superpose(reference, 0, atom_indices=indices)
        assert np.array_equal(reference.xyz, ref_copy)
        c = query.xyz[0].mean(0)
        assert np.all((80 < c) & (c < 120))
    # tests/test_rmsd.py:98-108: after superpose, plain displacement RMSD == QCP
    w = gen.walk(seed=4, n_frames=30, n_atoms=80)
    t = T(w.copy())
    qcp = md.rmsd(t, t, 0)
    t.superpose(t, 0)
    disp = np.sqrt(((t.xyz.astype(np.float64) - t.xyz[0]) ** 2).sum(2).mean(1))
    assert np.abs(qcp - disp).max() < 1.5e-5
    # tests/test_rmsd.py:319-334: 1-atom superpose stays finite
    t = T(rng.standard_normal((5, 20, 3)).astype(np.float32))
    t.superpose(t, frame=0, atom_indices=[0])
    assert np.all(np.isfinite(t.xyz))


def test_superpose_chunks_match_one_pass(gpu):
    """hr_superpose fits and rotates in chunks of HR_TUNE_SUPERPOSE_CHUNK_MIB.
    The boundaries change nothing: moved coordinates, rotations and RMSDs
    equal a single pass bit for bit, a ragged last chunk included, and the
    RMSDs match the FP64 oracle frame by frame."""
    import torch

    from mdtraj_amd import _lib, _device as dev

    F, N, frame = 500, 1000, 3  # 12,000 B per frame: 87 frames per MiB, 65 in the last chunk
    w = gen.walk(seed=9, n_frames=F, n_atoms=N)
    sel = np.arange(0, N, 3)
    idx = torch.from_numpy(sel.astype(np.int32)).cuda()
    x0 = torch.from_numpy(w).cuda()
    ref = x0[frame].clone()
    ws = dev.workspace(_lib.load().hr_superpose_workspace_bytes(F, len(sel)))
    got = {}
    try:
        for mib in (4096, 1):
            _lib.call("hr_set_tuning", _lib.HR_TUNE_SUPERPOSE_CHUNK_MIB, mib)
            x = x0.clone()
            rot = torch.empty(F, 9, device="cuda")
            out = torch.empty(F, device="cuda")
            _lib.call("hr_superpose", x.data_ptr(), F, N, None, N, idx.data_ptr(), len(sel),
                      ref.data_ptr(), idx.data_ptr(), 1, rot.data_ptr(), out.data_ptr(),
                      ws.data_ptr(), dev.stream())
            torch.cuda.synchronize()
            got[mib] = (x.cpu().numpy(), rot.cpu().numpy(), out.cpu().numpy())
    finally:
        _lib.call("hr_set_tuning", _lib.HR_TUNE_SUPERPOSE_CHUNK_MIB, 0)
    for a, b in zip(got[4096], got[1]):
        assert np.array_equal(a, b)
    assert np.abs(got[1][2] - fp64.rmsd(w, w, frame, sel)).max() < TOL_FP64
    assert np.abs(got[1][0] - fp64.superpose(w, w, frame, sel)).max() < 2e-5


def test_superpose_in_place(gpu):
    """The reference transforms the trajectory's own array
    (core/trajectory.py:1128 aliases self.xyz): a host array held elsewhere
    sees the result, and a device-resident trajectory keeps its memory."""
    import torch

    w = gen.walk(seed=9, n_frames=40, n_atoms=60)
    want = fp64.superpose(w, w, 3)
    a = w.copy()
    t = T(a)
    t.superpose(t, 3)
    assert t.xyz is a
    assert np.abs(a - want).max() < 2e-5
    d = torch.from_numpy(w.copy()).cuda()
    td = T(d)
    td.superpose(td, 3)
    assert td.xyz.data_ptr() == d.data_ptr()
    assert np.abs(d.cpu().numpy() - want).max() < 2e-5


def test_superpose_refinds_swap(gpu):
    # tests/test_rmsd.py:143-171
    rng = np.random.default_rng(52)
    normal = rng.standard_normal((1, 100, 3)).astype(np.float32)
    flipped = np.concatenate([normal[:, 50:], normal[:, :50]], axis=1)
    n, f = T(normal.copy()), T(flipped.copy())
    n.superpose(f, atom_indices=np.arange(0, 50), ref_atom_indices=np.arange(50, 100))
    assert np.abs(n.xyz - normal).max() < 1e-5
    f.superpose(n, atom_indices=np.arange(50, 100), ref_atom_indices=np.arange(0, 50))
    assert np.abs(f.xyz - flipped).max() < 1e-5
    n.superpose(f)
    assert not np.allclose(n.xyz, normal)


def test_getmultiple_atom_and_axis_major(gpu, golden):
    w = golden["walk_xyz"]
    c, g = fp64.center_and_trace(w)
    c = c.astype(np.float32)
    g = g.astype(np.float32)
    a = _rmsd.getMultipleRMSDs_atom_major(c, c, g, g, 6)
    axis = np.ascontiguousarray(c.transpose(0, 2, 1))
    b = _rmsd.getMultipleRMSDs_axis_major(axis, axis, g, g, 6)
    o = fp64.rmsd(c, c, 6)
    assert np.abs(a - o).max() < TOL_FP64
    assert np.abs(b - o).max() < TOL_FP64


def test_device_resident_returns_tensor(gpu, golden):
    w = golden["walk_xyz"]
    t = T(gpu.from_numpy(w.copy()).cuda())
    assert t.is_on_device
    got = md.rmsd(t, t, 1)
    assert got.is_cuda and got.dtype == gpu.float32
    assert np.abs(got.cpu().numpy() - fp64.rmsd(w, w, 1)).max() < TOL_FP64


def test_prepared_plan_vs_fp64(gpu):
    """mdtraj_amd.plan.OneVsAll (hr_rmsd_prepare_reference once, then
    hr_rmsd_one_vs_all_prepared per run): several references through one
    plan, an external reference with its own indices, self frame exactly 0."""
    from mdtraj_amd.plan import OneVsAll

    x = gen.walk(seed=21, n_frames=500, n_atoms=300)
    sel = gen.subset(5, 300, 220)
    plan = OneVsAll(x, sel)
    for f in (0, 123, 499):
        got = plan.set_reference_frame(f).run().cpu().numpy()
        want = fp64.rmsd(x, x, f, sel)
        assert np.abs(got - want).max() < TOL_FP64
        assert got[f] == 0.0
    other = gen.walk(seed=22, n_frames=3, n_atoms=400)
    rsel = np.sort(np.random.default_rng(3).choice(400, 220, replace=False))
    got = plan.set_reference(other[1], rsel).run(self_frame=-1).cpu().numpy()
    want = fp64.rmsd(x, other, 1, sel, rsel)
    assert np.abs(got - want).max() < TOL_FP64
