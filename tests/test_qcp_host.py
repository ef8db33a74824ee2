"""The kernels' QCP solver (mdtraj_amd/csrc/common.h: Newton on the
characteristic polynomial with a certified bracket, Jacobi on K where the
bracket is too wide), compiled for the host and checked against LAPACK's
symmetric eigensolver and Kabsch on generic, planar, collinear and
single-atom structures.  Collinear (two-atom) structures give K a double top
eigenvalue, where Newton on P alone stalls at about sqrt(eps) * lambda."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from mdtraj_amd import gen
from oracle import fp64

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(ROOT, "mdtraj_amd", "_build", "qcp_host.so")
_P = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def qcp():
    src = os.path.join(HERE, "qcp_host.hip")
    deps = [src, os.path.join(ROOT, "mdtraj_amd", "csrc", "common.h")]
    if not os.path.exists(LIB) or os.path.getmtime(LIB) < max(os.path.getmtime(d) for d in deps):
        os.makedirs(os.path.dirname(LIB), exist_ok=True)
        cmd = ["hipcc", "--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC", "-shared",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-I" + os.path.join(ROOT, "include"),
               "-I" + os.path.join(ROOT, "mdtraj_amd", "csrc"), src, "-o", LIB]
        try:
            subprocess.run(cmd, check=True, capture_output=True, timeout=600)
        except (OSError, subprocess.CalledProcessError) as e:
            pytest.skip(f"hipcc unavailable for the host build of the QCP solver: {e}")
    lib = ctypes.CDLL(LIB)
    lib.qcp_msd_host.restype = ctypes.c_double
    lib.qcp_msd_host.argtypes = [_P, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    lib.qcp_rot_host.restype = ctypes.c_double
    lib.qcp_rot_host.argtypes = [_P, ctypes.c_double, ctypes.c_double, _P]
    lib.qcp_msd2_host.restype = None
    lib.qcp_msd2_host.argtypes = [_P, ctypes.c_double, ctypes.c_double, _P, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_double, _P]
    lib.qcp_newton_resolved_host.restype = ctypes.c_int
    lib.qcp_newton_resolved_host.argtypes = [_P, ctypes.c_double, ctypes.c_double]
    return lib


def _sums(a, b):
    """Centred FP64 inputs of the solver, formed the way the kernels form them
    (raw sums, M = P - S_a S_b^T / N, G = Q - |S|^2 / N)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    n = a.shape[0]
    sa, sb = a.sum(0), b.sum(0)
    M = np.ascontiguousarray((a.T @ b - np.outer(sa, sb) / n).ravel())
    ga = float((a * a).sum() - sa @ sa / n)
    gb = float((b * b).sum() - sb @ sb / n)
    return M, ga, gb, n


def _structures(seed, count):
    rng = np.random.default_rng(seed)
    for t in range(count):
        kind = ("generic", "planar", "collinear", "two")[t % 4]
        n = {"generic": int(rng.integers(4, 40)), "planar": int(rng.integers(3, 12)),
             "collinear": int(rng.integers(2, 10)), "two": 2}[kind]
        scale = float(rng.choice([0.05, 1.0, 8.0]))
        a = rng.normal(size=(n, 3)) * scale
        if kind == "planar":
            a[:, 2] = 0.0
        elif kind in ("collinear", "two"):
            a = np.outer(rng.normal(size=n), rng.normal(size=3)) * scale
        R = gen.rotation_matrix_from_quaternion(gen.uniform_quaternion(rng, 1))[0]
        noise = float(rng.choice([0.0, 1e-4, 1e-2, 0.3]))
        b = a @ R.T + rng.normal(size=(n, 3)) * noise + rng.normal(size=3)
        yield kind, a.astype(np.float32), b.astype(np.float32)


def test_msd_matches_eigensolver_on_degenerate_spectra(qcp):
    worst = {}
    for kind, a, b in _structures(0, 2000):
        M, ga, gb, n = _sums(a, b)
        got = np.sqrt(qcp.qcp_msd_host(M.ctypes.data_as(_P), ga, gb, float(n)))
        want = fp64.rmsd(a[None], b[None], 0)[0]
        worst[kind] = max(worst.get(kind, 0.0), abs(got - want))
    # FP64 raw sums of float32 coordinates up to ~25 nm: r near 0 carries
    # sqrt(eps * G / N) noise, far below the 1e-5 nm bar
    assert all(v < 2e-6 for v in worst.values()), worst


def test_interleaved_pair_solve_equals_two_single_solves(qcp):
    """qcp_msd2 (the all-pairs epilogue's two interleaved solves) returns
    exactly what two qcp_msd calls return, Jacobi fallbacks included."""
    cases = [(_sums(a, b)) for _, a, b in _structures(3, 400)]
    for (Ma, gaa, gba, n), (Mb, gab, gbb, _) in zip(cases[0::2], cases[1::2]):
        out = np.zeros(2)
        qcp.qcp_msd2_host(Ma.ctypes.data_as(_P), gaa, gba, Mb.ctypes.data_as(_P), gab, gbb,
                          float(n), out.ctypes.data_as(_P))
        a = qcp.qcp_msd_host(Ma.ctypes.data_as(_P), gaa, gba, float(n))
        b = qcp.qcp_msd_host(Mb.ctypes.data_as(_P), gab, gbb, float(n))
        assert out[0] == a and out[1] == b


def test_rotation_is_proper_and_optimal(qcp):
    for kind, a, b in _structures(1, 1200):
        M, ga, gb, n = _sums(a, b)
        R = np.zeros(9)
        qcp.qcp_rot_host(M.ctypes.data_as(_P), ga, gb, R.ctypes.data_as(_P))
        R = R.reshape(3, 3)
        assert abs(np.linalg.det(R) - 1.0) < 1e-9
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-9)
        ac = a.astype(np.float64) - a.mean(0, dtype=np.float64)
        bc = b.astype(np.float64) - b.mean(0, dtype=np.float64)
        fit = np.sqrt(((ac @ R - bc) ** 2).sum() / len(a))
        want = fp64.rmsd(a[None], b[None], 0)[0]
        assert abs(fit - want) < 2e-6, (kind, fit, want)


def test_collinear_pairs_need_the_fallback(qcp):
    """A planted check on the certification: for two-atom structures Newton
    alone must report the root unresolved (so the Jacobi path is what makes
    the test above pass), while generic 3-D structures never need it."""
    unresolved = {"generic": 0, "two": 0}
    total = {"generic": 0, "two": 0}
    for kind, a, b in _structures(2, 800):
        if kind not in unresolved:
            continue
        M, ga, gb, _ = _sums(a, b)
        total[kind] += 1
        unresolved[kind] += 1 - qcp.qcp_newton_resolved_host(M.ctypes.data_as(_P), ga, gb)
    assert unresolved["generic"] == 0
    assert unresolved["two"] > total["two"] // 2


def test_walk_pairs_stay_on_the_newton_path(qcp):
    x = gen.walk(n_frames=60, n_atoms=300, seed=7)
    x = x[0] if isinstance(x, tuple) else x
    for i in range(0, 60, 7):
        for j in range(60):
            M, ga, gb, _ = _sums(x[i], x[j])
            assert qcp.qcp_newton_resolved_host(M.ctypes.data_as(_P), ga, gb) == 1


def test_isotropic_structures_where_the_start_bound_is_tight(qcp):
    """A cube's corners have isotropic inertia: against a rotated copy, M = c R
    has three equal singular values, so sqrt(3) |M|_F (one of the two upper
    bounds Newton starts from) is lambda_max itself.  The start, widened by
    1e-6, still lies above every root: the solve stays on the Newton path and
    matches the eigensolver.  Against an inverted copy (M = -c R) the top
    eigenvalue of K is triple (every half-turn fits equally well): that root
    takes the Jacobi fallback, and the RMSD still matches."""
    corners = 0.7 * np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)],
                             np.float64)
    rng = np.random.default_rng(11)
    R = gen.rotation_matrix_from_quaternion(gen.uniform_quaternion(rng, 40))
    for k in range(40):
        for sign in (1.0, -1.0):
            a = (corners + rng.uniform(-2, 2, 3)).astype(np.float32)
            b = (sign * corners @ R[k].T + rng.uniform(-2, 2, 3)).astype(np.float32)
            M, ga, gb, n = _sums(a, b)
            resolved = qcp.qcp_newton_resolved_host(M.ctypes.data_as(_P), ga, gb)
            assert resolved == (1 if sign > 0 else 0), (k, sign)
            got = np.sqrt(qcp.qcp_msd_host(M.ctypes.data_as(_P), ga, gb, float(n)))
            want = fp64.rmsd(a[None], b[None], 0)[0]
            assert abs(got - want) < 2e-6, (k, sign, got, want)


def test_single_atom_and_zero(qcp):
    for a in (np.zeros((1, 3), np.float32), np.ones((1, 3), np.float32),
              np.zeros((5, 3), np.float32)):
        M, ga, gb, n = _sums(a, a + 1.0)
        assert qcp.qcp_msd_host(M.ctypes.data_as(_P), ga, gb, float(n)) == 0.0
        R = np.zeros(9)
        qcp.qcp_rot_host(M.ctypes.data_as(_P), ga, gb, R.ctypes.data_as(_P))
        np.testing.assert_array_equal(R.reshape(3, 3), np.eye(3))
