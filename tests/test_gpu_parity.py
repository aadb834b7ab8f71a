"""GPU parity: the HIP path against the CPU FP64 oracle on the same inputs.

Setup (patch radius, lattice, centres, member lists, weight lists) must be
bit-identical; interpolated fields must agree to the FP64 tolerance of the
transfer (1e-10 relative L2 when the Wendland support equals the patch
radius; the reference-default support 0.1 gives worse-conditioned patches,
so those cases use 1e-8). The reference's own known-answer tests run
through all three drop-in surfaces (Python, C ABI, C++).
"""

import ctypes
import os
import subprocess

import numpy as np
import pytest

import pacman_amd as pacman
from pacman_amd import workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MESHES = os.path.join(ROOT, "tests", "golden", "meshes")
BASES = [0x10, 0x11, 0x12, 0x13, 0x14]

TOL_PATCH_RADIUS_SUPPORT = 1e-10   # support = patch radius
TOL_DEFAULT_SUPPORT = 1e-8         # reference default support 0.1: ceiling
EPS = np.finfo(np.float64).eps

# Wendland C0 .. C8 of p = d / support (rbf_functions.hpp:22-150), only to
# form the local matrices whose condition numbers bound the tolerance below.
_PHI = {
    0x10: lambda p: (1 - p) ** 2,
    0x11: lambda p: (1 - p) ** 4 * (4 * p + 1),
    0x12: lambda p: (1 - p) ** 6 * (35 * p * p + 18 * p + 3),
    0x13: lambda p: (1 - p) ** 8 * (32 * p ** 3 + 25 * p * p + 8 * p + 1),
    0x14: lambda p: (1 - p) ** 10 * (1287 * p ** 4 + 1350 * p ** 3
                                     + 630 * p * p + 150 * p + 15),
}


def default_support_tol(rbf, S, ora, support):
    """Tolerance of a transfer with the reference's absolute support 0.1:
    max(1e-10, 100 kappa eps), kappa the largest 2-norm condition number of
    the patches' Wendland matrices. Two FP64 eliminations of one patch
    system in different orders (blocked L D L^T with FMA contraction here,
    Doolittle in the oracle) differ by O(kappa eps) relative; the
    partition-of-unity blend is a convex combination of the local values, so
    the transfer inherits the worst patch's bound. The fixed 1e-8 stays a
    ceiling. Returns (tolerance, kappa)."""
    off, idx = ora.csr(0)
    kappa = 1.0
    phi = _PHI[rbf]
    for k in range(len(off) - 1):
        X = S[idx[off[k]:off[k + 1]]]
        if len(X) < 2:
            continue
        d = np.sqrt(((X[:, None, :] - X[None, :, :]) ** 2).sum(-1)) / support
        A = np.where(d < 1.0, phi(np.minimum(d, 1.0)), 0.0)
        kappa = max(kappa, np.linalg.cond(A))
    return min(TOL_DEFAULT_SUPPORT, max(TOL_PATCH_RADIUS_SUPPORT,
                                        100 * kappa * EPS)), kappa


def rel_l2(a, b):
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


def assert_setup_equal(plan, ora):
    assert plan.radius == ora.radius, (plan.radius, ora.radius)
    shape, nb = plan.lattice()
    oshape, _, _, _ = ora.lattice()
    assert nb == ora.num_candidates
    np.testing.assert_array_equal(shape, oshape)
    assert plan.num_patches == ora.num_patches
    c, cand = plan.centers()
    oc, ocand = ora.centers()
    np.testing.assert_array_equal(cand, ocand)
    np.testing.assert_array_equal(c, oc)
    for which in (0, 1, 2):
        off, idx = plan.csr(which)
        ooff, oidx = ora.csr(which)
        np.testing.assert_array_equal(off, ooff, err_msg=f"csr {which} offsets")
        np.testing.assert_array_equal(idx, oidx, err_msg=f"csr {which} indices")


def run_case(dim, rbf, S, T, vals, npc, support, oracle_mod, tol,
             overlap=0.15):
    print(f"\ncase dim={dim} rbf={hex(rbf)} N={len(S)} M={len(T)} npc={npc} "
          f"support={support}", flush=True)
    plan = pacman.rbf.Plan(dim, rbf, S, T, npc, overlap, support)
    ora = oracle_mod.Setup(dim, rbf, S, T, npc, overlap, support)
    assert_setup_equal(plan, ora)
    out = plan.apply(vals)
    ref = ora.apply(vals)
    err = rel_l2(out, ref)
    print(f"  K={plan.num_patches} r={plan.radius:.6g} rel_l2 vs oracle "
          f"{err:.3e}", flush=True)
    assert err <= tol
    return plan, ora, out, ref


@pytest.mark.parametrize("rbf", BASES)
def test_random_cloud_parity(rbf, oracle_mod):
    rng = np.random.default_rng(11)
    S = rng.uniform(0, 1, (3000, 3))
    T = rng.uniform(0, 1, (2500, 3))
    v = W.franke_3d(*S.T)
    run_case(3, rbf, S, T, v, 40, -1.0, oracle_mod, TOL_PATCH_RADIUS_SUPPORT)


def test_dense_targets_overflow_membership_rows(oracle_mod):
    """Targets 60x denser than the sources: every patch holds far more
    targets than the one-pass membership rows (256), so the lists come from
    the geometry again; they must still equal the oracle's."""
    rng = np.random.default_rng(12)
    S = rng.uniform(0, 1, (1500, 3))
    T = rng.uniform(0, 1, (90000, 3))
    v = W.franke_3d(*S.T)
    plan, _, _, _ = run_case(3, 0x11, S, T, v, 30, -1.0, oracle_mod,
                             TOL_PATCH_RADIUS_SUPPORT)
    off, _ = plan.csr(1)
    assert np.diff(off).max() > 256


def test_config1_parity(oracle_mod):
    S, T, npc, support = W.config("c1")
    v = W.franke_3d(*S.T)
    plan, ora, out, ref = run_case(3, 0x11, S, T, v, npc, support, oracle_mod,
                                   TOL_PATCH_RADIUS_SUPPORT)
    # accuracy against the analytic field
    exact = W.franke_3d(*T.T)
    print("  rel_l2 vs franke_3d", rel_l2(out, exact), rel_l2(ref, exact))
    assert rel_l2(out, exact) <= rel_l2(ref, exact) * (1 + 1e-6)


def test_config1_reference_defaults(oracle_mod, measured):
    S, T, _, _ = W.config("c1")
    v = W.franke_3d(*S.T)
    _, ora, out, ref = run_case(3, 0x11, S, T, v, 50, 0.1, oracle_mod,
                                TOL_DEFAULT_SUPPORT)
    tol, kappa = default_support_tol(0x11, S, ora, 0.1)
    err = rel_l2(out, ref)
    measured(f"c1 support 0.1 rel-L2 vs oracle (kappa {kappa:.2e})", err,
             tol)
    assert err <= tol


@pytest.mark.parametrize("dim", [1, 2])
def test_low_dimension_parity(dim, oracle_mod):
    rng = np.random.default_rng(5 + dim)
    n = 400 if dim == 1 else 2500
    S = rng.uniform(0, 1, (n, dim))
    T = rng.uniform(0, 1, (n // 2, dim))
    v = np.sin(3 * S).sum(axis=1)
    run_case(dim, 0x11, S, T, v, 20, -1.0, oracle_mod,
             TOL_PATCH_RADIUS_SUPPORT)


# ---- the reference's own known-answer tests --------------------------------

@pytest.mark.parametrize("rbf", BASES)
def test_reference_same_grid_python(rbf, oracle_mod, measured):
    """tests/test_cpp_rbf_interface.cpp:165-198: 20^3 grid, source = target,
    rel-L2 < 1e-8 for every Wendland basis."""
    P = W.regular_grid(20)
    v = W.franke_3d(*P.T)
    out = pacman.rbf.interpolate(3, pacman.execspaces.HIP, rbf, P, v, P)
    assert rel_l2(out, v) < 1e-8
    ref = oracle_mod.rbf_interpolate(3, rbf, P, v, P)
    tol, kappa = default_support_tol(rbf, P, oracle_mod.Setup(3, rbf, P, P),
                                     0.1)
    err = rel_l2(out, ref)
    measured(f"20^3 same grid {hex(rbf)} rel-L2 vs oracle "
             f"(kappa {kappa:.2e})", err, tol)
    assert err <= tol


def test_reference_same_grid_c_abi():
    """Same pin through pacman_rbf_interpolate_c (the Fortran shim's C ABI)."""
    lib = pacman._lib.load()
    P = W.regular_grid(20)
    v = W.franke_3d(*P.T)
    out = np.zeros(len(P))
    dp = ctypes.POINTER(ctypes.c_double)
    for rbf in BASES:
        rc = lib.pacman_rbf_interpolate_c(
            3, 4, rbf, P.ctypes.data_as(dp), len(P), v.ctypes.data_as(dp),
            P.ctypes.data_as(dp), len(P), out.ctypes.data_as(dp))
        assert rc == 0
        assert rel_l2(out, v) < 1e-8


def test_reference_same_grid_cpp():
    """tests/test_cpp_rbf_interface.cpp restated against include/interface.hpp
    (binary built by __graft_entry__.build())."""
    exe = os.path.join(ROOT, "tests", "cpp", "test_cpp_rbf_interface")
    if not os.path.exists(exe):
        pytest.skip("C++ test binary not built")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0


def test_reference_same_grid_fortran():
    """tests/test_fortran_rbf_interface.f90 restated against the Fortran
    module (binary built by __graft_entry__.build() when amdflang exists)."""
    exe = os.path.join(ROOT, "tests", "fortran", "test_fortran_rbf_interface")
    if not os.path.exists(exe):
        pytest.skip("Fortran test binary not built")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0


CASES = {
    "coarse_to_fine_0": ("0.03.npz", "0.003.npz"),
    "fine_to_coarse_0": ("0.003.npz", "0.03.npz"),
    "same_0": ("0.03.npz", "0.03.npz"),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("rbf", BASES)
def test_reference_npz_fixtures(case, rbf, oracle_mod, measured):
    """tests/test_rbf-pum.py:56-74,113-122: rel-L2 vs franke_3d < 1.5e-2."""
    a, b = CASES[case]
    sp = np.load(os.path.join(MESHES, a))["points"][:, :3].astype(np.float64)
    tp = np.load(os.path.join(MESHES, b))["points"][:, :3].astype(np.float64)
    v = W.franke_3d(*sp.T)
    out = pacman.rbf.interpolate(3, pacman.execspaces.HIP, rbf, sp, v, tp)
    exact = W.franke_3d(*tp.T)
    assert rel_l2(out, exact) < 1.5e-2
    ref = oracle_mod.rbf_interpolate(3, rbf, sp, v, tp)
    err = rel_l2(out, ref)
    tol, kappa = default_support_tol(rbf, sp,
                                     oracle_mod.Setup(3, rbf, sp, tp), 0.1)
    print(f"{case} {hex(rbf)} vs franke {rel_l2(out, exact):.3e} "
          f"vs oracle {err:.3e}")
    measured(f"npz {case} {hex(rbf)} rel-L2 vs oracle (kappa {kappa:.2e})",
             err, tol)
    assert err <= tol


def test_axis_drop_matches_oracle(oracle_mod, measured):
    """fine_to_coarse_0 has near-planar patches: the conditioning loop
    (serial_ops.hpp:108-139) must drop the same axes as the oracle."""
    sp = np.load(os.path.join(MESHES, "0.003.npz"))["points"][:, :3]
    tp = np.load(os.path.join(MESHES, "0.03.npz"))["points"][:, :3]
    sp = sp.astype(np.float64)
    tp = tp.astype(np.float64)
    v = W.franke_3d(*sp.T)
    plan = pacman.rbf.Plan(3, 0x11, sp, tp, 50, 0.15, 0.1)
    ora = oracle_mod.Setup(3, 0x11, sp, tp, 50, 0.15, 0.1)
    assert_setup_equal(plan, ora)
    out = plan.apply(v)
    ref, opv = ora.apply(v, return_pv=True)
    pv, status = plan.patch_info()
    print("patches with dropped axes:", int((opv < 4).sum()), "of", len(opv))
    assert (opv < 4).sum() > 0
    np.testing.assert_array_equal(pv, opv)
    assert not status.any()
    tol, kappa = default_support_tol(0x11, sp, ora, 0.1)
    err = rel_l2(out, ref)
    measured(f"npz fine->coarse axis drop rel-L2 vs oracle (kappa "
             f"{kappa:.2e})", err, tol)
    assert err <= tol


# ---- properties --------------------------------------------------------------

def test_multi_field_equals_single_fields(oracle_mod):
    rng = np.random.default_rng(3)
    S = rng.uniform(0, 1, (4000, 3))
    T = rng.uniform(0, 1, (3000, 3))
    V = W.trig_fields(S, 11, seed=4)      # 11 fields -> two launches
    plan = pacman.rbf.Plan(3, 0x11, S, T, 60, 0.15, -1.0)
    out = plan.apply(V)
    for f in (0, 7, 8, 10):
        one = plan.apply(V[f])
        np.testing.assert_array_equal(out[f], one)
    ref = oracle_mod.Setup(3, 0x11, S, T, 60, 0.15, -1.0).apply(V)
    assert rel_l2(out, ref) <= TOL_PATCH_RADIUS_SUPPORT


def test_linear_reproduction():
    """f = a + b.x is reproduced exactly (beta exact, residual 0, alpha 0)."""
    rng = np.random.default_rng(9)
    S = rng.uniform(0, 1, (20000, 3))
    T = rng.uniform(0.05, 0.95, (10000, 3))
    f = lambda P: 1.0 + 2.0 * P[:, 0] - 3.0 * P[:, 1] + 0.5 * P[:, 2]
    plan = pacman.rbf.Plan(3, 0x11, S, T, 80, 0.15, -1.0)
    out = plan.apply(f(S))
    off, _ = plan.csr(2)
    covered = np.diff(off) > 0
    assert np.max(np.abs(out[covered] - f(T)[covered])) < 1e-11


def test_device_tensors_match_host():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(21)
    S = rng.uniform(0, 1, (5000, 3))
    T = rng.uniform(0, 1, (4000, 3))
    v = W.franke_3d(*S.T)
    host = pacman.rbf.Plan(3, 0x12, S, T, 50, 0.15, -1.0).apply(v)
    dS = torch.from_numpy(S).cuda()
    dT = torch.from_numpy(T).cuda()
    plan = pacman.rbf.Plan(3, 0x12, dS, dT, 50, 0.15, -1.0)
    dev = plan.apply(torch.from_numpy(v).cuda())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dev.cpu().numpy(), host)


def test_deterministic_repeat():
    rng = np.random.default_rng(8)
    S = rng.uniform(0, 1, (6000, 3))
    T = rng.uniform(0, 1, (6000, 3))
    v = W.franke_3d(*S.T)
    a = pacman.rbf.Plan(3, 0x11, S, T, 70, 0.15, -1.0).apply(v)
    b = pacman.rbf.Plan(3, 0x11, S, T, 70, 0.15, -1.0).apply(v)
    np.testing.assert_array_equal(a, b)


def test_targets_outside_all_patches_are_zero():
    """Targets no patch covers keep 0 (src/common/transfer.hxx:109)."""
    rng = np.random.default_rng(2)
    S = rng.uniform(0, 1, (3000, 3))
    T = np.vstack([rng.uniform(0, 1, (100, 3)), rng.uniform(5, 6, (50, 3))])
    out = pacman.rbf.Plan(3, 0x11, S, T, 40, 0.15, -1.0).apply(
        W.franke_3d(*S.T))
    assert np.all(out[100:] == 0.0)
    assert np.all(out[:100] != 0.0)
