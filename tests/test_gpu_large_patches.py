"""Patches above 192 source points against the oracle.

The clouds are a uniform background plus a dense Gaussian cluster, so the
reference's single patch radius (FindRadius on the probes, which fall mostly
in the background) gives patches from a few dozen to several hundred points
inside the cluster. Patches above 264 points go through the persistent
workspace kernel (blocked L D L^T in device memory), in FP64 or in mixed
precision (FP32 factors, FP64 iterative refinement, FP64 fallback per
patch); the smaller ones through the register-resident kernel (16-wave
workgroups for the size bins above 10 tile rows).

Tolerances: rel-L2 1e-10 against the oracle with the patch radius as support
(the local matrices are well conditioned: cond(A) <= ~1e3, so FP64 round-off
stays near 1e-13); mixed precision converges to the same FP64 solution to
the refinement tolerance (1e-13 of |alpha|), so the same bound holds. The
measured errors are printed.
"""

import numpy as np
import pytest

import pacman_amd as pacman
from pacman_amd import workloads as W

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def clustered_cloud(n_bg, n_cl, sigma, dim, seed):
    """Uniform background plus a Gaussian cluster away from the box centre
    (the FindRadius probes then sit in the background)."""
    rng = np.random.default_rng(seed)
    bg = rng.uniform(0, 1, (n_bg, dim))
    cl = rng.normal(np.array([0.8, 0.25, 0.7])[:dim], sigma, (n_cl, dim))
    cl = cl[np.all((cl >= 0) & (cl <= 1), axis=1)]
    return np.ascontiguousarray(np.vstack([bg, cl]))


CASES = {
    # name: (n_bg, n_cl, sigma, dim, npc, seed); largest patches ~900,
    # ~1600 and ~930 points
    "3d": (6000, 1500, 0.06, 3, 30, 1),
    "3d-dense": (6000, 3000, 0.07, 3, 40, 1),
    "2d": (3000, 1500, 0.04, 2, 25, 1),
}


def _setup(name):
    n_bg, n_cl, sigma, dim, npc, seed = CASES[name]
    S = clustered_cloud(n_bg, n_cl, sigma, dim, seed)
    T = np.random.default_rng(seed + 100).uniform(0, 1, (5000, dim))
    return S, T, dim, npc


def _fields(S, F, seed=0):
    rng = np.random.default_rng(seed)
    k = rng.integers(1, 5, size=(F, S.shape[1]))
    return np.sin(S @ k.T + rng.uniform(0, 6, F)).T.copy()


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("rbf", [0x11, 0x13])
def test_large_patches_match_oracle(name, rbf, oracle_mod):
    S, T, dim, npc = _setup(name)
    V = _fields(S, 3, seed=len(name))
    o = oracle_mod.Setup(dim, rbf, S, T, npc, 0.15, -1.0)
    ref = o.apply(V)
    plan = pacman.rbf.Plan(dim, rbf, S, T, npc, 0.15, -1.0)
    assert plan.max_patch_size > 192 and plan.num_large_patches > 0
    off, _ = plan.csr(0)
    n = np.diff(off)
    got = plan.apply(V)
    err = rel_l2(got, ref)
    c = plan.counters()
    print(f"{name} rbf={hex(rbf)} K={plan.num_patches} large="
          f"{plan.num_large_patches} n in [{n.min()}, {n.max()}] fp64 "
          f"rel-L2 vs oracle {err:.2e} (tolerance 1e-10) counters {c}")
    assert err < 1e-10
    assert c["singular"] == 0
    plan.set_precision("mixed")
    mixed = plan.apply(V)
    c = plan.counters()
    errm = rel_l2(mixed, ref)
    print(f"  mixed rel-L2 vs oracle {errm:.2e}, vs fp64 "
          f"{rel_l2(mixed, got):.2e}; counters {c}")
    assert errm < 1e-10
    # the dense clusters give cond(A) up to ~1e9 (2-D) and ~1e7 (3-D): FP32
    # factors cannot be refined there and those patches fall back to FP64
    assert c["fallbacks"] <= plan.num_large_patches


def test_mixed_refinement_converges(oracle_mod):
    """With a support of half the patch radius the large patches are well
    conditioned (cond(A) ~ 1e2 - 1e3): FP32 factors plus FP64 refinement
    solve every one of them without a fallback, in a few steps each."""
    S, T, dim, npc = _setup("3d")
    V = _fields(S, 3, seed=9)
    r = pacman.rbf.Plan(dim, 0x11, S, T, npc, 0.15, -1.0).radius
    plan = pacman.rbf.Plan(dim, 0x11, S, T, npc, 0.15, 0.5 * r,
                           precision="mixed")
    got = plan.apply(V)
    c = plan.counters()
    ref = oracle_mod.Setup(dim, 0x11, S, T, npc, 0.15, 0.5 * r).apply(V)
    err = rel_l2(got, ref)
    L = plan.num_large_patches
    print(f"support r/2: {L} large patches, counters {c}, rel-L2 vs oracle "
          f"{err:.2e}")
    assert L > 0 and c["fallbacks"] == 0
    assert 2 * L <= c["refinements"] <= 4 * L
    assert err < 1e-10


def test_mixed_fallback_forced_from_host():
    """With no refinement steps allowed every large patch falls back to the
    FP64 workspace solve: the result equals the FP64 plan bitwise."""
    S, T, dim, npc = _setup("3d")
    V = _fields(S, 2, seed=5)
    plan = pacman.rbf.Plan(dim, 0x11, S, T, npc, 0.15, -1.0)
    fp64 = plan.apply(V)
    plan.set_precision("mixed", max_refinements=0)
    fb = plan.apply(V)
    c = plan.counters()
    print("forced fallback counters", c)
    assert c["fallbacks"] == plan.num_large_patches
    assert np.array_equal(fb, fp64)


@pytest.mark.parametrize("F", [1, 8, 11])
def test_large_patch_fields_and_determinism(F, oracle_mod):
    S, T, dim, npc = _setup("3d-dense")
    V = _fields(S, F, seed=F)
    plan = pacman.rbf.Plan(dim, 0x11, S, T, npc, 0.15, -1.0, precision="mixed")
    a = plan.apply(V)
    b = plan.apply(V)
    assert np.array_equal(a, b)            # run-to-run bitwise
    single = np.stack([plan.apply(V[f]) for f in range(min(F, 3))])
    err_fields = np.abs(single - a[: len(single)]).max() / np.abs(a).max()
    ref = oracle_mod.Setup(dim, 0x11, S, T, npc, 0.15, -1.0).apply(V)
    err = rel_l2(a, ref)
    print(f"F={F}: mixed rel-L2 vs oracle {err:.2e}; one field vs F fields "
          f"max rel {err_fields:.2e}")
    assert err < 1e-10
    assert err_fields < 1e-12


def test_large_patches_device_tensors_and_linear_field():
    import torch
    S, T, dim, npc = _setup("3d")
    dev = torch.device("cuda:0")
    lin = 0.3 + S @ np.array([1.0, -2.0, 0.5])
    plan = pacman.rbf.Plan(dim, 0x11, torch.from_numpy(S).to(dev),
                           torch.from_numpy(T).to(dev), npc, 0.15, -1.0)
    out = plan.apply(torch.from_numpy(lin).to(dev)).cpu().numpy()
    exact = 0.3 + T @ np.array([1.0, -2.0, 0.5])
    covered = np.diff(plan.csr(2)[0]) > 0
    err = np.abs(out - exact)[covered].max()
    print(f"linear field max abs error {err:.2e}")
    assert err < 1e-11


def test_drop_in_reference_defaults_no_longer_refuse():
    """The drop-in call (npc 50, overlap 0.15, support 0.1) on a clustered
    cloud used to stop at the 192-point limit; it now solves every patch and
    matches the oracle with the same defaults (support 0.1 gives worse
    conditioned patches: tolerance 1e-8, the measured error is printed)."""
    import oracle
    S = clustered_cloud(4000, 3000, 0.05, 3, 1)
    T = np.random.default_rng(19).uniform(0, 1, (3000, 3))
    v = W.franke_3d(*S.T)
    out = pacman.rbf.interpolate(3, pacman.execspaces.HIP, 0x11, S, v, T)
    ref = oracle.rbf_interpolate(3, 0x11, S, v, T)
    err = rel_l2(out, ref)
    print(f"drop-in clustered: rel-L2 vs oracle {err:.2e}")
    assert err < 1e-8


@pytest.mark.parametrize("large", [False, True])
def test_duplicate_sources_are_reported(large):
    """Duplicate source points make a local matrix singular (the reference
    returns inf / NaN silently): the drop-in call raises, the plan counts."""
    rng = np.random.default_rng(4)
    if large:
        S = clustered_cloud(6000, 1500, 0.06, 3, 4)
    else:
        S = rng.uniform(0, 1, (4000, 3))
    S = np.vstack([S, S[-40:]])           # 40 exact duplicates (cluster)
    T = rng.uniform(0, 1, (2000, 3))
    v = W.franke_3d(*S.T)
    with pytest.raises(RuntimeError, match="singular"):
        pacman.rbf.interpolate(3, pacman.execspaces.HIP, 0x11, S, v, T)
    plan = pacman.rbf.Plan(3, 0x11, S, T, 30, 0.15, -1.0)
    plan.apply(v)
    c = plan.counters()
    print("duplicates: large patches", plan.num_large_patches, c)
    assert c["singular"] > 0
    if large:
        assert plan.num_large_patches > 0
        plan.set_precision("mixed")
        plan.apply(v)
        c = plan.counters()
        print("  mixed:", c)
        assert c["singular"] > 0 and c["fallbacks"] > 0
