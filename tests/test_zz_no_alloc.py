"""pacman_rbf_plan_apply_noalloc: the allocation-free apply, through the C ABI.

Collected last on purpose (its scratch is shared process-wide on the device).

It must give the same bits as pacman_rbf_plan_apply (the same kernels run,
only the per-call scratch comes from the library's static device array), on
two calls in a row, the second one smaller (the array keeps the first call's
contents), and reject calls it cannot serve with a clear error instead of
allocating.
"""

import ctypes

import numpy as np
import pytest

import pacman_amd as pacman
from pacman_amd import _lib
from pacman_amd import workloads as W

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20


def _noalloc(plan, vals, F, M):
    import torch
    out = torch.full((F, M), np.nan, dtype=torch.float64, device="cuda:0")
    lib = _lib.load()
    rc = lib.pacman_rbf_plan_apply_noalloc(
        plan._h, ctypes.c_void_p(vals.data_ptr()), F,
        ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(0))
    torch.cuda.synchronize()
    return rc, out


def _alloc(plan, vals, F, M):
    import torch
    out = torch.full((F, M), np.nan, dtype=torch.float64, device="cuda:0")
    lib = _lib.load()
    rc = lib.pacman_rbf_plan_apply(
        plan._h, ctypes.c_void_p(vals.data_ptr()), F,
        ctypes.c_void_p(out.data_ptr()), 1, ctypes.c_void_p(0))
    torch.cuda.synchronize()
    assert rc == 0, _lib.last_error()
    return out


def _fields(S, F, seed):
    rng = np.random.default_rng(seed)
    k = rng.integers(1, 5, size=(F, 3))
    return np.ascontiguousarray(np.sin(S @ k.T + rng.uniform(0, 6, F)).T)


def _case(kind, seed):
    import torch
    dev = torch.device("cuda:0")
    if kind == "uniform-3f":
        S, T = W.jittered_grid(30, seed), W.uniform(25000, seed + 1)
        plan = pacman.rbf.Plan(3, 0x11, torch.from_numpy(S).to(dev),
                               torch.from_numpy(T).to(dev), 60, 0.15, -1.0)
        F = 3
    elif kind == "mixture-mixed-8f":
        S, T = W.gaussian_mixture(30000, seed), W.uniform(20000, seed + 1)
        plan = pacman.rbf.Plan.adaptive(3, 0x12, torch.from_numpy(S).to(dev),
                                        torch.from_numpy(T).to(dev), 40, 20,
                                        600, precision="mixed")
        F = 8
    else:  # small, one field
        S, T = W.jittered_grid(14, seed), W.uniform(3000, seed + 1)
        plan = pacman.rbf.Plan(3, 0x13, torch.from_numpy(S).to(dev),
                               torch.from_numpy(T).to(dev), 40, 0.15, -1.0)
        F = 1
    V = torch.from_numpy(_fields(S, F, seed + 2)).to(dev)
    return plan, V, F, len(T), S, T


@pytest.mark.parametrize("first", ["uniform-3f", "mixture-mixed-8f"])
def test_noalloc_matches_apply_on_two_calls(first):
    plan1, V1, F1, M1, S1, _ = _case(first, 3)
    ref1 = _alloc(plan1, V1, F1, M1)      # also prepares F1 on the plan
    rc, got1 = _noalloc(plan1, V1, F1, M1)
    assert rc == 0, _lib.last_error()
    d1 = float((got1 - ref1).abs().max())
    # second call: a smaller plan and input over the same scratch, which
    # still holds the first call's contents
    plan2, V2, F2, M2, _, _ = _case("small-1f", 7)
    ref2 = _alloc(plan2, V2, F2, M2)
    rc, got2 = _noalloc(plan2, V2, F2, M2)
    assert rc == 0, _lib.last_error()
    d2 = float((got2 - ref2).abs().max())
    print(f"{first}: max |noalloc - apply| = {d1:.1e} (F={F1}, M={M1}); "
          f"then small-1f {d2:.1e} (F={F2}, M={M2})")
    assert bool((got1 == ref1).all()) and bool((got2 == ref2).all())
    # and the first plan once more, after the smaller call
    rc, again = _noalloc(plan1, V1, F1, M1)
    assert rc == 0 and bool((again == ref1).all())


def test_noalloc_against_oracle(oracle_mod):
    import torch
    plan, V, F, M, S, T = _case("small-1f", 11)
    rc, got = _noalloc(plan, V, F, M)
    assert rc == 0, _lib.last_error()
    ref = oracle_mod.Setup(3, 0x13, S, T, 40, 0.15, -1.0).apply(
        V.cpu().numpy())
    g = got.cpu().numpy()
    err = np.linalg.norm(g - ref) / np.linalg.norm(ref)
    print(f"noalloc vs oracle: rel-L2 {err:.2e} (tolerance 1e-10)")
    assert err < 1e-10
    del torch


def test_noalloc_rejects_what_it_cannot_serve():
    import torch
    plan, V, F, M, S, _ = _case("uniform-3f", 5)
    # F = 2 was never applied with allocation on this plan: refused
    V2 = V[:2].contiguous()
    rc, _ = _noalloc(plan, V2, 2, M)
    assert rc == -1
    msg = _lib.last_error()
    print("unprepared:", msg)
    assert "allocating apply" in msg
    # more scratch than the static array holds: refused with the sizes
    N = len(S)
    Fbig = SCRATCH_BYTES // (8 * N) + 1
    Vb = torch.zeros((Fbig, N), dtype=torch.float64, device="cuda:0")
    rc, _ = _noalloc(plan, Vb, Fbig, M)
    assert rc == -1
    msg = _lib.last_error()
    print("oversize:", msg)
    assert "MiB of scratch" in msg and "64 MiB" in msg
    # a refused call leaves the entry point usable
    _alloc(plan, V, F, M)
    rc, got = _noalloc(plan, V, F, M)
    assert rc == 0
