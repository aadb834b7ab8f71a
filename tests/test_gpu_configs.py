"""The jittered-grid workloads at full size on the GPU.

* 4,096,000 jittered-grid sources (160^3, seed 2) -> 4,096,000 uniform
  targets (seed 3), 150 nodes per cluster, support = patch radius: radius,
  lattice, centres and all three CSR lists bit-exact against the oracle,
  and the transfer of franke_3d within rel-L2 1e-10 of the oracle's.
* 8,000,000 -> 8,000,000 (200^3 jittered, seed 9; uniform, seed 10), 150
  nodes per cluster: the repeated coupling apply of 32 fields per step
  (new fields each step) through the stored FP32 operator replayed as a HIP
  graph, against the per-step FP64 solve of the same plan (max relative
  difference 1e-5: the FP32 rounding of the operator moves an output by at
  most 2^-24 sum |op| |f|); the same path against the oracle on the
  sub-cube [0, 0.2)^3 of the workload.
* 1,000,000 -> 1,000,000 (100^3 jittered, seed 6; uniform, seed 7), 100
  nodes per cluster: gradients of 0.5 |u - g|^2 (g ~ N(0, 1), seed 8) with
  respect to the source values and the target coordinates through
  pacman_amd.autograd against central differences (step 1e-6) on 200 random
  entries each, relative 1e-6; the source-value gradient (the transpose
  apply) against the oracle's forward operator by the dot-product identity
  <g, Op f> = <Op^T g, f> on the sub-cube [0, 0.3)^3.
"""

import numpy as np
import pytest

import pacman_amd as pacman
from pacman_amd import workloads as W

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _sub_cube(S, T, edge):
    return (np.ascontiguousarray(S[np.all(S < edge, axis=1)]),
            np.ascontiguousarray(T[np.all(T < edge, axis=1)]))


def test_c2_full_size_setup_bit_exact_and_transfer(oracle_mod, measured):
    S, T, npc, sup = W.config("c2")
    print(f"c2: {len(S)} -> {len(T)}, npc {npc}", flush=True)
    plan = pacman.rbf.Plan(3, 0x11, _dev(S), _dev(T), npc, 0.15, sup)
    ora = oracle_mod.Setup(3, 0x11, S, T, npc, 0.15, sup)
    assert plan.radius == ora.radius
    shape, nb = plan.lattice()
    oshape, _, _, _ = ora.lattice()
    np.testing.assert_array_equal(shape, oshape)
    assert nb == ora.num_candidates
    assert plan.num_patches == ora.num_patches
    c, cand = plan.centers()
    oc, ocand = ora.centers()
    np.testing.assert_array_equal(cand, ocand)
    np.testing.assert_array_equal(c, oc)
    for which in (0, 1, 2):
        off, idx = plan.csr(which)
        ooff, oidx = ora.csr(which)
        np.testing.assert_array_equal(off, ooff)
        np.testing.assert_array_equal(idx, oidx)
    f = W.franke_3d(*S.T)
    got = plan.apply(_dev(f)).cpu().numpy()
    ref = ora.apply(f)
    err = rel_l2(got, ref)
    measured("c2 rel-L2 vs oracle", err, 1e-10)
    assert err < 1e-10


def _trig_dev(dS, F, seed):
    """sin(k . x + phi) on the device, k in [1, 6]^3, phi in [0, 2 pi)."""
    import torch
    rng = np.random.default_rng(seed)
    k = torch.from_numpy(rng.integers(1, 7, size=(F, 3)).astype(np.float64))
    phi = torch.from_numpy(rng.uniform(0.0, 2 * np.pi, size=F))
    return torch.sin(k.to(dS.device) @ dS.T + phi.to(dS.device)[:, None])


def test_c5_repeated_apply_fp32_operator_graph(measured):
    import torch
    S, T, npc, sup = W.config("c5")
    dS, dT = _dev(S), _dev(T)
    del S, T
    plan = pacman.rbf.Plan(3, 0x11, dS, dT, npc, 0.15, sup)
    F = 32
    vals = torch.empty((F, plan.num_sources), dtype=torch.float64,
                       device=dS.device)
    out = torch.empty((F, plan.num_targets), dtype=torch.float64,
                      device=dS.device)
    refs = []
    for step in range(3):  # the FP64 solve path, per step
        vals.copy_(_trig_dev(dS, F, 11 + step))
        refs.append(plan.apply(vals).cpu())
    plan.build_operator(precision="fp32")
    worst = 0.0
    for step in range(3):
        vals.copy_(_trig_dev(dS, F, 11 + step))
        plan.apply(vals, out=out, graph=True)
        torch.cuda.synchronize()
        got = out.cpu()
        eager = plan.apply(vals).cpu()
        assert torch.equal(got, eager)
        ref = refs[step]
        worst = max(worst, float((got - ref).abs().max() / ref.abs().max()))
    measured("c5 fp32 operator graph vs fp64 solve, max rel", worst, 1e-5)
    assert worst < 1e-5


def test_c5_sub_cube_operator_graph_vs_oracle(oracle_mod, measured):
    import torch
    S, T, npc, sup = W.config("c5")
    S, T = _sub_cube(S, T, 0.2)
    print(f"c5 sub-cube: {len(S)} -> {len(T)}", flush=True)
    F = 32
    V = np.stack([np.sin((1 + f % 6) * S[:, f % 3] + 0.2 * f)
                  for f in range(F)])
    ref = oracle_mod.Setup(3, 0x11, S, T, npc, 0.15, sup).apply(V)
    plan = pacman.rbf.Plan(3, 0x11, _dev(S), _dev(T), npc, 0.15, sup)
    plan.build_operator(precision="fp32")
    dV = _dev(V)
    out = torch.empty((F, len(T)), dtype=torch.float64, device="cuda:0")
    for _ in range(2):
        plan.apply(dV, out=out, graph=True)
        torch.cuda.synchronize()
    got = out.cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    measured("c5 sub-cube fp32 operator graph vs oracle, max rel", err, 1e-5)
    assert err < 1e-5


def test_c4_adjoint_finite_differences(measured):
    import torch
    S, T, npc, sup = W.config("c4")
    rng = np.random.default_rng(8)
    g = _dev(rng.normal(size=len(T)))
    f = _dev(W.franke_3d(*S.T)).requires_grad_()
    dS = _dev(S)
    tgt = _dev(T).requires_grad_()
    plan = pacman.rbf.Plan(3, 0x11, dS, tgt.detach(), npc, 0.15, sup)
    u = pacman.autograd.transfer(plan, f, tgt)
    loss = 0.5 * ((u - g) ** 2).sum()
    loss.backward()
    gF = f.grad.cpu().numpy()
    gT = tgt.grad.cpu().numpy()
    h = 1e-6
    pick = np.random.default_rng(9)
    # source values: the loss change of a perturbed source value, summed
    # per target as 0.5 (u+ - u-)(u+ + u- - 2 g) (no cancellation of the
    # whole loss)
    fd = np.empty(200)
    src = pick.choice(len(S), 200, replace=False)
    with torch.no_grad():
        fv = f.detach().clone()
        for i, e in enumerate(src):
            x0 = float(fv[e])
            fv[e] = x0 + h
            up = plan.apply(fv)
            fv[e] = x0 - h
            um = plan.apply(fv)
            fv[e] = x0
            fd[i] = float((0.5 * (up - um) * (up + um - 2 * g)).sum()) / (2 * h)
    ref = gF[src]
    err_s = np.abs(fd - ref).max() / np.abs(ref).max()
    measured("c4 source-value gradient vs central differences", err_s, 1e-6)
    # target coordinates: only target t moves (a new plan per perturbed
    # target set); the other terms of the loss are unchanged
    gh = g.cpu().numpy()
    fdt = np.empty(200)
    ent = pick.choice(T.size, 200, replace=False)
    with torch.no_grad():
        Tb = tgt.detach().clone()
        for i, e in enumerate(ent):
            t, a = divmod(int(e), 3)
            x0 = float(Tb[t, a])
            vals = []
            for sgn in (1.0, -1.0):
                Tb[t, a] = x0 + sgn * h
                p = pacman.rbf.Plan(3, 0x11, dS, Tb, npc, 0.15, sup)
                assert p.num_patches == plan.num_patches
                vals.append(float(p.apply(f.detach())[t]))
                del p
            Tb[t, a] = x0
            fdt[i] = (0.5 * (vals[0] - gh[t]) ** 2 -
                      0.5 * (vals[1] - gh[t]) ** 2) / (2 * h)
    reft = gT.reshape(-1)[ent]
    err_t = np.abs(fdt - reft).max() / np.abs(reft).max()
    measured("c4 target-coordinate gradient vs central differences", err_t,
             1e-6)
    assert err_s < 1e-6
    assert err_t < 1e-6


def test_c4_sub_cube_transpose_vs_oracle_operator(oracle_mod, measured):
    S, T, npc, sup = W.config("c4")
    S, T = _sub_cube(S, T, 0.3)
    rng = np.random.default_rng(10)
    ora = oracle_mod.Setup(3, 0x11, S, T, npc, 0.15, sup)
    plan = pacman.rbf.Plan(3, 0x11, _dev(S), _dev(T), npc, 0.15, sup)
    G = rng.normal(size=(3, len(T)))
    AtG = plan.apply_transpose(_dev(G)).cpu().numpy()
    worst = 0.0
    for j in range(3):
        fj = rng.normal(size=len(S))
        lhs = G @ ora.apply(fj)          # <g_i, Op f_j>, oracle forward
        rhs = AtG @ fj                   # <Op^T g_i, f_j>, GPU transpose
        worst = max(worst, float(np.abs(lhs - rhs).max() /
                                 np.abs(lhs).max()))
    measured("c4 sub-cube <g, Op f> - <Op^T g, f>, relative", worst, 1e-10)
    assert worst < 1e-10
