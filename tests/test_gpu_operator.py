"""Stored per-patch operator (Plan.build_operator) against the solve path.

The stored operator is built with the transpose chain on unit upstream
values; the transfer it gives must equal the plan's ordinary apply (the
forward chain) to FP64 round-off, for every basis, 2-D and 3-D, several
fields (up to 32 per pass over the operator, an FP64 MFMA contraction above
8 fields; 40 fields take two passes),
targets much denser than the sources
(many operator rows per patch), graph replays and after drop_operator.
The FP32-stored operator (FP64 sums) differs from the solve only by the
rounding of its entries: each output moves by at most 2^-24 sum_e |op_qe|
|v_e|, which is checked here as a relative max difference below 1e-5 (the
operator rows of these clouds have absolute sums of order 10; the measured
difference is printed).
"""

import numpy as np
import pytest

import pacman_amd as pacman
from pacman_amd import workloads as W

pytestmark = pytest.mark.gpu


def _rel_max(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("rbf,dim,F", [(0x11, 3, 1), (0x10, 3, 3),
                                       (0x14, 3, 1), (0x13, 2, 2),
                                       (0x11, 3, 2), (0x11, 3, 5),
                                       (0x11, 3, 8), (0x12, 3, 10),
                                       (0x11, 3, 20), (0x11, 3, 40)])
def test_operator_apply_equals_solve(rbf, dim, F):
    rng = np.random.default_rng(200 + rbf + dim + F)
    S = rng.uniform(0, 1, (6000, dim))
    T = rng.uniform(0, 1, (5000, dim))
    V = np.stack([np.sin((1 + i) * S[:, 0] + 0.2 * i) + S[:, -1] ** 2
                  for i in range(F)])
    plan = pacman.rbf.Plan(dim, rbf, S, T, 60, 0.15, -1.0)
    ref = plan.apply(V)
    plan.build_operator()
    assert plan.operator_bytes > 0
    got = plan.apply(V)
    err = _rel_max(got, ref)
    print(f"rbf={hex(rbf)} dim={dim} F={F} max rel {err:.2e}")
    assert err < 1e-10, err
    plan.drop_operator()
    assert plan.operator_bytes == 0
    assert np.array_equal(plan.apply(V), ref)


def test_operator_dense_targets_and_graph():
    import torch
    rng = np.random.default_rng(9)
    S = rng.uniform(0, 1, (2000, 3))
    T = rng.uniform(0, 1, (60000, 3))
    dev = torch.device("cuda:0")
    plan = pacman.rbf.Plan(3, 0x11, torch.from_numpy(S).to(dev),
                           torch.from_numpy(T).to(dev), 40, 0.15, -1.0)
    v = torch.from_numpy(W.trig_fields(S, 4, seed=3)).to(dev)
    ref = plan.apply(v).cpu().numpy()
    plan.build_operator()
    out = torch.empty_like(torch.from_numpy(ref)).to(dev)
    eager = plan.apply(v).cpu().numpy()
    assert _rel_max(eager, ref) < 1e-10
    for _ in range(3):
        plan.apply(v, out=out, graph=True)
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), eager)


def test_operator_at_1m_points():
    import torch
    S, T, npc, sup = W.config("c4")
    dev = torch.device("cuda:0")
    plan = pacman.rbf.Plan(3, 0x11, torch.from_numpy(S).to(dev),
                           torch.from_numpy(T).to(dev), npc, 0.15, sup)
    v = torch.from_numpy(W.franke_3d(*S.T)).to(dev)
    ref = plan.apply(v).cpu().numpy()
    plan.build_operator()
    got = plan.apply(v).cpu().numpy()
    assert _rel_max(got, ref) < 1e-10


@pytest.mark.parametrize("F", [1, 3, 11])
def test_operator_transpose_equals_solve(F):
    """The transpose apply (source-value gradient) through the stored
    operator equals the transpose chain's, and the pair keeps the
    dot-product identity."""
    rng = np.random.default_rng(300 + F)
    S = rng.uniform(0, 1, (5000, 3))
    T = rng.uniform(0, 1, (4000, 3))
    g = rng.normal(size=(F, len(T)))
    f = rng.normal(size=(F, len(S)))
    plan = pacman.rbf.Plan(3, 0x11, S, T, 60, 0.15, -1.0)
    ref = plan.apply_transpose(g)
    plan.build_operator()
    got = plan.apply_transpose(g)
    assert _rel_max(got, ref) < 1e-10
    lhs = np.sum(g * plan.apply(f))
    rhs = np.sum(got * f)
    assert abs(lhs - rhs) <= 1e-11 * (abs(lhs) + abs(rhs))


@pytest.mark.parametrize("rbf,F", [(0x11, 1), (0x12, 5), (0x11, 20)])
def test_operator_fp32_close_to_solve(rbf, F):
    rng = np.random.default_rng(500 + rbf + F)
    S = rng.uniform(0, 1, (6000, 3))
    T = rng.uniform(0, 1, (5000, 3))
    V = np.stack([np.sin((1 + i) * S[:, 0] + 0.2 * i) + S[:, -1] ** 2
                  for i in range(F)])
    plan = pacman.rbf.Plan(3, rbf, S, T, 60, 0.15, -1.0)
    ref = plan.apply(V)
    plan.build_operator()
    b64 = plan.operator_bytes
    plan.build_operator(precision="fp32")
    assert plan.operator_bytes * 2 == b64
    got = plan.apply(V)
    err = _rel_max(got, ref)
    print(f"fp32 operator rbf={hex(rbf)} F={F} max rel {err:.2e}")
    assert 0 < err < 1e-5, err
    # transpose through the FP32 operator
    g = rng.normal(size=(min(F, 8), len(T)))
    plan.drop_operator()
    tref = plan.apply_transpose(g)
    plan.build_operator(precision="fp32")
    terr = _rel_max(plan.apply_transpose(g), tref)
    assert terr < 1e-5, terr
    with pytest.raises(ValueError):
        plan.build_operator(precision="fp16")


@pytest.mark.parametrize("F,fp32", [(1, False), (6, False), (24, False),
                                    (24, True)])
def test_operator_apply_matches_oracle(F, fp32, oracle_mod):
    """The stored-operator apply (dot path up to 4 fields, FP64 MFMA
    contraction from 5) compared with the oracle directly."""
    rng = np.random.default_rng(300 + F)
    S = rng.uniform(0, 1, (5000, 3))
    T = rng.uniform(0, 1, (4500, 3))
    V = W.trig_fields(S, F, seed=F)
    ref = oracle_mod.Setup(3, 0x11, S, T, 60, 0.15, -1.0).apply(V)
    plan = pacman.rbf.Plan(3, 0x11, S, T, 60, 0.15, -1.0)
    plan.build_operator(precision="fp32" if fp32 else "fp64")
    got = plan.apply(V)
    err = _rel_max(got, ref)
    tol = 1e-5 if fp32 else 1e-10
    print(f"operator F={F} fp32={fp32}: max rel vs oracle {err:.2e} "
          f"(tolerance {tol:g})")
    assert err < tol
