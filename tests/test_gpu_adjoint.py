"""Adjoint (transpose apply) of the RBF-PUM transfer against independent pins.

* the transpose of the oracle's dense operator: Op is built column by column
  by applying the CPU oracle to unit vectors, and the GPU transpose apply
  must equal Op^T g (FP64 tolerance 1e-10, support = patch radius);
* the dot-product identity <g, Op f> = <Op^T g, f> with the GPU forward and
  transpose;
* torch autograd: d/df of 0.5 ||Op f - g||^2 against central finite
  differences (step 1e-6) on 200 random entries, relative 1e-6.
"""

import numpy as np
import pytest

import pacman_amd as pacman

pytestmark = pytest.mark.gpu


def cloud(seed, n, m, dim=3):
    rng = np.random.default_rng(seed)
    return rng.uniform(0, 1, (n, dim)), rng.uniform(0, 1, (m, dim))


# npc = 8: patches of about ten sources (one or two 16-row tile blocks), where
# the transpose chain's right-hand-side staging is the largest user of the
# solve kernel's scratch buffer
@pytest.mark.parametrize("rbf,dim,npc", [(0x11, 3, 40), (0x13, 3, 40),
                                         (0x11, 2, 40), (0x11, 3, 8)])
def test_transpose_equals_oracle_operator_transpose(oracle_mod, rbf, dim,
                                                    npc):
    S, T = cloud(3 + dim, 400, 300, dim)
    ora = oracle_mod.Setup(dim, rbf, S, T, npc, 0.15, -1.0)
    Op = ora.apply(np.eye(len(S)))            # (N, M): row j = Op e_j
    rng = np.random.default_rng(11)
    g = rng.normal(size=(2, len(T)))
    ref = g @ Op.T                            # Op^T g per field, (2, N)
    plan = pacman.rbf.Plan(dim, rbf, S, T, npc, 0.15, -1.0)
    got = plan.apply_transpose(g)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err < 1e-10, err


def test_dot_product_identity_large():
    S, T = cloud(17, 60000, 50000)
    rng = np.random.default_rng(5)
    f = rng.normal(size=(3, len(S)))
    g = rng.normal(size=(3, len(T)))
    plan = pacman.rbf.Plan(3, 0x11, S, T, 100, 0.15, -1.0)
    lhs = np.sum(g * plan.apply(f))
    rhs = np.sum(plan.apply_transpose(g) * f)
    assert abs(lhs - rhs) <= 1e-11 * (abs(lhs) + abs(rhs)), (lhs, rhs)


def test_autograd_matches_finite_differences():
    import torch
    dev = torch.device("cuda:0")
    S, T = cloud(23, 20000, 15000)
    rng = np.random.default_rng(8)
    g = torch.from_numpy(rng.normal(size=len(T))).to(dev)
    plan = pacman.rbf.Plan(3, 0x11, torch.from_numpy(S).to(dev),
                           torch.from_numpy(T).to(dev), 100, 0.15, -1.0)
    f = torch.from_numpy(rng.normal(size=len(S))).to(dev).requires_grad_()
    u = pacman.autograd.transfer(plan, f)
    loss = 0.5 * ((u - g) ** 2).sum()
    loss.backward()
    grad = f.grad.detach().cpu().numpy()

    def L(x):
        with torch.no_grad():
            return float(0.5 * ((plan.apply(x) - g) ** 2).sum())

    h = 1e-6
    idx = rng.choice(len(S), 200, replace=False)
    base = f.detach().clone()
    fd = np.empty(len(idx))
    for i, j in enumerate(idx):
        xp = base.clone()
        xp[j] += h
        xm = base.clone()
        xm[j] -= h
        fd[i] = (L(xp) - L(xm)) / (2 * h)
    err = np.abs(fd - grad[idx]).max() / np.abs(grad[idx]).max()
    assert err < 1e-6, err
