"""Gradient of the transfer with respect to the target coordinates.

* against the oracle's analytic gradient (itself pinned by central
  differences of the oracle's forward in tests/test_oracle.py): relative
  1e-9 for the five Wendland bases, 2-D and 3-D, 1, 3 and 10 fields (10
  exercises the accumulation over two solve launches);
* torch autograd through pacman_amd.autograd.transfer: d/dy of
  0.5 ||u - g||^2 against central differences (step 1e-6, a new plan per
  perturbed target set) on 200 random entries, relative 1e-6, together with
  the source-value gradient of the same backward.
"""

import numpy as np
import pytest

import pacman_amd as pacman

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


@pytest.mark.parametrize("rbf,dim,F", [(0x10, 3, 1), (0x11, 3, 3),
                                       (0x12, 3, 1), (0x13, 2, 2),
                                       (0x14, 3, 1), (0x11, 3, 10)])
def test_target_grad_equals_oracle(oracle_mod, rbf, dim, F):
    rng = np.random.default_rng(70 + rbf + dim + F)
    S = rng.uniform(0, 1, (3000, dim))
    T = rng.uniform(0, 1, (2500, dim))
    f = np.stack([np.sin((1 + i) * S[:, 0] + 0.3 * i) + S[:, -1] * i
                  for i in range(F)])
    g = rng.normal(size=(F, len(T)))
    ref = oracle_mod.Setup(dim, rbf, S, T, 60, 0.15, -1.0).target_grad(f, g)
    plan = pacman.rbf.Plan(dim, rbf, _dev(S), _dev(T), 60, 0.15, -1.0)
    u, state = plan.apply_with_state(_dev(f))
    got = plan.target_coord_grad(state, _dev(g)).cpu().numpy()
    assert got.shape == (len(T), dim)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err < 1e-9, err
    # the saved-state apply returns the plain apply's values
    plain = plan.apply(_dev(f)).cpu().numpy()
    assert np.array_equal(u.cpu().numpy(), plain)


def test_autograd_target_and_source_grads_match_finite_differences():
    import torch
    rng = np.random.default_rng(8)
    S = np.random.default_rng(6).uniform(0, 1, (20000, 3))
    T = np.random.default_rng(7).uniform(0, 1, (15000, 3))
    g = _dev(rng.normal(size=len(T)))
    from pacman_amd import workloads
    f = _dev(workloads.franke_3d(*S.T)).requires_grad_()
    dS = _dev(S)
    tgt = _dev(T).requires_grad_()
    plan = pacman.rbf.Plan(3, 0x11, dS, tgt.detach(), 100, 0.15, -1.0)
    u = pacman.autograd.transfer(plan, f, tgt)
    loss = 0.5 * ((u - g) ** 2).sum()
    loss.backward()
    gT = tgt.grad.cpu().numpy()
    gF = f.grad.cpu().numpy()
    assert np.array_equal(gF, plan.apply_transpose(u.detach() - g).cpu()
                          .numpy())

    gh = g.cpu().numpy()

    def L(Tp, t):
        # only target t moved: the other terms of the loss are unchanged
        with torch.no_grad():
            p = pacman.rbf.Plan(3, 0x11, dS, _dev(Tp), 100, 0.15, -1.0)
            assert p.num_patches == plan.num_patches
            ut = float(p.apply(f.detach())[t])
        return 0.5 * (ut - gh[t]) ** 2

    h = 1e-6
    picks = rng.choice(T.size, 200, replace=False)
    fd = np.empty(len(picks))
    for i, e in enumerate(picks):
        t, a = divmod(int(e), 3)
        Tp = T.copy()
        Tp[t, a] += h
        Tm = T.copy()
        Tm[t, a] -= h
        fd[i] = (L(Tp, t) - L(Tm, t)) / (2 * h)
    ref = gT.reshape(-1)[picks]
    err = np.abs(fd - ref).max() / np.abs(ref).max()
    assert err < 1e-6, err
