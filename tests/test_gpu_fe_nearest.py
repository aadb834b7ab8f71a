"""GPU parity of the nearest-source transfer (FE method NEAREST_NEAREST).

The HIP grid search must pick exactly the oracle's source for every target
(same squared distances, ties to the lowest source index), so values and
status are compared bit for bit. Cases: 1-D/2-D/3-D random clouds with
duplicated sources (exact ties), targets outside the source hull, a target
cloud equal to the source cloud (the reference's same-mesh FE check,
tests/test_fe.py:73-102, copies the values), clustered sources, no sources.
"""

import numpy as np
import pytest

import pacman_amd as pacman

pytestmark = pytest.mark.gpu

NN = pacman.fe.methods.NEAREST_NEAREST


def run(dim, S, v, T):
    return pacman.fe.interpolate(dim, pacman.execspaces.HIP, NN, S, v, None,
                                 None, None, T)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_nearest_bit_exact_random(oracle_mod, dim):
    rng = np.random.default_rng(100 + dim)
    S = rng.uniform(0, 1, (5000, dim))
    S[2500:2600] = S[:100]                       # exact ties
    T = np.concatenate([rng.uniform(-0.5, 1.5, (4000, dim)), S[:200]])
    v = rng.normal(size=len(S))
    out, st = run(dim, S, v, T)
    ref, rst = oracle_mod.fe_nearest(dim, S, v, T)
    np.testing.assert_array_equal(out, ref)
    np.testing.assert_array_equal(st, rst)
    assert not st.any()


def test_nearest_same_mesh_copies_values(oracle_mod):
    rng = np.random.default_rng(5)
    S = rng.uniform(0, 1, (20000, 3))
    v = rng.normal(size=len(S))
    out, _ = run(3, S, v, S)
    np.testing.assert_array_equal(out, v)


def test_nearest_clustered_and_far_targets(oracle_mod):
    rng = np.random.default_rng(9)
    S = np.concatenate([rng.normal(0.2, 0.01, (3000, 3)),
                        rng.normal(0.8, 0.05, (3000, 3))])
    T = np.concatenate([rng.uniform(-2, 3, (3000, 3)),
                        rng.normal(0.2, 0.02, (1000, 3))])
    v = rng.normal(size=len(S))
    out, _ = run(3, S, v, T)
    ref, _ = oracle_mod.fe_nearest(3, S, v, T)
    np.testing.assert_array_equal(out, ref)


def test_nearest_no_sources_gives_zero():
    T = np.random.default_rng(1).uniform(0, 1, (100, 3))
    out, st = run(3, np.zeros((0, 3)), np.zeros(0), T)
    assert not out.any() and not st.any()


def test_nearest_device_tensors_match_host(oracle_mod):
    import torch
    rng = np.random.default_rng(21)
    S = rng.uniform(0, 1, (8000, 3))
    T = rng.uniform(0, 1, (6000, 3))
    v = rng.normal(size=len(S))
    dev = torch.device("cuda:0")
    out, st = pacman.fe.nearest_device(
        3, torch.from_numpy(S).to(dev), torch.from_numpy(v).to(dev),
        torch.from_numpy(T).to(dev))
    ref, _ = oracle_mod.fe_nearest(3, S, v, T)
    np.testing.assert_array_equal(out.cpu().numpy(), ref)
    assert not st.cpu().numpy().any()
