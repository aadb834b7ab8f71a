"""HIP-graph replay of the apply: every replay equals the eager apply.

The graph is replayed several times with eager work (other applies, torch
ops, synchronisations) between the replays; new contents written into the
captured input buffer are picked up by the next replay.
"""

import numpy as np
import pytest

import pacman_amd as pacman
from pacman_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nfields", [1, 3, 10])
def test_graph_replay_matches_eager(nfields):
    import torch
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(31)
    S = rng.uniform(0, 1, (6000, 3))
    T = rng.uniform(0, 1, (5000, 3))
    V = W.trig_fields(S, nfields, seed=5)
    V2 = W.trig_fields(S, nfields, seed=6)
    dS, dT = torch.from_numpy(S).to(dev), torch.from_numpy(T).to(dev)
    dV, dV2 = torch.from_numpy(V).to(dev), torch.from_numpy(V2).to(dev)
    plan = pacman.rbf.Plan(3, 0x11, dS, dT, 60, 0.15, -1.0)
    eager = plan.apply(dV).clone()
    eager2 = plan.apply(dV2).clone()
    torch.cuda.synchronize()
    out = torch.full_like(eager, float("nan"))
    for rep in range(4):
        plan.apply(dV, out=out, graph=True)
        torch.cuda.synchronize()
        assert torch.equal(out, eager), rep
        other = plan.apply(dV2)               # eager work between replays
        assert torch.equal(other, eager2)
        _ = float((dV * 2.0).sum())
        torch.cuda.synchronize()
    dV.copy_(dV2)                             # new contents, same buffer
    plan.apply(dV, out=out, graph=True)
    torch.cuda.synchronize()
    assert torch.equal(out, eager2)


def test_device_cache_reuse_gives_identical_results():
    """Plans built after others were destroyed reuse their cached device
    blocks (stale contents) and must give bitwise the same transfer."""
    import gc
    from pacman_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(31)
    S = rng.uniform(0, 1, (30000, 3))
    T = rng.uniform(0, 1, (20000, 3))
    f = rng.normal(size=(2, len(S)))
    first = pacman.rbf.Plan(3, 0x11, S, T, 80, 0.15, -1.0).apply(f)
    gc.collect()
    assert lib.pacman_device_cache_bytes() > 0
    for _ in range(3):
        again = pacman.rbf.Plan(3, 0x11, S, T, 80, 0.15, -1.0).apply(f)
        assert np.array_equal(first, again)
    gc.collect()
    assert lib.pacman_release_device_cache() == 0
    assert lib.pacman_device_cache_bytes() == 0
    fresh = pacman.rbf.Plan(3, 0x11, S, T, 80, 0.15, -1.0).apply(f)
    assert np.array_equal(first, fresh)
