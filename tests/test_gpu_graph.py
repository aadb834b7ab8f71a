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


_NO_CACHE_SCRIPT = r"""
import sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import pacman_amd as pacman
from pacman_amd import _lib
rng = np.random.default_rng(31)
S = rng.uniform(0, 1, (30000, 3))
T = rng.uniform(0, 1, (20000, 3))
f = rng.normal(size=(2, len(S)))
outs = []
for _ in range(3):  # plans built and destroyed back to back
    outs.append(pacman.rbf.Plan(3, 0x11, S, T, 80, 0.15, -1.0).apply(f))
    assert _lib.load().pacman_device_cache_bytes() == 0
assert all(np.array_equal(outs[0], o) for o in outs)
np.save(sys.argv[2], outs[0])
print("no-cache ok", flush=True)
"""


def test_setup_and_apply_without_device_cache(tmp_path):
    """With PACMAN_AMD_NO_MEMORY_CACHE=1 set before the GPU is initialised
    (a fresh process), every temporary goes back to hipFree as soon as it is
    released, while kernels may still read it: hipFree waits for the device,
    so the transfer must equal the cached run's."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = str(tmp_path / "nocache.npy")
    env = dict(os.environ, PACMAN_AMD_NO_MEMORY_CACHE="1")
    r = subprocess.run([sys.executable, "-c", _NO_CACHE_SCRIPT, root, out],
                       env=env, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    rng = np.random.default_rng(31)
    S = rng.uniform(0, 1, (30000, 3))
    T = rng.uniform(0, 1, (20000, 3))
    f = rng.normal(size=(2, len(S)))
    cached = pacman.rbf.Plan(3, 0x11, S, T, 80, 0.15, -1.0).apply(f)
    assert np.array_equal(np.load(out), cached)


def test_graph_recaptured_after_buffers_grow():
    """An eager apply with more fields reallocates the plan's scratch buffers;
    the next graph apply with the old field count must capture again (a
    replay of the first graph would write through freed pointers)."""
    import torch
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(32)
    S = rng.uniform(0, 1, (5000, 3))
    T = rng.uniform(0, 1, (4000, 3))
    V1 = W.trig_fields(S, 1, seed=2)
    V8 = W.trig_fields(S, 8, seed=3)
    dS, dT = torch.from_numpy(S).to(dev), torch.from_numpy(T).to(dev)
    dV1, dV8 = torch.from_numpy(V1).to(dev), torch.from_numpy(V8).to(dev)
    plan = pacman.rbf.Plan(3, 0x11, dS, dT, 60, 0.15, -1.0)
    eager1 = plan.apply(dV1).clone()
    out = torch.empty_like(eager1)
    plan.apply(dV1, out=out, graph=True)          # captured at F = 1
    torch.cuda.synchronize()
    assert torch.equal(out, eager1)
    eager8 = plan.apply(dV8).clone()              # grows the buffers
    out.fill_(float("nan"))
    plan.apply(dV1, out=out, graph=True)          # must capture again
    torch.cuda.synchronize()
    assert torch.equal(out, eager1)
    assert torch.equal(plan.apply(dV8), eager8)
    plan.apply_transpose(torch.ones_like(eager1))  # adjoint buffers
    plan.apply(dV1, out=out, graph=True)
    torch.cuda.synchronize()
    assert torch.equal(out, eager1)


@pytest.mark.parametrize("nfields", [1, 12])
def test_graph_replay_matches_oracle(nfields, oracle_mod):
    """Replays compared with the oracle itself (not only the eager apply)."""
    import torch
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(33)
    S = rng.uniform(0, 1, (5000, 3))
    T = rng.uniform(0, 1, (4000, 3))
    V = W.trig_fields(S, nfields, seed=4)
    ref = oracle_mod.Setup(3, 0x11, S, T, 60, 0.15, -1.0).apply(V)
    plan = pacman.rbf.Plan(3, 0x11, torch.from_numpy(S).to(dev),
                           torch.from_numpy(T).to(dev), 60, 0.15, -1.0)
    dV = torch.from_numpy(V).to(dev)
    out = torch.empty((nfields, len(T)), dtype=torch.float64, device=dev)
    for rep in range(3):
        out.fill_(float("nan"))
        plan.apply(dV, out=out, graph=True)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        print(f"graph replay {rep} F={nfields}: rel-L2 vs oracle {err:.2e}")
        assert err < 1e-10
