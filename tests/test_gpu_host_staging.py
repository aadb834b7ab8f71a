"""Host-array entry points above the staging threshold.

Host arrays larger than 1 MiB reach the device through the library's pinned
staging buffers: 8 lanes, each moving a contiguous range in 4 MiB chunks
through two buffers of its own. These cases use array sizes that are not
multiples of the chunk or lane ranges, and compare every host-array surface
bit for bit with the same call on device tensors (which involves no host
copy at all): plan creation, apply with several fields, the transpose, the
drop-in transfer (whose value copy runs beside the setup) and the MLS plan.
"""

import numpy as np
import pytest

import pacman_amd as pacman
from pacman_amd import workloads as W

pytestmark = pytest.mark.gpu

NS, NT = 400_003, 350_001   # 9.6 MB / 8.4 MB of coordinates


def _clouds(seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(0, 1, (NS, 3)), rng.uniform(0, 1, (NT, 3))


def test_plan_host_arrays_match_device_tensors():
    torch = pytest.importorskip("torch")
    S, T = _clouds(31)
    V = np.stack([W.franke_3d(*S.T), np.sin(3 * S[:, 0]), S[:, 1] * S[:, 2]])
    host = pacman.rbf.Plan(3, 0x11, S, T, 60, 0.15, -1.0)
    out_h = host.apply(V)
    dS, dT = torch.from_numpy(S).cuda(), torch.from_numpy(T).cuda()
    dev = pacman.rbf.Plan(3, 0x11, dS, dT, 60, 0.15, -1.0)
    out_d = dev.apply(torch.from_numpy(V).cuda())
    torch.cuda.synchronize()
    assert out_h.shape == (3, NT)
    np.testing.assert_array_equal(out_d.cpu().numpy(), out_h)
    # transpose: host target gradients in, host source gradients out
    G = np.stack([np.cos(2 * T[:, 2]), T[:, 0], np.ones(NT)])
    tr_h = host.apply_transpose(G)
    tr_d = dev.apply_transpose(torch.from_numpy(G).cuda())
    torch.cuda.synchronize()
    assert tr_h.shape == (3, NS)
    np.testing.assert_array_equal(tr_d.cpu().numpy(), tr_h)


def test_drop_in_transfer_matches_device_plan():
    """pacman.rbf.interpolate on host arrays (values copied beside the
    setup) equals a device-tensor plan with the drop-in's fixed parameters
    (50 nodes per cluster, overlap 0.15, support 0.1)."""
    torch = pytest.importorskip("torch")
    S, T = _clouds(32)
    v = W.franke_3d(*S.T)
    out_h = pacman.rbf.interpolate(3, 4, 0x11, S, v, T)
    plan = pacman.rbf.Plan(3, 0x11, torch.from_numpy(S).cuda(),
                           torch.from_numpy(T).cuda(), 50, 0.15, 0.1)
    out_d = plan.apply(torch.from_numpy(v).cuda())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out_d.cpu().numpy(), out_h)


def test_mls_plan_host_arrays_match_device_tensors():
    torch = pytest.importorskip("torch")
    S, T = _clouds(33)
    v = W.franke_3d(*S.T)
    out_h = pacman.mls.Plan(3, S, T).apply(v)
    dplan = pacman.mls.Plan(3, torch.from_numpy(S).cuda(),
                            torch.from_numpy(T).cuda())
    out_d = dplan.apply(torch.from_numpy(v).cuda())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out_d.cpu().numpy(), out_h)
