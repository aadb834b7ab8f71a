"""Seeded synthetic point clouds and fields for tests and benchmarks.

Recipes: numpy ``default_rng(seed)`` (PCG64). A jittered grid is
``linspace(0, 1, g)`` per axis in row-major order (last axis fastest) plus
``U(-0.3 h, 0.3 h)`` per coordinate with ``h = 1 / (g - 1)``; targets are
``U(0, 1)^3``. ``franke_3d`` is the reference's analytic test field
(tests/franke_functions.py:118-165).
"""

from __future__ import annotations

import hashlib

import numpy as np


def jittered_grid(g: int, seed: int, jitter: float = 0.3) -> np.ndarray:
    ax = np.linspace(0.0, 1.0, g)
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    h = 1.0 / (g - 1)
    rng = np.random.default_rng(seed)
    P += rng.uniform(-jitter * h, jitter * h, size=P.shape)
    return np.ascontiguousarray(P)


def regular_grid(g: int) -> np.ndarray:
    """g^3 grid on [0,1]^3 in the order of the reference's C++ test
    (tests/test_cpp_rbf_interface.cpp:62-83: x slowest, z fastest)."""
    step = 1.0 / (g - 1)
    i = np.arange(g) * step
    X, Y, Z = np.meshgrid(i, i, i, indexing="ij")
    return np.ascontiguousarray(np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1))


def uniform(m: int, seed: int, dim: int = 3) -> np.ndarray:
    return np.random.default_rng(seed).uniform(0.0, 1.0, size=(m, dim))


def franke_3d(x, y, z):
    return (0.75 * np.exp(-((9 * x - 2) ** 2 + (9 * y - 2) ** 2
                            + (9 * z - 2) ** 2) / 4.0)
            + 0.75 * np.exp(-((9 * x + 1) ** 2 / 49.0 + (9 * y + 1) / 10.0
                              + (9 * z + 1) / 10.0))
            + 0.5 * np.exp(-((9 * x - 7) ** 2 + (9 * y - 3) ** 2
                             + (9 * z - 5) ** 2) / 4.0)
            - 0.2 * np.exp(-((9 * x - 4) ** 2 + (9 * y - 7) ** 2
                             + (9 * z - 5) ** 2)))


def trig_fields(P: np.ndarray, nfields: int, seed: int) -> np.ndarray:
    """sin(k . x + phi) with integer k in [1, 6]^3 and phi in [0, 2 pi)."""
    rng = np.random.default_rng(seed)
    k = rng.integers(1, 7, size=(nfields, 3)).astype(np.float64)
    phi = rng.uniform(0.0, 2 * np.pi, size=nfields)
    return np.sin(k @ P.T + phi[:, None])


def sha256(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


# Named workloads: (sources, targets, nodes per cluster, support radius).
# support <= 0 means "the patch radius" (patches hold about npc points and
# the Wendland support equals the patch radius).
def config(name: str):
    if name == "c1":      # 36k jittered -> 50k uniform, npc 100
        return jittered_grid(33, 0), uniform(50_000, 1), 100, -1.0
    if name == "c2":      # 4.1M jittered -> 4.1M uniform, npc 150
        return jittered_grid(160, 2), uniform(4_096_000, 3), 150, -1.0
    if name == "c4":      # 1M jittered -> 1M uniform, npc 100
        return jittered_grid(100, 6), uniform(1_000_000, 7), 100, -1.0
    if name == "c5":      # 8M jittered -> 8M uniform, npc 150
        return jittered_grid(200, 9), uniform(8_000_000, 10), 150, -1.0
    raise KeyError(name)


def gaussian_mixture(n: int, seed: int, components: int = 12,
                     sigma=(0.02, 0.15)) -> np.ndarray:
    """n points from a Gaussian mixture in [0, 1]^3: means U[0, 1]^3,
    isotropic sigma ~ U[sigma], weights ~ Dirichlet(1); points outside the
    cube are drawn again (component and position) until inside.
    Draw order: means, sigmas, weights, then rounds of (components,
    normals) for the points still outstanding."""
    rng = np.random.default_rng(seed)
    means = rng.uniform(0.0, 1.0, (components, 3))
    sig = rng.uniform(sigma[0], sigma[1], components)
    w = rng.dirichlet(np.ones(components))
    out = np.empty((n, 3))
    todo = np.arange(n)
    while len(todo):
        c = rng.choice(components, size=len(todo), p=w)
        p = means[c] + sig[c][:, None] * rng.normal(size=(len(todo), 3))
        ok = np.all((p >= 0.0) & (p <= 1.0), axis=1)
        out[todo[ok]] = p[ok]
        todo = todo[~ok]
    return np.ascontiguousarray(out)


def mixture_fields(P: np.ndarray, nfields: int = 8, seed: int = 12):
    """franke_3d followed by nfields - 1 seeded sin(k . x + phi) fields."""
    F = np.empty((nfields, len(P)))
    F[0] = franke_3d(*P.T)
    if nfields > 1:
        F[1:] = trig_fields(P, nfields - 1, seed)
    return F


# The uneven-cloud workload: 4.1M mixture sources (seed 4) -> 4.1M uniform
# targets (seed 5); patches from the adaptive builder.
def mixture_config(n: int = 4_096_000):
    return gaussian_mixture(n, 4), uniform(n, 5)
