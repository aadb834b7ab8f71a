"""Pins for the CPU FP64 oracle (runs without a GPU).

Each oracle stage is checked against something it does not share code with:
hand-computed values, textbook library routines (numpy / scipy), brute
force in numpy, mathematical invariants of the transfer, and the
reference's own known-answer tests (tests/test_cpp_rbf_interface.cpp and
tests/test_rbf-pum.py of the reference).
"""

import os

import numpy as np
import pytest
import scipy.linalg

from pacman_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MESHES = os.path.join(ROOT, "tests", "golden", "meshes")
BASES = [0x10, 0x11, 0x12, 0x13, 0x14]


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


# ---- Wendland kernels (rbf_functions.hpp:22-150) ---------------------------

def test_wendland_hand_values(oracle_mod):
    w = oracle_mod.wendland
    # phi(0): 1, 1, 3, 1, 15 (the kernels are not normalised)
    assert [w(f, 0.0, 1.0) for f in BASES] == [1.0, 1.0, 3.0, 1.0, 15.0]
    # p = 1/2, computed by hand from the closed forms
    half = [0.25, 0.1875, 0.32421875, 0.0595703125, 0.48504638671875]
    for f, v in zip(BASES, half):
        assert w(f, 0.5, 1.0) == pytest.approx(v, rel=1e-15)
    # radiusInv scales the argument: phi(r; 1/delta) = phi(r / delta; 1)
    for f in BASES:
        assert w(f, 0.03, 1 / 0.1) == pytest.approx(w(f, 0.3, 1.0), rel=1e-14)
    # support mask: zero for p >= 1 and for negative r
    for f in BASES:
        assert w(f, 1.0, 1.0) == 0.0
        assert w(f, 1.5, 1.0) == 0.0
        assert w(f, -0.1, 1.0) == 0.0


@pytest.mark.parametrize("f,order", [(0x11, 2), (0x12, 4), (0x13, 6),
                                     (0x14, 8)])
def test_wendland_smoothness_at_support(f, order, oracle_mod):
    """C^{2k} kernels vanish at p = 1 to order 2k + 1: phi(1 - e) = O(e^{2k+2})
    for the C2..C8 members ((1-p)^{l} factors with l = 4, 6, 8, 10)."""
    w = oracle_mod.wendland
    e1, e2 = 1e-2, 5e-3
    ratio = w(f, 1 - e1, 1.0) / w(f, 1 - e2, 1.0)
    assert np.log2(ratio) == pytest.approx(order + 2, abs=0.1)


@pytest.mark.parametrize("f", BASES)
def test_wendland_positive_definite_in_3d(f, oracle_mod):
    rng = np.random.default_rng(1)
    P = rng.uniform(0, 1, (60, 3))
    D = np.sqrt(((P[:, None, :] - P[None, :, :]) ** 2).sum(-1))
    A = np.vectorize(lambda r: oracle_mod.wendland(f, r, 1 / 0.6))(D)
    assert np.linalg.eigvalsh(A).min() > 0


# ---- dense per-patch algebra vs textbook library routines -------------------

def test_lu_solve_matches_scipy(oracle_mod):
    rng = np.random.default_rng(2)
    for n in (1, 5, 37, 120):
        B = rng.normal(size=(n, n))
        A = B @ B.T + n * np.eye(n)
        b = rng.normal(size=n)
        x = oracle_mod.lu_solve(A, b)
        ref = scipy.linalg.solve(A, b)
        assert np.linalg.norm(x - ref) <= 1e-12 * np.linalg.norm(ref)


def test_lstsq_matches_numpy(oracle_mod):
    rng = np.random.default_rng(3)
    for n, p in ((10, 4), (150, 4), (150, 3), (7, 2), (3, 1)):
        Q = np.hstack([np.ones((n, 1)), rng.uniform(0.4, 0.6, (n, p - 1))])
        b = rng.normal(size=n)
        beta = oracle_mod.lstsq(Q, b)
        ref = np.linalg.lstsq(Q, b, rcond=None)[0]
        assert np.allclose(beta, ref, rtol=1e-9, atol=1e-9)


def test_singular_values_match_numpy(oracle_mod):
    rng = np.random.default_rng(4)
    for n, p in ((50, 4), (6, 3), (200, 2)):
        Q = rng.normal(size=(n, p))
        Q[:, -1] = Q[:, 0] * 1e-7 + Q[:, -1] * 1e-9  # ill-conditioned column
        s = oracle_mod.singular_values(Q)
        ref = np.linalg.svd(Q, compute_uv=False)
        assert np.allclose(s, ref, rtol=1e-8, atol=1e-14 * ref[0])


# ---- FindRadius, lattice, tag/snap/prune, membership -------------------------

def _sq(a, b):
    d = b - a
    return (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]


def pad3(P):
    out = np.zeros((len(P), 3))
    out[:, : P.shape[1]] = P
    return out


def find_radius_numpy(S, dim, npc):
    """clusters_radius.hpp:22-125 recomputed by brute force in numpy."""
    S3 = pad3(S)
    lo, hi = np.zeros(3), np.zeros(3)
    lo[:dim], hi[:dim] = S.min(0), S.max(0)
    c = np.zeros(3)
    c[:dim] = (lo[:dim] + hi[:dim]) / 2.0
    probes = [c.copy()]
    minus, plus = [], []
    for ax in range(dim):
        s = c.copy()
        ln = abs(hi[ax] - lo[ax])
        s[ax] -= ln * 0.25
        minus.append(s.copy())
        s[ax] += ln * 0.5
        plus.append(s.copy())
    probes += minus + plus
    k = min(npc, len(S))
    m = []
    for q in probes:
        nn = int(np.argmin(_sq(q, S3)))   # first index among ties
        d2 = np.sort(_sq(S3[nn], S3))
        m.append(d2[k - 1])
    m.sort()
    return np.sqrt(m[dim]), lo, hi


@pytest.mark.parametrize("dim,n,npc", [(3, 1500, 30), (2, 800, 20),
                                       (1, 300, 10), (3, 8000, 100)])
def test_find_radius_and_lattice_bruteforce(dim, n, npc, oracle_mod):
    rng = np.random.default_rng(dim * 100 + n)
    S = rng.uniform(0, 1, (n, dim))
    T = rng.uniform(0, 1, (n // 2, dim))
    o = oracle_mod.Setup(dim, 0x11, S, T, npc, 0.15, -1.0)
    r, lo, hi = find_radius_numpy(S, dim, npc)
    assert o.radius == r
    # lattice size: prod(ceil(max(1, edge / spacing)) + 1)
    spacing = np.sqrt(4.0 / dim) * r * (1 - 0.15)
    nb = 1
    for a in range(dim):
        nb *= int(np.ceil(max(1.0, abs(hi[a] - lo[a]) / spacing))) + 1
    assert o.num_candidates == nb


def test_membership_bruteforce(oracle_mod):
    rng = np.random.default_rng(7)
    S = rng.uniform(0, 1, (2000, 3))
    T = rng.uniform(0, 1, (1500, 3))
    o = oracle_mod.Setup(3, 0x11, S, T, 40, 0.15, -1.0)
    r = o.radius
    C, cand = o.centers()
    assert np.all(np.diff(cand) > 0)
    # every centre is a source point
    key = {tuple(p) for p in S}
    assert all(tuple(c) in key for c in C)
    for which, P, rad in ((0, S, r), (1, T, r - 1e-14)):
        off, idx = o.csr(which)
        for k in range(len(C)):
            d = np.sqrt(_sq(C[k], P))
            np.testing.assert_array_equal(idx[off[k]:off[k + 1]],
                                          np.nonzero(d <= rad)[0])
    off, idx = o.csr(2)
    for t in range(len(T)):
        d = np.sqrt(_sq(T[t], C))
        np.testing.assert_array_equal(idx[off[t]:off[t + 1]],
                                      np.nonzero(d <= r)[0])
    # Shepard weights sum to one on every covered target
    w = o.weights()
    sums = np.add.reduceat(w, off[:-1][np.diff(off) > 0])
    assert np.allclose(sums, 1.0, rtol=0, atol=1e-14)


def _psi_c2(p):
    """Wendland C2, (1 - p)^4 (4 p + 1) on [0, 1), closed form."""
    q = np.where(p < 1.0, 1.0 - p, 0.0)
    return q ** 4 * (4.0 * p + 1.0)


def _weights_numpy(T, C, off, idx, scale, kernel):
    """Shepard weights of every (target, patch) entry of the per-target lists:
    kernel(|y - c| * (1 / scale)) / sum over the target's list."""
    out = np.zeros(len(idx))
    for t in range(len(T)):
        ks = idx[off[t]:off[t + 1]]
        if len(ks) == 0:
            continue
        p = np.sqrt(_sq(pad3(T[t:t + 1])[0], pad3(C[ks]))) * (1.0 / scale)
        v = kernel(p)
        out[off[t]:off[t + 1]] = v / v.sum()
    return out


@pytest.mark.parametrize("dim,rbf", [(3, 0x12), (2, 0x13), (3, 0x10)])
def test_shepard_weights_bruteforce(dim, rbf, oracle_mod):
    """PU weights (solver.hpp:128-169): w_k(y) = psi(|y - c_k| / r) / sum,
    psi = Wendland C2 scaled by the patch radius r (interpolator.hpp:86,91),
    whatever the interpolation basis and its support. The basis here is not
    C2 and the support (0.1) is not r, so using phi instead of psi, or the
    support instead of r, fails the comparison (checked below)."""
    rng = np.random.default_rng(40 + dim)
    S = rng.uniform(0, 1, (1500, dim))
    T = rng.uniform(0, 1, (900, dim))
    o = oracle_mod.Setup(dim, rbf, S, T, 30, 0.15, 0.1)
    C, _ = o.centers()
    C = C[:, :dim]
    r = o.radius
    off, idx = o.csr(2)
    w = o.weights()
    assert len(w) == len(idx) and len(idx) > 2 * len(T)
    ref = _weights_numpy(T, C, off, idx, r, _psi_c2)
    np.testing.assert_allclose(w, ref, rtol=1e-14, atol=1e-16)
    # planted mistakes: the basis phi instead of psi, the support instead of r
    phi = np.vectorize(lambda p: oracle_mod.wendland(rbf, p, 1.0))
    wrong_phi = _weights_numpy(T, C, off, idx, r, phi)
    wrong_delta = _weights_numpy(T, C, off, idx, 0.1,
                                 lambda p: _psi_c2(np.minimum(p, 1.0 - 1e-12)))
    for bad in (wrong_phi, wrong_delta):
        assert np.max(np.abs(bad - w)) > 1e-3


def create_clusters_numpy(S, T, dim, r, overlap=0.15, variant=None):
    """CreateClusters (clustering.hpp:137-270) restated by brute force:
    lattice from the source box (last axis fastest), kill candidates whose
    nearest source, then nearest target, is farther than r (both from the
    lattice position), snap survivors to their nearest source (lowest index
    among ties), the literal sequential prune of the host loop (:226-254),
    re-tag against the targets from the snapped position (:256-262), stable
    compaction. `variant` plants a step-order mistake (tests only)."""
    S3, T3 = pad3(S), pad3(T)
    lo, hi = S.min(0), S.max(0)
    spacing = np.sqrt(4.0 / dim) * r * (1.0 - overlap)
    shape, dist = [], []
    for a in range(dim):
        edge = abs(hi[a] - lo[a])
        shp = int(np.ceil(max(1.0, edge / spacing)))
        dist.append(edge / shp)
        shape.append(shp + 1)
    nb = int(np.prod(shape))
    coords = np.zeros((nb, dim), np.int64)
    rem = np.arange(nb)
    for a in range(dim - 1, -1, -1):
        coords[:, a] = rem % shape[a]
        rem //= shape[a]
    cand = np.zeros((nb, 3))
    for a in range(dim):
        cand[:, a] = lo[a] + dist[a] * coords[:, a].astype(np.float64)
    r2 = r * r
    alive = np.ones(nb, bool)

    def nearest(c, P3):
        d2 = _sq(c, P3)
        i = int(np.argmin(d2))        # first index among ties
        return i, d2[i]

    for L in range(nb):
        ls = cand[L].copy()
        i, d2 = nearest(ls, S3)
        if d2 > r2:
            alive[L] = False
            continue
        if variant == "tag_targets_after_snap":
            cand[L] = S3[i]
        _, d2t = nearest(cand[L], T3)
        if d2t > r2:
            alive[L] = False
            continue
        cand[L] = S3[i]
    thr = 0.4 * min(dist)
    thr *= thr
    offs = []
    for t in range(3 ** dim):
        o, tt = [0] * dim, t
        for d in range(dim):
            o[d] = tt % 3 - 1
            tt //= 3
        if any(o):
            idx, stride = 0, 1
            for d in range(dim - 1, -1, -1):
                idx += o[d] * stride
                stride *= shape[d]
            offs.append(idx)
    for i in range(nb):
        if not alive[i]:
            continue
        for off in offs:
            j = i + off
            if 0 <= j < nb and alive[j] and _sq(cand[i], cand[j]) < thr:
                alive[i] = False
                break
    if variant != "no_retag":
        for L in range(nb):
            if not alive[L]:
                continue
            pos = cand[L]
            if variant == "retag_from_lattice":
                pos = np.zeros(3)
                for a in range(dim):
                    pos[a] = lo[a] + dist[a] * float(coords[L, a])
            if nearest(pos, T3)[1] > r2:
                alive[L] = False
    keep = np.nonzero(alive)[0]
    return cand[keep], keep


def _cluster_clouds():
    rng = np.random.default_rng(77)
    g = np.linspace(0.0, 1.0, 6)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    grid = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    clus = np.vstack([rng.normal(c, 0.04, (120, 2))
                      for c in rng.uniform(0.2, 0.8, (3, 2))])
    return [
        ("1d", rng.uniform(0, 1, (70, 1)), rng.uniform(-0.1, 1.1, (50, 1)), 6),
        ("2d", rng.uniform(0, 1, (180, 2)), rng.uniform(0, 1, (90, 2)), 10),
        ("2d-clustered", clus, rng.uniform(0, 1, (150, 2)), 12),
        ("3d", rng.uniform(0, 1, (400, 3)), rng.uniform(0, 1, (250, 3)), 15),
        # a regular grid: nearest-source ties everywhere (lowest index wins)
        ("3d-grid", grid, grid[::2] + 0.05, 8),
        ("3d-sparse-targets", rng.uniform(0, 1, (500, 3)),
         rng.uniform(0.3, 0.7, (40, 3)), 12),
    ]


def test_create_clusters_bruteforce(oracle_mod):
    """Centres and candidate indices of the oracle equal the brute-force
    restatement exactly, on 1-D/2-D/3-D clouds including a regular grid;
    each planted step-order mistake changes the result on some cloud."""
    misses = {"tag_targets_after_snap": 0, "retag_from_lattice": 0,
              "no_retag": 0}
    for name, S, T, npc in _cluster_clouds():
        dim = S.shape[1]
        r, _, _ = find_radius_numpy(S, dim, npc)
        o = oracle_mod.Setup(dim, 0x11, S, T, npc, 0.15, -1.0)
        assert o.radius == r, name
        C, cand = o.centers()
        refC, refCand = create_clusters_numpy(S, T, dim, r)
        np.testing.assert_array_equal(cand, refCand, err_msg=name)
        np.testing.assert_array_equal(C, refC, err_msg=name)
        for v in misses:
            bc, bi = create_clusters_numpy(S, T, dim, r, variant=v)
            if len(bi) != len(cand) or np.any(bi != cand) or np.any(bc != C):
                misses[v] += 1
    assert all(m > 0 for m in misses.values()), misses


@pytest.mark.parametrize("seed", range(4))
def test_grid_queries_equal_bruteforce(seed, oracle_mod):
    rng = np.random.default_rng(seed)
    S = rng.uniform(0, 1, (3000, 3)) ** (1 + seed)   # uneven densities
    T = rng.uniform(0, 1, (2000, 3))
    a = oracle_mod.Setup(3, 0x11, S, T, 35, 0.15, -1.0)
    b = oracle_mod.Setup(3, 0x11, S, T, 35, 0.15, -1.0, brute=True)
    np.testing.assert_array_equal(a.centers()[0], b.centers()[0])
    for which in (0, 1, 2):
        for x, y in zip(a.csr(which), b.csr(which)):
            np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("cloud", ["random", "grid", "clustered", "plane"])
def test_parallel_prune_equals_sequential_loop(cloud, oracle_mod):
    """The reference prunes centres in a sequential host loop
    (clustering.hpp:230-254); the parallel rule must keep the same set."""
    rng = np.random.default_rng(12)
    if cloud == "random":
        S = rng.uniform(0, 1, (5000, 3))
    elif cloud == "grid":
        S = W.regular_grid(17)
    elif cloud == "clustered":
        S = np.vstack([rng.normal(c, 0.05, (1500, 3))
                       for c in rng.uniform(0, 1, (4, 3))])
    else:
        S = rng.uniform(0, 1, (4000, 3))
        S[:, 2] = 0.5 + 1e-9 * rng.normal(size=len(S))
    T = rng.uniform(0, 1, (3000, 3))
    for npc in (15, 60):
        a = oracle_mod.Setup(3, 0x11, S, T, npc, 0.15, -1.0)
        b = oracle_mod.Setup(3, 0x11, S, T, npc, 0.15, -1.0,
                             sequential_prune=True)
        np.testing.assert_array_equal(a.centers()[1], b.centers()[1])


# ---- whole operator ----------------------------------------------------------

def operator_numpy(S, T, vals, o, delta, basis):
    """u(y) = sum_k w_k(y) [E_k A_k^{-1} (b - Q_k beta) + V_k beta] with
    scipy/numpy dense algebra on the oracle's own patch lists."""
    C, _ = o.centers()
    soff, sidx = o.csr(0)
    toff, tidx = o.csr(1)
    woff, widx = o.csr(2)
    w = o.weights()
    phi = np.vectorize(lambda r: __import__("oracle").wendland(basis, r,
                                                               1 / delta))
    out = np.zeros(len(T))
    for k in range(len(C)):
        src = sidx[soff[k]:soff[k + 1]]
        tgt = tidx[toff[k]:toff[k + 1]]
        X, Y = S[src], T[tgt]
        # conditioning loop of serial_ops.hpp:108-139 with numpy's SVD
        active = [0, 1, 2]
        while len(active) > 0:
            Q = np.hstack([np.ones((len(X), 1)), X[:, active]])
            s = np.linalg.svd(Q, compute_uv=False)
            s = np.concatenate([s, np.zeros(Q.shape[1] - len(s))])
            if s[0] / max(s[-1], 1e-14) <= 1e5:
                break
            rng_ = np.ptp(X[:, active], axis=0)
            active.pop(int(np.argmin(rng_)))
        Q = np.hstack([np.ones((len(X), 1)), X[:, active]])
        V = np.hstack([np.ones((len(Y), 1)), Y[:, active]])
        b = vals[src]
        beta = np.linalg.lstsq(Q, b, rcond=None)[0]
        A = phi(np.sqrt(((X[:, None] - X[None]) ** 2).sum(-1)))
        alpha = scipy.linalg.solve(A, b - Q @ beta)
        E = phi(np.sqrt(((Y[:, None] - X[None]) ** 2).sum(-1)))
        t = E @ alpha + V @ beta
        for j, y in enumerate(tgt):
            lst = widx[woff[y]:woff[y + 1]]
            out[y] += w[woff[y] + int(np.nonzero(lst == k)[0][0])] * t[j]
    return out


def test_operator_matches_dense_numpy(oracle_mod):
    rng = np.random.default_rng(13)
    S = rng.uniform(0, 1, (1200, 3))
    T = rng.uniform(0, 1, (800, 3))
    v = W.franke_3d(*S.T)
    o = oracle_mod.Setup(3, 0x12, S, T, 30, 0.15, -1.0)
    ours = o.apply(v)
    ref = operator_numpy(S, T, v, o, o.radius, 0x12)
    assert rel_l2(ours, ref) < 1e-11


def test_linear_fields_reproduced(oracle_mod):
    rng = np.random.default_rng(14)
    S = rng.uniform(0, 1, (3000, 3))
    T = rng.uniform(0.1, 0.9, (1000, 3))
    f = lambda P: 0.3 - P[:, 0] + 2 * P[:, 1] + 4 * P[:, 2]
    out = oracle_mod.rbf_interpolate(3, 0x11, S, f(S), T, 50, 0.15, -1.0)
    assert np.max(np.abs(out - f(T))) < 1e-12


@pytest.mark.parametrize("rbf", BASES)
def test_reference_same_grid(rbf, oracle_mod):
    """tests/test_cpp_rbf_interface.cpp:165-198 of the reference: 20^3 grid,
    source = target, rel-L2 < 1e-8 for each basis."""
    P = W.regular_grid(20)
    v = W.franke_3d(*P.T)
    assert rel_l2(oracle_mod.rbf_interpolate(3, rbf, P, v, P), v) < 1e-8


CASES = {"coarse_to_fine_0": ("0.03.npz", "0.003.npz"),
         "fine_to_coarse_0": ("0.003.npz", "0.03.npz"),
         "same_0": ("0.03.npz", "0.03.npz")}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("rbf", BASES)
def test_reference_npz_fixtures(case, rbf, oracle_mod):
    """tests/test_rbf-pum.py:56-74,113-122 of the reference (< 1.5e-2)."""
    a, b = CASES[case]
    sp = np.load(os.path.join(MESHES, a))["points"][:, :3].astype(np.float64)
    tp = np.load(os.path.join(MESHES, b))["points"][:, :3].astype(np.float64)
    out = oracle_mod.rbf_interpolate(3, rbf, sp, W.franke_3d(*sp.T), tp)
    assert rel_l2(out, W.franke_3d(*tp.T)) < 1.5e-2


def test_fine_to_coarse_exercises_axis_drop(oracle_mod):
    sp = np.load(os.path.join(MESHES, "0.003.npz"))["points"][:, :3]
    tp = np.load(os.path.join(MESHES, "0.03.npz"))["points"][:, :3]
    o = oracle_mod.Setup(3, 0x11, sp.astype(float), tp.astype(float))
    _, pv = o.apply(W.franke_3d(*sp.T.astype(float)), return_pv=True)
    assert (pv < 4).any()


def test_coplanar_patch_drops_normal_axis(oracle_mod):
    """A flat point set must lose exactly its normal axis (z)."""
    rng = np.random.default_rng(15)
    S = rng.uniform(0, 1, (2000, 3))
    S[:, 2] = 0.25
    T = S[:300].copy()
    o = oracle_mod.Setup(3, 0x11, S, T, 40, 0.15, -1.0)
    out, pv = o.apply(S[:, 0] + S[:, 1], return_pv=True)
    assert np.all(pv == 3)
    assert np.allclose(out, T[:, 0] + T[:, 1], atol=1e-10)


# ---- FE NEAREST_NEAREST ------------------------------------------------------

@pytest.mark.parametrize("dim", [1, 2, 3])
def test_fe_nearest_matches_brute_force(oracle_mod, dim):
    """Grid search == numpy brute force (lowest index wins ties), status 0."""
    rng = np.random.default_rng(40 + dim)
    S = rng.uniform(0, 1, (700, dim))
    S[350:360] = S[10:20]               # duplicated sources: exact ties
    T = np.concatenate([rng.uniform(-0.3, 1.3, (400, dim)), S[5:25]])
    v = rng.normal(size=len(S))
    out, st = oracle_mod.fe_nearest(dim, S, v, T)
    d2 = ((T[:, None, :] - S[None]) ** 2).sum(-1)
    ref = v[np.argmin(d2, axis=1)]      # argmin returns the first minimum
    np.testing.assert_array_equal(out, ref)
    assert not st.any()
    out_b, _ = oracle_mod.fe_nearest(dim, S, v, T, brute=True)
    np.testing.assert_array_equal(out, out_b)


def test_fe_nearest_same_points_and_empty(oracle_mod):
    """Same source and target cloud copies the values; no sources -> 0."""
    rng = np.random.default_rng(7)
    S = rng.uniform(0, 1, (300, 3))
    v = rng.normal(size=300)
    out, _ = oracle_mod.fe_nearest(3, S, v, S)
    np.testing.assert_array_equal(out, v)
    out0, st0 = oracle_mod.fe_nearest(3, np.zeros((0, 3)), np.zeros(0), S)
    assert not out0.any() and not st0.any()


@pytest.mark.parametrize("rbf,dim", [(0x10, 3), (0x11, 3), (0x12, 3),
                                     (0x13, 2), (0x14, 3)])
def test_target_grad_matches_finite_differences(oracle_mod, rbf, dim):
    """The oracle's analytic target-coordinate gradient against central
    differences of its own forward transfer (a new setup per perturbed
    target set, step 1e-6): pins the Wendland derivatives, the polynomial
    part and the derivative of the Shepard weights."""
    rng = np.random.default_rng(40 + rbf + dim)
    S = rng.uniform(0, 1, (500, dim))
    T = rng.uniform(0.02, 0.98, (400, dim))
    f = np.stack([np.sin(3 * S[:, 0]) + S[:, -1] ** 2,
                  np.cos(2 * S.sum(1))])
    g = rng.normal(size=(2, len(T)))
    args = (40, 0.15, -1.0)
    base = oracle_mod.Setup(dim, rbf, S, T, *args)
    grad = base.target_grad(f, g)
    h = 1e-6
    picks = rng.choice(len(T) * dim, 40, replace=False)
    fd = np.empty(len(picks))
    for i, e in enumerate(picks):
        t, a = divmod(int(e), dim)
        val = []
        for sgn in (1.0, -1.0):
            Tp = T.copy()
            Tp[t, a] += sgn * h
            o = oracle_mod.Setup(dim, rbf, S, Tp, *args)
            assert o.num_patches == base.num_patches
            val.append(float(np.sum(g[:, t] * o.apply(f)[:, t])))
        fd[i] = (val[0] - val[1]) / (2 * h)
    ref = grad.reshape(-1)[picks]
    err = np.abs(fd - ref).max() / np.abs(ref).max()
    assert err < 1e-6, err


def test_target_grad_of_linear_field_is_its_slope(oracle_mod):
    """A linear field is reproduced exactly, so d/dy of <g, u> is g times the
    slope at every target covered by a patch."""
    rng = np.random.default_rng(9)
    S = rng.uniform(0, 1, (600, 3))
    T = rng.uniform(0.05, 0.95, (300, 3))
    slope = np.array([0.7, -1.3, 2.1])
    f = 0.4 + S @ slope
    g = rng.normal(size=len(T))
    o = oracle_mod.Setup(3, 0x11, S, T, 40, 0.15, -1.0)
    grad = o.target_grad(f, g)
    assert np.allclose(grad, g[:, None] * slope[None, :], rtol=0,
                       atol=1e-9 * np.abs(g).max() * 3)


# ---- MLS -------------------------------------------------------------------

def _mls_basis(s, dim):
    cols = ([np.ones(len(s))] + [s[:, a] for a in range(dim)]
            + [s[:, a] * s[:, b] for a in range(dim) for b in range(a, dim)])
    return np.stack(cols, 1)


def _mls_numpy(S, T, f, dim):
    """Independent MLS: brute-force k nearest (lexicographic (d^2, index)),
    weighted least squares by numpy.linalg.lstsq."""
    k = (dim + 1) * (dim + 2) // 2
    idx = np.empty((len(T), k), np.int64)
    out = np.empty(len(T))
    for t in range(len(T)):
        d2 = ((S - T[t]) ** 2).sum(1)
        order = np.lexsort((np.arange(len(S)), d2))[:k]
        idx[t] = order
        s = S[order] - T[t]
        d = np.sqrt((s ** 2).sum(1))
        rho = 1.1 * max(d.max(), np.finfo(float).eps)
        w = 1 - d / rho
        B = w[:, None] * _mls_basis(s / rho, dim)
        out[t] = np.linalg.lstsq(B, w * f[order], rcond=None)[0][0]
    return idx, out


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_mls_matches_numpy_least_squares(oracle_mod, dim):
    rng = np.random.default_rng(90 + dim)
    S = rng.uniform(0, 1, (3000, dim))
    T = rng.uniform(0, 1, (400, dim))
    f = np.sin(3 * S[:, 0]) + np.cos(2 * S[:, -1]) + S.sum(1) ** 3
    idx, coef = oracle_mod.mls_coefficients(dim, S, T)
    ridx, ref = _mls_numpy(S, T, f, dim)
    assert np.array_equal(idx, ridx)
    got = np.sum(coef * f[idx], 1)
    assert np.abs(got - ref).max() < 1e-10 * np.abs(ref).max()


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_mls_reproduces_quadratics(oracle_mod, dim):
    rng = np.random.default_rng(95 + dim)
    S = rng.uniform(0, 1, (5000, dim))
    T = rng.uniform(0, 1, (1000, dim))
    c = rng.normal(size=(dim + 1) * (dim + 2) // 2)
    q = lambda X: _mls_basis(X, dim) @ c
    got = oracle_mod.mls_interpolate(dim, S, q(S), T)
    assert np.abs(got - q(T)).max() < 1e-9


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_mls_reference_same_grid(oracle_mod, dim):
    """The reference's known-answer test: a 20^D grid onto itself, rel-L2
    below 1e-4 against the Franke field (tests/test_mls.py:86-94,
    tests/test_cpp_mls_interface.cpp:15,156-166)."""
    ax = [np.linspace(0.0, 1.0, 20)] * dim
    G = np.stack([g.ravel() for g in np.meshgrid(*ax, indexing="ij")], 1)
    fr = {1: oracle_mod.franke_1d, 2: oracle_mod.franke_2d,
          3: oracle_mod.franke_3d}[dim]
    f = fr(*G.T)
    got = oracle_mod.mls_interpolate(dim, G, f, G)
    assert np.linalg.norm(got - f) / np.linalg.norm(f) < 1e-4
    # and the independent least squares agrees on the grid's rank-deficient
    # neighbourhoods
    _, ref = _mls_numpy(G, G[::7], f, dim)
    assert np.abs(got[::7] - ref).max() < 1e-10


# ---- adaptive patches (density-aware builder; not in the reference) --------

def _quantize(x, lo, ext):
    t = ((x - lo) / ext) * 65536.0
    q = np.where(t > 0.0, t, 0.0)
    q = np.where(q >= 65535.0, 65535.0, q)
    return q.astype(np.uint64)      # truncation, as the C cast


def _prefix(j, D, d):
    """Morton prefix of depth d of integer cell coordinates j (n, D)."""
    p = np.zeros(len(j), np.uint64)
    for t in range(d):
        a = t % D
        ba = (d + D - 1 - a) // D
        bit = np.uint64(ba - 1 - t // D)
        p = (p << np.uint64(1)) | ((j[:, a] >> bit) & np.uint64(1))
    return p


def adaptive_numpy(S, T, dim, leaf, nmin, nmax, ov=0.15, variant=None):
    """The adaptive builder restated by brute force in numpy: Morton codes of
    16-bit-quantised coordinates over the box of both sets, binary tree over
    the code bits (split cells with > leaf sources; keep cells with sources
    or targets; leaves with targets are patches), centre = cell centre,
    r = max((1 + ov) * half diagonal, distance to the nmin-th nearest
    source), halve patches holding > nmax sources (children with targets),
    order by cell. `variant` plants a mistake (tests only)."""
    D, DB = dim, 16 * dim
    lo = np.minimum(S.min(0), T.min(0))
    ext = np.maximum(S.max(0), T.max(0)) - lo
    ext = np.where(ext > 0, ext, 1.0)
    qs = np.stack([_quantize(S[:, a], lo[a], ext[a]) for a in range(D)], 1)
    qt = np.stack([_quantize(T[:, a], lo[a], ext[a]) for a in range(D)], 1)
    cs, ct = np.sort(_prefix(qs, D, DB)), np.sort(_prefix(qt, D, DB))

    def count(codes, p, d):
        a, b = p << (DB - d), (p + 1) << (DB - d)
        return int(np.searchsorted(codes, b) - np.searchsorted(codes, a))

    leaves, stack = [], [(0, 0)]
    while stack:
        p, d = stack.pop()
        if count(cs, p, d) > leaf and d < DB:
            for h in (0, 1):
                c = (2 * p + h, d + 1)
                if count(cs, *c) or count(ct, *c):
                    stack.append(c)
        elif count(ct, p, d):
            leaves.append((p, d))
    S3 = pad3(S)
    kq = min(nmin - (1 if variant == "knn_off_by_one" else 0), len(S))
    done, todo = [], leaves
    while todo:
        nxt = []
        for p, d in todo:
            cen, s = np.zeros(3), 0.0
            for a in range(D):
                ba = (d + D - 1 - a) // D
                ia = 0
                for t in range(a, d, D):
                    ia = (ia << 1) | ((p >> (d - 1 - t)) & 1)
                h = ext[a] * 2.0 ** -ba
                cen[a] = lo[a] + (float(ia) + 0.5) * h
                s = s + h * h
            grow = 1.0 if variant == "no_overlap" else 1.0 + ov
            rc = grow * (0.5 * np.sqrt(s))
            d2 = _sq(cen, S3)
            dk = np.sqrt(np.partition(d2, kq - 1)[kq - 1])
            r = max(rc, dk)
            n = int(np.sum(np.sqrt(d2) <= r))
            if n > nmax and d < DB and variant != "no_split":
                nxt += [c for c in ((2 * p, d + 1), (2 * p + 1, d + 1))
                        if count(ct, *c)]
            else:
                done.append(((p << (DB - d)), cen, r))
        todo = nxt
    done.sort(key=lambda x: x[0])
    return (np.array([x[1] for x in done]), np.array([x[2] for x in done]),
            np.array([x[0] for x in done], np.uint64))


def _adaptive_clouds():
    rng = np.random.default_rng(5)
    return [
        ("3d-mixture", W.gaussian_mixture(6000, 4), W.uniform(3000, 5), 40,
         20, 300),
        ("3d-dense-cap", W.gaussian_mixture(5000, 9), W.uniform(2000, 8), 60,
         15, 150),
        ("2d", rng.normal(0.5, 0.12, (3000, 2)), rng.uniform(0, 1, (1500, 2)),
         30, 12, 200),
        ("1d", rng.uniform(0, 1, (800, 1)) ** 3, rng.uniform(0, 1, (400, 1)),
         10, 6, 60),
    ]


def test_adaptive_builder_bruteforce(oracle_mod):
    """Centres, radii and the cell order of the oracle's adaptive patches
    equal the numpy restatement exactly; members and weights by brute
    force; each planted mistake changes the patches on some cloud."""
    misses = {"no_overlap": 0, "knn_off_by_one": 0, "no_split": 0}
    for name, S, T, leaf, nmin, nmax in _adaptive_clouds():
        dim = S.shape[1]
        o = oracle_mod.AdaptiveSetup(dim, 0x11, S, T, leaf, nmin, nmax)
        C, cand = o.centers()
        r, delta, _ = o.patches()
        refC, refR, refKey = adaptive_numpy(S, T, dim, leaf, nmin, nmax)
        np.testing.assert_array_equal(cand.astype(np.uint64), refKey,
                                      err_msg=name)
        np.testing.assert_array_equal(C[:, :dim], refC[:, :dim], err_msg=name)
        np.testing.assert_array_equal(r, refR, err_msg=name)
        np.testing.assert_array_equal(delta, r)       # support scale 1
        n = np.diff(o.csr(0)[0])
        assert n.min() >= min(nmin, len(S)) and n.max() <= nmax, name
        # members: sources within r_k, targets within r_k - 1e-14
        for which, P, eps in ((0, S, 0.0), (1, T, 1e-14)):
            off, idx = o.csr(which)
            for k in range(0, len(C), max(1, len(C) // 40)):
                d = np.sqrt(_sq(C[k], pad3(P)))
                np.testing.assert_array_equal(idx[off[k]:off[k + 1]],
                                              np.nonzero(d <= r[k] - eps)[0])
        # every target is covered; weights psi(|y - c_k| / r_k) / sum
        off, idx = o.csr(2)
        assert np.all(np.diff(off) > 0), name
        w = o.weights()
        for t in range(0, len(T), max(1, len(T) // 60)):
            ks = idx[off[t]:off[t + 1]]
            d = np.sqrt(_sq(pad3(T[t:t + 1])[0], pad3(C[ks, :dim])))
            np.testing.assert_array_equal(ks, np.nonzero(
                np.sqrt(_sq(pad3(T[t:t + 1])[0], pad3(C[:, :dim])))
                <= r)[0])
            v = _psi_c2(d * (1.0 / r[ks]))
            np.testing.assert_allclose(w[off[t]:off[t + 1]], v / v.sum(),
                                       rtol=1e-14, atol=1e-16)
        for v in misses:
            bc, br, bk = adaptive_numpy(S, T, dim, leaf, nmin, nmax, variant=v)
            if (len(bk) != len(refKey) or np.any(bk != refKey)
                    or np.any(br != refR)):
                misses[v] += 1
    assert all(m > 0 for m in misses.values()), misses


def test_adaptive_morton_code_matches_numpy(oracle_mod):
    rng = np.random.default_rng(3)
    for dim in (1, 2, 3):
        P = rng.uniform(-1, 2, (50, dim))
        lo, ext = P.min(0), P.max(0) - P.min(0)
        q = np.stack([_quantize(P[:, a], lo[a], ext[a])
                      for a in range(dim)], 1)
        ref = _prefix(q, dim, 16 * dim)
        got = [oracle_mod.morton_code(p, dim, lo, ext) for p in P]
        assert [int(x) for x in ref] == got


def test_adaptive_operator_matches_dense_numpy(oracle_mod):
    """The adaptive transfer (per-patch support and PU radius) against a
    dense numpy evaluation of sum_k w_k [E_k A_k^-1 (b - Q beta) + V beta]
    on the oracle's own lists, and exact reproduction of linear fields."""
    S = W.gaussian_mixture(3000, 4)
    T = W.uniform(1500, 5)
    o = oracle_mod.AdaptiveSetup(3, 0x13, S, T, 40, 20, 300, 0.15, 0.8)
    f = W.franke_3d(*S.T)
    got = o.apply(f)
    C, _ = o.centers()
    r, delta, _ = o.patches()
    soff, sidx = o.csr(0)
    toff, tidx = o.csr(1)
    woff, widx = o.csr(2)
    w = o.weights()
    phi = np.vectorize(lambda p: oracle_mod.wendland(0x13, p, 1.0))
    ref = np.zeros(len(T))
    for k in range(len(C)):
        src, tgt = sidx[soff[k]:soff[k + 1]], tidx[toff[k]:toff[k + 1]]
        X, Y = S[src], T[tgt]
        Q = np.hstack([np.ones((len(X), 1)), X])
        beta = np.linalg.lstsq(Q, f[src], rcond=None)[0]
        A = phi(np.sqrt(((X[:, None] - X[None]) ** 2).sum(-1)) / delta[k])
        alpha = scipy.linalg.solve(A, f[src] - Q @ beta)
        E = phi(np.sqrt(((Y[:, None] - X[None]) ** 2).sum(-1)) / delta[k])
        loc = E @ alpha + np.hstack([np.ones((len(Y), 1)), Y]) @ beta
        for j, t in enumerate(tgt):
            e = woff[t] + int(np.nonzero(widx[woff[t]:woff[t + 1]] == k)[0][0])
            ref[t] += w[e] * loc[j]
    assert rel_l2(got, ref) < 1e-11
    lin = 1.0 + S @ np.array([0.5, -1.0, 2.0])
    assert np.abs(o.apply(lin) - (1.0 + T @ np.array([0.5, -1.0, 2.0]))
                  ).max() < 1e-11
