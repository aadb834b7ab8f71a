"""CPU FP64 oracle for the RBF-PUM transfer path (test infrastructure only).

This package checks results; it is never the thing measured or shipped.
Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import it. The product library
(``pacman_amd``) never loads it.

The C++ restatements live in ``oracle/rbf_pum_oracle.cpp`` (follows
``src/rbf_pum/**`` of the reference, file:line citations in that file) and
``oracle/mls_oracle.cpp`` (the MLS transfer of ``src/mls/``). This
module is a thin ctypes wrapper plus the reference's analytic test field.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_ip = ctypes.POINTER(ctypes.c_int)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_wendland.restype = ctypes.c_double
        L.oracle_wendland.argtypes = [ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double]
        L.oracle_singular_values.argtypes = [_dp, ctypes.c_int, ctypes.c_int,
                                             _dp]
        L.oracle_lstsq.argtypes = [_dp, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.oracle_lu_solve.argtypes = [_dp, ctypes.c_int, _dp]
        L.oracle_setup_create.restype = ctypes.c_void_p
        L.oracle_setup_create.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
            ctypes.c_double, _dp, ctypes.c_int64, _dp, ctypes.c_int64,
            ctypes.c_int, ctypes.c_int]
        L.oracle_setup_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_setup_radius.restype = ctypes.c_double
        L.oracle_setup_radius.argtypes = [ctypes.c_void_p]
        L.oracle_setup_min_margin.restype = ctypes.c_double
        L.oracle_setup_min_margin.argtypes = [ctypes.c_void_p]
        L.oracle_setup_num_candidates.restype = ctypes.c_int64
        L.oracle_setup_num_candidates.argtypes = [ctypes.c_void_p]
        L.oracle_setup_num_patches.restype = ctypes.c_int64
        L.oracle_setup_num_patches.argtypes = [ctypes.c_void_p]
        L.oracle_setup_lattice.argtypes = [ctypes.c_void_p, _i64p, _dp, _dp,
                                           _dp]
        L.oracle_setup_probes.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.oracle_setup_centers.argtypes = [ctypes.c_void_p, _dp, _i64p]
        L.oracle_setup_csr_size.restype = ctypes.c_int64
        L.oracle_setup_csr_size.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.oracle_setup_csr.argtypes = [ctypes.c_void_p, ctypes.c_int, _i64p,
                                       _i64p]
        L.oracle_setup_weights.argtypes = [ctypes.c_void_p, _dp]
        L.oracle_setup_apply.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int,
                                         ctypes.c_int64, ctypes.c_int64, _dp,
                                         _ip]
        L.oracle_rbf_interpolate.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
            ctypes.c_double, _dp, ctypes.c_int64, _dp, _dp, ctypes.c_int64,
            _dp]
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_mls_coefficients.argtypes = [
            ctypes.c_int, _dp, ctypes.c_int64, _dp, ctypes.c_int64, _i64p,
            _dp]
        L.oracle_setup_target_grad.argtypes = [ctypes.c_void_p, _dp,
                                               ctypes.c_int, _dp, _dp]
        L.oracle_adaptive_create.restype = ctypes.c_void_p
        L.oracle_adaptive_create.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
            ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp,
            ctypes.c_int64, _dp, ctypes.c_int64]
        L.oracle_adaptive_patches.argtypes = [ctypes.c_void_p, _dp, _dp, _ip]
        L.oracle_morton_code.restype = ctypes.c_uint64
        L.oracle_morton_code.argtypes = [_dp, ctypes.c_int, _dp, _dp]
        L.oracle_fe_nearest.argtypes = [
            ctypes.c_int, _dp, ctypes.c_int64, _dp, _dp, ctypes.c_int64,
            ctypes.c_int, _dp, _ip]
        _lib = L
    return _lib


def _ptr(a: np.ndarray, t=_dp):
    return a.ctypes.data_as(t)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n: int) -> int:
    return lib().oracle_set_threads(int(n))


def wendland(rbf_function: int, r: float, radius_inv: float) -> float:
    """phi(r) of rbf_functions.hpp:22-150 (selector 0x10..0x14 or 0..4)."""
    return lib().oracle_wendland(int(rbf_function), float(r),
                                 float(radius_inv))


def singular_values(Q) -> np.ndarray:
    Q = _f64(Q)
    n, p = Q.shape
    S = np.zeros(p)
    assert lib().oracle_singular_values(_ptr(Q), n, p, _ptr(S)) == 0
    return S


def lstsq(Q, b) -> np.ndarray:
    Q = _f64(Q)
    b = _f64(b)
    n, p = Q.shape
    beta = np.zeros(p)
    assert lib().oracle_lstsq(_ptr(Q), n, p, _ptr(b), _ptr(beta)) == 0
    return beta


def lu_solve(A, b) -> np.ndarray:
    A = _f64(A)
    x = _f64(b).copy()
    lib().oracle_lu_solve(_ptr(A), A.shape[0], _ptr(x))
    return x


class Setup:
    """Stage-by-stage oracle state (radius, lattice, centres, CSR lists)."""

    def __init__(self, dim, rbf_function, source_points, target_points,
                 npc=50, overlap=0.15, support=0.1, brute=False,
                 sequential_prune=False):
        self.dim = int(dim)
        self.src = _f64(source_points).reshape(-1, self.dim)
        self.tgt = _f64(target_points).reshape(-1, self.dim)
        self.ns, self.nt = self.src.shape[0], self.tgt.shape[0]
        self._h = lib().oracle_setup_create(
            self.dim, int(rbf_function), int(npc), float(overlap),
            float(support), _ptr(self.src), self.ns, _ptr(self.tgt), self.nt,
            int(bool(brute)), int(bool(sequential_prune)))
        if not self._h:
            raise ValueError("oracle rejected the parameters")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().oracle_setup_destroy(h)
            self._h = None

    @property
    def radius(self) -> float:
        return lib().oracle_setup_radius(self._h)

    @property
    def min_margin(self) -> float:
        return lib().oracle_setup_min_margin(self._h)

    @property
    def num_candidates(self) -> int:
        return lib().oracle_setup_num_candidates(self._h)

    @property
    def num_patches(self) -> int:
        return lib().oracle_setup_num_patches(self._h)

    def lattice(self):
        shape = np.zeros(3, np.int64)
        sp, lo, hi = np.zeros(3), np.zeros(3), np.zeros(3)
        lib().oracle_setup_lattice(self._h, _ptr(shape, _i64p), _ptr(sp),
                                   _ptr(lo), _ptr(hi))
        return shape, sp, lo, hi

    def probes(self):
        s = np.zeros(21)
        m = np.zeros(7)
        lib().oracle_setup_probes(self._h, _ptr(s), _ptr(m))
        return s.reshape(7, 3), m

    def centers(self):
        K = self.num_patches
        c = np.zeros((K, 3))
        cand = np.zeros(K, np.int64)
        lib().oracle_setup_centers(self._h, _ptr(c), _ptr(cand, _i64p))
        return c, cand

    def csr(self, which: int):
        """which: 0 source members, 1 target members, 2 per-target centres."""
        n = self.num_patches if which in (0, 1) else self.nt
        off = np.zeros(n + 1, np.int64)
        idx = np.zeros(lib().oracle_setup_csr_size(self._h, which), np.int64)
        lib().oracle_setup_csr(self._h, which, _ptr(off, _i64p),
                               _ptr(idx, _i64p))
        return off, idx

    def weights(self):
        w = np.zeros(lib().oracle_setup_csr_size(self._h, 2))
        lib().oracle_setup_weights(self._h, _ptr(w))
        return w

    def apply(self, values, k_begin=0, k_end=-1, return_pv=False):
        v = _f64(values)
        F = 1 if v.ndim == 1 else v.shape[0]
        v = v.reshape(F, self.ns)
        out = np.zeros((F, self.nt))
        pv = np.zeros(self.num_patches, np.int32)
        lib().oracle_setup_apply(self._h, _ptr(v), F, int(k_begin), int(k_end),
                                 _ptr(out), _ptr(pv, _ip))
        out = out[0] if F == 1 and np.ndim(values) == 1 else out
        return (out, pv) if return_pv else out

    def target_grad(self, values, target_grad):
        """d/dy of sum_f <target_grad_f, u_f> at every target: (nt, dim)."""
        v = _f64(values)
        F = 1 if v.ndim == 1 else v.shape[0]
        v = v.reshape(F, self.ns)
        g = _f64(target_grad).reshape(F, self.nt)
        out = np.zeros((self.nt, self.dim))
        rc = lib().oracle_setup_target_grad(self._h, _ptr(v), F, _ptr(g),
                                            _ptr(out))
        if rc != 0:
            raise ValueError("oracle_setup_target_grad failed")
        return out


class AdaptiveSetup(Setup):
    """Oracle state of the density-aware patch builder (not in the
    reference; rbf_pum_oracle.cpp, runAdaptiveSetup): patch centres, per-
    patch radii and supports, CSR lists, weights, apply."""

    def __init__(self, dim, rbf_function, source_points, target_points,
                 leaf_points=40, min_points=20, max_points=600,
                 overlap=0.15, support_scale=1.0):
        self.dim = int(dim)
        self.src = _f64(source_points).reshape(-1, self.dim)
        self.tgt = _f64(target_points).reshape(-1, self.dim)
        self.ns, self.nt = self.src.shape[0], self.tgt.shape[0]
        self._h = lib().oracle_adaptive_create(
            self.dim, int(rbf_function), int(leaf_points), int(min_points),
            int(max_points), float(overlap), float(support_scale),
            _ptr(self.src), self.ns, _ptr(self.tgt), self.nt)
        if not self._h:
            raise ValueError("oracle rejected the parameters")

    def patches(self):
        """(radius, support, depth) per patch."""
        K = self.num_patches
        r, d = np.zeros(K), np.zeros(K)
        depth = np.zeros(K, np.int32)
        lib().oracle_adaptive_patches(self._h, _ptr(r), _ptr(d),
                                      _ptr(depth, _ip))
        return r, d, depth


def morton_code(x, dim, lo, ext) -> int:
    """The builder's 16-bit-per-axis Morton code of one point."""
    x = _f64(x)
    lo = _f64(lo)
    ext = _f64(ext)
    return int(lib().oracle_morton_code(_ptr(x), int(dim), _ptr(lo),
                                        _ptr(ext)))


def rbf_interpolate(dim, rbf_function, source_points, source_values,
                    target_points, npc=50, overlap=0.15, support=0.1):
    """Oracle counterpart of pacman.rbf.interpolate with explicit params."""
    src = _f64(source_points).reshape(-1, dim)
    tgt = _f64(target_points).reshape(-1, dim)
    vals = _f64(source_values)
    out = np.zeros(tgt.shape[0])
    rc = lib().oracle_rbf_interpolate(
        int(dim), int(rbf_function), int(npc), float(overlap), float(support),
        _ptr(src), src.shape[0], _ptr(vals), _ptr(tgt), tgt.shape[0],
        _ptr(out))
    if rc != 0:
        raise ValueError("oracle rejected the parameters")
    return out


def mls_coefficients(dim, source_points, target_points):
    """k nearest sources of every target and their MLS coefficients:
    (idx (nt, k) int64, coef (nt, k))."""
    S = _f64(source_points).reshape(-1, dim)
    T = _f64(target_points).reshape(-1, dim)
    k = (dim + 1) * (dim + 2) // 2
    idx = np.zeros((len(T), k), np.int64)
    coef = np.zeros((len(T), k))
    rc = lib().oracle_mls_coefficients(int(dim), _ptr(S), len(S), _ptr(T),
                                       len(T), _ptr(idx, _i64p), _ptr(coef))
    if rc != k:
        raise ValueError("oracle_mls_coefficients: bad dimension or fewer "
                         "sources than the basis size")
    return idx, coef


def mls_interpolate(dim, source_points, source_values, target_points):
    """Oracle counterpart of pacman.MLS.interpolate (one or more fields)."""
    idx, coef = mls_coefficients(dim, source_points, target_points)
    v = _f64(source_values)
    return np.sum(coef * v[..., idx], axis=-1)


def fe_nearest(dim, source_points, source_values, target_points,
               brute=False):
    """Oracle of the FE NEAREST_NEAREST transfer: (values, status)."""
    src = _f64(source_points).reshape(-1, dim)
    tgt = _f64(target_points).reshape(-1, dim)
    vals = _f64(source_values)
    out = np.zeros(tgt.shape[0])
    st = np.zeros(tgt.shape[0], np.int32)
    rc = lib().oracle_fe_nearest(int(dim), _ptr(src), src.shape[0],
                                 _ptr(vals), _ptr(tgt), tgt.shape[0],
                                 int(bool(brute)), _ptr(out), _ptr(st, _ip))
    if rc != 0:
        raise ValueError("oracle rejected the parameters")
    return out, st


def franke_1d(x):
    """Franke-1D test field (tests/franke_functions.py:29-69; that file's
    second definition of franke_1d, the one in effect)."""
    x = np.asarray(x, dtype=np.float64)
    return (0.75 * np.exp(-((9 * x - 2) ** 2) / 4)
            + 0.75 * np.exp(-((9 * x + 1) ** 2) / 49)
            + 0.5 * np.exp(-((9 * x - 7) ** 2) / 4)
            - 0.2 * np.exp(-((9 * x - 4) ** 2)))


def franke_2d(x, y):
    """Franke-2D test field (tests/franke_functions.py:71-116)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    return (0.75 * np.exp(-((9 * x - 2) ** 2 + (9 * y - 2) ** 2) / 4)
            + 0.75 * np.exp(-((9 * x + 1) ** 2) / 49 - (9 * y + 1) / 10)
            + 0.5 * np.exp(-((9 * x - 7) ** 2 + (9 * y - 3) ** 2) / 4)
            - 0.2 * np.exp(-((9 * x - 4) ** 2) - (9 * y - 7) ** 2))


def franke_3d(x, y, z):
    """Franke-3D test field (tests/franke_functions.py:118-165)."""
    return (0.75 * np.exp(-((9 * x - 2) ** 2 + (9 * y - 2) ** 2
                            + (9 * z - 2) ** 2) / 4.0)
            + 0.75 * np.exp(-((9 * x + 1) ** 2 / 49.0 + (9 * y + 1) / 10.0
                              + (9 * z + 1) / 10.0))
            + 0.5 * np.exp(-((9 * x - 7) ** 2 + (9 * y - 3) ** 2
                             + (9 * z - 5) ** 2) / 4.0)
            - 0.2 * np.exp(-((9 * x - 4) ** 2 + (9 * y - 7) ** 2
                             + (9 * z - 5) ** 2)))
