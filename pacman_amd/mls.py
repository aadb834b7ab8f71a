"""Moving-least-squares transfer: the drop-in ``pacman.MLS`` submodule.

``interpolate`` mirrors ``pacman.MLS.interpolate``
(src/pybindings/bindings.cpp:124-132, src/pybindings/mls_bindings.cpp:64-105):
same argument names and order, the same ValueError checks, float64 output.
An unsupported execution space raises ValueError instead of ending the
process, and a bad dimension raises RuntimeError.

The method is the reference's (ArborX MovingLeastSquares defaults,
src/mls/interpolator.hpp:70-74): each target takes the value at itself of
the weighted least-squares quadratic through its k nearest sources
(k = 3, 6, 10 in dimension 1, 2, 3), weights (1 - d / (1.1 d_max))^2.

``Plan`` keeps the neighbours and weights of every target (setup once) and
transfers any number of fields by a gather (apply), from numpy arrays or
float64 torch tensors on the GPU.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from . import execspaces as _es
from .rbf import _as_f64, _ptr


def interpolate(space_dimension, execspace, source_points, source_values,
                target_points):
    """MLS transfer of source_values to target_points (float64 array)."""
    sp = _as_f64(source_points)
    sv = _as_f64(source_values)
    tp = _as_f64(target_points)
    if sp.ndim != 2:
        raise ValueError("Source points must be a 2-d numpy array!")
    if tp.ndim != 2:
        raise ValueError("Target points must be a 2-d numpy array!")
    if sp.shape[1] != tp.shape[1]:
        raise ValueError(
            "Dimensions mismatch between source points and target "
            f"points!\nProvided source points dimension: {sp.shape[1]}"
            f"\nProvided target points dimension: {tp.shape[1]}\n")
    if sp.shape[0] != sv.shape[0]:
        raise ValueError(
            "Each source point must have its corresponding value in the "
            f"source values!\nNumber of source points: {sp.shape[0]}"
            f"\nNumber of source values: {sv.shape[0]}\n")
    if int(execspace) not in _es._NAMES:
        raise ValueError("execspace: Serial, OpenMP, Threads, Cuda, HIP, SYCL")
    if _es.name(execspace) not in _es.available():
        raise ValueError(
            f"execution space {_es.name(execspace)} is not enabled in this "
            f"build (available: {', '.join(_es.available())})")
    if int(space_dimension) not in (1, 2, 3):
        raise RuntimeError(
            "The dimension of the points can only be: 1, 2 or 3.\n")
    if sp.shape[1] != int(space_dimension):
        raise ValueError(
            f"space_dimension={space_dimension} but the points have "
            f"{sp.shape[1]} coordinates")
    lib = _lib.load()
    out = np.zeros(tp.shape[0], dtype=np.float64)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = lib.pacman_mls_interpolate_c(
        int(space_dimension), int(execspace), sp.ctypes.data_as(dp),
        sp.shape[0], sv.ctypes.data_as(dp), tp.ctypes.data_as(dp),
        tp.shape[0], out.ctypes.data_as(dp))
    if rc != 0:
        raise RuntimeError(_lib.last_error())
    return out


class Plan:
    """Persistent MLS plan: neighbours and weights once, then any number of
    fields per apply. Points are numpy arrays (host) or float64 torch
    tensors on the GPU, both in the same place."""

    def __init__(self, space_dimension, source_points, target_points,
                 stream=None):
        self._lib = _lib.load()
        self._h = ctypes.c_void_p()
        self.dim = int(space_dimension)
        src, tgt = source_points, target_points
        if not hasattr(src, "data_ptr"):
            src, tgt = _as_f64(src), _as_f64(tgt)
        else:
            src, tgt = src.contiguous(), tgt.contiguous()
            if (str(src.dtype) != "torch.float64"
                    or str(tgt.dtype) != "torch.float64"):
                raise ValueError("points must be float64")
        self.num_sources = int(src.shape[0])
        self.num_targets = int(tgt.shape[0])
        ps, dev = _ptr(src)
        pt, devt = _ptr(tgt)
        if dev != devt:
            raise ValueError("source and target points must both be on the "
                             "host or both on the device")
        rc = self._lib.pacman_mls_plan_create(
            ctypes.byref(self._h), self.dim, ps, self.num_sources, pt,
            self.num_targets, dev, ctypes.c_void_p(stream or 0))
        if rc != 0:
            raise RuntimeError(_lib.last_error())
        self.k = int(self._lib.pacman_mls_plan_k(self._h))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.pacman_mls_plan_destroy(h)
            self._h = ctypes.c_void_p()

    def apply(self, source_values, out=None, stream=None):
        """values (F, N) or (N,) -> (F, M) or (M,)."""
        v = source_values
        squeeze = len(v.shape) == 1
        F = 1 if squeeze else int(v.shape[0])
        if hasattr(v, "data_ptr"):
            import torch
            v = v.contiguous().to(torch.float64)
            if out is None:
                out = torch.empty((F, self.num_targets), dtype=torch.float64,
                                  device=v.device)
        else:
            v = _as_f64(v)
            if out is None:
                out = np.zeros((F, self.num_targets))
        pv, dev = _ptr(v)
        po, devo = _ptr(out)
        if dev != devo:
            raise ValueError("values and output must live in the same place")
        rc = self._lib.pacman_mls_plan_apply(self._h, pv, F, po, dev,
                                             ctypes.c_void_p(stream or 0))
        if rc != 0:
            raise RuntimeError(_lib.last_error())
        return out.reshape(-1) if squeeze else out

    def neighbours(self):
        """(idx (M, k) int32 source indices, ascending distance; coef (M, k))."""
        idx = np.zeros((self.num_targets, self.k), np.int32)
        coef = np.zeros((self.num_targets, self.k))
        rc = self._lib.pacman_mls_plan_neighbours(
            self._h, ctypes.c_void_p(idx.ctypes.data),
            ctypes.c_void_p(coef.ctypes.data))
        if rc != 0:
            raise RuntimeError(_lib.last_error())
        return idx, coef
