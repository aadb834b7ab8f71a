"""Nearest-source transfer: the ``pacman.fe`` submodule for NEAREST_NEAREST.

``interpolate`` mirrors ``pacman.fe.interpolate``
(src/pybindings/bindings.cpp:94-112, src/pybindings/fe_bindings.cpp:131-178):
same argument names and order, the same ValueError checks, and the
``(target_values, transfer_status)`` result. This build serves the method
``methods.NEAREST_NEAREST`` (src/finite_elements/FTNearest.hpp:28-52): each
target takes the value of its nearest source (ties: lowest source index) and
status 0 (OUTSIDE). The connectivity arguments are accepted and not read; the
other FE methods raise ValueError. A bad dimension raises RuntimeError
instead of throwing a C++ pointer (fe_bindings.cpp:174).
"""

from __future__ import annotations

import ctypes
import types

import numpy as np

from . import _lib
from . import execspaces as _es

methods = types.SimpleNamespace(
    NEAREST_NEAREST=0xF0, INTERP_CLAMP=0xF1, INTERP_NEAREST=0xF2,
    INTERP_ZEROFILL=0xF3, INTERP_EXTRAP=0xF4)

status = types.SimpleNamespace(OUTSIDE=0, NEAREST=1, INTER=2, EXTRAP=3,
                               CLAMP=4)


def _as_f64(a) -> np.ndarray:
    if hasattr(a, "detach"):  # torch tensor
        a = a.detach().cpu().numpy()
    return np.ascontiguousarray(a, dtype=np.float64)


def interpolate(space_dimension, execspace, method, source_points,
                source_values, conn_val, conn_off, cell_types,
                target_points):
    """FE transfer of source_values to target_points: (values, status)."""
    del conn_val, conn_off, cell_types  # not read by NEAREST_NEAREST
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
    if int(method) != methods.NEAREST_NEAREST:
        raise ValueError(f"FE method {int(method)} is not enabled in this "
                         "build (available: NEAREST_NEAREST)")
    if int(space_dimension) not in (1, 2, 3):
        raise RuntimeError(
            "The dimension of the points can only be: 1, 2 or 3.\n")
    if sp.shape[1] != int(space_dimension):
        raise ValueError(
            f"space_dimension {int(space_dimension)} does not match the "
            f"points' {sp.shape[1]} columns")
    m = tp.shape[0]
    out = np.zeros(m, np.float64)
    st = np.zeros(m, np.int32)
    lib = _lib.load()
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    rc = lib.pacman_fe_interpolate_c(
        int(space_dimension), int(execspace), int(method),
        sp.ctypes.data_as(dp), sp.shape[0], sv.ctypes.data_as(dp), None, 0,
        None, 0, None, tp.ctypes.data_as(dp), m, out.ctypes.data_as(dp),
        st.ctypes.data_as(ip))
    if rc != 0:
        raise RuntimeError(_lib.last_error())
    return out, st


def nearest_device(space_dimension, source_points, source_values,
                   target_points):
    """NEAREST_NEAREST on CUDA (HIP) torch tensors, result on the device.

    source_points (N, dim), source_values (N,), target_points (M, dim),
    float64 device tensors. Returns (values (M,) float64, status (M,) int32).
    """
    import torch

    sp, sv, tp = source_points, source_values, target_points
    for t in (sp, sv, tp):
        if not (t.is_cuda and t.dtype == torch.float64):
            raise ValueError("expected float64 device tensors")
    sp, sv, tp = sp.contiguous(), sv.contiguous(), tp.contiguous()
    dim = int(space_dimension)
    if sp.ndim != 2 or tp.ndim != 2 or sp.shape[1] != dim or \
            tp.shape[1] != dim or sv.shape[0] != sp.shape[0]:
        raise ValueError("inconsistent shapes")
    m = tp.shape[0]
    out = torch.zeros(m, dtype=torch.float64, device=tp.device)
    st = torch.zeros(m, dtype=torch.int32, device=tp.device)
    stream = torch.cuda.current_stream(tp.device).cuda_stream
    rc = _lib.load().pacman_fe_nearest_device(
        dim, sp.data_ptr(), sp.shape[0], sv.data_ptr(), tp.data_ptr(), m,
        out.data_ptr(), st.data_ptr(), stream)
    if rc != 0:
        raise RuntimeError(_lib.last_error())
    return out, st
