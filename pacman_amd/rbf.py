"""RBF-PUM transfer: the drop-in ``pacman.rbf`` submodule plus a plan API.

``interpolate`` mirrors ``pacman.rbf.interpolate``
(src/pybindings/bindings.cpp:75-85, src/pybindings/rbf_bindings.cpp:74-120):
same argument names and order, the same ValueError checks, float64 output.
Two deliberate differences: an unsupported execution space or basis raises
ValueError instead of ending the process (rbf_bindings.hpp:50-53), and a bad
dimension raises RuntimeError instead of throwing a C++ pointer
(rbf_bindings.cpp:117).

``Plan`` exposes the setup/apply split with explicit parameters, device
(torch) or host (numpy) arrays and F fields per apply.
"""

from __future__ import annotations

import ctypes
import types

import numpy as np

from . import _lib
from . import execspaces as _es

functions = types.SimpleNamespace(
    WENDLANDC0=0x10, WENDLANDC2=0x11, WENDLANDC4=0x12, WENDLANDC6=0x13,
    WENDLANDC8=0x14)

_BASES = {0x10, 0x11, 0x12, 0x13, 0x14}

# Reference parameters of the drop-in call (src/rbf_pum/interpolator.hxx:39-41)
DEFAULT_NODES_PER_CLUSTER = 50
DEFAULT_RELATIVE_OVERLAP = 0.15
DEFAULT_SUPPORT_RADIUS = 0.1


def _as_f64(a) -> np.ndarray:
    if hasattr(a, "detach"):  # torch tensor
        a = a.detach().cpu().numpy()
    return np.ascontiguousarray(a, dtype=np.float64)


def _check_selectors(space_dimension, execspace, rbf_function):
    if int(execspace) not in _es._NAMES:
        raise ValueError("execspace: Serial, OpenMP, Threads, Cuda, HIP, SYCL")
    if _es.name(execspace) not in _es.available():
        raise ValueError(
            f"execution space {_es.name(execspace)} is not enabled in this "
            f"build (available: {', '.join(_es.available())})")
    if int(rbf_function) not in _BASES:
        raise ValueError("function: WendlandC0, WendlandC2, WendlandC4, "
                         "WendlandC6, WendlandC8")
    if int(space_dimension) not in (1, 2, 3):
        raise RuntimeError(
            "The dimension of the points can only be: 1, 2 or 3.\n")


def interpolate(space_dimension, execspace, rbf_function, source_points,
                source_values, target_points):
    """RBF-PUM transfer of source_values to target_points (float64 array)."""
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
    _check_selectors(space_dimension, execspace, rbf_function)
    if sp.shape[1] != int(space_dimension):
        raise ValueError(
            f"space_dimension={space_dimension} but the points have "
            f"{sp.shape[1]} coordinates")
    lib = _lib.load()
    out = np.zeros(tp.shape[0], dtype=np.float64)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = lib.pacman_rbf_interpolate_c(
        int(space_dimension), int(execspace), int(rbf_function),
        sp.ctypes.data_as(dp), sp.shape[0], sv.ctypes.data_as(dp),
        tp.ctypes.data_as(dp), tp.shape[0], out.ctypes.data_as(dp))
    if rc != 0:
        raise RuntimeError(_lib.last_error())
    return out


def _ptr(a):
    """(address, on_device) of a numpy array or a torch tensor."""
    if hasattr(a, "data_ptr"):
        return ctypes.c_void_p(a.data_ptr()), int(a.is_cuda)
    return ctypes.c_void_p(a.ctypes.data), 0


def _check_points(p, dim, name):
    if len(p.shape) != 2 or int(p.shape[1]) != dim:
        raise ValueError(f"{name} must have shape (n, {dim}), got "
                         f"{tuple(p.shape)}")


def _fields_in(v, n, name, convert=True):
    """(contiguous float64 fields, F, squeeze) of a (F, n) or (n,) array or
    tensor; raises ValueError on a length or rank mismatch. A device tensor
    of another dtype is converted unless convert is False (graph replay keeps
    the caller's buffer)."""
    if len(v.shape) not in (1, 2) or int(v.shape[-1]) != n:
        raise ValueError(f"{name} must have shape (F, {n}) or ({n},), got "
                         f"{tuple(v.shape)}")
    squeeze = len(v.shape) == 1
    F = 1 if squeeze else int(v.shape[0])
    if F < 1:
        raise ValueError(f"{name} holds no field")
    if hasattr(v, "data_ptr"):
        import torch
        if v.dtype != torch.float64:
            if not convert:
                raise ValueError(f"{name} must be float64")
            v = v.to(torch.float64)
        v = v.contiguous()
    else:
        v = _as_f64(v)
    return v, F, squeeze


def _out_for(out, F, m, like, squeeze):
    """A new (F, m) float64 output in the place of `like`, or the caller's
    `out` after checking its shape, dtype, contiguity and place."""
    if out is None:
        if hasattr(like, "data_ptr"):
            import torch
            return torch.empty((F, m), dtype=torch.float64, device=like.device)
        return np.zeros((F, m))
    want = (m,) if squeeze else (F, m)
    if tuple(out.shape) not in (want, (F, m)):
        raise ValueError(f"out must have shape {want}, got {tuple(out.shape)}")
    if hasattr(out, "data_ptr"):
        import torch
        if out.dtype != torch.float64 or not out.is_contiguous():
            raise ValueError("out must be a contiguous float64 tensor")
    else:
        if (not isinstance(out, np.ndarray) or out.dtype != np.float64
                or not out.flags.c_contiguous or not out.flags.writeable):
            raise ValueError("out must be a writeable C-contiguous float64 "
                             "array")
    return out


class Plan:
    """Persistent RBF-PUM plan (setup once, apply to any number of fields).

    Points may be numpy arrays (host) or float64 torch tensors on the GPU;
    both point sets must live in the same place. ``support_radius <= 0``
    uses the patch radius as the Wendland support.
    """

    def __init__(self, space_dimension, rbf_function, source_points,
                 target_points, nodes_per_cluster=DEFAULT_NODES_PER_CLUSTER,
                 relative_overlap=DEFAULT_RELATIVE_OVERLAP,
                 support_radius=DEFAULT_SUPPORT_RADIUS, stream=None,
                 precision="fp64", max_refinements=6, _adaptive=None):
        self._lib = _lib.load()
        self._h = ctypes.c_void_p()
        self.dim = int(space_dimension)
        if int(rbf_function) not in _BASES:
            raise ValueError("unsupported rbf_function")
        src, tgt = source_points, target_points
        if hasattr(src, "data_ptr") != hasattr(tgt, "data_ptr"):
            raise ValueError("source and target points must both be numpy "
                             "arrays or both torch tensors")
        if not hasattr(src, "data_ptr"):
            src, tgt = _as_f64(src), _as_f64(tgt)
        else:
            src, tgt = src.contiguous(), tgt.contiguous()
            if str(src.dtype) != "torch.float64" or str(tgt.dtype) != "torch.float64":
                raise ValueError("points must be float64")
        if self.dim not in (1, 2, 3):
            raise RuntimeError(
                "The dimension of the points can only be: 1, 2 or 3.\n")
        _check_points(src, self.dim, "source_points")
        _check_points(tgt, self.dim, "target_points")
        self._keep = (src, tgt)
        self.num_sources = int(src.shape[0])
        self.num_targets = int(tgt.shape[0])
        ps, dev = _ptr(src)
        pt, devt = _ptr(tgt)
        if dev != devt:
            raise ValueError("source and target points must both be on the "
                             "host or both on the device")
        if _adaptive is None:
            rc = self._lib.pacman_rbf_plan_create(
                ctypes.byref(self._h), self.dim, int(rbf_function),
                int(nodes_per_cluster), float(relative_overlap),
                float(support_radius), ps, self.num_sources, pt,
                self.num_targets, dev, ctypes.c_void_p(stream or 0))
        else:
            leaf, nmin, nmax, scale = _adaptive
            rc = self._lib.pacman_rbf_plan_create_adaptive(
                ctypes.byref(self._h), self.dim, int(rbf_function), int(leaf),
                int(nmin), int(nmax), float(relative_overlap), float(scale),
                ps, self.num_sources, pt, self.num_targets, dev,
                ctypes.c_void_p(stream or 0))
        if rc != 0:
            raise RuntimeError(_lib.last_error())
        self.on_device = bool(dev)
        if precision != "fp64" or max_refinements != 6:
            self.set_precision(precision, max_refinements)

    _PRECISIONS = {"fp64": 0, "mixed": 1, "auto": 2}

    @classmethod
    def adaptive(cls, space_dimension, rbf_function, source_points,
                 target_points, leaf_points=40, min_points=20, max_points=600,
                 relative_overlap=DEFAULT_RELATIVE_OVERLAP, support_scale=1.0,
                 stream=None, precision="fp64", max_refinements=6):
        """Plan on density-aware patches (no reference counterpart): a
        binary tree over Morton codes halves cells holding more than
        leaf_points sources; every leaf with targets is a patch centred on
        its cell with radius max((1 + overlap) half-diagonal, distance to
        the min_points-th nearest source); patches above max_points sources
        are halved again; support = support_scale * radius."""
        return cls(space_dimension, rbf_function, source_points,
                   target_points, 1, relative_overlap, -1.0, stream,
                   precision, max_refinements,
                   _adaptive=(leaf_points, min_points, max_points,
                              support_scale))

    def patch_radii(self):
        r = np.zeros(self.num_patches)
        rc = self._lib.pacman_rbf_plan_patch_radii(
            self._h, r.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        if rc != 0:
            raise RuntimeError(_lib.last_error())
        return r

    def set_precision(self, precision="fp64", max_refinements=6):
        """Arithmetic of the patches above 264 points (the workspace solve):
        "fp64"; "mixed" (FP32 assembly and factorisation, FP64 iterative
        refinement with at most max_refinements steps, FP64 fallback per
        patch); or "auto" (per field count, whichever of the two the
        library's size-bin table measured faster, pacman_amd/autotune.py).
        Smaller patches are solved in FP64 either way."""
        if precision not in self._PRECISIONS:
            raise ValueError('precision must be "fp64", "mixed" or "auto"')
        rc = self._lib.pacman_rbf_plan_set_precision(
            self._h, self._PRECISIONS[precision], int(max_refinements))
        if rc != 0:
            raise RuntimeError(_lib.last_error())
        self.precision = precision

    def counters(self, stream=None) -> dict:
        """Counters of the last apply (synchronises): patches found
        numerically singular (duplicate sources; their values are not
        meaningful), refinement steps summed over patches, patches that fell
        back from FP32 refinement to FP64."""
        c = (ctypes.c_int * 3)()
        rc = self._lib.pacman_rbf_plan_apply_counters(
            self._h, c, ctypes.c_void_p(stream or 0))
        if rc != 0:
            raise RuntimeError(_lib.last_error())
        return {"singular": c[0], "refinements": c[1], "fallbacks": c[2]}

    def time_solve(self, values, size_bin, waves=8, ws_precision="fp64",
                   reps=5, stream=None) -> float:
        """Size-bin tuning hook: median milliseconds of `reps` launches of
        one class of the per-patch solve on this plan's patches, for the
        fields of `values` (device tensor, F x N, F <= 8): register bin
        `size_bin` >= 1 with `waves` waves per workgroup, or with
        size_bin = 0 the workspace path in `ws_precision` ("fp64" or
        "mixed"). -1 when the plan has no such patch or variant."""
        v, F, _ = _fields_in(values, self.num_sources, "values")
        pv, dev = _ptr(v)
        if not dev:
            raise ValueError("tuning needs device values")
        if F > 8:
            raise ValueError("tuning takes at most 8 fields")
        ms = ctypes.c_double(0.0)
        rc = self._lib.pacman_rbf_plan_time_solve(
            self._h, pv, F, int(size_bin), int(waves),
            {"fp64": 0, "mixed": 1}[ws_precision], int(reps),
            ctypes.byref(ms), ctypes.c_void_p(stream or 0))
        if rc != 0:
            raise RuntimeError(_lib.last_error())
        return float(ms.value)

    @property
    def num_large_patches(self) -> int:
        return int(self._lib.pacman_rbf_plan_num_large_patches(self._h))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.pacman_rbf_plan_destroy(h)
            self._h = ctypes.c_void_p()

    # -- apply -------------------------------------------------------------
    def apply(self, source_values, out=None, stream=None, graph=False):
        """Transfers F fields: values (F, N) or (N,) -> (F, M) or (M,).

        graph=True (device tensors only) replays the launch sequence as a HIP
        graph captured on the first call for these buffers: for repeated
        transfers that reuse the same value and output tensors.
        """
        v, F, squeeze = _fields_in(source_values, self.num_sources,
                                   "source_values", convert=not graph)
        out = _out_for(out, F, self.num_targets, v, squeeze)
        pv, dev = _ptr(v)
        po, devo = _ptr(out)
        if dev != devo:
            raise ValueError("values and output must live in the same place")
        if graph:
            if not dev:
                raise ValueError("graph replay needs device tensors")
            rc = self._lib.pacman_rbf_plan_apply_graph(
                self._h, pv, F, po, ctypes.c_void_p(stream or 0))
        else:
            rc = self._lib.pacman_rbf_plan_apply(self._h, pv, F, po, dev,
                                                 ctypes.c_void_p(stream or 0))
        if rc != 0:
            raise RuntimeError(_lib.last_error())
        return out.reshape(-1) if squeeze else out

    def apply_transpose(self, target_grad, out=None, stream=None):
        """Transpose of apply: (F, M) or (M,) -> (F, N) or (N,).

        The gradient of a loss with respect to the source values, given its
        gradient with respect to the target values.
        """
        g, F, squeeze = _fields_in(target_grad, self.num_targets,
                                   "target_grad")
        out = _out_for(out, F, self.num_sources, g, squeeze)
        pg, dev = _ptr(g)
        po, devo = _ptr(out)
        if dev != devo:
            raise ValueError("gradients must live in the same place")
        rc = self._lib.pacman_rbf_plan_apply_transpose(
            self._h, pg, F, po, dev, ctypes.c_void_p(stream or 0))
        if rc != 0:
            raise RuntimeError(_lib.last_error())
        return out.reshape(-1) if squeeze else out

    def apply_with_state(self, source_values, stream=None):
        """apply() that also returns the solver state the target-coordinate
        gradient needs: (values (F, M) or (M,), state). Device tensors only
        (the state is a float64 tensor on the same device)."""
        import torch
        v = source_values
        if not (hasattr(v, "data_ptr") and v.is_cuda):
            raise ValueError("apply_with_state needs device tensors")
        v, F, squeeze = _fields_in(v, self.num_sources, "source_values")
        out = torch.empty((F, self.num_targets), dtype=torch.float64,
                          device=v.device)
        n = int(self._lib.pacman_rbf_plan_state_size(self._h, F))
        if n < 0:
            raise RuntimeError(_lib.last_error())
        state = torch.empty(max(n, 1), dtype=torch.float64, device=v.device)
        rc = self._lib.pacman_rbf_plan_apply_state(
            self._h, v.data_ptr(), F, out.data_ptr(), state.data_ptr(),
            ctypes.c_void_p(stream or 0))
        if rc != 0:
            raise RuntimeError(_lib.last_error())
        return (out.reshape(-1) if squeeze else out), state

    def target_coord_grad(self, state, target_grad, stream=None):
        """Gradient of sum <target_grad, u> with respect to the target
        coordinates (M, dim), the patch structure held fixed; `state` from
        apply_with_state with the same field count. Device tensors only."""
        import torch
        g = target_grad
        if not (hasattr(g, "data_ptr") and g.is_cuda):
            raise ValueError("target_coord_grad needs device tensors")
        g, F, _ = _fields_in(g, self.num_targets, "target_grad")
        n = int(self._lib.pacman_rbf_plan_state_size(self._h, F))
        if (not hasattr(state, "data_ptr") or state.dtype != torch.float64
                or not state.is_contiguous() or state.device != g.device):
            raise ValueError("state must be a contiguous float64 tensor on "
                             "the device of target_grad")
        if state.numel() < n:
            raise ValueError("state is smaller than this field count needs")
        out = torch.empty((self.num_targets, self.dim), dtype=torch.float64,
                          device=g.device)
        rc = self._lib.pacman_rbf_plan_target_coord_grad(
            self._h, state.data_ptr(), F, g.data_ptr(), out.data_ptr(),
            ctypes.c_void_p(stream or 0))
        if rc != 0:
            raise RuntimeError(_lib.last_error())
        return out

    def build_operator(self, stream=None, precision="fp64"):
        """Keep the weight-scaled local operator of every patch in device
        memory; later applies read it instead of solving (see
        operator_bytes for the memory it takes). precision "fp32" stores it
        rounded to FP32 (half the memory and apply traffic, FP64 sums,
        about 1e-7 relative difference from "fp64")."""
        if precision not in ("fp64", "fp32"):
            raise ValueError('precision must be "fp64" or "fp32"')
        fn = (self._lib.pacman_rbf_plan_build_operator if precision == "fp64"
              else self._lib.pacman_rbf_plan_build_operator_fp32)
        rc = fn(self._h, ctypes.c_void_p(stream or 0))
        if rc != 0:
            raise RuntimeError(_lib.last_error())

    def drop_operator(self):
        self._lib.pacman_rbf_plan_drop_operator(self._h)

    @property
    def operator_bytes(self) -> int:
        return int(self._lib.pacman_rbf_plan_operator_bytes(self._h))

    # -- introspection -----------------------------------------------------
    @property
    def radius(self) -> float:
        return self._lib.pacman_rbf_plan_radius(self._h)

    @property
    def support(self) -> float:
        return self._lib.pacman_rbf_plan_support(self._h)

    @property
    def num_patches(self) -> int:
        return int(self._lib.pacman_rbf_plan_num_patches(self._h))

    @property
    def num_pairs(self) -> int:
        return int(self._lib.pacman_rbf_plan_num_pairs(self._h))

    @property
    def max_patch_size(self) -> int:
        return int(self._lib.pacman_rbf_plan_max_patch_size(self._h))

    def lattice(self):
        shape = (ctypes.c_longlong * 3)()
        nb = ctypes.c_longlong()
        self._lib.pacman_rbf_plan_lattice(self._h, shape, ctypes.byref(nb))
        return np.array(shape[:], dtype=np.int64), int(nb.value)

    def centers(self):
        K = self.num_patches
        c = np.zeros((K, 3))
        cand = np.zeros(K, np.int64)
        rc = self._lib.pacman_rbf_plan_centers(
            self._h, c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            cand.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)))
        if rc != 0:
            raise RuntimeError(_lib.last_error())
        return c, cand

    def csr(self, which: int):
        rows = self.num_patches if which in (0, 1) else self.num_targets
        size = self._lib.pacman_rbf_plan_csr_size(self._h, int(which))
        if size < 0:
            raise RuntimeError(_lib.last_error())
        off = np.zeros(rows + 1, np.int64)
        idx = np.zeros(max(size, 1), np.int64)
        llp = ctypes.POINTER(ctypes.c_longlong)
        rc = self._lib.pacman_rbf_plan_csr(self._h, int(which),
                                           off.ctypes.data_as(llp),
                                           idx.ctypes.data_as(llp))
        if rc != 0:
            raise RuntimeError(_lib.last_error())
        return off, idx[:size]

    PHASES = ("grids", "radius", "centres", "membership", "setup",
              "gather", "solve", "blend", "apply", "poly", "workspace",
              "tiles", "eval")
    WORK = ("patches", "sum_n", "sum_n2", "sum_n3", "sum_nm", "pairs",
            "max_n", "large", "large_sum_n3", "large_sum_n2", "large_sum_nm",
            "heavy_sum_nm", "tile_sum_n3", "tile_sum_n2", "tile_sum_nm")

    def phase_times(self) -> dict:
        """Milliseconds per phase (setup: host clock; apply: HIP events)."""
        t = np.zeros(len(self.PHASES))
        c = self._lib.pacman_rbf_plan_phase_times(
            self._h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(t))
        if c < 0:
            raise RuntimeError(_lib.last_error())
        return dict(zip(self.PHASES, t.tolist()))

    def work(self) -> dict:
        w = np.zeros(len(self.WORK))
        c = self._lib.pacman_rbf_plan_work(
            self._h, w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(w))
        if c < 0:
            raise RuntimeError(_lib.last_error())
        return dict(zip(self.WORK, w.tolist()))

    def patch_info(self):
        K = self.num_patches
        pv = np.zeros(K, np.int32)
        st = np.zeros(K, np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        rc = self._lib.pacman_rbf_plan_patch_info(self._h, pv.ctypes.data_as(ip),
                                                  st.ctypes.data_as(ip))
        if rc != 0:
            raise RuntimeError(_lib.last_error())
        return pv, st
