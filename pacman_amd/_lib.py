"""ctypes loader for libpacman_rbf.so (the HIP implementation).

The library is built in-tree by ``pacman_amd.build``. There is no CPU
fallback: if the library is missing, loading raises.
"""

from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpacman_rbf.so")

_dp = ctypes.POINTER(ctypes.c_double)
_llp = ctypes.POINTER(ctypes.c_longlong)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p

# (name, restype, argtypes) for every symbol the public headers declare.
SIGNATURES = {
    # include/fortran_interface.h
    "pacman_kokkos_initialize_c": (None, []),
    "pacman_kokkos_finalize_c": (None, []),
    "pacman_best_execspace_c": (ctypes.c_int, []),
    "pacman_rbf_interpolate_c": (ctypes.c_int, [
        ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp,
        ctypes.c_int, _dp]),
    "pacman_fe_interpolate_c": (ctypes.c_int, [
        ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp,
        _ip, ctypes.c_int, _ip, ctypes.c_int, _ip, _dp, ctypes.c_int, _dp,
        _ip]),
    # include/pacman_rbf_plan.h
    "pacman_rbf_plan_create": (ctypes.c_int, [
        ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.c_int,
        ctypes.c_double, ctypes.c_double, _vp, ctypes.c_longlong, _vp,
        ctypes.c_longlong, ctypes.c_int, _vp]),
    "pacman_rbf_plan_apply": (ctypes.c_int, [
        _vp, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp]),
    "pacman_rbf_plan_apply_transpose": (ctypes.c_int, [
        _vp, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp]),
    "pacman_rbf_plan_apply_graph": (ctypes.c_int, [
        _vp, _vp, ctypes.c_int, _vp, _vp]),
    "pacman_rbf_plan_apply_noalloc": (ctypes.c_int, [
        _vp, _vp, ctypes.c_int, _vp, _vp]),
    "pacman_rbf_plan_state_size": (ctypes.c_longlong, [_vp, ctypes.c_int]),
    "pacman_rbf_plan_apply_state": (ctypes.c_int, [
        _vp, _vp, ctypes.c_int, _vp, _vp, _vp]),
    "pacman_rbf_plan_target_coord_grad": (ctypes.c_int, [
        _vp, _vp, ctypes.c_int, _vp, _vp, _vp]),
    "pacman_rbf_plan_destroy": (ctypes.c_int, [_vp]),
    "pacman_rbf_plan_radius": (ctypes.c_double, [_vp]),
    "pacman_rbf_plan_support": (ctypes.c_double, [_vp]),
    "pacman_rbf_plan_num_patches": (ctypes.c_longlong, [_vp]),
    "pacman_rbf_plan_num_pairs": (ctypes.c_longlong, [_vp]),
    "pacman_rbf_plan_max_patch_size": (ctypes.c_int, [_vp]),
    "pacman_rbf_plan_num_large_patches": (ctypes.c_longlong, [_vp]),
    "pacman_rbf_plan_create_adaptive": (ctypes.c_int, [
        ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.c_int,
        ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _vp,
        ctypes.c_longlong, _vp, ctypes.c_longlong, ctypes.c_int, _vp]),
    "pacman_rbf_plan_patch_radii": (ctypes.c_int, [_vp, _dp]),
    "pacman_rbf_plan_set_precision": (ctypes.c_int, [_vp, ctypes.c_int,
                                                     ctypes.c_int]),
    "pacman_rbf_plan_time_solve": (ctypes.c_int, [
        _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
        ctypes.c_int, _dp, _vp]),
    "pacman_rbf_plan_apply_counters": (ctypes.c_int, [_vp, _ip, _vp]),
    "pacman_rbf_plan_lattice": (ctypes.c_int, [_vp, _llp, _llp]),
    "pacman_rbf_plan_centers": (ctypes.c_int, [_vp, _dp, _llp]),
    "pacman_rbf_plan_csr_size": (ctypes.c_longlong, [_vp, ctypes.c_int]),
    "pacman_rbf_plan_csr": (ctypes.c_int, [_vp, ctypes.c_int, _llp, _llp]),
    "pacman_rbf_plan_patch_info": (ctypes.c_int, [_vp, _ip, _ip]),
    "pacman_rbf_plan_phase_times": (ctypes.c_int, [_vp, _dp, ctypes.c_int]),
    "pacman_rbf_plan_work": (ctypes.c_int, [_vp, _dp, ctypes.c_int]),
    "pacman_fe_nearest_device": (ctypes.c_int, [
        ctypes.c_int, _vp, ctypes.c_longlong, _vp, _vp, ctypes.c_longlong,
        _vp, _vp, _vp]),
    "pacman_mls_interpolate_c": (ctypes.c_int, [
        ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp,
        ctypes.c_int, _vp]),
    "pacman_mls_plan_create": (ctypes.c_int, [
        _vp, ctypes.c_int, _vp, ctypes.c_longlong, _vp, ctypes.c_longlong,
        ctypes.c_int, _vp]),
    "pacman_mls_plan_apply": (ctypes.c_int, [
        _vp, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp]),
    "pacman_mls_plan_k": (ctypes.c_int, [_vp]),
    "pacman_mls_plan_neighbours": (ctypes.c_int, [_vp, _vp, _vp]),
    "pacman_mls_plan_destroy": (ctypes.c_int, [_vp]),
    "pacman_rbf_plan_build_operator": (ctypes.c_int, [_vp, _vp]),
    "pacman_rbf_plan_build_operator_fp32": (ctypes.c_int, [_vp, _vp]),
    "pacman_rbf_plan_drop_operator": (ctypes.c_int, [_vp]),
    "pacman_rbf_plan_operator_bytes": (ctypes.c_longlong, [_vp]),
    "pacman_release_device_cache": (ctypes.c_int, []),
    "pacman_device_cache_bytes": (ctypes.c_longlong, []),
    "pacman_rbf_last_error": (ctypes.c_char_p, []),
}

_lib = None


def load() -> ctypes.CDLL:
    """Loads the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        # torch bundles its own libamdhip64.so.7; whichever copy is loaded
        # first serves the whole process. Load torch's first so torch and
        # this library share one HIP runtime (torch fails to initialise on
        # the other copy).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with "
                "`python -m pacman_amd.build` (hipcc, gfx950)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def last_error() -> str:
    msg = load().pacman_rbf_last_error()
    return msg.decode() if msg else ""
