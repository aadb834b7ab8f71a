"""Builds libpacman_rbf.so (HIP kernels + C ABI) in-tree for gfx950.

Run as ``python -m pacman_amd.build`` or through ``__graft_entry__.build()``.
Each translation unit is compiled by hipcc with its own floating-point
contraction policy:

* rbf_geometry.hip, rbf_plan.hip, capi.hip, fe_nearest.hip, mls.hip:
  ``-ffp-contract=off`` - distances, lattice coordinates and probe points
  must round exactly as written so the patch membership and nearest
  neighbours are bit-identical to the CPU restatement;
* rbf_solve.hip, rbf_grad.hip and the size-bin units rbf_solve_tile_a.hip
  ... rbf_solve_tile_g.hip: FMA contraction allowed (dense arithmetic);
  rbf_solve.hip also keeps MFMA accumulators in ordinary VGPRs
  (-amdgpu-mfma-vgpr-form: the stored-operator contraction otherwise copies
  them to and from AGPRs every step; 16 fields 7.9 -> 7.4 ms per apply).

All HIP code is compiled with the packed-FP32 instructions disabled.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpacman_rbf.so")
ARCH = os.environ.get("PACMAN_AMD_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
]

UNITS = {
    "rbf_geometry.hip": ["-ffp-contract=off"],
    "rbf_plan.hip": ["-ffp-contract=off"],
    "rbf_adaptive.hip": ["-ffp-contract=off"],
    "capi.hip": ["-ffp-contract=off"],
    "fe_nearest.hip": ["-ffp-contract=off"],
    "mls.hip": ["-ffp-contract=off"],
    "dev_alloc.hip": [],
    "host_copy.hip": [],
    "rbf_solve.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
    "rbf_grad.hip": [],
    "rbf_solve_ws.hip": [],
    "rbf_solve_tile_a.hip": [],
    "rbf_solve_tile_b.hip": [],
    "rbf_solve_tile_c.hip": [],
    "rbf_solve_tile_d.hip": [],
    "rbf_solve_tile_e.hip": [],
    "rbf_solve_tile_f.hip": [],
    "rbf_solve_tile_g.hip": [],
}


def hipcc() -> str:
    for c in (os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin",
                           "hipcc"), shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found; the HIP build needs ROCm")


def _stale(src: str, obj: str) -> bool:
    """True when obj is older than src or any header of csrc/."""
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [os.path.join(CSRC, src)] + [
        os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".hpp")]
    deps += [os.path.join(HERE, "..", "include", h)
             for h in os.listdir(os.path.join(HERE, "..", "include"))]
    return any(os.path.getmtime(d) > t for d in deps)


def _compile(src: str, extra: list[str], incremental: bool = False) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    if incremental and not _stale(src, obj):
        return obj
    cmd = [hipcc(), *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False, incremental: bool = False) -> str:
    """Compiles every unit (incremental=True: only units older than their
    source or a header) and links the library."""
    os.makedirs(BUILD, exist_ok=True)
    workers = max(1, min(len(UNITS), os.cpu_count() or 1))
    # The solve units take longest: start them first.
    order = sorted(UNITS.items(), key=lambda kv: not kv[0].startswith(
        "rbf_solve_tile"))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(lambda kv: _compile(*kv, incremental), order))
    tmp = LIB + ".tmp"
    rocm_lib = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs,
           "-L" + rocm_lib, "-lrocprofiler-sdk-roctx",
           "-Wl,-rpath," + rocm_lib, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True, incremental="--incremental" in sys.argv)
    sys.exit(0)
