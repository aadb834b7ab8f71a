"""Execution-space selectors (src/pybindings/bindings.cpp:38-73).

This build provides the HIP execution space only; the other selectors are
kept so callers written for the reference module keep importing.
"""

SERIAL = 0x00
OPENMP = 0x01
THREADS = 0x02
CUDA = 0x03
HIP = 0x04
SYCL = 0x05

_NAMES = {SERIAL: "SERIAL", OPENMP: "OPENMP", THREADS: "THREADS",
          CUDA: "CUDA", HIP: "HIP", SYCL: "SYCL"}


def available() -> list:
    """Compiled execution spaces (the reference lists its Kokkos build)."""
    return ["HIP"]


def name(execspace: int) -> str:
    return _NAMES.get(int(execspace), f"<unknown {execspace}>")
