"""MI355X-native RBF-PUM field transfer with the reference's Python API.

Drop-in for the reference module ``pacman`` on its RBF-PUM path
(src/pybindings/bindings.cpp:17-93): ``initialize``, ``finalize``,
``execspaces`` and ``rbf`` (``interpolate``, ``functions``), so
``import pacman_amd as pacman`` keeps callers of that path working. ``fe``
serves the nearest-source FE method (NEAREST_NEAREST) on the same search,
``MLS`` (``mls``) the moving-least-squares transfer;
``autograd.transfer`` makes a plan's transfer differentiable with respect to
the source values and the target coordinates. The
compute runs in hand-written HIP kernels for gfx950 inside
``libpacman_rbf.so``; there is no CPU fallback.
"""

from . import execspaces
from . import fe
from . import mls
from . import rbf
from . import _lib

MLS = mls  # the reference's submodule name (bindings.cpp:124)

try:  # the differentiable transfer needs torch
    from . import autograd
except ImportError:  # pragma: no cover
    autograd = None

__all__ = ["autograd", "execspaces", "fe", "mls", "MLS", "rbf", "initialize",
           "finalize"]


def initialize() -> None:
    """Counterpart of pacman.initialize (bindings.cpp:26-30): device init."""
    _lib.load().pacman_kokkos_initialize_c()


def finalize() -> None:
    """Counterpart of pacman.finalize (bindings.cpp:31-35)."""
    _lib.load().pacman_kokkos_finalize_c()
