"""``import pacman``: the reference's module name for this library.

Callers of the reference's Python binding (src/pybindings/bindings.cpp:17-134,
e.g. tests/test_rbf-pum.py, which does ``import pacman`` and calls
``pacman.rbf.interpolate`` / ``pacman.execspaces.available()``) run unchanged:
this package is ``pacman_amd`` under that name, and ``pacman.rbf``,
``pacman.execspaces``, ``pacman.fe``, ``pacman.mls`` / ``pacman.MLS`` are the
same module objects as their ``pacman_amd`` counterparts.
"""

import sys as _sys

import pacman_amd as _impl
from pacman_amd import (MLS, autograd, execspaces, fe, finalize, initialize,
                        mls, rbf)

for _name in ("execspaces", "rbf", "fe", "mls", "autograd"):
    _mod = getattr(_impl, _name, None)
    if _mod is not None:
        _sys.modules[__name__ + "." + _name] = _mod

__all__ = list(_impl.__all__)
