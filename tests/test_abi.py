"""C-ABI / drop-in boundary checks that need no GPU: the library loads, it
exports every symbol the public headers declare, the constants equal the
reference's, and argument validation fails the way the reference does
(before any device work)."""

import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import pacman_amd as pacman
from pacman_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def header_functions():
    names = set()
    for h in os.listdir(INCLUDE):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(pacman_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def exported():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_library_exports_every_declared_symbol():
    syms = exported()
    declared = header_functions()
    assert len(declared) >= 18
    missing = sorted(declared - syms)
    assert not missing, missing
    # the C++ drop-in PACMAN::rbf_interpolate(int, uchar, uchar, double*, int,
    # double*, double*, int)
    assert "_ZN6PACMAN15rbf_interpolateEihhPdiS0_S0_i" in syms


def test_ctypes_signature_table_covers_headers():
    assert header_functions() == set(_lib.SIGNATURES)
    lib = _lib.load()
    for name in _lib.SIGNATURES:
        getattr(lib, name)


def test_constants_match_reference():
    # src/common/execspaces.hpp:17-24, src/interface.hpp:25-31,
    # src/fortran_interface.h:33-49
    es = pacman.execspaces
    assert (es.SERIAL, es.OPENMP, es.THREADS, es.CUDA, es.HIP, es.SYCL) == \
        (0, 1, 2, 3, 4, 5)
    f = pacman.rbf.functions
    assert (f.WENDLANDC0, f.WENDLANDC2, f.WENDLANDC4, f.WENDLANDC6,
            f.WENDLANDC8) == (0x10, 0x11, 0x12, 0x13, 0x14)
    hdr = open(os.path.join(INCLUDE, "fortran_interface.h")).read()
    for name, val in (("PACMAN_HIP", "0x04"), ("PACMAN_SERIAL", "0x00"),
                      ("PACMAN_WENDLANDC0", "0x10"),
                      ("PACMAN_WENDLANDC8", "0x14")):
        assert re.search(rf"{name}\s*=\s*{val}", hdr)
    assert es.available() == ["HIP"]


P2 = np.zeros((4, 2))
P3 = np.zeros((4, 3))
V4 = np.zeros(4)


@pytest.mark.parametrize("args,exc", [
    ((3, 4, 0x11, np.zeros(12), V4, P3), ValueError),       # 1-d points
    ((3, 4, 0x11, P3, V4, np.zeros(12)), ValueError),       # 1-d targets
    ((3, 4, 0x11, P3, V4, P2), ValueError),                 # dim mismatch
    ((3, 4, 0x11, P3, np.zeros(5), P3), ValueError),        # count mismatch
    ((3, 0, 0x11, P3, V4, P3), ValueError),                 # SERIAL absent
    ((3, 9, 0x11, P3, V4, P3), ValueError),                 # unknown space
    ((3, 4, 0x33, P3, V4, P3), ValueError),                 # unknown basis
    ((4, 4, 0x11, np.zeros((4, 4)), V4, np.zeros((4, 4))), RuntimeError),
])
def test_python_argument_validation(args, exc):
    with pytest.raises(exc):
        pacman.rbf.interpolate(*args)


def test_c_abi_error_codes():
    lib = _lib.load()
    dp = ctypes.POINTER(ctypes.c_double)
    p = P3.ctypes.data_as(dp)
    v = V4.ctypes.data_as(dp)
    out = np.zeros(4)
    o = out.ctypes.data_as(dp)
    # bad dimension, execution space not in this build, unknown basis
    assert lib.pacman_rbf_interpolate_c(4, 4, 0x11, p, 4, v, p, 4, o) == -1
    assert lib.pacman_rbf_interpolate_c(3, 1, 0x11, p, 4, v, p, 4, o) == -1
    assert "not enabled" in _lib.last_error()
    assert lib.pacman_rbf_interpolate_c(3, 4, 0x42, p, 4, v, p, 4, o) == -1
    # null pointers with positive counts (src/rbf_interface.cpp:109-114)
    assert lib.pacman_rbf_interpolate_c(3, 4, 0x11, None, 4, v, p, 4, o) == -1
    assert "sourcePoints is null" in _lib.last_error()
    assert lib.pacman_rbf_interpolate_c(3, 4, 0x11, p, 4, None, p, 4, o) == -1
    assert lib.pacman_rbf_interpolate_c(3, 4, 0x11, p, 4, v, None, 4, o) == -1
    # zero targets: nothing to compute, success
    assert lib.pacman_rbf_interpolate_c(3, 4, 0x11, p, 4, v, p, 0, o) == 0


def test_plan_create_validation():
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.pacman_rbf_plan_create(ctypes.byref(h), 3, 0x99, 50, 0.15, 0.1,
                                      None, 0, None, 0, 0, None) == -1
    assert "RBF function" in _lib.last_error()
    assert lib.pacman_rbf_plan_create(ctypes.byref(h), 0, 0x11, 50, 0.15, 0.1,
                                      None, 0, None, 0, 0, None) == -1
    assert not h.value


def test_fe_constants_and_validation():
    # src/common/types.hpp:29-45, src/fortran_interface.h:51-59
    m = pacman.fe.methods
    assert (m.NEAREST_NEAREST, m.INTERP_CLAMP, m.INTERP_NEAREST,
            m.INTERP_ZEROFILL, m.INTERP_EXTRAP) == (0xF0, 0xF1, 0xF2, 0xF3,
                                                    0xF4)
    hdr = open(os.path.join(INCLUDE, "fortran_interface.h")).read()
    assert re.search(r"PACMAN_FE_NEAREST_NEAREST\s*=\s*0xF0", hdr)
    assert re.search(r"PACMAN_MLS\s*=\s*0xF6", hdr)
    with pytest.raises(ValueError):   # methods other than NEAREST_NEAREST
        pacman.fe.interpolate(3, 4, m.INTERP_CLAMP, P3, V4, None, None, None,
                              P3)
    with pytest.raises(ValueError):   # dimension mismatch
        pacman.fe.interpolate(3, 4, m.NEAREST_NEAREST, P3, V4, None, None,
                              None, P2)
    with pytest.raises(RuntimeError):
        pacman.fe.interpolate(4, 4, m.NEAREST_NEAREST, np.zeros((4, 4)), V4,
                              None, None, None, np.zeros((4, 4)))
    lib = _lib.load()
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    p = P3.ctypes.data_as(dp)
    v = V4.ctypes.data_as(dp)
    out = np.zeros(4)
    st = np.zeros(4, np.int32)
    o = out.ctypes.data_as(dp)
    s = st.ctypes.data_as(ip)
    assert lib.pacman_fe_interpolate_c(3, 4, 0xF1, p, 4, v, None, 0, None, 0,
                                       None, p, 4, o, s) == -1
    assert "not enabled" in _lib.last_error()
    assert lib.pacman_fe_interpolate_c(3, 1, 0xF0, p, 4, v, None, 0, None, 0,
                                       None, p, 4, o, s) == -1
    assert lib.pacman_fe_interpolate_c(3, 4, 0xF0, None, 4, v, None, 0, None,
                                       0, None, p, 4, o, s) == -1
    assert "sourcePoints is null" in _lib.last_error()
    assert lib.pacman_fe_interpolate_c(3, 4, 0xF0, p, 4, v, None, 0, None, 0,
                                       None, p, 0, o, s) == 0


def test_pacman_alias_is_the_same_package():
    """``import pacman`` (the reference's module name) gives this library."""
    import pacman as ref_name
    import pacman.rbf
    assert ref_name.rbf is pacman.rbf is pacman_amd_rbf()
    assert ref_name.execspaces.available() == ["HIP"]
    assert ref_name.rbf.functions.WENDLANDC8 == 0x14
    assert ref_name.MLS is ref_name.mls


def pacman_amd_rbf():
    import pacman_amd
    return pacman_amd.rbf


def test_plan_point_shapes_validated():
    rng = np.random.default_rng(0)
    S, T = rng.uniform(size=(50, 3)), rng.uniform(size=(40, 3))
    with pytest.raises(ValueError):
        pacman.rbf.Plan(2, 0x11, S, T)  # 3 columns, dimension 2
    with pytest.raises(ValueError):
        pacman.rbf.Plan(3, 0x11, S, T[:, :2])
    with pytest.raises(ValueError):
        pacman.rbf.Plan(3, 0x11, S.ravel(), T)
    with pytest.raises(RuntimeError):
        pacman.rbf.Plan(4, 0x11, rng.uniform(size=(50, 4)),
                        rng.uniform(size=(40, 4)))


def test_plan_field_and_output_checks():
    """The helpers every Plan method runs before passing pointers to C."""
    from pacman_amd.rbf import _fields_in, _out_for
    v, F, sq = _fields_in(np.arange(10.0), 10, "v")
    assert (F, sq) == (1, True) and v.dtype == np.float64
    v, F, sq = _fields_in(np.zeros((3, 10), np.float32), 10, "v")
    assert (F, sq) == (3, False) and v.dtype == np.float64
    for bad in (np.zeros(9), np.zeros((2, 11)), np.zeros((2, 2, 10)),
                np.zeros((0, 10))):
        with pytest.raises(ValueError):
            _fields_in(bad, 10, "v")
    like = np.zeros((2, 10))
    assert _out_for(None, 2, 7, like, False).shape == (2, 7)
    _out_for(np.zeros((2, 7)), 2, 7, like, False)
    _out_for(np.zeros(7), 1, 7, like, True)
    for bad in (np.zeros((2, 6)), np.zeros((2, 7), np.float32),
                np.zeros((7, 2)).T, np.zeros((3, 7))):
        with pytest.raises(ValueError):
            _out_for(bad, 2, 7, like, False)
    torch = pytest.importorskip("torch")
    with pytest.raises(ValueError):  # graph replay keeps the caller's buffer
        _fields_in(torch.zeros(3, 10, dtype=torch.float32), 10, "v",
                   convert=False)
    with pytest.raises(ValueError):
        _out_for(torch.zeros(2, 7, dtype=torch.float32), 2, 7, like, False)
