"""The C ABI library loads on the CPU and exports every entry point that
include/sembafdtd.h declares; creating a grid without a gfx950 GPU fails
loudly instead of falling back to the CPU."""
import ctypes as C
import os
import re

import pytest

import semba_fdtd_amd as sf
from semba_fdtd_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "sembafdtd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fdtd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_hot_path():
    syms = declared_symbols()
    for s in ("fdtd_create", "fdtd_advance_e", "fdtd_advance_h", "fdtd_advance_e_yee",
              "fdtd_advance_e_cpml", "fdtd_advance_e_dispersive", "fdtd_advance_h_cpml",
              "fdtd_run", "fdtd_set_cpml", "fdtd_set_edispersive"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_fortran_module_binds_the_header():
    """The iso_c_binding module (the Fortran callers' side) binds every entry
    point except the ones returning C strings, which it wraps as
    fdtd_error_message."""
    text = open(os.path.join(ROOT, "semba_fdtd_amd", "fortran", "sembafdtd_gpu.F90")).read()
    bound = set(re.findall(r'name="(fdtd_[a-z0-9_]+)"', text))
    want = set(declared_symbols()) - {"fdtd_batch_last_error"}
    assert not (want - bound), sorted(want - bound)


def test_abi_version():
    assert sf.load_library().fdtd_abi_version() == 1


def test_create_without_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = sf.load_library()
    d = _lib.FdtdGridDesc(8, 8, 8, 0, 0, 0, None)
    h = C.c_void_p()
    rc = lib.fdtd_create(C.byref(d), C.byref(h))
    assert rc == -2 and not h.value
    assert b"device" in lib.fdtd_last_error(None)
    with pytest.raises(sf.FdtdError):
        sf.Solver(sf.cavity_problem(n=8, nsteps=4))


def test_bad_arguments_are_rejected_without_a_device():
    lib = sf.load_library()
    assert lib.fdtd_create(None, None) == -1
    d = _lib.FdtdGridDesc(0, 8, 8, 0, 0, 0, None)
    h = C.c_void_p()
    assert lib.fdtd_create(C.byref(d), C.byref(h)) == -1
    d = _lib.FdtdGridDesc(8, 8, 8, 7, 0, 0, None)
    assert lib.fdtd_create(C.byref(d), C.byref(h)) == -1
