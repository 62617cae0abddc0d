"""The C-ABI library loads on a machine without a GPU and exports every entry
point include/*.h declares; host-only parts (mesh generator, option
defaults, error reporting) work without a device."""
import ctypes as C
import glob
import os
import re

import pytest

import gpufem_amd as gf
from gpufem_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(gf_[A-Za-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_functions():
    assert len(declared_functions()) >= 25


@pytest.mark.parametrize("name", declared_functions())
def test_symbol_exported_and_bound(name):
    lib = gf.load()
    assert hasattr(lib, name), name
    assert name in _lib.SIGNATURES, "python binding missing for %s" % name


def test_no_device_reports_error_not_crash():
    lib = gf.load()
    h = C.c_void_p()
    code = lib.gf_create(C.byref(h), 0)
    if code == gf.GF_OK:            # a GPU is present: nothing to check here
        lib.gf_destroy(h)
        pytest.skip("device present")
    assert code < 0
    assert lib.gf_last_error(None)


def test_solve_opts_defaults():
    o = gf.solve_opts()
    assert o.precond == gf.GF_PRECOND_JACOBI and o.precision == gf.GF_PREC_FP64
    assert o.rtol == 1e-8 and o.maxit > 0


def test_grid_rejects_bad_arguments():
    with pytest.raises(gf.GpufemError):
        gf.Grid(99, 4, 4)
    with pytest.raises(gf.GpufemError):
        gf.Grid(gf.GF_BRICK8, 4, 4, 0)


def test_python_constants_match_header():
    """Sizes the bindings allocate for the library to fill must match the header."""
    text = open(os.path.join(ROOT, "include", "gpufem.h")).read()
    defines = dict(re.findall(r"#define (GF_[A-Z0-9_]+) \(?(-?\d+)\)?", text))
    assert int(defines["GF_SPMV_VARIANTS"]) == _lib.GF_SPMV_VARIANTS
    assert int(defines["GF_ASM_VARIANTS"]) == _lib.GF_ASM_VARIANTS
    for name in ("GF_OK", "GF_ERR_ARG", "GF_ERR_NOCONV", "GF_Q4", "GF_BRICK20", "GF_PREC_MIXED",
                 "GF_SCHED_PERSISTENT"):
        assert int(defines[name]) == getattr(_lib, name), name
