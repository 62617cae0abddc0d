"""The C++ host layer (include/gpufem.hpp) compiles against the C-ABI, and on
a GPU its test program -- written like the reference's FemTest/AxyGPUTest
suites -- passes."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "host_mirror_test.cpp")
BIN = os.path.join(ROOT, "tests", "cpp", "host_mirror_test")


def compile_to(out):
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
                    SRC, "-L" + os.path.join(ROOT, "gpufem_amd"), "-lgpufem",
                    "-Wl,-rpath," + os.path.join(ROOT, "gpufem_amd"), "-o", out], check=True)


def test_host_layer_compiles():
    with tempfile.TemporaryDirectory() as tmp:
        compile_to(os.path.join(tmp, "t"))


@pytest.mark.gpu
def test_host_layer_runs_reference_style_suite():
    assert os.path.exists(BIN), "build() compiles tests/cpp/host_mirror_test"
    out = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    for name in ["macro_test_2dstiffness_diag_nonzero", "macro_test_3dstiffness_diag_nonzero",
                 "macro_test_stiffness_constraint", "solve_q8_plate_jacobi",
                 "Ell_diagonal_Ax_y_GPU", "Ell_CGGPU"]:
        assert "PASS " + name in out.stdout
