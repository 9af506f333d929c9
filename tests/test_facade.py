"""The C++ libigl-signature facade (include/gwn_igl.h) compiles against the C
ABI and, on the GPU, reproduces the reference's known answers: through its
generic conversion layer (facade_test.cpp, a column-major mock matrix) and
through its Eigen-signature overloads (eigen_facade_test.cpp, compiled
against the mock Eigen/Core in tests/cpp/mock_eigen since Eigen is absent)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
LIBDIR = os.path.join(ROOT, "libigl_amd")


def _build(name, extra=()):
    src, exe = os.path.join(CPP, name + ".cpp"), os.path.join(CPP, name)
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"), *extra, src, "-o",
           exe, "-L" + LIBDIR, "-lgwn", "-Wl,-rpath," + LIBDIR]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


EIGEN = ("-I" + os.path.join(CPP, "mock_eigen"),)


def test_facade_compiles_and_links():
    assert os.path.exists(_build("facade_test"))
    assert os.path.exists(_build("eigen_facade_test", EIGEN))


@pytest.mark.gpu
@pytest.mark.parametrize("name,extra", [("facade_test", ()), ("eigen_facade_test", EIGEN)])
def test_facade_runs_on_gpu(name, extra):
    exe = _build(name, extra)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and f"{name}: OK" in r.stdout
