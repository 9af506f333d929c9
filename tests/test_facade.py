"""The C++ libigl-signature facade (include/gwn_igl.h) compiles against the C
ABI and, on the GPU, reproduces the reference's known answers."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "facade_test.cpp")
EXE = os.path.join(ROOT, "tests", "cpp", "facade_test")
LIBDIR = os.path.join(ROOT, "libigl_amd")


def _build():
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-I" + os.path.join(ROOT, "include"), SRC, "-o", EXE,
           "-L" + LIBDIR, "-lgwn", "-Wl,-rpath," + LIBDIR]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return EXE


def test_facade_compiles_and_links():
    assert os.path.exists(_build())


@pytest.mark.gpu
def test_facade_runs_on_gpu():
    exe = _build()
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "facade_test: OK" in r.stdout
