"""The C++ adaptors (include/kc/kc.hpp) compile against the C ABI and, on a
GPU, reproduce the reference interfaces' results."""
import os
import subprocess

import pytest

from tests.conftest import ROOT

SRC = os.path.join(ROOT, "tests", "cpp", "adaptor_check.cpp")
PKG = os.path.join(ROOT, "ktcountermemorygreedy_amd")


def test_adaptor_header_compiles():
    subprocess.check_call(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), SRC])


@pytest.mark.gpu
def test_adaptor_runs(tmp_path):
    exe = str(tmp_path / "adaptor_check")
    subprocess.check_call(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"), SRC, "-o", exe,
                           "-L", PKG, "-l:libkc.so", "-Wl,-rpath," + PKG])
    out = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip().endswith("ok"), out.stdout
