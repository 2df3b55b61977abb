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


def test_guard_maps_exceptions_to_status(tmp_path):
    """The C-ABI exception guard turns allocation and thread-start failures
    into KC_ENOMEM and passes ordinary returns through (host only)."""
    src = os.path.join(ROOT, "tests", "cpp", "guard_check.cpp")
    exe = str(tmp_path / "guard_check")
    subprocess.check_call(["g++", "-std=c++17", "-Wall", "-Werror", "-O1", "-I", os.path.join(PKG, "csrc"), src,
                           "-o", exe])
    out = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout


def test_admission_ledger(tmp_path):
    """The serving loop's HBM admission (kc_admit.h): reservations are
    returned exactly once, a request refused before it ran releases nothing,
    and resident bytes return to their start after the tables are dropped
    (host only)."""
    src = os.path.join(ROOT, "tests", "cpp", "admit_check.cpp")
    exe = str(tmp_path / "admit_check")
    subprocess.check_call(["g++", "-std=c++17", "-Wall", "-Werror", "-O1", "-I", os.path.join(PKG, "csrc"), src,
                           "-o", exe])
    out = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout


@pytest.mark.gpu
def test_adaptor_runs(tmp_path):
    exe = str(tmp_path / "adaptor_check")
    subprocess.check_call(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"), SRC, "-o", exe,
                           "-L", PKG, "-l:libkc.so", "-Wl,-rpath," + PKG])
    d = tmp_path / "tmpdir"
    d.mkdir()
    env = dict(os.environ, KC_ADAPTOR_DIR=str(d))
    out = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120, env=env)
    assert out.returncode == 0 and out.stdout.strip().endswith("ok"), out.stdout
