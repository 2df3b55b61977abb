"""Builds libkc.so in-tree for gfx950 (hipcc) and returns its path.

Run as ``python -m ktcountermemorygreedy_amd.build`` or via
``__graft_entry__.build()``.  Always recompiles.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libkc.so")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 ops stay off for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
HIP_SOURCES = [
    "kc_count.hip",
    "kc_sort.hip",
    "kc_part.hip",
    "kc_dump.hip",
    "kc_batch.hip",
    "kc_tools.hip",
    "kc_noalloc.hip",
]
CXX_SOURCES = ["kc_host.cpp", "kc_server.cpp"]
# command-line tools over libkc.so, with the reference tools' argv
TOOLS = ["histogram_tool", "dump_tool"]
BIN = os.path.join(PKG, "bin")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build step failed: " + " ".join(cmd))


def build(jobs: int = 4) -> str:
    os.makedirs(OUT, exist_ok=True)
    objs, procs = [], []
    for src in HIP_SOURCES:
        obj = os.path.join(OUT, src + ".o")
        objs.append(obj)
        procs.append(["hipcc", *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    for src in CXX_SOURCES:
        obj = os.path.join(OUT, src + ".o")
        objs.append(obj)
        procs.append(["g++", "-O3", "-std=c++17", "-fPIC", "-Wall", "-c", os.path.join(CSRC, src), "-o", obj])
    running = [subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True) for c in procs]
    failed = []
    for c, p in zip(procs, running):
        out, _ = p.communicate()
        out = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l)
        if p.returncode != 0:
            failed.append((c, out))
    if failed:
        for c, out in failed:
            sys.stderr.write(" ".join(c) + "\n" + out + "\n")
        raise RuntimeError("libkc build failed")
    tmp = LIB + ".tmp"
    _run(["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lz", "-pthread"])
    os.replace(tmp, LIB)
    os.makedirs(BIN, exist_ok=True)
    for tool in TOOLS:
        exe = os.path.join(BIN, tool)
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-o", exe + ".tmp", os.path.join(PKG, "tools", tool + ".cpp"),
              "-L", PKG, "-l:libkc.so", "-Wl,-rpath,$ORIGIN/.."])
        os.replace(exe + ".tmp", exe)
    return LIB


if __name__ == "__main__":
    print(build())
