"""Drop-in at the file level: the engine's temp_dir export read by the
reference's own dump_tool and histogram_tool (built from the reference
sources into oracle/_ref) gives the reference's goldens byte for byte."""
import hashlib
import os
import subprocess

import pytest

from tests.conftest import ROOT
from tests.kc_testlib import gpu_count, load_json, read_golden

pytestmark = pytest.mark.gpu
REF = os.path.join(ROOT, "oracle", "_ref")


def _tools():
    d, h = os.path.join(REF, "dump_tool_ref"), os.path.join(REF, "histogram_tool_ref")
    if not (os.path.exists(d) and os.path.exists(h)):
        pytest.skip("reference tools not built (oracle/_ref needs the reference sources)")
    return d, h


def run_tools(tmp, hist_max):
    d, h = _tools()
    subprocess.run([d, "precise", tmp + "/", "8", "4", tmp + "/dump.txt", "1", "4294967295", "--sort"],
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, check=True, timeout=300)
    subprocess.run([h, "precise", tmp + "/", "8", "4", tmp + "/hist.txt", "1", str(hist_max)],
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, check=True, timeout=300)
    with open(tmp + "/dump.txt", "rb") as f:
        dump = f.read()
    with open(tmp + "/hist.txt", "rb") as f:
        hist = f.read()
    return dump, hist


@pytest.mark.parametrize("bloom", [False, True], ids=["plain", "bloom"])
def test_reference_tools_read_our_export(ctx, tmp_path, bloom):
    from ktcountermemorygreedy_amd import workloads as W
    _tools()
    g = load_json("count_5mbp.json")
    seq = W.COUNT_5MBP.reads(fastq=False).tobytes()
    t = gpu_count(ctx, seq, 31, g["count_max"], bloom=bloom)
    t.export_dir(str(tmp_path))
    t.close()
    dump, hist = run_tools(str(tmp_path), 10000)
    assert hashlib.sha256(dump).hexdigest() == g["dump_sha256"]
    assert hashlib.sha256(hist).hexdigest() == g["hist_sha256"]


@pytest.mark.parametrize("case", [c for c in load_json("tiny_cases.json") if c["input"] != "tiny_long.fastq"],
                         ids=lambda c: "%s-k%d-c%d" % (c["input"], c["k"], c["count_max"]))
def test_reference_tools_tiny(ctx, tmp_path, case):
    from ktcountermemorygreedy_amd import kc
    _tools()
    seq = kc.fastq_frame(read_golden(case["input"]))
    t = gpu_count(ctx, seq, case["k"], case["count_max"], bloom=True)
    t.export_dir(str(tmp_path))
    t.close()
    dump, hist = run_tools(str(tmp_path), case["hist_max"])
    assert hashlib.sha256(dump).hexdigest() == case["dump_sha256"]
    assert hist == read_golden(case["hist"])
