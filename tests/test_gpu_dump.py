"""Device dump (dump_tool precise --sort text built on the GPU) against the
reference's goldens, byte for byte."""
import hashlib

import numpy as np
import pytest

from oracle import oracle as O
from tests.kc_testlib import gpu_count, load_json, read_golden, views_from_export

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("path", ["direct", "partitioned"])
@pytest.mark.parametrize("bloom", [False, True], ids=["plain", "bloom"])
@pytest.mark.parametrize("case", load_json("tiny_cases.json"),
                         ids=lambda c: "%s-k%d-c%d" % (c["input"], c["k"], c["count_max"]))
def test_tiny_device_dump(ctx, case, bloom, path):
    from ktcountermemorygreedy_amd import kc
    seq = kc.fastq_frame(read_golden(case["input"]))
    t = gpu_count(ctx, seq, case["k"], case["count_max"], bloom=bloom, path=path)
    dump = t.dump()
    assert hashlib.sha256(dump).hexdigest() == case["dump_sha256"]
    # a second dump sees the same table (the merge flags are cleared)
    assert t.dump() == dump
    t.close()


@pytest.mark.parametrize("bloom", [False, True], ids=["plain", "bloom"])
def test_count_5mbp_device_dump(ctx, tmp_path, bloom):
    from ktcountermemorygreedy_amd import workloads as W
    g = load_json("count_5mbp.json")
    seq = W.COUNT_5MBP.reads(fastq=False).tobytes()
    t = gpu_count(ctx, seq, 31, g["count_max"], bloom=bloom)
    out = str(tmp_path / "dump.txt")
    t.dump_file(out)
    with open(out, "rb") as f:
        assert hashlib.sha256(f.read()).hexdigest() == g["dump_sha256"]
    t.close()


@pytest.mark.parametrize("lo,hi", [(2, 4294967295), (1, 1), (3, 7), (50, 100)])
def test_dump_ranges_match_tool_rules(ctx, lo, hi):
    """min/max filters follow dump_tool's rules on the exported view (oracle
    merge of the same export) for both splits."""
    from ktcountermemorygreedy_amd import kc
    seq = kc.fastq_frame(read_golden("tiny_repeats.fastq")) + kc.fastq_frame(read_golden("tiny_mixed.fastq"))
    for bloom in (False, True):
        t = gpu_count(ctx, seq, 31, 255, bloom=bloom)
        high, low = t.export_device()
        nw = 1
        hk, hc = O.parse_high(high, 31, 255)
        lk = O.unpack(low, 31) if low else np.zeros((0, nw), np.uint64)
        mk, mc = O.merge(hk, hc, lk, nw, 255, lo, hi)
        assert t.dump(lo, hi) == O.dump_text(mk, mc, 31)
        t.close()
