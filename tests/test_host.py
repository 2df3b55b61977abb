"""Host-side product code (framing, synthetic generator) on the CPU."""
import hashlib

import numpy as np

from oracle import oracle as O
from tests.kc_testlib import load_json, read_golden


def test_framing_matches_oracle():
    from ktcountermemorygreedy_amd import kc
    for case in load_json("tiny_cases.json"):
        fq = read_golden(case["input"])
        assert kc.fastq_frame(fq) == O.fastq_frame(fq)


def test_frame_file_plain_and_gzip(tmp_path):
    import gzip
    from ktcountermemorygreedy_amd import kc
    fq = read_golden("tiny_mixed.fastq")
    p = tmp_path / "a.fastq"
    p.write_bytes(fq)
    g = tmp_path / "a.fastq.gz"
    g.write_bytes(gzip.compress(fq))
    want = O.fastq_frame(fq)
    assert kc.fastq_frame_file(str(p)).tobytes() == want
    assert kc.fastq_frame_file(str(g)).tobytes() == want


def test_generator_is_deterministic_and_pinned():
    from ktcountermemorygreedy_amd import workloads as W
    g = load_json("count_5mbp.json")
    a = W.COUNT_5MBP.reads(fastq=True, threads=3)
    b = W.COUNT_5MBP.reads(fastq=True, threads=8)
    assert hashlib.sha256(a.tobytes()).hexdigest() == g["fastq_sha256"]
    assert np.array_equal(a, b)
    framed = W.COUNT_5MBP.reads(fastq=False)
    assert framed.tobytes() == O.fastq_frame(a.tobytes())


def test_uneven_reads_lengths():
    from ktcountermemorygreedy_amd import workloads as W
    rs = W.UNEVEN_50MBP.scaled(2000)
    g = W.ReadSet("g", 3, 2_000_000, 13, 0).genome()
    seq = rs.reads(genome=g).tobytes()
    lens = np.array([len(l) for l in seq.split(b"\n")[:-1]])
    assert len(lens) == 2000
    short = lens == 150
    assert 0.75 < short.mean() < 0.85
    assert 6000 < lens[~short].mean() < 10000
