"""Host-side product code (framing, synthetic generator) on the CPU."""
import hashlib

import numpy as np
import pytest

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


def test_merge_sorted_dumps(tmp_path):
    """k-way merge of key-sorted dump texts (the multi-pass dump merge):
    equals the sorted union; empty inputs and a missing final newline."""
    from ktcountermemorygreedy_amd import kc
    rng = np.random.default_rng(7)
    keys = sorted({"".join(rng.choice(list("ACGT"), 21)) for _ in range(3000)})
    parts = [[], [], [], []]
    for i, key in enumerate(keys):
        parts[int(rng.integers(0, 3))].append("%s\t%d\n" % (key, i % 97 + 1))
    paths = []
    for j, lines in enumerate(parts):  # part 3 stays empty
        p = tmp_path / ("p%d.txt" % j)
        text = "".join(lines)
        if j == 1 and text:
            text = text[:-1]
        p.write_text(text)
        paths.append(str(p))
    out = tmp_path / "merged.txt"
    kc.merge_sorted_dumps(paths, str(out))
    want = "".join("%s\t%d\n" % (key, i % 97 + 1) for i, key in enumerate(keys))
    assert out.read_text() == want
    with pytest.raises(kc.KCError):
        kc.merge_sorted_dumps([str(tmp_path / "missing.txt")], str(out))
