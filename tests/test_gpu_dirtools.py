"""The on-disk format for caller keys and the reference's tools over a
temp_dir, on the device.

* kc_export_low / kc_export_high (ExportWriter.h:296-313,
  FinalDrainWriter.h:71-146): packed bytes equal the oracle's packing (pinned
  by test_export_reader.cpp's vectors) and, with a filter, the records a
  table with the same filter exports (that table's views equal the reference
  run with those filters, tests/test_gpu_filters.py).
* kc_histogram_dir / kc_dump_dir and the two command-line tools
  (ktcountermemorygreedy_amd/bin/{histogram_tool,dump_tool}, the reference
  tools' argv): over a temp_dir the reference's own counter wrote, over the
  committed exports of tiny_counts (plain and FIRST splits, with ranges), and
  over this engine's export, the output bytes equal the reference tools'."""
import hashlib
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import ROOT
from tests.kc_testlib import gpu_count, load_json, read_golden

pytestmark = pytest.mark.gpu
REF = os.path.join(ROOT, "oracle", "_ref")
BIN = os.path.join(ROOT, "ktcountermemorygreedy_amd", "bin")


def _sha(b):
    return hashlib.sha256(b).hexdigest()


def _ref(name):
    p = os.path.join(REF, name)
    if not os.path.exists(p):
        pytest.skip("reference binaries not built (oracle/_ref needs the reference sources)")
    return p


def _run(args, timeout=300):
    r = subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=timeout)
    assert r.returncode == 0, (args, r.stdout[-2000:])


def _read(path):
    with open(path, "rb") as f:
        return f.read()


@pytest.mark.parametrize("k", [15, 21, 31, 32, 33, 45, 63])
def test_export_low_high_pack(ctx, k):
    rng = np.random.default_rng(k)
    nw = O.words_for_k(k)
    n = 50_000
    keys = rng.integers(0, 2**63, size=(n, nw), dtype=np.uint64) * np.uint64(2) + \
        rng.integers(0, 2, size=(n, nw), dtype=np.uint64)
    pad = 64 * nw - 2 * k  # zero the padding bits below the key
    for w in range(nw - 1, -1, -1):
        b = min(pad, 64)
        if b > 0:
            keys[:, w] &= np.uint64((~((1 << b) - 1)) & 0xFFFFFFFFFFFFFFFF) if b < 64 else np.uint64(0)
        pad -= b
    assert ctx.export_low(keys, k) == O.pack(keys, k)
    counts = rng.integers(1, 1000, n, dtype=np.uint64).astype(np.uint32)
    for cm in (255, 65535, 0xFFFFFF, 0xFFFFFFFF):
        c = np.minimum(counts, cm).astype(np.uint32)
        hk, hc = O.parse_high(ctx.export_high(keys, c, k, count_max=cm), k, cm)
        assert np.array_equal(hk, keys) and np.array_equal(hc, c)  # input order kept


@pytest.mark.parametrize("filt", [(0, 0xFFFFFFFF), (2, 0xFFFFFFFF), (3, 6), (5, 50)])
def test_export_high_filter_matches_table_export(ctx, filt):
    from ktcountermemorygreedy_amd import kc
    seq = kc.fastq_frame(read_golden("tiny_counts.fastq"))
    uniq, cnt = O.count(O.extract(seq, 31), 1)
    t = kc.Table(ctx, 31, capacity=1 << 12, filter_min=filt[0], filter_max=filt[1], max_distinct=1024)
    d = ctx.upload(np.frombuffer(seq, dtype=np.uint8))
    t.count_seq(d, len(seq))
    ctx.sync()
    high, _ = t.export_device()
    rep = cnt >= 2  # the plain split: keys seen once go to low.bin
    mine = ctx.export_high(uniq[rep], cnt[rep].astype(np.uint32), 31, 0xFFFFFFFF, *filt)
    recs = lambda b: sorted(b[i:i + 12] for i in range(0, len(b), 12))  # noqa: E731
    assert recs(mine) == recs(high)
    # the u32 wrap: a count of 2^32-1 is left out whenever filter_min >= 1
    top = ctx.export_high(uniq[:2], np.array([0xFFFFFFFF, 7], np.uint32), 31, 0xFFFFFFFF, *filt)
    assert len(top) // 12 == (filt[0] == 0) + (filt[0] <= 8 and filt[1] >= 7)
    t.close()


def _write_dir(d, k, count_max, high, low):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "infos.bin"), "wb") as f:
        f.write(np.array([k, count_max], np.uint32).tobytes())
    with open(os.path.join(d, "high.bin"), "wb") as f:
        f.write(high)
    with open(os.path.join(d, "low.bin"), "wb") as f:
        f.write(low)


@pytest.mark.parametrize("split", ["plain", "first"])
def test_tools_on_committed_splits(ctx, tmp_path, split):
    """The CLIs and the library calls over the two committed exports of
    tiny_counts reproduce the reference tools' bytes for every range."""
    g = load_json("tiny_counts.json")
    d = str(tmp_path / split)
    _write_dir(d, 31, g["count_max"], read_golden("tiny_counts.%s.high.bin" % split),
               read_golden("tiny_counts.%s.low.bin" % split))
    for r, want in g["splits"][split]["dump"].items():
        lo, hi = r.split("-")
        out = str(tmp_path / ("d%s.txt" % r))
        _run([os.path.join(BIN, "dump_tool"), "precise", d, "4", "2", out, lo, hi, "--sort"])
        assert _read(out).decode() == want, r
        ctx.dump_dir(d, out, int(lo), int(hi))
        assert _read(out).decode() == want, r
    for r, want in g["splits"][split]["hist"].items():
        lo, hi = r.split("-")
        out = str(tmp_path / ("h%s.txt" % r))
        _run([os.path.join(BIN, "histogram_tool"), "precise", d, "4", "2", out, lo, hi])
        assert _read(out).decode() == want, r
        assert O.histogram_text(ctx.histogram_dir(d, int(lo), int(hi)), int(lo), int(hi)).decode() == want


def _reference_dir(tmp, fastq_path, k, count_max):
    d = str(tmp)
    os.makedirs(d, exist_ok=True)
    _run([_ref("kmc_ref"), fastq_path, str(k), "6", "4", d + "/", "1024", "2", "4294967295", str(count_max)])
    return d


def _compare_tools(d, tmp_path, hist_range, dump_ranges):
    """Our CLIs against the reference tools on the same temp_dir, every mode."""
    for mode in ("precise", "approximate"):
        lo, hi = hist_range
        a, b = str(tmp_path / "ha.txt"), str(tmp_path / "hb.txt")
        _run([_ref("histogram_tool_ref"), mode, d + "/", "8", "4", a, str(lo), str(hi)])
        _run([os.path.join(BIN, "histogram_tool"), mode, d + "/", "8", "4", b, str(lo), str(hi)])
        assert _read(a) == _read(b), ("histogram", mode)
        for lo, hi in dump_ranges:
            a, b = str(tmp_path / "da.txt"), str(tmp_path / "db.txt")
            _run([_ref("dump_tool_ref"), mode, d + "/", "8", "4", a, str(lo), str(hi), "--sort"])
            _run([os.path.join(BIN, "dump_tool"), mode, d + "/", "8", "4", b, str(lo), str(hi), "--sort"])
            assert _read(a) == _read(b), ("dump", mode, lo, hi)
            # unsorted: the same lines in another order
            _run([os.path.join(BIN, "dump_tool"), mode, d + "/", "8", "4", b, str(lo), str(hi)])
            assert sorted(_read(a).splitlines()) == sorted(_read(b).splitlines()), ("dump unsorted", mode, lo, hi)


@pytest.mark.parametrize("case", [("tiny_mixed.fastq", 31, 0xFFFFFFFF), ("tiny_mixed.fastq", 63, 0xFFFFFFFF),
                                  ("tiny_repeats.fastq", 31, 255), ("tiny_counts.fastq", 31, 65535)],
                         ids=lambda c: "%s-k%d-c%d" % c)
def test_cli_on_reference_written_dir(tmp_path, case):
    name, k, cm = case
    d = _reference_dir(tmp_path / "ref", os.path.join(ROOT, "tests", "golden", name), k, cm)
    _compare_tools(d, tmp_path, (1, 500), [(1, 0xFFFFFFFF), (2, 7), (1, 1)])


def test_cli_on_reference_written_count_5mbp(tmp_path):
    """The count_5mbp workload counted by the reference: the device tools
    reproduce the reference tools' dump (283 MB) and histogram, which also
    equal the committed goldens."""
    from ktcountermemorygreedy_amd import workloads as W
    g = load_json("count_5mbp.json")
    fq = str(tmp_path / "in.fastq")
    W.COUNT_5MBP.reads(fastq=True).tofile(fq)
    d = _reference_dir(tmp_path / "ref", fq, 31, 0xFFFFFFFF)
    out = str(tmp_path / "dump.txt")
    _run([os.path.join(BIN, "dump_tool"), "precise", d, "8", "4", out, "1", "4294967295", "--sort"])
    assert _sha(_read(out)) == g["dump_sha256"]
    _run([os.path.join(BIN, "histogram_tool"), "precise", d, "8", "4", out, "1", "10000"])
    assert _sha(_read(out)) == g["hist_sha256"]
    _compare_tools(d, tmp_path, (1, 10000), [(2, 100)])


@pytest.mark.parametrize("bloom", [False, True], ids=["plain", "bloom"])
def test_dir_tools_on_our_export(ctx, tmp_path, bloom):
    from ktcountermemorygreedy_amd import kc
    from ktcountermemorygreedy_amd import workloads as W
    g = load_json("count_5mbp.json")
    seq = W.COUNT_5MBP.reads(fastq=False).tobytes()
    t = gpu_count(ctx, seq, 31, g["count_max"], bloom=bloom)
    d = str(tmp_path / "ours")
    os.makedirs(d)
    t.export_dir(d)
    t.close()
    out = str(tmp_path / "dump.txt")
    ctx.dump_dir(d, out)
    assert _sha(_read(out)) == g["dump_sha256"]
    assert _sha(O.histogram_text(ctx.histogram_dir(d, 1, 10000), 1, 10000)) == g["hist_sha256"]
    # a loaded export answers lookups like the table it came from
    lt = kc.Table.load_dir(ctx, d)
    keys = O.extract(seq[:200_000], 31)
    uniq, _ = O.count(keys, 1)
    full_u, full_c = O.count(O.extract(seq, 31), 1)
    want = full_c[np.searchsorted(full_u[:, 0], uniq[:, 0])]
    assert np.array_equal(lt.query(uniq).astype(np.uint64), want)
    lt.close()
    shutil.rmtree(d)
