"""Non-default filters against the reference binaries' output on the same
FASTQ (tests/golden/tiny_counts.*: keys at counts 1..8 and 50).

* The counter's high.bin filter, Tree argv 7-8 (NewKmerTree.h:1303-1311,
  FinalDrainWriter.h:95-104): the table exported with filter_min /
  filter_max gives the reference's dump and histogram bytes.
* The tools' min/max ranges (dump_tool.cpp:717-718, 884-892;
  histogram_tool's pools): the device dump and histogram over a range equal
  the reference tools' output over the same split -- the plain split (every
  key seen once in low.bin, the others in high.bin) for a table without the
  Bloom split, and the FIRST split (every key in low.bin, count - 1 in
  high.bin) that the Bloom split gives on this input."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.kc_testlib import load_json, read_golden, views_from_export

pytestmark = pytest.mark.gpu


def _table(ctx, bloom, path, **kw):
    from ktcountermemorygreedy_amd import kc
    seq = kc.fastq_frame(read_golden("tiny_counts.fastq"))
    t = kc.Table(ctx, 31, count_max=0xFFFFFFFF, capacity=1 << 12, max_distinct=1024, bloom=bloom, path=path, **kw)
    d = ctx.upload(np.frombuffer(seq, dtype=np.uint8))
    t.count_seq(d, len(seq))
    ctx.sync()
    d.free()
    return t


@pytest.mark.parametrize("path", ["direct", "partitioned"])
@pytest.mark.parametrize("filt", ["3-6", "5-50"])
def test_counter_filter_golden(ctx, filt, path):
    g = load_json("tiny_counts.json")
    fmin, fmax = map(int, filt.split("-"))
    want = g["filters"][filt]
    t = _table(ctx, False, path, filter_min=fmin, filter_max=fmax)
    assert t.stats()["errors"] == 0
    assert t.dump(1, 0xFFFFFFFF, sorted=True).decode() == want["dump"]
    assert O.histogram_text(t.histogram(1, 60), 1, 60).decode() == want["hist"]
    high, low = t.export_device()
    dump, hist = views_from_export(high, low, 31, 0xFFFFFFFF, 60)
    assert (dump.decode(), hist.decode()) == (want["dump"], want["hist"])
    t.close()


@pytest.mark.parametrize("path", ["direct", "partitioned"])
@pytest.mark.parametrize("split", ["plain", "first"])
def test_tool_ranges_golden(ctx, split, path):
    g = load_json("tiny_counts.json")
    # 2^20 Bloom blocks: no key's first occurrence is taken for a second one
    # (a Bloom false positive would move that key out of the FIRST split)
    t = _table(ctx, split == "first", path, **({"bloom_blocks": 1 << 20} if split == "first" else {}))
    st = t.stats()
    assert st["errors"] == 0
    if split == "first":
        assert st["low"] == 270  # every key's first occurrence is on the FIRST list
    # the export is the split the goldens were made from (as sets of records)
    high, low = t.export_device()
    for got, name, rec in ((high, "high", 12), (low, "low", 8)):
        ref = read_golden("tiny_counts.%s.%s.bin" % (split, name))
        as_set = lambda b: sorted(b[i:i + rec] for i in range(0, len(b), rec))  # noqa: E731
        assert as_set(got) == as_set(ref), name
    for r, want in g["splits"][split]["dump"].items():
        lo, hi = map(int, r.split("-"))
        assert t.dump(lo, hi, sorted=True).decode() == want, r
    for r, want in g["splits"][split]["hist"].items():
        lo, hi = map(int, r.split("-"))
        assert O.histogram_text(t.histogram(lo, hi), lo, hi).decode() == want, r
    t.close()
