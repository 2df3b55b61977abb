"""The headline workload's generator on the GPU: a repeat-rich genome (500 x
1 kb segments, 30% of the genome as 1%-diverged copies) read at 11x, and the
same with 20 segments copied ~60 times each (hot keys in the hundreds, as in
repeat_100mbp), counted with the Bloom singleton split on a table sized by
the planner -- on both counting paths, eagerly and through replays of a
captured graph.  Every distinct key's merged count is compared with the
oracle's (read-only lookups), and the device dump and histogram with the
bytes the reference binaries wrote for the same FASTQ
(FastqClassifier.h:508-690 and BloomFilter.h:233-273 decide the split; the
merged views do not depend on it).  At full size (20M reads, 2.4e9 31-mers)
one hash slice of 64 is counted and compared key by key with the oracle's
count of that slice."""
import hashlib

import numpy as np
import pytest

from oracle import oracle as O
from tests.kc_testlib import load_json, views_from_export

pytestmark = pytest.mark.gpu

_CASES = {}


def _case(name):
    """(framed bytes, windows, distinct keys, their counts), cached per name."""
    if name not in _CASES:
        from ktcountermemorygreedy_amd import workloads as W
        w = {"repeat_4mbp": W.REPEAT_4MBP, "repeat_4mbp_hot": W.REPEAT_4MBP_HOT}[name]
        seq = w.reads().tobytes()
        keys = O.extract(seq, 31)
        uniq, cnt = O.count(keys, 1)
        _CASES[name] = (seq, keys.shape[0], uniq, cnt)
    return _CASES[name]


def _sha(b):
    return hashlib.sha256(b).hexdigest()


@pytest.mark.parametrize("launch", ["eager", "graph"])
@pytest.mark.parametrize("path", ["direct", "partitioned"])
@pytest.mark.parametrize("name", ["repeat_4mbp", "repeat_4mbp_hot"])
def test_repeat_parity(ctx, name, path, launch):
    from ktcountermemorygreedy_amd import kc
    print("repeat parity:", name, path, launch, flush=True)
    g = load_json("repeat_4mbp.json")[name]
    seq, n_windows, uniq, cnt = _case(name)
    assert n_windows == g["kmers"] and uniq.shape[0] == g["distinct"]
    d = ctx.upload(np.frombuffer(seq, dtype=np.uint8))
    kw, plan = kc.plan_table(ctx, d, 31, len(seq), bloom=True, sample_slices=8)
    t = kc.Table(ctx, 31, path=path, **kw)
    b = kc.Batch(t) if launch == "graph" else None
    for rep in range(1 if b is None else 3):  # graph replays, eager checks in between
        t.clear()
        if b is None:
            t.count_seq(d, len(seq))
        else:
            b.launch_device(d, len(seq))
        ctx.sync()
        st = t.stats()
        assert st["errors"] == 0 and st["kmers"] == n_windows, (rep, st)
        assert 0 < st["low"] <= uniq.shape[0]
        got = t.query(uniq).astype(np.uint64)
        assert np.array_equal(got, cnt), rep
        assert t.stats() == st  # lookups change nothing
        assert _sha(t.dump(1, 0xFFFFFFFF, sorted=True)) == g["dump_sha256"], rep
        assert _sha(O.histogram_text(t.histogram(1, 10000), 1, 10000)) == g["hist_sha256"], rep
    if b is not None:
        assert b.info()["launches"] == 3 and b.info()["graphs_captured"] == 1
        b.close()
    if launch == "eager" and path == "partitioned":
        # the exported files, merged with the tools' rules
        high, low = t.export_device()
        dump, hist = views_from_export(high, low, 31, 0xFFFFFFFF)
        assert _sha(dump) == g["dump_sha256"] and _sha(hist) == g["hist_sha256"]
    t.close()
    d.free()


def test_repeat_100mbp_slice_exact(ctx):
    """The full headline input: slice 0 of 64 (kc_table_set_pass) counted with
    the Bloom split on, every key of the slice exact against the oracle."""
    from ktcountermemorygreedy_amd import kc
    from ktcountermemorygreedy_amd import workloads as W
    print("repeat_100mbp: generating 20M reads", flush=True)
    seq = W.REPEAT_100MBP.reads(threads=16)
    print("repeat_100mbp: oracle slice 0/64", flush=True)
    keys = O.extract_slice(seq, 31, 6, 0, threads=16)
    uniq, cnt = O.count(keys, 1)
    print("repeat_100mbp: slice windows %d distinct %d max %d" % (keys.shape[0], uniq.shape[0], cnt.max()),
          flush=True)
    d = ctx.upload(seq)
    cap = 1 << int(np.ceil(np.log2(2 * uniq.shape[0])))
    t = kc.Table(ctx, 31, capacity=cap, max_distinct=uniq.shape[0], bloom=True)
    t.set_pass(0, 64)
    t.count_seq(d, seq.size)
    ctx.sync()
    st = t.stats()
    print("repeat_100mbp:", st, flush=True)
    assert st["errors"] == 0 and st["kmers"] == keys.shape[0]
    got = t.query(uniq).astype(np.uint64)
    assert np.array_equal(got, cnt)
    bins = t.histogram(1, 10000)
    want = np.bincount(np.minimum(cnt, 10000).astype(np.int64), minlength=10001)[1:]
    want[-1] -= int((cnt > 10000).sum())  # counts above the range are not binned
    assert np.array_equal(bins, want)
    t.close()
    d.free()
