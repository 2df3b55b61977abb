"""Scatter guards of the partitioned pass: each scatter writes a key at its
digit region's offset plus its rank, the region sizes coming from a
separate count pass (the reference's tree counts a block's prefixes before
it copies them, NewKmerTree.h:600-614, 972-981).  A count that does not match
the keys the scatter sees must not make it write into a neighbouring region:
the guard drops such writes and the call reports KC_EFULL.  The mismatch is
forced from the host with the test-only KC_TEST_SKEW_COUNTS switch (read at
table creation), which adds -1 / +1 to one count of the front (1 / 2) or the
fine (3 / 4) level; every write stays inside the key arrays."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _seq():
    from ktcountermemorygreedy_amd import workloads as W
    g = W.ReadSet("g", 3, 300_000, 13, 0).genome()
    return W.COUNT_5MBP.scaled(3000).reads(genome=g).tobytes()


@pytest.mark.parametrize("source", ["bytes", "keys"])
@pytest.mark.parametrize("skew", [0, 1, 2, 3, 4])
def test_scatter_guard_reports_mismatch(ctx, monkeypatch, skew, source):
    from ktcountermemorygreedy_amd import kc
    seq = _seq()
    if skew:
        monkeypatch.setenv("KC_TEST_SKEW_COUNTS", str(skew))
    # 1024 partitions: a coarse and a fine scatter level
    t = kc.Table(ctx, 31, capacity=1 << 23, max_distinct=len(seq), path="partitioned")
    monkeypatch.delenv("KC_TEST_SKEW_COUNTS", raising=False)
    keys = O.extract(seq, 31)
    if source == "bytes":
        d = ctx.upload(np.frombuffer(seq, dtype=np.uint8))
        t.count_seq(d, len(seq))
        ctx.sync()
    else:
        t.increment(keys)
    st = t.stats()
    if skew == 0:
        assert st["errors"] == 0
        uniq, cnt = O.count(keys, 1)
        assert np.array_equal(t.query(uniq).astype(np.uint64), cnt)
    else:
        assert st["errors"] > 0, st
        with pytest.raises(kc.KCError) as e:
            t.export_device()
        assert e.value.status == kc.KC_EFULL
    t.close()
