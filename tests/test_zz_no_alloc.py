"""kc_table_histogram_noalloc (no counterpart in the reference; for callers
that cannot allocate device memory) against kc_table_histogram on the same
tables, two calls in a row with the second on a smaller input, and its
argument limit.  Runs alone in a fresh process and after the other tests."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _table(ctx, n_reads, bloom):
    from ktcountermemorygreedy_amd import kc
    from ktcountermemorygreedy_amd import workloads as W
    g = W.REPEAT_4MBP_HOT.genome()
    seq = W.REPEAT_4MBP_HOT.scaled(n_reads).reads(genome=g).tobytes()
    t = kc.Table(ctx, 31, max_distinct=len(seq), bloom=bloom)
    d = ctx.upload(np.frombuffer(seq, dtype=np.uint8))
    t.count_seq(d, len(seq))
    ctx.sync()
    d.free()
    return t


@pytest.mark.parametrize("bloom", [False, True], ids=["plain", "bloom"])
def test_noalloc_histogram_matches(ctx, bloom):
    from ktcountermemorygreedy_amd import kc
    big = _table(ctx, 200_000, bloom)
    small = _table(ctx, 3_000, bloom)
    for t, lo, hi in ((big, 1, 10000), (small, 2, 300), (big, 5, 65540), (small, 1, 1)):
        want = t.histogram(lo, hi)
        got = t.histogram_noalloc(lo, hi)
        assert np.array_equal(got, want), (lo, hi)
        assert want.sum() > 0
    with pytest.raises(kc.KCError) as e:
        big.histogram_noalloc(1, 65537)  # one bin more than the array holds
    assert e.value.status == kc.KC_EINVAL
    big.close()
    small.close()
