"""Front end of the partitioned pass (kc_tune.front_mode): the coarse scatter
into fixed-capacity regions without a count pass (speculative), and the
exact redo in the same launch sequence when a region overflows.

Counts are compared with the oracle's exact counts of the same reads through
the table's statistics, point queries and histogram, for every front mode;
kc_table_front_redos tells which branch ran.  Skewed input (one read
repeated, so a few coarse digits take every key) must overflow and be redone;
reads of a random genome must fit.  Graph replays alternate the two, so one
captured graph takes both branches.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

EXACT, SPEC = 1, 2


def _random_reads(seed=3, n=6000):
    from ktcountermemorygreedy_amd import workloads as W
    g = W.ReadSet("g", seed, 400_000, 13, 0).genome()
    return W.COUNT_5MBP.scaled(n).reads(genome=g).tobytes()


def _skewed_reads(seed=5, n=6000):
    """One 150-base read repeated n times (both strands' k-mers of one read):
    ~120 distinct keys, so a handful of coarse regions take every key."""
    rng = np.random.default_rng(seed)
    read = bytes(rng.choice(np.frombuffer(b"ACGT", np.uint8), 150)) + b"\n"
    return read * n


def _check_exact(t, seq, k):
    nw = O.words_for_k(k)
    keys = O.extract(seq, k)
    uniq, cnt = O.count(keys, nw)
    st = t.stats()
    assert st["errors"] == 0 and st["kmers"] == keys.shape[0], st
    assert np.array_equal(t.query(uniq).astype(np.uint64), cnt)
    bins = t.histogram(1, 1000)  # counts above the range are left out (histogram_tool's range rule)
    want = np.bincount(cnt[cnt <= 1000].astype(np.int64), minlength=1001)[1:]
    assert np.array_equal(bins, want)


def _table(ctx, k, bloom, mode, seq):
    from ktcountermemorygreedy_amd import kc
    # 2^23 slots: 2^11 (Bloom, k <= 32) to 2^13 partitions, two scatter levels
    t = kc.Table(ctx, k, capacity=1 << 23, bloom=bloom, max_distinct=len(seq), path="partitioned")
    t.set_tune(front_mode=mode)
    return t


def _count(ctx, t, seq):
    d = ctx.upload(np.frombuffer(seq, dtype=np.uint8))
    t.count_seq(d, len(seq))
    ctx.sync()
    d.free()


@pytest.mark.parametrize("mode", [EXACT, SPEC], ids=["exact", "spec"])
@pytest.mark.parametrize("bloom", [False, True], ids=["plain", "bloom"])
@pytest.mark.parametrize("k", [21, 31, 63])
def test_random_reads_fit(ctx, k, bloom, mode):
    seq = _random_reads()
    t = _table(ctx, k, bloom, mode, seq)
    _count(ctx, t, seq)
    _check_exact(t, seq, k)
    assert t.front_redos() == 0
    t.close()


@pytest.mark.parametrize("bloom", [False, True], ids=["plain", "bloom"])
@pytest.mark.parametrize("k", [31, 63])
def test_skewed_reads_redo_exactly(ctx, k, bloom):
    seq = _skewed_reads()
    t = _table(ctx, k, bloom, SPEC, seq)
    _count(ctx, t, seq)
    _check_exact(t, seq, k)
    assert t.front_redos() == 1
    # a second call on top: random reads fit, the redo count stays
    more = _random_reads(seed=11, n=3000)
    _count(ctx, t, more)
    _check_exact(t, seq + more, k)
    assert t.front_redos() == 1
    # the exact mode never speculates
    t.clear()
    t.set_tune(front_mode=EXACT)
    _count(ctx, t, seq)
    _check_exact(t, seq, k)
    assert t.front_redos() == 0
    t.close()


def test_multi_pass_slices_speculative(ctx):
    """Hash-slice passes (kc_table_set_pass) through the speculative front:
    each slice holds exactly the oracle's keys of that slice."""
    seq = _random_reads(seed=17)
    t = _table(ctx, 31, True, SPEC, seq)
    for p in range(4):
        keys = O.extract_slice(seq, 31, 2, p)
        uniq, cnt = O.count(keys, 1)
        t.clear()
        t.set_pass(p, 4)
        _count(ctx, t, seq)
        st = t.stats()
        assert st["errors"] == 0 and st["kmers"] == keys.shape[0], (p, st)
        assert np.array_equal(t.query(uniq).astype(np.uint64), cnt)
        assert t.front_redos() == 0
    t.close()


@pytest.mark.parametrize("k", [31, 63])
def test_graph_replays_take_both_branches(ctx, k):
    """One captured graph per input pointer: replays over a random and a
    skewed batch (cleared in between) with eager counting between them; the
    skewed replays redo, the random ones do not, every replay is exact."""
    from ktcountermemorygreedy_amd import kc
    rnd, skw = _random_reads(seed=23), _skewed_reads(seed=29, n=12000)
    n = max(len(rnd), len(skw))
    buf = np.zeros(n + 16, np.uint8)
    t = kc.Table(ctx, k, capacity=1 << 23, bloom=True, max_distinct=n, path="partitioned")
    t.set_tune(front_mode=SPEC)
    b = kc.Batch(t, batch_bytes=2 << 20, path="partitioned")
    d = ctx.upload(buf)
    for rep in range(3):
        for seq, skewed in ((rnd, False), (skw, True)):
            d.upload(np.frombuffer(seq, dtype=np.uint8))
            t.clear()
            b.launch_device(d, len(seq))
            b.sync()
            _check_exact(t, seq, k)
            assert t.front_redos() == int(skewed), (rep, skewed)
            # eager counting between the replays
            e = _table(ctx, k, True, SPEC, seq)
            _count(ctx, e, seq)
            _check_exact(e, seq, k)
            e.close()
    assert b.info()["graphs_captured"] <= 2
    b.close()
    d.free()
    t.close()
