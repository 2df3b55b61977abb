"""Batched counting (kc_batch: uneven chunks packed into staging slots, one
captured-graph launch per slot) and the launch-shape autotuner.

Every batch result is compared with the oracle's exact counts of the same
reads (the chunks concatenated, a newline after each chunk that lacks one),
through the table's statistics, point queries and histogram.  Graph replays
are checked one by one, with the program's own eager work and
synchronisations between them.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _reads(seed, n_short, n_long, genome_len=400_000):
    from ktcountermemorygreedy_amd import kc
    g = kc.synth_genome(seed, genome_len)
    short = kc.synth_reads(g, seed + 1, n_short, 150, True, 0.005, 1e-4).tobytes()
    rng = np.random.default_rng(seed)
    longs = []
    for i in range(n_long):
        L = int(rng.lognormal(8.92, 0.36))
        p = int(rng.integers(0, genome_len - L))
        longs.append(g[p:p + L].tobytes() + b"\n")
    return short, b"".join(longs)


def _chunks(seq: bytes, rng, max_len):
    """Cuts a framed stream into uneven chunks at line ends; some chunks lose
    their final newline (the batch restores it)."""
    lines = seq.split(b"\n")[:-1]
    out, i = [], 0
    while i < len(lines):
        j = i + int(rng.integers(1, 40))
        c = b"\n".join(lines[i:j]) + b"\n"
        if len(c) > max_len:
            c = lines[i] + b"\n"
            j = i + 1
        if rng.random() < 0.3:
            c = c[:-1]
        out.append(c)
        i = j
    return out


def _joined(chunks):
    return b"".join(c if c.endswith(b"\n") else c + b"\n" for c in chunks)


def _check_exact(t, seq, k):
    nw = O.words_for_k(k)
    keys = O.extract(seq, k)
    uniq, cnt = O.count(keys, nw)
    st = t.stats()
    assert st["errors"] == 0 and st["kmers"] == keys.shape[0], st
    got = t.query(uniq).astype(np.uint64)
    assert np.array_equal(got, cnt)
    bins = t.histogram(1, 1000)
    want = np.bincount(np.minimum(cnt, 1000).astype(np.int64), minlength=1001)[1:]
    assert np.array_equal(bins, want)


@pytest.mark.parametrize("path", ["direct", "partitioned"])
@pytest.mark.parametrize("bloom", [False, True], ids=["plain", "bloom"])
@pytest.mark.parametrize("k", [21, 31, 63])
def test_batch_host_chunks_exact(ctx, k, bloom, path):
    from ktcountermemorygreedy_amd import kc
    short, long_ = _reads(5, 6000, 60)
    rng = np.random.default_rng(k)
    chunks = _chunks(short + long_, rng, 200_000)
    # one chunk larger than a slot: counted in pieces with a halo
    chunks.insert(len(chunks) // 2, long_ * 3)
    seq = _joined(chunks)
    print("batch: k=%d bloom=%s path=%s chunks=%d bytes=%d" % (k, bloom, path, len(chunks), len(seq)), flush=True)
    t = kc.Table(ctx, k, capacity=1 << 22, bloom=bloom, max_distinct=len(seq))
    b = kc.Batch(t, batch_bytes=1 << 20, queue_depth=3, path=path)
    for c in chunks:
        b.add(c)
    b.flush()
    b.sync()
    info = b.info()
    assert info["path"] == (1 if path == "direct" else 2)
    assert info["bytes"] >= len(seq) and info["pieces"] >= 2 and info["launches"] > 3
    assert info["graphs_captured"] <= 3
    b.close()
    _check_exact(t, seq, k)
    t.close()


@pytest.mark.parametrize("path", ["direct", "partitioned"])
def test_batch_add_chunks_matches_add(ctx, path):
    """kc_batch_add_chunks over one buffer packs the same batches as
    kc_batch_add per chunk (same launches, pieces and bytes) and counts the
    same keys; spans of >= 16 MB take the multithreaded copy, and the input
    holds empty chunks, chunks without a final newline and one chunk larger
    than a slot."""
    from ktcountermemorygreedy_amd import kc
    g = kc.synth_genome(41, 2_000_000)
    seq = kc.synth_reads(g, 42, 200_000, 150, True, 0.005, 1e-4).tobytes()
    rng = np.random.default_rng(43)
    lines = seq.split(b"\n")[:-1]
    chunks, i = [], 0
    while i < len(lines):
        j = i + (130_000 if len(chunks) == 20 else int(rng.integers(1, 400)))
        c = b"\n".join(lines[i:j]) + b"\n"
        if i < len(lines) // 2 and rng.random() < 0.02:  # later spans run to a full slot
            c = c[:-1]
        chunks.append(c)
        if rng.random() < 0.01:
            chunks.append(b"")
        i = j
    buf = np.frombuffer(b"".join(chunks), dtype=np.uint8)
    ends = np.cumsum([len(c) for c in chunks]).astype(np.uint64)
    print("add_chunks: path=%s chunks=%d bytes=%d" % (path, len(chunks), buf.size), flush=True)
    ta = kc.Table(ctx, 31, capacity=1 << 25, bloom=True, max_distinct=buf.size)
    tc = kc.Table(ctx, 31, capacity=1 << 25, bloom=True, max_distinct=buf.size)
    ba = kc.Batch(ta, batch_bytes=16 << 20, queue_depth=3, path=path)
    bc = kc.Batch(tc, batch_bytes=16 << 20, queue_depth=3, path=path)
    for c in chunks:
        ba.add(c)
    bc.add_chunks(buf, ends)
    for b in (ba, bc):
        b.flush()
        b.sync()
    ia, ic = ba.info(), bc.info()
    for f in ("chunks", "pieces", "launches", "bytes"):
        assert ia[f] == ic[f], (f, ia, ic)
    assert ic["pieces"] >= 2 and ic["chunks"] == sum(1 for c in chunks if c)
    ba.close()
    bc.close()
    sa, sc = ta.stats(), tc.stats()
    # (table keys before the FIRST-list fold follow the Bloom decisions,
    # which depend on the order concurrent waves reach the filter)
    assert sc["errors"] == 0 and sa["kmers"] == sc["kmers"]
    assert sc["kmers"] == O.extract(_joined(chunks), 31).shape[0]
    assert np.array_equal(ta.histogram(1, 2000), tc.histogram(1, 2000))
    keys = O.extract(seq[:200_000], 31)[:, 0].copy()
    assert np.array_equal(ta.query(keys), tc.query(keys))
    with pytest.raises(kc.KCError):  # ends must not decrease
        b = kc.Batch(tc, batch_bytes=1 << 20, path=path)
        try:
            b.add_chunks(buf, np.array([100, 50], dtype=np.uint64))
        finally:
            b.close()
    ta.close()
    tc.close()


@pytest.mark.parametrize("path", ["direct", "partitioned"])
def test_batch_graph_replays_match_eager(ctx, path):
    """The same slot graphs replayed over different batches, with eager
    counting and synchronisations in between: every replay equals the eager
    count of the same bytes."""
    from ktcountermemorygreedy_amd import kc
    rng = np.random.default_rng(11)
    tb = kc.Table(ctx, 31, capacity=1 << 22, bloom=True, max_distinct=1 << 21)
    te = kc.Table(ctx, 31, capacity=1 << 22, bloom=True, max_distinct=1 << 21)
    b = kc.Batch(tb, batch_bytes=4 << 20, queue_depth=2, path=path)
    for rep in range(4):
        short, long_ = _reads(20 + rep, 3000 + 1000 * rep, 20 * rep)
        chunks = _chunks(short + long_, rng, 100_000)
        tb.clear()
        for c in chunks:
            b.add(c)
        b.flush()
        b.sync()
        seq = _joined(chunks)
        te.clear()
        d = ctx.upload(np.frombuffer(seq, dtype=np.uint8))
        te.count_seq(d, len(seq))
        ctx.sync()
        sb, se = tb.stats(), te.stats()
        assert sb["kmers"] == se["kmers"] and sb["errors"] == 0, (rep, sb, se)
        assert np.array_equal(tb.histogram(1, 2000), te.histogram(1, 2000)), rep
        keys = O.extract(seq[:50_000], 31)[:, 0].copy()
        assert np.array_equal(tb.query(keys), te.query(keys)), rep
    assert b.info()["graphs_captured"] <= 2
    b.close()
    tb.close()
    te.close()


@pytest.mark.parametrize("path", ["direct", "partitioned"])
@pytest.mark.parametrize("k", [31, 63])
def test_batch_device_launch_replays(ctx, k, path):
    """One graph launch over a device-resident batch, replayed three times
    (table cleared in between) and once over a shorter prefix."""
    from ktcountermemorygreedy_amd import kc
    short, long_ = _reads(7, 8000, 40)
    seq = short + long_
    d = ctx.upload(np.frombuffer(seq, dtype=np.uint8))
    t = kc.Table(ctx, k, capacity=1 << 22, bloom=True, max_distinct=len(seq))
    b = kc.Batch(t, batch_bytes=1 << 20, path=path)
    for rep in range(3):
        t.clear()
        b.launch_device(d, len(seq))
        b.sync()
        _check_exact(t, seq, k)
    cut = seq[:len(short) // 2]
    cut = cut[:cut.rindex(b"\n") + 1]
    t.clear()
    b.launch_device(d, len(cut))
    b.sync()
    _check_exact(t, cut, k)
    assert b.info()["graphs_captured"] == 1
    # after a clear the partitioned graph writes the lazily cleared table
    # whole; the second launch adds to it with the graph that reads it
    t.clear()
    b.launch_device(d, len(seq))
    b.launch_device(d, len(seq))
    b.sync()
    _check_exact(t, seq + seq, k)
    assert b.info()["graphs_captured"] == (2 if path == "partitioned" else 1)
    b.close()
    t.close()


def test_autotune_picks_a_working_shape(ctx):
    from ktcountermemorygreedy_amd import kc
    short, long_ = _reads(9, 20000, 100, genome_len=1_000_000)
    seq = short + long_
    d = ctx.upload(np.frombuffer(seq, dtype=np.uint8))
    t = kc.Table(ctx, 31, capacity=1 << 22, bloom=True, max_distinct=len(seq))
    tune, ms = t.autotune(d, len(seq), reps=1)
    print("autotune:", tune, ms, flush=True)
    assert ms > 0 and tune["path"] in (1, 2)
    if tune["path"] == 2:
        assert tune["front_blocks"] > 0 and tune["fine_slices"] in (2, 4, 8, 16)
    assert t.stats()["kmers"] == 0  # left empty
    t.count_seq(d, len(seq))
    ctx.sync()
    _check_exact(t, seq, 31)
    # the tuned shape also drives batches
    t.clear()
    b = kc.Batch(t, batch_bytes=8 << 20)
    info = b.info()
    assert info["path"] == tune["path"]
    if tune["path"] == 2:
        assert info["front_blocks"] == tune["front_blocks"] and info["fine_slices"] == tune["fine_slices"]
    b.launch_device(d, len(seq))
    b.sync()
    _check_exact(t, seq, 31)
    b.close()
    t.close()


@pytest.mark.parametrize("front_blocks,fine_slices", [(64, 1), (256, 3), (4096, 16), (1024, 64)])
def test_partitioned_shapes_exact(ctx, front_blocks, fine_slices):
    """Counting is exact for any front grid / slice count, including slices
    that are empty and grids larger than the number of tiles."""
    from ktcountermemorygreedy_amd import kc
    short, long_ = _reads(13, 5000, 30)
    seq = short + long_
    t = kc.Table(ctx, 31, capacity=1 << 23, bloom=True, max_distinct=len(seq), path="partitioned")
    t.set_tune(front_blocks=front_blocks, fine_slices=fine_slices)
    d = ctx.upload(np.frombuffer(seq, dtype=np.uint8))
    t.count_seq(d, len(seq))
    ctx.sync()
    _check_exact(t, seq, 31)
    t.close()


def test_batch_small_launches_many_partitions(ctx):
    """Many small launches into a table of many partitions: each partition
    sees a few hundred keys per launch, so only a few waves of its counting
    block are active.  The FIRST list must still fill its segments evenly
    (a tight list capacity would overflow otherwise); results equal the
    eager count."""
    from ktcountermemorygreedy_amd import kc
    short, long_ = _reads(31, 60000, 200, genome_len=4_000_000)
    seq = short + long_
    d = ctx.upload(np.frombuffer(seq, dtype=np.uint8))
    te = kc.Table(ctx, 31, capacity=1 << 24, bloom=True, max_distinct=len(seq))
    te.count_seq(d, len(seq))
    ctx.sync()
    se = te.stats()
    assert se["errors"] == 0
    tb = kc.Table(ctx, 31, capacity=1 << 24, bloom=True, max_distinct=len(seq),
                  low_capacity=int(se["low"] * 1.05))
    b = kc.Batch(tb, batch_bytes=1 << 20, queue_depth=3, path="partitioned")
    bounds = np.flatnonzero(np.frombuffer(seq, dtype=np.uint8) == 10)[::50] + 1
    start = 0
    for end in bounds.tolist() + [len(seq)]:
        if end > start:
            b.add(seq[start:end])
        start = end
    b.flush()
    b.sync()
    assert b.info()["launches"] >= 10
    b.close()
    sb = tb.stats()
    print("eager", se, "batched", sb, flush=True)
    assert sb["errors"] == 0 and sb["kmers"] == se["kmers"]
    assert np.array_equal(tb.histogram(1, 5000), te.histogram(1, 5000))
    te.close()
    tb.close()
