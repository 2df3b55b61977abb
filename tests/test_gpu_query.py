"""Point lookups are read-only, as FlatConcurrentHashMap::find_prepared
(FlatConcurrentHashMap.h:142-205) is: on a Bloom-split table a lookup adds
one for a key on the FIRST list (the merge rule of
LowFrequencyQueryThreadPool.h:213-232) through a hash index of the list, and
neither the table, the list nor the counters change.  So counting, looking
up and exporting interleave in any order with the same results as counting
alone."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.kc_testlib import views_from_export

pytestmark = pytest.mark.gpu


def _halves():
    from ktcountermemorygreedy_amd import workloads as W
    g = W.ReadSet("g", 3, 2_000_000, 13, 0).genome()
    seq = W.UNEVEN_50MBP.scaled(2500).reads(genome=g).tobytes()
    cut = seq.index(b"\n", len(seq) // 2) + 1
    return seq[:cut], seq[cut:]


@pytest.mark.parametrize("path", ["direct", "partitioned"])
@pytest.mark.parametrize("k", [31, 63])
def test_count_query_count_query_export(ctx, k, path):
    from ktcountermemorygreedy_amd import kc
    a, b = _halves()
    nw = O.words_for_k(k)
    ua, ca = O.count(O.extract(a, k), nw)
    uab, cab = O.count(O.extract(a + b, k), nw)
    absent = uab.copy()
    absent[:, nw - 1] |= np.uint64(1)  # a padding bit set: never a stored key

    def new_table():
        return kc.Table(ctx, k, capacity=1 << 23, bloom=True, max_distinct=uab.shape[0], path=path)

    da = ctx.upload(np.frombuffer(a, dtype=np.uint8))
    db = ctx.upload(np.frombuffer(b, dtype=np.uint8))
    t = new_table()
    t.count_seq(da, len(a))
    ctx.sync()
    st = t.stats()
    assert np.array_equal(t.query(ua).astype(np.uint64), ca)
    assert np.array_equal(t.query(ua).astype(np.uint64), ca)  # the index is reused
    assert int(t.query(absent).sum()) == 0
    assert t.stats() == st
    t.count_seq(db, len(b))  # counting goes on after lookups
    ctx.sync()
    assert np.array_equal(t.query(uab).astype(np.uint64), cab)
    high, low = t.export_device()
    dump, hist = views_from_export(high, low, k, 0xFFFFFFFF, 1000)
    assert dump == O.dump_text(uab, cab, k)
    # the same export bytes' views as a table that was never looked up
    u = new_table()
    u.count_seq(da, len(a))
    u.count_seq(db, len(b))
    ctx.sync()
    h2, l2 = u.export_device()
    assert views_from_export(h2, l2, k, 0xFFFFFFFF, 1000) == (dump, hist)
    assert np.array_equal(t.histogram(1, 1000), u.histogram(1, 1000))
    # a cleared table looks up nothing
    t.clear()
    assert int(t.query(uab).sum()) == 0
    for x in (t, u):
        x.close()
    da.free()
    db.free()


def test_query_saturates_with_first_list(ctx):
    """count_max applies to the merged count (table + FIRST list)."""
    from ktcountermemorygreedy_amd import kc
    reads = b"".join(b"ACGTTGCAAGGCTTACCGATGCAATCGGATCCAT\n" for _ in range(9))
    t = kc.Table(ctx, 31, count_max=5, capacity=1 << 12, bloom=True, max_distinct=64)
    d = ctx.upload(np.frombuffer(reads, dtype=np.uint8))
    t.count_seq(d, len(reads))
    ctx.sync()
    uniq, cnt = O.count(O.extract(reads, 31), 1)
    assert (cnt == 9).all() and t.stats()["low"] == uniq.shape[0]
    assert (t.query(uniq) == 5).all()
    t.close()


@pytest.mark.parametrize("path", ["direct", "partitioned"])
@pytest.mark.parametrize("k", [31, 63])
def test_query_on_another_context(ctx, k, path):
    """kc_table_query_on: lookups on a second context's stream after the
    table's context was synchronized, and a count enqueued on the table's
    stream right after them waits for them."""
    from ktcountermemorygreedy_amd import kc
    a, b = _halves()
    nw = O.words_for_k(k)
    ua, ca = O.count(O.extract(a, k), nw)
    uab, cab = O.count(O.extract(a + b, k), nw)
    other = kc.Context(0)
    t = kc.Table(ctx, k, capacity=1 << 23, bloom=True, max_distinct=uab.shape[0], path=path)
    da = ctx.upload(np.frombuffer(a, dtype=np.uint8))
    db = ctx.upload(np.frombuffer(b, dtype=np.uint8))
    t.count_seq(da, len(a))
    ctx.sync()
    got = t.query(ua, ctx=other)
    t.count_seq(db, len(b))
    assert np.array_equal(got.astype(np.uint64), ca)
    assert np.array_equal(t.query(uab, ctx=other).astype(np.uint64), cab)
    assert np.array_equal(t.query(uab).astype(np.uint64), cab)
    t.close()
    da.free()
    db.free()
    other.close()


def test_query_on_mapped_host_memory(ctx):
    """A lookup that reads its keys from and writes its counts to page-locked
    host memory in place (kc_host_device_ptr), as the serving loop does."""
    import ctypes as C
    from ktcountermemorygreedy_amd import kc
    a, _ = _halves()
    ua, ca = O.count(O.extract(a, 31), 1)
    t = kc.Table(ctx, 31, capacity=1 << 22, bloom=True, max_distinct=ua.shape[0])
    da = ctx.upload(np.frombuffer(a, dtype=np.uint8))
    t.count_seq(da, len(a))
    ctx.sync()
    other = kc.Context(0)
    n = ua.shape[0]
    hk = kc.PinnedBuffer(other, 8 * n)
    hc = kc.PinnedBuffer(other, 4 * n)
    hk.array[:] = np.ascontiguousarray(ua[:, 0]).view(np.uint8)
    hc.array[:] = 0xFF
    dk, dc = C.c_void_p(), C.c_void_p()
    kc.check(kc.lib().kc_host_device_ptr(other.h, C.c_void_p(hk.ptr), C.byref(dk)), "kc_host_device_ptr")
    kc.check(kc.lib().kc_host_device_ptr(other.h, C.c_void_p(hc.ptr), C.byref(dc)), "kc_host_device_ptr")
    kc.check(kc.lib().kc_table_query_on(t.h, other.h, dk, n, dc), "kc_table_query_on")
    other.sync()
    assert np.array_equal(hc.array.view(np.uint32).astype(np.uint64), ca)
    hk.free()
    hc.free()
    t.close()
    da.free()
    other.close()
