"""Parity of the HIP counting path against the reference's goldens.

Every case runs the engine through the C ABI (libkc.so) on the GPU, exports
the reference's high/low format, and compares the merged views (dump --sort
text, histogram text) byte for byte with what the reference binaries
produced for the same input.  The device histogram kernel is compared too.
"""
import hashlib

import numpy as np
import pytest

from oracle import oracle as O
from tests.kc_testlib import gpu_count, gpu_views, load_json, read_golden, views_from_export

pytestmark = pytest.mark.gpu


def _tiny_cases():
    return load_json("tiny_cases.json")


@pytest.mark.parametrize("path", ["direct", "partitioned"])
@pytest.mark.parametrize("bloom", [False, True], ids=["plain", "bloom"])
@pytest.mark.parametrize("case", _tiny_cases(), ids=lambda c: "%s-k%d-c%d" % (c["input"], c["k"], c["count_max"]))
def test_tiny_parity(ctx, case, bloom, path):
    from ktcountermemorygreedy_amd import kc
    seq = kc.fastq_frame(read_golden(case["input"]))
    dump, dev_hist, file_hist, st = gpu_views(ctx, seq, case["k"], case["count_max"], bloom=bloom,
                                              hist_max=case["hist_max"], path=path)
    assert st["kmers"] == O.extract(seq, case["k"]).shape[0]
    assert hashlib.sha256(dump).hexdigest() == case["dump_sha256"]
    want_hist = read_golden(case["hist"])
    assert file_hist == want_hist
    assert dev_hist == want_hist


@pytest.mark.parametrize("path", ["direct", "partitioned"])
@pytest.mark.parametrize("mode", ["plain", "bloom", "host_staged"])
def test_count_5mbp_parity(ctx, mode, path):
    from ktcountermemorygreedy_amd import workloads as W
    g = load_json("count_5mbp.json")
    seq = W.COUNT_5MBP.reads(fastq=False).tobytes()
    dump, dev_hist, file_hist, st = gpu_views(ctx, seq, 31, g["count_max"], bloom=(mode == "bloom"),
                                              host_path=(mode == "host_staged"), path=path)
    assert st["kmers"] == g["kmers"]
    assert hashlib.sha256(dump).hexdigest() == g["dump_sha256"]
    assert hashlib.sha256(dev_hist).hexdigest() == g["hist_sha256"]
    assert hashlib.sha256(file_hist).hexdigest() == g["hist_sha256"]
    if mode != "bloom":
        assert st["distinct"] == g["distinct"]


@pytest.mark.parametrize("k", [21, 31, 63])
def test_uneven_reads_counts_match_oracle(ctx, k):
    """Long and short reads in one launch: exact multiset vs the oracle."""
    from ktcountermemorygreedy_amd import workloads as W
    g = W.ReadSet("g", 3, 3_000_000, 13, 0).genome()
    seq = W.UNEVEN_50MBP.scaled(1500).reads(genome=g).tobytes()
    nw = O.words_for_k(k)
    keys = O.extract(seq, k)
    uniq, cnt = O.count(keys, nw)
    t = gpu_count(ctx, seq, k, 0xFFFFFFFF, max_distinct=len(seq))
    st = t.stats()
    assert st["errors"] == 0 and st["kmers"] == keys.shape[0] and st["distinct"] == uniq.shape[0]
    got = t.query(uniq).astype(np.uint64)
    assert np.array_equal(got, cnt)
    # absent keys read 0
    absent = uniq.copy()
    absent[:, nw - 1] |= np.uint64(1)  # padding bit set: never a stored key
    if 64 * nw - 2 * k > 0:
        assert int(t.query(absent).sum()) == 0
    t.close()


@pytest.mark.parametrize("capacity", [1 << 20, 1 << 21, 1 << 23, 1 << 29])
@pytest.mark.parametrize("bloom", [False, True], ids=["plain", "bloom"])
@pytest.mark.parametrize("k", [31, 63])
def test_partition_levels_match_oracle(ctx, capacity, bloom, k):
    """Partitioned path with 1, <=256 (one scatter level) and >256 (two
    levels) partitions, up to 2^18 partitions (2^29 slots: 256 to 1024 fine
    digits, the scatter forms that recompute digits): exact counts vs the
    oracle."""
    from ktcountermemorygreedy_amd import workloads as W
    g = W.ReadSet("g", 3, 400_000, 13, 0).genome()
    seq = W.COUNT_5MBP.scaled(6000).reads(genome=g).tobytes()
    nw = O.words_for_k(k)
    keys = O.extract(seq, k)
    uniq, cnt = O.count(keys, nw)
    t = gpu_count(ctx, seq, k, 0xFFFFFFFF, bloom=bloom, capacity=capacity, path="partitioned")
    st = t.stats()
    assert st["errors"] == 0 and st["kmers"] == keys.shape[0]
    got = t.query(uniq).astype(np.uint64)
    assert np.array_equal(got, cnt)
    bins = t.histogram(1, 1000)
    want = np.bincount(np.minimum(cnt, 1000).astype(np.int64), minlength=1001)[1:]
    assert np.array_equal(bins, want)
    t.close()


def test_partitioned_increment_keys(ctx):
    from ktcountermemorygreedy_amd import kc
    rng = np.random.default_rng(19)
    keys = (rng.integers(0, 2**62, 200_000, dtype=np.uint64) << np.uint64(2))
    rep = np.repeat(keys, rng.integers(1, 30, keys.size))
    rng.shuffle(rep)
    t = kc.Table(ctx, 31, capacity=1 << 22, path="partitioned")
    t.increment(rep)
    uniq, cnt = O.count(rep.reshape(-1, 1), 1)
    assert np.array_equal(t.query(uniq).astype(np.uint64), cnt)
    assert t.stats()["distinct"] == keys.size
    t.close()


def test_increment_and_query_saturation(ctx):
    from ktcountermemorygreedy_amd import kc
    rng = np.random.default_rng(9)
    keys = (rng.integers(0, 2**62, 5000, dtype=np.uint64) << np.uint64(2))
    rep = np.repeat(keys, rng.integers(1, 400, keys.size))
    rng.shuffle(rep)
    t = kc.Table(ctx, 31, count_max=255, capacity=1 << 14)
    t.increment(rep)
    t.increment(keys, np.full(keys.size, 7, np.uint32))
    got = t.query(keys)
    uniq, cnt = O.count(rep.reshape(-1, 1), 1)
    order = np.argsort(uniq[:, 0])
    want = dict(zip(uniq[:, 0].tolist(), (cnt + 7).tolist()))
    assert [min(want[int(x)], 255) for x in keys] == got.tolist()
    assert t.stats()["distinct"] == keys.size
    del order
    t.close()


def test_table_full_reports_error(ctx):
    from ktcountermemorygreedy_amd import kc
    t = kc.Table(ctx, 31, capacity=1024)
    keys = (np.arange(5000, dtype=np.uint64) << np.uint64(2))
    t.increment(keys)
    st = t.stats()
    assert st["errors"] > 0
    with pytest.raises(kc.KCError) as e:
        t.export_device()
    assert e.value.status == kc.KC_EFULL
    t.close()


def test_clear_resets(ctx):
    from ktcountermemorygreedy_amd import kc
    seq = kc.fastq_frame(read_golden("tiny_mixed.fastq"))
    t = gpu_count(ctx, seq, 31, 0xFFFFFFFF, bloom=True)
    a = t.stats()
    t.clear()
    assert t.stats()["kmers"] == 0
    d = ctx.upload(np.frombuffer(seq, dtype=np.uint8))
    t.count_seq(d, len(seq))
    b = t.stats()
    assert a == b
    t.close()


@pytest.mark.parametrize("bloom", [False, True], ids=["plain", "bloom"])
def test_partitioned_pieces_exact(ctx, bloom):
    """An input larger than the tune's piece_bytes is counted in pieces that
    overlap by 64 bytes; every window is counted exactly once."""
    from ktcountermemorygreedy_amd import kc
    from ktcountermemorygreedy_amd import workloads as W
    g = W.ReadSet("g", 3, 1_000_000, 13, 0).genome()
    seq = W.UNEVEN_50MBP.scaled(2000).reads(genome=g).tobytes()
    keys = O.extract(seq, 31)
    uniq, cnt = O.count(keys, 1)
    t = kc.Table(ctx, 31, capacity=1 << 23, bloom=bloom, max_distinct=len(seq), path="partitioned")
    t.set_tune(piece_bytes=(1 << 18) + 48)  # rounded down to whole 16-byte groups
    d = ctx.upload(np.frombuffer(seq, dtype=np.uint8))
    t.count_seq(d, len(seq))
    ctx.sync()
    st = t.stats()
    assert len(seq) > 4 * (1 << 18)
    assert st["errors"] == 0 and st["kmers"] == keys.shape[0]
    assert np.array_equal(t.query(uniq).astype(np.uint64), cnt)
    t.close()


@pytest.mark.parametrize("path", ["direct", "partitioned"])
@pytest.mark.parametrize("k", [31, 63])
def test_first_list_spill_exact(ctx, k, path):
    """A FIRST list far too small (about 1/10 of the FIRST occurrences): full
    segments spill their FIRSTs into the table, the merged counts and the
    exported views stay exact and no error is reported."""
    from ktcountermemorygreedy_amd import kc
    from ktcountermemorygreedy_amd import workloads as W
    g = W.ReadSet("g", 3, 4_000_000, 13, 0).genome()
    seq = W.UNEVEN_50MBP.scaled(6000).reads(genome=g).tobytes()
    nw = O.words_for_k(k)
    keys = O.extract(seq, k)
    uniq, cnt = O.count(keys, nw)
    # the list keeps 4096 keys of slack per segment (~4.2M keys) whatever
    # its nominal size: this input has about twice as many FIRSTs
    t = kc.Table(ctx, k, capacity=1 << 24, bloom=True, max_distinct=uniq.shape[0],
                 low_capacity=uniq.shape[0] // 10, path=path)
    d = ctx.upload(np.frombuffer(seq, dtype=np.uint8))
    t.count_seq(d, len(seq))
    ctx.sync()
    st = t.stats()
    assert st["errors"] == 0 and st["kmers"] == keys.shape[0]
    assert st["low"] < 0.75 * uniq.shape[0]  # a large part of the FIRSTs was spilled
    bins = t.histogram(1, 1000)
    want = np.bincount(np.minimum(cnt, 1000).astype(np.int64), minlength=1001)[1:]
    assert np.array_equal(bins, want)
    high, low = t.export_device()
    dump, hist = views_from_export(high, low, k, 0xFFFFFFFF, 1000)
    assert dump == O.dump_text(uniq, cnt, k)
    assert hist == O.histogram_text(want, 1, 1000)
    got = t.query(uniq).astype(np.uint64)
    assert np.array_equal(got, cnt)
    t.close()
    d.free()


@pytest.mark.parametrize("filter_min", [0, 1, 2])
def test_export_filter_wraps_at_u32_max(ctx, filter_min):
    """The counter's export filter tests count + 1 < filter_min in u32
    (NewKmerTree.h:1305, FinalDrainWriter.h:101): a count saturated at
    2^32 - 1 wraps to 0, so its record is left out of high.bin whenever
    filter_min >= 1 (the reference default is 2) and kept at 0.  Lookups
    are not filtered."""
    from ktcountermemorygreedy_amd import kc
    keys = np.array([11, 22, 33, 44], dtype=np.uint64) << np.uint64(2)
    t = kc.Table(ctx, 31, count_max=0xFFFFFFFF, capacity=1 << 12, filter_min=filter_min)
    t.increment(keys, np.array([0xFFFFFFFF, 0xFFFFFFFE, 7, 0xFFFFFFFF], np.uint32))
    t.increment(keys[3:], np.array([9], np.uint32))  # 2^32 + 8: saturates at 2^32 - 1
    high, low = t.export_device()
    hk, hc = O.parse_high(high, 31, 0xFFFFFFFF)
    got = dict(zip(hk[:, 0].tolist(), hc.tolist()))
    want = {22 << 2: 0xFFFFFFFE, 33 << 2: 7}
    if filter_min == 0:
        want.update({11 << 2: 0xFFFFFFFF, 44 << 2: 0xFFFFFFFF})
    assert got == want and low == b""
    assert t.dump().count(b"\n") == len(want)
    assert t.query(keys).tolist() == [0xFFFFFFFF, 0xFFFFFFFE, 7, 0xFFFFFFFF]
    t.close()


@pytest.mark.parametrize("path", ["direct", "partitioned"])
def test_count_on_context_leaving_cus(path):
    """A context whose stream leaves 16 CUs to other streams (the serving
    loop's job contexts) counts the 5 Mbp reads to the same golden dump."""
    from ktcountermemorygreedy_amd import kc
    from ktcountermemorygreedy_amd import workloads as W
    g = load_json("count_5mbp.json")
    seq = W.COUNT_5MBP.reads(fastq=False).tobytes()
    c = kc.Context(0, left_cus=16)
    try:
        dump, dev_hist, _, st = gpu_views(c, seq, 31, g["count_max"], path=path)
    finally:
        c.close()
    assert st["kmers"] == g["kmers"]
    assert hashlib.sha256(dump).hexdigest() == g["dump_sha256"]
    assert hashlib.sha256(dev_hist).hexdigest() == g["hist_sha256"]
