"""Keys of k = 64..128 on the device (kmer<2> at k = 64 and kmer<4> above as
the caller sees them; four-word slots with a claim-state word inside the
table): counts, the device dump and histogram, the exported files, lookups,
multi-pass slices, batches with halos, temp_dir tools and the record sort.

Expected values come from the oracle, which is pinned at k = 64 by the
reference's dump and histogram goldens, and above by a count done on
strings and the reference's histograms (tests/test_oracle.py; the
reference's dump text at k > 64 and its counts at k = 96 are defective).
At k = 64 the device views are also compared with the reference's files
directly."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.kc_testlib import gpu_count, load_json, read_golden, views_from_export

pytestmark = pytest.mark.gpu

CASES = load_json("tiny_wide.json")


def _oracle_counts(seq: bytes, k: int):
    nw = O.words_for_k(k)
    keys = O.extract(seq, k)
    uniq, cnt = O.count(keys, nw)
    return keys, uniq, cnt


def _absent_keys(uniq, k, n, seed):
    """Valid k = k keys (zero padding) that are not in uniq."""
    nw = O.words_for_k(k)
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, 2**63, size=(4 * n, nw), dtype=np.uint64) * np.uint64(2) + \
        rng.integers(0, 2, size=(4 * n, nw), dtype=np.uint64)
    used = (k - 1) // 32 + 1
    keys[:, used:] = 0
    rem = k - 32 * (used - 1)
    if rem < 32:
        keys[:, used - 1] &= np.uint64(((1 << (2 * rem)) - 1) << (64 - 2 * rem))
    have = {tuple(x) for x in uniq.tolist()}
    keep = [i for i, x in enumerate(keys.tolist()) if tuple(x) not in have]
    return keys[keep[:n]]


@pytest.mark.parametrize("bloom", [False, True], ids=["plain", "bloom"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "k%d-c%d" % (c["k"], c["count_max"]))
def test_tiny_wide_views(ctx, case, bloom):
    from ktcountermemorygreedy_amd import kc
    k, cm, hm = case["k"], case["count_max"], case["hist_max"]
    fq = read_golden(case["input"])
    seq = kc.fastq_frame(fq)
    print("wide views: k=%d count_max=%d bloom=%s" % (k, cm, bloom), flush=True)
    want_dump, want_hist, n_kmers = O.count_fastq_views(fq, k, cm, 1, hm)
    t = gpu_count(ctx, seq, k, cm, bloom=bloom)
    st = t.stats()
    assert st["errors"] == 0 and st["kmers"] == n_kmers, st
    dump = t.dump()
    same = dump == want_dump  # (no pytest diff of large texts)
    assert same
    same = t.dump() == dump  # the merge flags are cleared
    assert same
    assert O.histogram_text(t.histogram(1, hm), 1, hm) == want_hist
    high, low = t.export_device()
    same = views_from_export(high, low, k, cm, hm) == (want_dump, want_hist)
    assert same
    if k == 64:  # the reference's own files
        same = dump == read_golden(case["dump"])
        assert same
        assert O.histogram_text(t.histogram(1, hm), 1, hm) == read_golden(case["hist"])
    _, uniq, cnt = _oracle_counts(seq, k)
    assert np.array_equal(t.query(uniq).astype(np.uint64), np.minimum(cnt, cm))
    absent = _absent_keys(uniq, k, 500, k)
    assert not t.query(absent).any()
    t.close()


@pytest.mark.parametrize("bloom", [False, True], ids=["plain", "bloom"])
@pytest.mark.parametrize("k", [64, 65, 96, 128])
def test_wide_reads_exact(ctx, k, bloom):
    """3 Mbp of reads with errors and Ns from both strands of a 200 kb
    genome: every k-mer count, through lookups, the device histogram and
    the device dump."""
    from ktcountermemorygreedy_amd import kc
    g = kc.synth_genome(11, 200_000)
    seq = kc.synth_reads(g, 12, 20_000, 150, True, 0.005, 1e-3).tobytes()
    print("wide reads: k=%d bloom=%s bytes=%d" % (k, bloom, len(seq)), flush=True)
    keys, uniq, cnt = _oracle_counts(seq, k)
    t = gpu_count(ctx, seq, k, 0xFFFFFFFF, bloom=bloom)
    st = t.stats()
    assert st["errors"] == 0 and st["kmers"] == keys.shape[0], st
    assert np.array_equal(t.query(uniq).astype(np.uint64), cnt)
    want = np.bincount(np.minimum(cnt, 1000).astype(np.int64), minlength=1001)[1:]
    assert np.array_equal(t.histogram(1, 1000), want)
    same = t.dump() == O.dump_text(uniq, cnt.astype(np.uint32), k)
    assert same
    t.close()


@pytest.mark.parametrize("k", [64, 100])
def test_wide_pass_slices(ctx, k):
    """Multi-pass slices at k = 64 / 100: pass p counts exactly the keys whose
    XXH3_64 over the caller's key bytes (16 at k = 64, 32 above) has top bits
    p (oracle extract_slice)."""
    from ktcountermemorygreedy_amd import kc
    g = kc.synth_genome(13, 100_000)
    seq = kc.synth_reads(g, 14, 6000, 150, True, 0.005, 1e-3).tobytes()
    nw = O.words_for_k(k)
    t = kc.Table(ctx, k, capacity=1 << 20, max_distinct=len(seq))
    d = ctx.upload(np.frombuffer(seq, dtype=np.uint8))
    total = 0
    for p in range(4):
        t.clear()
        t.set_pass(p, 4)
        t.count_seq(d, len(seq))
        ctx.sync()
        mine = O.extract_slice(seq, k, 2, p)
        uniq, cnt = O.count(mine, nw)
        st = t.stats()
        assert st["errors"] == 0 and st["kmers"] == mine.shape[0] and st["distinct"] == uniq.shape[0], (p, st)
        assert np.array_equal(t.query(uniq).astype(np.uint64), cnt)
        total += st["kmers"]
    assert total == O.extract(seq, k).shape[0]
    d.free()
    t.close()


@pytest.mark.parametrize("k", [64, 127])
def test_wide_batch_with_halo(ctx, k):
    """Batched host chunks, one of them larger than a staging slot (counted
    in pieces with a 128-byte halo): exact counts."""
    from ktcountermemorygreedy_amd import kc
    from tests.test_gpu_batch import _check_exact, _chunks, _joined, _reads
    short, long_ = _reads(5, 4000, 60)
    rng = np.random.default_rng(k)
    chunks = _chunks(short + long_, rng, 200_000)
    chunks.insert(len(chunks) // 2, long_ * 3)
    seq = _joined(chunks)
    print("wide batch: k=%d chunks=%d bytes=%d" % (k, len(chunks), len(seq)), flush=True)
    t = kc.Table(ctx, k, capacity=1 << 22, max_distinct=len(seq))
    b = kc.Batch(t, batch_bytes=1 << 20, queue_depth=3, path="partitioned")
    for c in chunks:
        b.add(c)
    b.flush()
    b.sync()
    info = b.info()
    print(info, flush=True)
    assert info["path"] == 1  # four-word keys count on the direct path
    assert info["pieces"] >= 2
    b.close()
    _check_exact(t, seq, k)
    t.close()


@pytest.mark.parametrize("with_low", [True, False], ids=["precise", "approximate"])
@pytest.mark.parametrize("k", [64, 100])
def test_wide_dir_round_trip(ctx, tmp_path, k, with_low):
    """export_dir -> load_dir / histogram_dir / dump_dir at k = 64 / 100:
    the files hold the oracle's packing and the tools' views equal the
    oracle's merge of them."""
    from ktcountermemorygreedy_amd import kc
    fq = read_golden("tiny_wide.fastq")
    seq = kc.fastq_frame(fq)
    cm = 255
    t = gpu_count(ctx, seq, k, cm, bloom=True)
    out = tmp_path / "dir"
    out.mkdir()
    t.export_dir(str(out) + "/")
    high, low = (out / "high.bin").read_bytes(), (out / "low.bin").read_bytes()
    dh, dl = t.export_device()
    t.close()
    nw = O.words_for_k(k)
    # the files hold the records of the device export (compared as sets)
    same_high = len(dh) == len(high) and sorted(O.parse_high(dh, k, cm)[0].tolist()) == \
        sorted(O.parse_high(high, k, cm)[0].tolist())
    same_low = len(dl) == len(low) and sorted(O.unpack(dl, k).tolist()) == sorted(O.unpack(low, k).tolist())
    assert same_high and same_low
    hk, hc = O.parse_high(high, k, cm)
    lk = O.unpack(low, k) if (low and with_low) else np.zeros((0, nw), np.uint64)
    mk, mc = O.merge(hk, hc, lk, nw, cm)
    bins = O.histogram(hk, hc, lk, nw, cm, 1, 300)
    assert np.array_equal(ctx.histogram_dir(str(out), 1, 300, precise=with_low), bins)
    ctx.dump_dir(str(out), str(tmp_path / "dump.txt"), precise=with_low)
    want = O.dump_text(mk, mc, k)
    same = (tmp_path / "dump.txt").read_bytes() == want
    assert same
    t2 = kc.Table.load_dir(ctx, str(out), with_low=with_low)
    same = t2.dump() == want
    assert same
    assert np.array_equal(t2.query(mk).astype(np.uint64), mc.astype(np.uint64))
    t2.close()


@pytest.mark.parametrize("k", [64, 65, 96, 128])
def test_wide_export_records(ctx, k):
    """kc_export_low / kc_export_high on caller keys: the oracle's packing
    (ExportWriter.h:296-313), with the counter's filter."""
    rng = np.random.default_rng(k)
    uniq = _absent_keys(np.zeros((0, O.words_for_k(k)), np.uint64), k, 3000, k + 1)
    counts = rng.integers(1, 256, size=uniq.shape[0]).astype(np.uint32)
    same = ctx.export_low(uniq, k) == O.pack(uniq, k)
    assert same
    high = ctx.export_high(uniq, counts, k, count_max=255, filter_min=5, filter_max=200)
    hk, hc = O.parse_high(high, k, 255)
    keep = (counts + 1 >= 5) & (counts <= 200)  # FinalDrainWriter.h:95-104
    assert np.array_equal(hk, uniq[keep]) and np.array_equal(hc, counts[keep])


@pytest.mark.parametrize("k,bits", [(128, 12), (100, 8), (65, 12), (96, 8)])
@pytest.mark.parametrize("packed", [True, False], ids=["packed", "padded"])
def test_wide_sort_vs_oracle(ctx, k, bits, packed):
    """kc_sort_records with four-word keys (36- or 40-byte records) against
    the oracle's restatement of the reference sorts: stable, on the same bit
    range."""
    from tests.test_gpu_sort import device_sort
    n = 60_000
    keys = _absent_keys(np.zeros((0, 4), np.uint64), k, n, k)
    n = keys.shape[0]
    keys[n // 3: n // 3 + 5000] = keys[:5000]  # equal keys: stability
    counts = np.arange(n, dtype=np.uint32)
    want = O.sort_records12(keys, counts, k, 10) if bits == 12 else O.sort_records8(keys, counts, k)
    got = device_sort(ctx, keys, counts, 4, k, bits, 36 if packed else 40)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    recs = np.concatenate([keys.view(np.uint32).reshape(n, 8), counts.reshape(n, 1)], axis=1)
    bins = ctx.records_histogram(ctx.upload(recs), n, 4, 36, 1, n)
    assert np.array_equal(bins, np.bincount(counts[counts >= 1] - 1, minlength=n))
