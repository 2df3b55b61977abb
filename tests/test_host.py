"""Host-side product code (framing, synthetic generator) on the CPU."""
import hashlib

import numpy as np
import pytest

from oracle import oracle as O
from tests.kc_testlib import load_json, read_golden


def test_framing_matches_oracle():
    from ktcountermemorygreedy_amd import kc
    for case in load_json("tiny_cases.json"):
        fq = read_golden(case["input"])
        assert kc.fastq_frame(fq) == O.fastq_frame(fq)


def test_frame_file_plain_and_gzip(tmp_path):
    import gzip
    from ktcountermemorygreedy_amd import kc
    fq = read_golden("tiny_mixed.fastq")
    p = tmp_path / "a.fastq"
    p.write_bytes(fq)
    g = tmp_path / "a.fastq.gz"
    g.write_bytes(gzip.compress(fq))
    want = O.fastq_frame(fq)
    assert kc.fastq_frame_file(str(p)).tobytes() == want
    assert kc.fastq_frame_file(str(g)).tobytes() == want


def _uneven_fastq(n_reads, seed=5):
    """FASTQ of reads of uneven length (both short and long), so member runs
    end in every record line and mid-line."""
    from ktcountermemorygreedy_amd import workloads as W
    rs = W.UNEVEN_50MBP.scaled(n_reads)
    g = W.ReadSet("g", seed, 2_000_000, 13, 0).genome()
    return rs.reads(fastq=True, genome=g).tobytes()


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_bgzf_parallel_framing_equals_stream(tmp_path, threads):
    """A BGZF file (many member runs) framed by the parallel reader equals
    the framing of the decompressed bytes, and the stream reader's."""
    from ktcountermemorygreedy_amd import kc
    from ktcountermemorygreedy_amd import workloads as W
    fq = _uneven_fastq(3000)
    p = tmp_path / "r.fastq.gz"
    W.write_bgzf(str(p), fq, level=1, threads=4)
    want = O.fastq_frame(fq)
    got, st = kc.fastq_frame_files([str(p)], threads=threads)
    assert got.tobytes() == want
    assert st["input_bytes"] == len(fq) and st["framed_bytes"] == len(want)
    if threads > 1:  # the file was cut into runs of members
        assert p.stat().st_size > 4 << 20 and st["chunks"] > 4
    assert kc.fastq_frame_file(str(p)).tobytes() == want


def test_bgzf_missing_final_newline_and_mixed_files(tmp_path):
    """A BGZF file whose last record lacks its final newline, next to a plain
    gzip and a plain file: each file's reads come out whole and in order."""
    import gzip
    from ktcountermemorygreedy_amd import kc
    from ktcountermemorygreedy_amd import workloads as W
    a = _uneven_fastq(2500, seed=7)
    a = a[:a.rstrip(b"\n").rfind(b"\n") + 1] + b"+\n" + b"I" * 10  # cut the last quality line short, no newline
    b = _uneven_fastq(300, seed=9)
    c = _uneven_fastq(200, seed=11)
    pa, pb, pc = tmp_path / "a.fq.gz", tmp_path / "b.fq.gz", tmp_path / "c.fq"
    W.write_bgzf(str(pa), a, level=1)
    pb.write_bytes(gzip.compress(b, 1))
    pc.write_bytes(c)
    got, st = kc.fastq_frame_files([str(pa), str(pb), str(pc)], threads=6)
    want_lines = sorted(l for x in (a, b, c) for l in O.fastq_frame(x).split(b"\n")[:-1])
    assert sorted(got.tobytes().split(b"\n")[:-1]) == want_lines
    # file a alone keeps its order
    got_a, _ = kc.fastq_frame_files([str(pa)], threads=6)
    fa = O.fastq_frame(a)
    assert got_a.tobytes() == (fa if fa.endswith(b"\n") else fa + b"\n")


@pytest.mark.parametrize("threads", [1, 2, 5])
def test_gzip_stream_with_framing_helper(tmp_path, threads):
    """A single-member gzip file (no members to split) with spare threads:
    one thread inflates into a ring of blocks while a helper frames them; the
    framed reads equal the stream framing, also without a final newline and
    next to a second file that takes a thread of its own."""
    import gzip
    from ktcountermemorygreedy_amd import kc
    a = _uneven_fastq(1500, seed=21)
    a = a[:-1]  # no final newline
    b = _uneven_fastq(200, seed=23)
    pa, pb = tmp_path / "a.fastq.gz", tmp_path / "b.fastq.gz"
    pa.write_bytes(gzip.compress(a, 1))
    pb.write_bytes(gzip.compress(b, 1))
    got, st = kc.fastq_frame_files([str(pa)], threads=threads)
    want = O.fastq_frame(a)
    assert got.tobytes() == (want if want.endswith(b"\n") else want + b"\n")
    assert st["input_bytes"] == len(a)
    got2, _ = kc.fastq_frame_files([str(pa), str(pb)], threads=threads)
    lines = sorted(got2.tobytes().split(b"\n")[:-1])
    assert lines == sorted(O.fastq_frame(a).split(b"\n")[:-1] + O.fastq_frame(b).split(b"\n")[:-1])


def _gzip_fixed_codes(data):
    """A gzip member of fixed-Huffman blocks only (no dynamic block to find)."""
    import zlib
    c = zlib.compressobj(6, zlib.DEFLATED, 31, 9, zlib.Z_FIXED)
    return c.compress(data) + c.flush()


@pytest.mark.parametrize("kind", ["gzip6", "parallel_writer", "fixed_codes", "members", "binary"])
def test_gzip_member_in_shares(tmp_path, monkeypatch, kind):
    """One gzip member cut into shares inflated on several threads (shares of
    64 KB here): the framed reads equal the one-stream reader's, for a gzip -6
    member, the parallel writer's sync-flushed blocks, fixed-code blocks (no
    dynamic block to start from: every share decoded again in order), two
    members (the second through zlib) and bytes that are not text."""
    import gzip
    import os
    from ktcountermemorygreedy_amd import kc
    from ktcountermemorygreedy_amd import workloads as W
    fq = _uneven_fastq(2000, seed=31)
    p = tmp_path / "r.fastq.gz"
    if kind == "gzip6":
        p.write_bytes(gzip.compress(fq, 6))
    elif kind == "parallel_writer":
        W.write_gzip_parallel(str(p), fq, level=6, threads=4, block=256 << 10)
    elif kind == "fixed_codes":
        p.write_bytes(_gzip_fixed_codes(fq))
    elif kind == "members":
        h = len(fq) // 3
        h = fq.index(b"\n@", h) + 1
        p.write_bytes(gzip.compress(fq[:h], 6) + gzip.compress(fq[h:], 1))
    else:
        fq = os.urandom(1 << 20) + fq[:1 << 20]
        p.write_bytes(gzip.compress(fq, 6))
    want = O.fastq_frame(fq)
    want = want if not want or want.endswith(b"\n") else want + b"\n"
    one, _ = kc.fastq_frame_files([str(p)], threads=1)
    assert one.tobytes() == want
    monkeypatch.setenv("KC_TEST_GZIP_CHUNK", str(64 << 10))
    got, st = kc.fastq_frame_files([str(p)], threads=6)
    assert got.tobytes() == want
    assert st["input_bytes"] == len(fq) and st["threads"] == 6


def test_gzip_member_in_shares_errors(tmp_path, monkeypatch):
    """A member whose bytes no longer match its CRC, and a member cut short:
    KC_EIO, as gzread reports them."""
    import gzip
    from ktcountermemorygreedy_amd import kc
    fq = _uneven_fastq(1500, seed=33)
    z = bytearray(gzip.compress(fq, 6))
    monkeypatch.setenv("KC_TEST_GZIP_CHUNK", str(64 << 10))
    bad = bytearray(z)
    bad[-8] ^= 0x01  # the stored CRC-32
    (tmp_path / "crc.fastq.gz").write_bytes(bytes(bad))
    (tmp_path / "cut.fastq.gz").write_bytes(bytes(z[:len(z) * 2 // 3]))
    for name in ("crc.fastq.gz", "cut.fastq.gz"):
        with pytest.raises(kc.KCError) as e:
            kc.fastq_frame_files([str(tmp_path / name)], threads=4)
        assert e.value.status == kc.KC_EIO


def test_bgzf_corrupt_member_is_an_error(tmp_path):
    """A member whose data no longer matches its CRC fails with KC_EIO, as
    gzread does; nothing hangs."""
    from ktcountermemorygreedy_amd import kc
    from ktcountermemorygreedy_amd import workloads as W
    fq = _uneven_fastq(3000, seed=13)
    p = tmp_path / "bad.fastq.gz"
    W.write_bgzf(str(p), fq, level=1)
    raw = bytearray(p.read_bytes())
    # flip one byte inside a member's deflate data in the middle of the file
    at = len(raw) // 2
    at = raw.rfind(b"\x1f\x8b\x08\x04", 0, at) + 18 + 100
    raw[at] ^= 0x55
    p.write_bytes(bytes(raw))
    with pytest.raises(kc.KCError):
        kc.fastq_frame_files([str(p)], threads=4)


def test_gzip_detected_by_suffix_as_the_reader_does(tmp_path):
    """FastqReader.h:423,575: a ".gz" name is inflated, any other file is read
    as it is (gzip bytes without the suffix are framed as text), in both
    host readers."""
    import gzip
    from ktcountermemorygreedy_amd import kc
    fq = read_golden("tiny_mixed.fastq")
    z = gzip.compress(fq)
    p = tmp_path / "no_suffix.fastq"
    p.write_bytes(z)
    want = O.fastq_frame(z)
    assert kc.fastq_frame_file(str(p)).tobytes() == want
    got, _ = kc.fastq_frame_files([str(p)], threads=2)
    assert got.tobytes() == (want if not want or want.endswith(b"\n") else want + b"\n")
    g = tmp_path / "plain_named.fastq.gz"
    g.write_bytes(fq)  # a plain file with the suffix: zlib passes it through
    assert kc.fastq_frame_file(str(g)).tobytes() == O.fastq_frame(fq)


def test_generator_is_deterministic_and_pinned():
    from ktcountermemorygreedy_amd import workloads as W
    g = load_json("count_5mbp.json")
    a = W.COUNT_5MBP.reads(fastq=True, threads=3)
    b = W.COUNT_5MBP.reads(fastq=True, threads=8)
    assert hashlib.sha256(a.tobytes()).hexdigest() == g["fastq_sha256"]
    assert np.array_equal(a, b)
    framed = W.COUNT_5MBP.reads(fastq=False)
    assert framed.tobytes() == O.fastq_frame(a.tobytes())


def test_uneven_reads_lengths():
    from ktcountermemorygreedy_amd import workloads as W
    rs = W.UNEVEN_50MBP.scaled(2000)
    g = W.ReadSet("g", 3, 2_000_000, 13, 0).genome()
    seq = rs.reads(genome=g).tobytes()
    lens = np.array([len(l) for l in seq.split(b"\n")[:-1]])
    assert len(lens) == 2000
    short = lens == 150
    assert 0.75 < short.mean() < 0.85
    assert 6000 < lens[~short].mean() < 10000


def test_merge_sorted_dumps(tmp_path):
    """k-way merge of key-sorted dump texts (the multi-pass dump merge):
    equals the sorted union; empty inputs and a missing final newline."""
    from ktcountermemorygreedy_amd import kc
    rng = np.random.default_rng(7)
    keys = sorted({"".join(rng.choice(list("ACGT"), 21)) for _ in range(3000)})
    parts = [[], [], [], []]
    for i, key in enumerate(keys):
        parts[int(rng.integers(0, 3))].append("%s\t%d\n" % (key, i % 97 + 1))
    paths = []
    for j, lines in enumerate(parts):  # part 3 stays empty
        p = tmp_path / ("p%d.txt" % j)
        text = "".join(lines)
        if j == 1 and text:
            text = text[:-1]
        p.write_text(text)
        paths.append(str(p))
    out = tmp_path / "merged.txt"
    kc.merge_sorted_dumps(paths, str(out))
    want = "".join("%s\t%d\n" % (key, i % 97 + 1) for i, key in enumerate(keys))
    assert out.read_text() == want
    with pytest.raises(kc.KCError):
        kc.merge_sorted_dumps([str(tmp_path / "missing.txt")], str(out))


def test_plan_sizes_reproduces_device_plans():
    """kc_plan_sizes is the sizing half of kc_plan_table: fed the scaled
    totals the device probe reported on two MI355X workloads it returns the
    device bytes of those plans (Bloom blocks of 8 bytes), and those sizes
    held what the full count then found (table keys, FIRST occurrences).
    Totals and outcomes are those runs'."""
    from ktcountermemorygreedy_amd import kc
    runs = [  # k, distinct, repeated, device_bytes, measured table keys, FIRST occurrences
        (31, 391599168, 101114048, 9732416840, 101989117, 391114715),    # 20M x 150 bp over 100 Mbp
        (63, 1836491392, 40757888, 37296572096, 58607250, 1818313084),  # 1.16M uneven reads over 50 Mbp
    ]
    for k, distinct, repeated, dev_bytes, table_keys, firsts in runs:
        kw, plan = kc.plan_sizes(k, distinct, repeated, bloom=True)
        assert plan["device_bytes"] == dev_bytes and plan["passes"] == 1
        cap, bloom = kw["capacity"], kw["bloom_blocks"]
        assert cap & (cap - 1) == 0 and bloom & (bloom - 1) == 0
        assert table_keys <= 0.625 * cap
        assert firsts <= kw["low_capacity"]
        assert distinct <= 8 * bloom
        slot = 16 if k <= 32 else 32
        assert dev_bytes == cap * slot + 8 * bloom + 8 * (k > 32) * kw["low_capacity"] + 8 * kw["low_capacity"]


def test_plan_sizes_passes_and_budget():
    from ktcountermemorygreedy_amd import kc
    # no Bloom: smallest power-of-two table at <= 62.5% load
    kw, plan = kc.plan_sizes(31, 1 << 26, 1 << 26)
    cap = kw["capacity"]
    assert cap == 1 << 27 and plan["device_bytes"] == cap * 16 and plan["passes"] == 1
    assert (1 << 26) <= 0.625 * cap and (1 << 26) > 0.3125 * cap
    # a budget of half the bytes takes two passes, each of half the keys
    kw2, plan2 = kc.plan_sizes(31, 1 << 26, 1 << 26, hbm_budget=plan["device_bytes"] // 2)
    assert plan2["passes"] == 2 and kw2["capacity"] == cap // 2
    assert plan2["device_bytes"] <= plan["device_bytes"] // 2
    # passes never decrease as the budget shrinks, and every plan fits
    last = 1
    for budget in (1 << 36, 1 << 34, 1 << 32, 1 << 30, 1 << 28):
        kw, p = kc.plan_sizes(63, 1 << 30, 1 << 26, bloom=True, hbm_budget=budget)
        assert p["passes"] >= last and p["device_bytes"] <= budget
        assert kw["max_distinct"] * p["passes"] >= 1 << 30
        last = p["passes"]
    with pytest.raises(kc.KCError):  # even 256 passes do not fit
        kc.plan_sizes(31, 1 << 34, 1 << 30, bloom=True, hbm_budget=1 << 20)
    # four-word keys (k = 64..128) take 64-byte slots
    kw4, p4 = kc.plan_sizes(100, 1 << 20, 1 << 20)
    assert p4["device_bytes"] == kw4["capacity"] * 64
    with pytest.raises(kc.KCError):
        kc.plan_sizes(129, 1000, 10)


def test_context_close_frees_its_buffers(monkeypatch):
    """Closing a Context frees the device and pinned buffers it allocated
    (once, before the context itself), not the views into them; freeing a
    buffer after its context closed is a no-op.  Runs against a stand-in
    for libkc that records the calls."""
    import ctypes as C
    from ktcountermemorygreedy_amd import kc
    calls, keep = [], []

    class Fake:
        def kc_ctx_create(self, dev, out):
            out._obj.value = 0x10
            return 0

        def kc_dev_alloc(self, h, n, out):
            out._obj.value = 0x1000 * (len(calls) + 1)
            calls.append(("alloc", out._obj.value))
            return 0

        def kc_dev_free(self, h, p):
            calls.append(("free", p.value))
            return 0

        def kc_host_alloc_pinned(self, h, n, out):
            buf = (C.c_uint8 * n)()
            keep.append(buf)
            out._obj.value = C.addressof(buf)
            return 0

        def kc_host_free_pinned(self, h, p):
            calls.append(("pinned_free", p.value))
            return 0

        def kc_ctx_destroy(self, h):
            calls.append(("destroy",))

    fake = Fake()
    monkeypatch.setattr(kc, "lib", lambda: fake)
    ctx = kc.Context(0)
    a = kc.DeviceBuffer(ctx, 64)
    b = kc.DeviceBuffer(ctx, 64)
    v = a.view(0, 16)
    p = kc.PinnedBuffer(ctx, 32)
    a_addr, b_addr, pin_addr = a.ptr, b.ptr, p.ptr
    b.free()  # freed by its owner first: not again at close
    ctx.close()
    frees = [c for c in calls if c[0] in ("free", "pinned_free")]
    assert sorted(frees) == sorted([("free", b_addr), ("free", a_addr), ("pinned_free", pin_addr)])
    assert calls[-1] == ("destroy",) and calls.count(("destroy",)) == 1
    assert a.ptr == 0 and p.ptr == 0 and v.ptr == a_addr  # the view owns nothing
    n = len(calls)
    a.free()
    p.free()
    v.free()
    ctx.close()
    assert len(calls) == n


def test_batch_add_chunks_binding(monkeypatch):
    """Batch.add_chunks hands the buffer, the chunk ends as uint64 and their
    count to kc_batch_add_chunks in one call, takes bytes or arrays, and
    rejects ends past the buffer before calling.  Runs against a stand-in
    for libkc that records the calls."""
    import ctypes as C
    from ktcountermemorygreedy_amd import kc
    seen = []

    class Fake:
        def kc_ctx_create(self, dev, out):
            out._obj.value = 0x10
            return 0

        def kc_table_create(self, ctx, cfg, out):
            out._obj.value = 0x20
            return 0

        def kc_batch_create(self, t, tune, out):
            out._obj.value = 0x30
            return 0

        def kc_batch_add_chunks(self, b, base, ends, n):
            e = np.ctypeslib.as_array((C.c_uint64 * n).from_address(ends.value)) if n else np.zeros(0, np.uint64)
            data = C.string_at(base.value, int(e[-1])) if n else b""
            seen.append((b.value, data, e.tolist()))
            return 0

        def kc_batch_destroy(self, b):
            pass

        def kc_table_destroy(self, t):
            pass

        def kc_ctx_destroy(self, h):
            pass

    fake = Fake()
    monkeypatch.setattr(kc, "lib", lambda: fake)
    ctx = kc.Context(0)
    t = kc.Table(ctx, 31, capacity=1 << 16)
    b = kc.Batch(t)
    b.add_chunks(b"ACGT\nAC\nGGT", [5, 8, 8, 11])
    b.add_chunks(np.frombuffer(b"TTTT\n", dtype=np.uint8), np.array([2, 5], dtype=np.int64))
    assert seen == [(0x30, b"ACGT\nAC\nGGT", [5, 8, 8, 11]), (0x30, b"TTTT\n", [2, 5])]
    with pytest.raises(ValueError):
        b.add_chunks(b"ACGT\n", [3, 9])
    assert len(seen) == 2
    ctx.close()
