"""Streaming ingest of FASTQ(.gz) files (kc_count_fastq_files): host threads
inflate and frame files in parallel into a kc_batch.  Counts equal the
oracle's over the files' framed reads; edge files (empty, no final newline,
plain and gzip mixed) included."""
import gzip

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _fastq(reads):
    out = []
    for i, r in enumerate(reads):
        out.append(b"@r%d\n%s\n+\n%s\n" % (i, r, b"I" * len(r)))
    return b"".join(out)


@pytest.mark.parametrize("threads", [1, 3])
def test_ingest_files_exact(ctx, tmp_path, threads):
    from ktcountermemorygreedy_amd import kc
    g = kc.synth_genome(17, 300_000)
    framed = kc.synth_reads(g, 18, 30_000, 150, True, 0.005, 1e-4).tobytes()
    reads = framed.split(b"\n")[:-1]
    rng = np.random.default_rng(5)
    cuts = np.sort(rng.choice(np.arange(1, len(reads)), 5, replace=False))
    parts = np.split(np.arange(len(reads)), cuts)
    paths, want = [], []
    for i, idx in enumerate(parts):
        fq = _fastq([reads[j] for j in idx])
        if i == 2:
            fq = fq[:-1]  # no final newline
        p = tmp_path / ("p%d.fastq%s" % (i, ".gz" if i % 2 else ""))
        p.write_bytes(gzip.compress(fq) if i % 2 else fq)
        paths.append(str(p))
        want.append(kc.fastq_frame(fq))
    (tmp_path / "empty.fastq.gz").write_bytes(gzip.compress(b""))
    paths.append(str(tmp_path / "empty.fastq.gz"))
    seq = b"".join(w if w.endswith(b"\n") else w + b"\n" for w in want)
    t = kc.Table(ctx, 31, capacity=1 << 22, bloom=True, max_distinct=len(seq))
    st = kc.count_fastq_files(t, paths, threads=threads, batch_bytes=1 << 20, queue_depth=3)
    print("ingest", st, flush=True)
    assert st["files"] == len(paths) and st["framed_bytes"] >= len(seq) - 1
    keys = O.extract(seq, 31)
    uniq, cnt = O.count(keys, 1)
    s = t.stats()
    assert s["errors"] == 0 and s["kmers"] == keys.shape[0]
    assert np.array_equal(t.query(uniq).astype(np.uint64), cnt)
    t.close()


def test_ingest_missing_file_reports_io_error(ctx, tmp_path):
    from ktcountermemorygreedy_amd import kc
    t = kc.Table(ctx, 31, capacity=1 << 16)
    with pytest.raises(kc.KCError) as e:
        kc.count_fastq_files(t, [str(tmp_path / "nope.fastq.gz")], threads=2, batch_bytes=1 << 20)
    assert e.value.status == kc.KC_EIO
    t.close()


@pytest.mark.parametrize("k", [31, 63])
def test_ingest_bgzf_single_file_parallel(ctx, tmp_path, k):
    """One BGZF file (many member runs) counted with 8 host threads sharing
    it, next to a plain gzip file: counts equal the oracle's over the framed
    reads; the same file streamed by one thread counts the same."""
    from ktcountermemorygreedy_amd import kc
    from ktcountermemorygreedy_amd import workloads as W
    rs = W.UNEVEN_50MBP.scaled(4000)
    g = W.ReadSet("g", 5, 2_000_000, 13, 0).genome()
    fq = rs.reads(fastq=True, genome=g).tobytes()
    more = _fastq([b"ACGTTGCA" * 20, b"N" * 40, b"ACGT" * 50])
    pa, pb = tmp_path / "a.fastq.gz", tmp_path / "b.fastq.gz"
    W.write_bgzf(str(pa), fq, level=1, threads=4)
    pb.write_bytes(gzip.compress(more, 1))
    seq = kc.fastq_frame(fq) + kc.fastq_frame(more)
    nw = O.words_for_k(k)
    keys = O.extract(seq, k)
    uniq, cnt = O.count(keys, nw)
    for threads in (8, 1):
        t = kc.Table(ctx, k, capacity=1 << 23, bloom=True, max_distinct=len(seq))
        st = kc.count_fastq_files(t, [str(pa), str(pb)], threads=threads, batch_bytes=4 << 20, queue_depth=3)
        print("ingest", threads, st, flush=True)
        assert st["framed_bytes"] == len(seq) and st["input_bytes"] == len(fq) + len(more)
        if threads > 1:
            assert st["chunks"] > 4  # the BGZF file went in member runs
        s = t.stats()
        assert s["errors"] == 0 and s["kmers"] == keys.shape[0], s
        assert np.array_equal(t.query(uniq).astype(np.uint64), cnt)
        t.close()
