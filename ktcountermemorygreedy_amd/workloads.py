"""Seeded synthetic workloads the engine is measured and checked on.

Each workload is fully determined by its seeds; the generator is the
C++ routine in csrc/kc_host.cpp (SplitMix64 streams, one per read).

  count_5mbp      5 Mbp uniform genome (seed 1); 200k x 150 bp reads from both
                  strands, 0.5% substitutions, 0.01% N, Phred33 20..40; k=31
  repeat_100mbp   100 Mbp genome (seed 2) with 30% of its bases written as
                  diverged (1%) copies of 500 random 1 kb segments; 20M x
                  150 bp reads (30x), 0.5% substitutions; k=31, Bloom split on
  uneven_50mbp    50 Mbp genome (seed 3); 80% 150 bp reads (0.5% error) and
                  20% long reads (lognormal mean 8 kb, sd 3 kb, 5% error) by
                  read count, 40x coverage; k = 21, 31, 63
  repeat_4mbp     repeat_100mbp's generator on a 4 Mbp genome (same seeds,
                  500 x 1 kb segments, 30% copies, 1% divergence), 300k reads
                  (11x): the repeat-rich parity case
  repeat_4mbp_hot the same with 20 segments, so each is copied about 60 times
                  as in repeat_100mbp (hot keys counted in the hundreds)
"""
from __future__ import annotations

import math
from dataclasses import dataclass

from . import kc


@dataclass
class ReadSet:
    name: str
    genome_seed: int
    genome_len: int
    read_seed: int
    n_reads: int
    read_len: int = 150
    sub_rate: float = 0.005
    n_rate: float = 0.0
    n_segments: int = 0
    segment_len: int = 0
    repeat_fraction: float = 0.0
    segment_divergence: float = 0.0
    long_fraction: float = 0.0
    long_mean: float = 0.0
    long_sd: float = 0.0
    long_sub_rate: float = 0.0

    def lognormal(self):
        if not self.long_fraction:
            return 0.0, 0.0
        s2 = math.log(1.0 + (self.long_sd / self.long_mean) ** 2)
        return math.log(self.long_mean) - s2 / 2.0, math.sqrt(s2)

    def scaled(self, n_reads, name=None):
        d = dict(self.__dict__)
        d["n_reads"] = n_reads
        d["name"] = name or "%s_%d" % (self.name, n_reads)
        return ReadSet(**d)

    def genome(self):
        return kc.synth_genome(self.genome_seed, self.genome_len, self.n_segments, self.segment_len,
                               self.repeat_fraction, self.segment_divergence)

    def reads(self, fastq=False, genome=None, threads=0, out=None):
        g = self.genome() if genome is None else genome
        mu, sigma = self.lognormal()
        return kc.synth_reads(g, self.read_seed, self.n_reads, self.read_len, True, self.sub_rate,
                              self.n_rate, self.long_fraction, mu, sigma, self.long_sub_rate,
                              min_len=100, fastq=fastq, threads=threads, out=out)


COUNT_5MBP = ReadSet("count_5mbp", 1, 5_000_000, 11, 200_000, n_rate=1e-4)
REPEAT_100MBP = ReadSet("repeat_100mbp", 2, 100_000_000, 12, 20_000_000, n_segments=500,
                        segment_len=1000, repeat_fraction=0.30, segment_divergence=0.01)
UNEVEN_50MBP = ReadSet("uneven_50mbp", 3, 50_000_000, 13, 1_163_000, long_fraction=0.2,
                       long_mean=8000.0, long_sd=3000.0, long_sub_rate=0.05)
REPEAT_4MBP = ReadSet("repeat_4mbp", 2, 4_000_000, 12, 300_000, n_segments=500, segment_len=1000,
                      repeat_fraction=0.30, segment_divergence=0.01)
REPEAT_4MBP_HOT = ReadSet("repeat_4mbp_hot", 2, 4_000_000, 12, 300_000, n_segments=20, segment_len=1000,
                          repeat_fraction=0.30, segment_divergence=0.01)


# BGZF, the blocked gzip of htslib's bgzip (SAM/BAM format specification,
# section 4.1): gzip members of at most 65280 input bytes, each carrying its
# own size in a "BC" extra subfield, then an empty member as end-of-file
# marker.  Any gzip reader inflates it as one stream; kc_count_fastq_files
# inflates its members on several threads.
BGZF_BLOCK = 65280
BGZF_EOF = bytes.fromhex("1f8b08040000000000ff0600424302001b0003000000000000000000")


def bgzf_member(data, level=6) -> bytes:
    import struct
    import zlib
    c = zlib.compressobj(level, zlib.DEFLATED, -15)
    body = c.compress(data) + c.flush()
    size = 18 + len(body) + 8
    head = b"\x1f\x8b\x08\x04\x00\x00\x00\x00\x00\xff\x06\x00BC\x02\x00" + struct.pack("<H", size - 1)
    return head + body + struct.pack("<II", zlib.crc32(data) & 0xFFFFFFFF, len(data))


def write_bgzf(path, data, level=6, threads=8):
    """Writes data (bytes or a uint8 array) to path as BGZF, compressing the
    blocks on `threads` threads (zlib releases the GIL); returns the file size."""
    from concurrent.futures import ThreadPoolExecutor
    mv = memoryview(data).cast("B")
    spans = [(i, min(i + BGZF_BLOCK, len(mv))) for i in range(0, len(mv), BGZF_BLOCK)]
    n = 0
    with open(path, "wb") as f, ThreadPoolExecutor(max(1, threads)) as ex:
        for i in range(0, len(spans), 4096):  # bounded memory: 4096 blocks at a time
            for blk in ex.map(lambda s: bgzf_member(mv[s[0]:s[1]], level), spans[i:i + 4096]):
                f.write(blk)
                n += len(blk)
        f.write(BGZF_EOF)
    return n + len(BGZF_EOF)


def write_gzip_parallel(path, data, level=6, threads=8, block=4 << 20):
    """Writes data as ONE gzip member (a single deflate stream, as gzip -6
    writes it, so readers must inflate it in order), compressing 4 MB blocks
    on `threads` threads the way pigz does: each block primed with the 32 KB
    before it as dictionary and ended by a sync flush, the last one by the
    final flush; the CRC-32 runs over the whole input.  Returns the file size."""
    import struct
    import zlib
    from concurrent.futures import ThreadPoolExecutor
    mv = memoryview(data).cast("B")
    n = len(mv)
    spans = [(i, min(i + block, n)) for i in range(0, n, block)] or [(0, 0)]

    def comp(sp):
        a, b = sp
        c = zlib.compressobj(level, zlib.DEFLATED, -15, 9, zlib.Z_DEFAULT_STRATEGY,
                             *([bytes(mv[max(0, a - 32768):a])] if a else []))
        return c.compress(mv[a:b]) + c.flush(zlib.Z_FINISH if b == n else zlib.Z_SYNC_FLUSH)

    size = 0
    with open(path, "wb") as f, ThreadPoolExecutor(max(1, threads)) as ex:
        f.write(b"\x1f\x8b\x08\x00\x00\x00\x00\x00\x00\xff")
        for i in range(0, len(spans), 64):  # bounded memory: 64 blocks at a time
            for part in ex.map(comp, spans[i:i + 64]):
                f.write(part)
                size += len(part)
        f.write(struct.pack("<II", zlib.crc32(mv) & 0xFFFFFFFF, n & 0xFFFFFFFF))
    return size + 18
