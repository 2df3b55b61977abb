"""Writes the golden fixtures under tests/golden/ by running the reference
binaries built from source (oracle/_ref, see oracle/Makefile).

Run in the build container (it needs the reference tree to build _ref):
    make -C oracle ref && python tests/golden/make_goldens.py

Outputs (all data, no reference source):
  tiny_<name>.fastq / .k<k>.c<count_max>.dump.txt / .hist.txt
  count_5mbp.json          sha256 of the reference dump/histogram of the
                           seeded 5 Mbp workload plus its k-mer counts
  repeat_4mbp.json         the same for the repeat-rich workloads
                           repeat_4mbp and repeat_4mbp_hot
  tiny_counts.*            keys at counts 1..8 and 50: Tree run with non-default
                           filter_min / filter_max (argv 7-8), and dump_tool /
                           histogram_tool min/max ranges over two exports of it
                           (the plain split, and every key in low.bin with
                           count - 1 in high.bin), listed in tiny_counts.json
  tiny_wide.*              k = 64..128 (kmer<2> and kmer<4>) views of reads
                           from a small genome, listed in tiny_wide.json
                           (dump text at k = 64 only)
  hash_kat.txt             hashes and Bloom probes of fixed keys
  sort_kat_<nw>_<bits>.npz records before/after the reference sorts
"""
from __future__ import annotations

import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")
sys.path.insert(0, ROOT)


def ref_views(fastq_path, k, count_max, hist_max=10000, threads=6):
    """Runs Tree + dump_tool + histogram_tool; returns (dump bytes, hist bytes, log).
    The reference occasionally fails a run (its tools exit(-1)); such a run
    is repeated, up to three attempts, and the error is shown."""
    for attempt in range(3):
        try:
            return _ref_views_once(fastq_path, k, count_max, hist_max, threads)
        except subprocess.CalledProcessError as e:
            sys.stderr.write("reference run failed (attempt %d): %s\n%s\n" % (attempt, e.cmd, e.output))
    raise RuntimeError("reference run failed three times")


def _ref_views_once(fastq_path, k, count_max, hist_max, threads):
    tmp = tempfile.mkdtemp(prefix="kcgold_")
    try:
        log = subprocess.run([os.path.join(REF, "kmc_ref"), fastq_path, str(k), str(threads), "4",
                              tmp + "/", "1024", "2", "4294967295", str(count_max)],
                             stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, check=True).stdout
        subprocess.run([os.path.join(REF, "dump_tool_ref"), "precise", tmp + "/", str(threads), "4",
                        tmp + "/dump.txt", "1", "4294967295", "--sort"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, check=True)
        subprocess.run([os.path.join(REF, "histogram_tool_ref"), "precise", tmp + "/", str(threads), "4",
                        tmp + "/hist.txt", "1", str(hist_max)],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, check=True)
        with open(tmp + "/dump.txt", "rb") as f:
            dump = f.read()
        with open(tmp + "/hist.txt", "rb") as f:
            hist = f.read()
        return dump, hist, log
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def tiny_inputs():
    """Small FASTQ texts exercising the framing and window-reset rules."""
    rng = np.random.default_rng(20240601)

    def rand_seq(n, alphabet="ACGT"):
        return "".join(alphabet[i] for i in rng.integers(0, len(alphabet), n))

    out = {}
    # mixed case, U, N, other bytes, reads shorter than / equal to k
    recs = [
        rand_seq(150), rand_seq(31), rand_seq(30), rand_seq(64),
        rand_seq(40).lower(), rand_seq(50, "ACGU"), rand_seq(120, "ACGTN"),
        "ACGT" * 20, "A" * 70, "T" * 70, rand_seq(33) + "N" + rand_seq(33),
        rand_seq(90, "ACGTacgtRY"), rand_seq(63), rand_seq(100) + "\r",
    ]
    out["mixed"] = "".join("@r%d extra text\n%s\n+\n%s\n" % (i, s, "I" * len(s)) for i, s in enumerate(recs))
    # one read much longer than the reference's 32 KiB reader block
    long_read = rand_seq(100_000)
    out["long"] = "@long\n%s\n+\n%s\n@short\n%s\n+\n%s\n" % (long_read, "5" * len(long_read),
                                                             rand_seq(200), "5" * 200)
    # highly repetitive reads: few distinct k-mers with large counts
    motif = rand_seq(45)
    out["repeats"] = "".join("@m%d\n%s\n+\n%s\n" % (i, motif * 3, "I" * 135) for i in range(400))
    # palindromic k-mers (equal to their reverse complement) at even k
    pal = "ACGTACGTAATTACGTACGT"  # revcomp == itself
    pal_reads = [pal + rand_seq(40) for _ in range(10)] * 2
    out["palin"] = "".join("@p%d\n%s\n+\n%s\n" % (i, r, "I" * len(r)) for i, r in enumerate(pal_reads))
    # empty sequence lines and a record missing its final newline; one
    # record repeats so that high.bin is not empty (dump_tool refuses an
    # empty high.bin, dump_tool.cpp collect_high_files)
    dup = rand_seq(50)
    out["edges"] = "@e0\n\n+\n\n@d0\n%s\n+\n%s\n@d1\n%s\n+\n%s\n@e1\n%s\n+\n%s\n@e2\n%s\n+\n%s" % (
        dup, "I" * 50, dup, "I" * 50, rand_seq(80), "I" * 80, rand_seq(60), "I" * 60)
    return out


TINY_CASES = [
    # (input, k, count_max)
    ("mixed", 31, 0xFFFFFFFF), ("mixed", 21, 0xFFFFFFFF), ("mixed", 63, 0xFFFFFFFF),
    ("mixed", 32, 0xFFFFFFFF), ("mixed", 33, 0xFFFFFFFF), ("mixed", 15, 255),
    ("long", 31, 0xFFFFFFFF), ("long", 63, 0xFFFFFFFF),
    ("repeats", 31, 255), ("repeats", 31, 65535), ("repeats", 31, 0xFFFFFFFF), ("repeats", 41, 300),
    ("palin", 20, 0xFFFFFFFF), ("palin", 31, 0xFFFFFFFF),
    ("edges", 31, 0xFFFFFFFF), ("edges", 25, 0xFFFFFF),
]


def write_tiny():
    cases = []
    for name, text in tiny_inputs().items():
        with open(os.path.join(HERE, "tiny_%s.fastq" % name), "w") as f:
            f.write(text)
    for name, k, cm in TINY_CASES:
        path = os.path.join(HERE, "tiny_%s.fastq" % name)
        dump, hist, _ = ref_views(path, k, cm, hist_max=500)
        base = "tiny_%s.k%d.c%d" % (name, k, cm)
        case = {"input": "tiny_%s.fastq" % name, "k": k, "count_max": cm, "hist": base + ".hist.txt",
                "hist_max": 500, "dump_sha256": hashlib.sha256(dump).hexdigest(), "dump_bytes": len(dump)}
        if len(dump) <= 200_000:  # larger dumps are kept as digests only
            with open(os.path.join(HERE, base + ".dump.txt"), "wb") as f:
                f.write(dump)
            case["dump"] = base + ".dump.txt"
        with open(os.path.join(HERE, base + ".hist.txt"), "wb") as f:
            f.write(hist)
        cases.append(case)
    with open(os.path.join(HERE, "tiny_cases.json"), "w") as f:
        json.dump(cases, f, indent=1)


def tiny_wide_fastq():
    """Reads for k = 64..128 (kmer<2> at 64, kmer<4> above): 240 reads of
    150-260 bases from both strands of a 6 kb random genome (12x, so long
    k-mers repeat), some lower case, some with an N or a U, and a few reads
    shorter than k."""
    rng = np.random.default_rng(20251204)
    genome = "".join("ACGT"[i] for i in rng.integers(0, 4, 6000))
    comp = {"A": "T", "C": "G", "G": "C", "T": "A"}
    recs = []
    for i in range(240):
        n = int(rng.integers(150, 261))
        at = int(rng.integers(0, len(genome) - n))
        r = genome[at:at + n]
        if rng.random() < 0.5:
            r = "".join(comp[c] for c in reversed(r))
        if i % 17 == 0:
            j = int(rng.integers(0, n))
            r = r[:j] + "N" + r[j + 1:]
        if i % 23 == 0:
            r = r.lower()
        if i % 29 == 0:
            r = r.replace("T", "U", 3)
        recs.append(r)
    recs += [genome[:60], genome[100:227], genome[300:428]]
    return "".join("@w%d\n%s\n+\n%s\n" % (i, r, "I" * len(r)) for i, r in enumerate(recs))


WIDE_CASES = [(64, 0xFFFFFFFF), (65, 0xFFFFFFFF), (96, 0xFFFFFFFF), (100, 255), (127, 0xFFFFFFFF),
              (128, 0xFFFFFFFF)]


def write_tiny_wide():
    """tiny_wide.fastq and the reference views at k = 64..128 (tiny_wide.json).
    k = 64: dump and histogram files.  k > 64: the histogram file; of the
    dump only its size, line count, hash and first lines are kept, since the
    reference's dump --sort does not spell k-mers longer than 64 bases (its
    sort key holds 128 bits) and at k = 96 its counter splits counts."""
    path = os.path.join(HERE, "tiny_wide.fastq")
    with open(path, "w") as f:
        f.write(tiny_wide_fastq())
    cases = []
    for k, cm in WIDE_CASES:
        dump, hist, _ = ref_views(path, k, cm, hist_max=500)
        base = "tiny_wide.k%d.c%d" % (k, cm)
        case = {"input": "tiny_wide.fastq", "k": k, "count_max": cm, "hist_max": 500,
                "hist": base + ".hist.txt", "dump_sha256": hashlib.sha256(dump).hexdigest(),
                "dump_bytes": len(dump), "dump_lines": dump.count(b"\n")}
        if k <= 64:
            case["dump"] = base + ".dump.txt"
            with open(os.path.join(HERE, case["dump"]), "wb") as f:
                f.write(dump)
        else:
            case["dump_head"] = dump.decode().splitlines()[:8]
        with open(os.path.join(HERE, case["hist"]), "wb") as f:
            f.write(hist)
        cases.append(case)
    with open(os.path.join(HERE, "tiny_wide.json"), "w") as f:
        json.dump(cases, f, indent=1)


def write_count_5mbp():
    from ktcountermemorygreedy_amd import workloads as W
    rec = workload_golden(W.COUNT_5MBP)
    with open(os.path.join(HERE, "count_5mbp.json"), "w") as f:
        json.dump(rec, f, indent=1)


def write_repeat():
    from ktcountermemorygreedy_amd import workloads as W
    recs = {w.name: workload_golden(w) for w in (W.REPEAT_4MBP, W.REPEAT_4MBP_HOT)}
    with open(os.path.join(HERE, "repeat_4mbp.json"), "w") as f:
        json.dump(recs, f, indent=1)


def workload_golden(w):
    fq = w.reads(fastq=True)
    with tempfile.NamedTemporaryFile(suffix=".fastq", delete=False) as f:
        f.write(fq.tobytes())
        path = f.name
    try:
        dump, hist, log = ref_views(path, 31, 0xFFFFFFFF)
    finally:
        os.unlink(path)
    kmers = [int(l.split(":")[1]) for l in log.splitlines() if l.startswith("Total read k-mer count")][0]
    rec = {
        "workload": w.name, "k": 31, "count_max": 4294967295,
        "fastq_sha256": hashlib.sha256(fq.tobytes()).hexdigest(),
        "kmers": kmers, "distinct": dump.count(b"\n"),
        "dump_sha256": hashlib.sha256(dump).hexdigest(), "dump_bytes": len(dump),
        "hist_sha256": hashlib.sha256(hist).hexdigest(),
        "hist_head": hist.decode().splitlines()[:40],
    }
    return rec


def tiny_counts_fastq():
    """258 reads: three random 40-base reads at each count 1..8 and 50 (ten
    31-mers each), shuffled."""
    rng = np.random.default_rng(77)

    def rs(n):
        return "".join("ACGT"[i] for i in rng.integers(0, 4, n))

    recs = []
    for r in list(range(1, 9)) + [50]:
        for _ in range(3):
            s = rs(40)
            recs += [s] * r
    rng.shuffle(recs)
    return "".join("@c%d\n%s\n+\n%s\n" % (i, s, "I" * len(s)) for i, s in enumerate(recs))


FILTERS = [(3, 6), (5, 50)]
DUMP_RANGES = [(1, 1), (2, 6), (3, 50), (7, 4294967295)]
HIST_RANGES = [(1, 8), (2, 6), (5, 60)]


def _tools_on(d, threads=6):
    """dump --sort and histogram outputs of an export dir over the ranges."""
    out = {"dump": {}, "hist": {}}
    for lo, hi in DUMP_RANGES:
        subprocess.run([os.path.join(REF, "dump_tool_ref"), "precise", d + "/", str(threads), "4", d + "/o.txt",
                        str(lo), str(hi), "--sort"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, check=True)
        with open(d + "/o.txt") as f:
            out["dump"]["%d-%d" % (lo, hi)] = f.read()
    for lo, hi in HIST_RANGES:
        subprocess.run([os.path.join(REF, "histogram_tool_ref"), "precise", d + "/", str(threads), "4",
                        d + "/o.txt", str(lo), str(hi)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, check=True)
        with open(d + "/o.txt") as f:
            out["hist"]["%d-%d" % (lo, hi)] = f.read()
    return out


def write_tiny_counts():
    from oracle import oracle as O
    text = tiny_counts_fastq()
    fq_path = os.path.join(HERE, "tiny_counts.fastq")
    with open(fq_path, "w") as f:
        f.write(text)
    rec = {"input": "tiny_counts.fastq", "k": 31, "count_max": 4294967295, "filters": {}, "splits": {}}
    # the counter's own filter (Tree argv 7-8); merged views over the full range
    for fmin, fmax in FILTERS:
        tmp = tempfile.mkdtemp(prefix="kcgold_")
        try:
            subprocess.run([os.path.join(REF, "kmc_ref"), fq_path, "31", "6", "4", tmp + "/", "1024", str(fmin),
                            str(fmax), "4294967295"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, check=True)
            subprocess.run([os.path.join(REF, "dump_tool_ref"), "precise", tmp + "/", "6", "4", tmp + "/d.txt", "1",
                            "4294967295", "--sort"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, check=True)
            subprocess.run([os.path.join(REF, "histogram_tool_ref"), "precise", tmp + "/", "6", "4", tmp + "/h.txt",
                            "1", "60"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, check=True)
            with open(tmp + "/d.txt") as f:
                dump = f.read()
            with open(tmp + "/h.txt") as f:
                hist = f.read()
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
        rec["filters"]["%d-%d" % (fmin, fmax)] = {"dump": dump, "hist": hist}
    # two exports of the same counts, and the tools' range rules over each
    seq = O.fastq_frame(text.encode())
    uniq, cnt = O.count(O.extract(seq, 31), 1)
    infos, high, low = O.split_export(uniq, cnt, 31, 0xFFFFFFFF)
    # FIRST split: every key in low.bin, count - 1 in high.bin for the keys
    # seen at least twice (what the Bloom split writes when no first
    # occurrence is mistaken for a second one)
    rep = cnt >= 2
    kb = np.frombuffer(O.pack(uniq[rep], 31), dtype=np.uint8).reshape(-1, 8)
    cb = (cnt[rep] - 1).astype(np.uint32).view(np.uint8).reshape(-1, 4)
    high_f = np.concatenate([kb, cb], axis=1).tobytes()
    low_f = O.pack(uniq, 31)
    for name, h, l in (("plain", high, low), ("first", high_f, low_f)):
        with open(os.path.join(HERE, "tiny_counts.%s.high.bin" % name), "wb") as f:
            f.write(h)
        with open(os.path.join(HERE, "tiny_counts.%s.low.bin" % name), "wb") as f:
            f.write(l)
        tmp = tempfile.mkdtemp(prefix="kcgold_")
        try:
            for fn, b in (("infos.bin", infos), ("high.bin", h), ("low.bin", l)):
                with open(os.path.join(tmp, fn), "wb") as f:
                    f.write(b)
            rec["splits"][name] = _tools_on(tmp)
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
    with open(os.path.join(HERE, "tiny_counts.json"), "w") as f:
        json.dump(rec, f, indent=1)


def write_hash_kat():
    out = subprocess.run([os.path.join(REF, "hash_kat_ref"), "7", "48"], stdout=subprocess.PIPE,
                         text=True, check=True).stdout
    lines = [l for l in out.splitlines() if l[:1].isdigit() and len(l.split()) > 5]
    with open(os.path.join(HERE, "hash_kat.txt"), "w") as f:
        f.write("# nw words... xxh3_64 xxh3_128_lo xxh3_128_hi rapidhash komihash "
                "concurrent_map_hash hash_func bloom_block(cap=2^20) bloom_mask\n")
        f.write("\n".join(lines) + "\n")


def run_sort_ref(keys, counts, nw, k, cpb, bits):
    n = keys.shape[0]
    hdr = np.array([nw, k, cpb, bits], dtype=np.uint32).tobytes() + np.array([n], dtype=np.uint64).tobytes()
    rec = np.zeros((n, nw * 8 + 4), dtype=np.uint8)
    rec[:, :nw * 8] = keys.view(np.uint8).reshape(n, nw * 8)
    rec[:, nw * 8:] = counts.view(np.uint8).reshape(n, 4)
    r = subprocess.run([os.path.join(REF, "sort_kat_ref")], input=hdr + rec.tobytes(),
                       stdout=subprocess.PIPE, check=True)
    out = np.frombuffer(r.stdout, dtype=np.uint8).reshape(n, nw * 8 + 4)
    return out[:, :nw * 8].copy().view(np.uint64).reshape(n, nw), out[:, nw * 8:].copy().view(np.uint32).reshape(n)


def write_sort_kat():
    rng = np.random.default_rng(4)
    for nw, k, bits in [(1, 31, 8), (1, 31, 12), (1, 28, 8), (1, 21, 12), (2, 63, 8), (2, 63, 12), (2, 45, 8),
                        (2, 41, 12)]:
        n = 6000
        total = 64 * nw
        keys = rng.integers(0, 2**63, size=(n, nw), dtype=np.uint64) * np.uint64(2) + \
            rng.integers(0, 2, size=(n, nw), dtype=np.uint64)
        # clear padding below 2k bits (in the 64*nw-bit number)
        pad = total - 2 * k
        for w in range(nw - 1, -1, -1):
            b = min(pad, 64)
            if b > 0:
                keys[:, w] &= np.uint64(~((1 << b) - 1) & 0xFFFFFFFFFFFFFFFF)
            pad -= b
            if pad <= 0:
                break
        # shared 10-base prefix (top 20 bits of word 0) on half the records,
        # duplicates of sort keys to exercise stability
        keys[: n // 2, 0] = (keys[: n // 2, 0] & np.uint64((1 << 44) - 1)) | np.uint64(0xABCDE << 44)
        keys[n // 2: n // 2 + 500] = keys[:500]
        counts = np.arange(n, dtype=np.uint32)
        sk, sc = run_sort_ref(keys, counts, nw, k, 10, bits)
        np.savez(os.path.join(HERE, "sort_kat_%d_%d_%d.npz" % (nw, k, bits)), keys=keys, counts=counts,
                 sorted_keys=sk, sorted_counts=sc, nw=nw, k=k, bits=bits, cpb=10)


if __name__ == "__main__":
    write_hash_kat()
    write_sort_kat()
    write_tiny()
    write_count_5mbp()
    write_repeat()
    write_tiny_counts()
    write_tiny_wide()
    print("goldens written to", HERE)
