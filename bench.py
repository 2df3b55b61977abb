"""Throughput of the canonical k-mer counting path on one MI355X.

One step = one pass of the hot path over the whole read set, which is
already resident in HBM as the framed "seq\\n" byte stream: clear the table,
fused parse + canonical 31-mers + Bloom singleton split + table count, then
pack the reference's high/low export records into device buffers.
value = k-mers counted per second over the timed steps.

Default workload: repeat_100mbp (100 Mbp genome with 30% diverged repeat
copies, 20M x 150 bp reads at 30x, 0.5% error, k=31, Bloom split on).
The CPU baseline times the reference binary (oracle/_ref/kmc_ref, built from
the reference sources) on a bounded prefix of the same reads, on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# one kc_table_count_seq call on the partitioned path: these kernels, in order
# (k_front_count_seq only when the speculative coarse regions overflow)
COUNT_KERNELS = ("count pipeline: k_spec_regions + k_front_scatter_seq_staged [+ exact redo: k_front_count_seq + "
                 "k_front_scatter_seq_staged] + k_fine_count_dig + k_fine_scatter_staged + k_count_parts (timed "
                 "together with HIP events on the table's stream)")
# the same call on the direct path (four-word keys, k = 64..128)
DIRECT_KERNELS = "direct path: k_count_seq (fused parse + Bloom + count with global atomics per k-mer)"


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="repeat_100mbp",
                    choices=["repeat_100mbp", "count_5mbp", "uneven_50mbp", "export_sort", "serving", "ingest_gz",
                             "ingest_bgzf", "tools_5mbp"])
    ap.add_argument("--k", type=int, default=31, help="k-mer length (uneven_50mbp is quoted at 21, 31 and 63)")
    ap.add_argument("--serve-seconds", type=float, default=5.0, help="serving: length of the schedule")
    ap.add_argument("--records-log2", type=int, default=30, help="export_sort: 2^N records")
    ap.add_argument("--reads", type=int, default=0, help="override the read count (smaller runs)")
    ap.add_argument("--no-bloom", action="store_true")
    ap.add_argument("--cpu-sample-reads", type=int, default=0,
                    help="reads timed on the CPU (0: 4M, or 300k for uneven_50mbp, about 10-20 s of CPU work)")
    ap.add_argument("--copy-probe-bytes", type=int, default=1 << 30,
                    help="one device-to-device copy of this size before the timed region, the "
                         "reference dispatch for scaling rocprof FETCH_SIZE/WRITE_SIZE counters")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--launch", default="auto", choices=["auto", "eager", "graph"],
                    help="counting call: eager stream launches or one captured-graph launch (kc_batch); "
                         "auto = graph for uneven_50mbp")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--passes", type=int, default=1,
                    help="also time multi-pass counting (a power of two): the same input counted in P hash-slice "
                         "passes into a table of 1/P the slots (the out-of-HBM fallback); a 'multipass' object")
    ap.add_argument("--gz-format", default="bgzf", choices=["bgzf", "gzip"],
                    help="ingest_bgzf: the one file as BGZF (members inflated on all host threads) or as one "
                         "plain gzip member as gzip -6 writes it (its deflate stream decoded in shares on all host "
                         "threads)")
    ap.add_argument("--no-host-batch", action="store_true",
                    help="skip the host-batched (PCIe-inclusive) line of uneven_50mbp")
    return ap.parse_args()


# Host threads the CPU baselines use: nproc (the CPUs this process may run
# on), at most 16 -- the host CPU share of one GPU on the GPU boxes this runs
# on -- and at least 6, the reference counter's minimum (main.cpp:556-560).
def host_threads():
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count() or 6
    return max(6, min(16, nproc))


def host_info():
    """nproc, the CPU model and the RAM of this host (recorded beside every
    CPU baseline)."""
    info = {"nproc": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None,
            "cpu_count": os.cpu_count()}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    info["cpu_model"] = line.split(":", 1)[1].strip()
                    break
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemTotal"):
                    info["ram_gb"] = round(int(line.split()[1]) / 1e6, 1)
                    break
    except OSError:
        pass
    info["threads_used"] = host_threads()
    return info


def cpu_baseline(ws, sample_reads, k, genome):
    """Reference Tree binary on the first sample_reads reads (FASTQ)."""
    from ktcountermemorygreedy_amd import kc  # noqa: F401
    ref = os.path.join(ROOT, "oracle", "_ref", "kmc_ref")
    threads = host_threads()
    sample = ws.scaled(sample_reads)
    fq = sample.reads(fastq=True, genome=genome)
    if os.path.exists(ref):
        tmp = tempfile.mkdtemp(prefix="kcbench_")
        path = os.path.join(tmp, "sample.fastq")
        with open(path, "wb") as f:
            f.write(fq.tobytes())
        os.makedirs(os.path.join(tmp, "out"))
        t0 = time.time()
        out = subprocess.run([ref, path, str(k), str(threads), "16", os.path.join(tmp, "out") + "/", "1024", "2",
                              "4294967295", "4294967295"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                             text=True, timeout=600).stdout
        wall = time.time() - t0
        subprocess.run(["rm", "-rf", tmp])
        kmers = int(re.search(r"Total read k-mer count: (\d+)", out).group(1))
        us = int(re.search(r"Total elapsed us: (\d+)", out).group(1))
        return {"value": kmers / (us * 1e-6), "unit": "k-mers/s", "cores": threads, "kind": "reference",
                "host": host_info(),
                "sample": "first %d reads of the workload as FASTQ (%d k-mers), Tree <fastq> %d %d 16 <tmp> "
                          "1024 2 4294967295 4294967295; elapsed %.2f s (wall %.2f s)"
                          % (sample_reads, kmers, k, threads, us * 1e-6, wall)}
    # no reference build on this host: time the oracle port (single thread)
    from oracle import oracle as O
    seq = O.fastq_frame(fq.tobytes())
    t0 = time.time()
    keys = O.extract(seq, k)
    O.count(keys, O.words_for_k(k))
    dt = time.time() - t0
    return {"value": keys.shape[0] / dt, "unit": "k-mers/s", "cores": 1, "kind": "port",
            "sample": "first %d reads, oracle extract+count" % sample_reads}


def measured_traffic(kernel, workload, k=None):
    """HBM bytes per launch of the kernel from the committed rocprof --pmc
    summary (profiles/traffic.json), when it covers this kernel, workload
    and k (entries "kernel/workload/k<k>", or "kernel" for one workload)."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            data = json.load(f)
        for key in ("%s/%s/k%s" % (kernel, workload, k), kernel):
            rec = data.get(key)
            if rec and rec["workload"] == workload and rec.get("k", k) == k:
                return rec["bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        pass
    return None


def sort_cpu_baseline(n_sample):
    """Reference radix_sort_export_records_12bit (single thread, harness built
    from the reference header) on n_sample records of the same law."""
    import numpy as np
    ref = os.path.join(ROOT, "oracle", "_ref", "sort_kat_ref")
    rng = np.random.default_rng(4)
    keys = (rng.integers(0, 2**62, n_sample, dtype=np.uint64) << np.uint64(2)).reshape(-1, 1)
    counts = np.minimum(np.floor(rng.random(n_sample) ** -5.0), 2**32 - 1).astype(np.uint32)
    if os.path.exists(ref):
        hdr = np.array([1, 31, 10, 12], np.uint32).tobytes() + np.array([n_sample], np.uint64).tobytes()
        rec = np.zeros((n_sample, 12), np.uint8)
        rec[:, :8] = keys.view(np.uint8).reshape(-1, 8)
        rec[:, 8:] = counts.view(np.uint8).reshape(-1, 4)
        t0 = time.time()
        subprocess.run([ref], input=hdr + rec.tobytes(), stdout=subprocess.DEVNULL, check=True)
        dt = time.time() - t0
        return {"value": n_sample / dt, "unit": "records/s", "cores": 1, "kind": "reference",
                "sample": "%d records through radix_sort_export_records_12bit<1,10> (reference header, "
                          "one thread), wall time incl. pipe I/O %.2f s" % (n_sample, dt)}
    from oracle import oracle as O
    t0 = time.time()
    O.sort_records12(keys, counts, 31, 10)
    dt = time.time() - t0
    return {"value": n_sample / dt, "unit": "records/s", "cores": 1, "kind": "port",
            "sample": "%d records, oracle 12-bit LSD" % n_sample}


def bench_sort(args):
    """Export/sort workload: 2^N distinct k=31 keys with Zipf(1.2) counts as
    12-byte records; one step = stable radix sort on bits [2,44) (the
    12-bit sort's range) + abundance histogram 1..10000."""
    import numpy as np

    from ktcountermemorygreedy_amd import kc
    n = 1 << args.records_log2
    ctx = kc.Context(0)
    d, s = ctx.alloc(12 * n), ctx.alloc(12 * n)
    src = ctx.alloc(12 * n)
    ctx.synth_records(src, n, 12, seed=4, zipf_s=1.2)
    ctx.sync()

    def step(slot=None):
        kc.check(kc.lib().kc_copy_d2d(ctx.h, kc._ptr(d), kc._ptr(src), 12 * n), "reset records")
        if slot is not None:
            ctx.record(slot)
        in_s = ctx.sort_records(d, s, n, 1, 31, 10, 12, 12)
        bins = ctx.records_histogram(s if in_s else d, n, 1, 12, 1, 10000)
        if slot is not None:
            ctx.record(slot + 1)
        return in_s, bins

    for _ in range(args.warmup):
        step()
    ms = []
    for i in range(args.steps):
        in_s, bins = step(2 * (i % 32))
        ms.append(ctx.elapsed_ms(2 * (i % 32), 2 * (i % 32) + 1))
    out = (s if in_s else d).numpy(np.uint32, 3 * min(n, 1 << 24)).reshape(-1, 3)
    k = out[:, 0].astype(np.uint64) | (out[:, 1].astype(np.uint64) << np.uint64(32))
    sk = (k >> np.uint64(2)) & np.uint64((1 << 42) - 1)
    step_s = float(np.mean(ms)) * 1e-3
    digit_bits = 9  # the device sort's widest digit (kSortDigitBits)
    passes = (42 + digit_bits - 1) // digit_bits
    alg = n * (2 * 12 * 4 + 4)
    line = {
        "metric": "export records sorted/s (stable, 12-bit sort key range) + abundance histogram",
        "value": n / step_s, "unit": "records/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": step_s * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic",
        "config": {"workload": "export_sort_2^%d" % args.records_log2, "records": n, "record_bytes": 12,
                   "key_bits": [2, 44], "digit_bits": digit_bits, "passes": passes, "hist_bins": [1, 10000]},
        "roofline": {"bound": "hbm", "kernel": "sort pipeline", "achieved": alg / step_s / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": alg / step_s / 1e9 / HBM_PEAK_GBS,
                     "traffic": None, "alg_bytes_per_launch": alg,
                     "alg_bytes_rule": "n*(2*12*4+4): 4 passes of 12-bit digits as in the reference; this "
                                       "build runs %d passes of at most %d bits" % (passes, digit_bits)},
        "cpu_baseline": None if args.no_cpu_baseline else sort_cpu_baseline(1 << 24),
        "checks": {"prefix_sorted": bool(np.all(sk[1:] >= sk[:-1])),
                   "hist_mass": int(bins.sum())},
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def bench_ingest_bgzf(args):
    """End-to-end from ONE gzip FASTQ file: the whole repeat_100mbp read set
    (default 20M reads) as a single .fastq.gz (level 6), counted by
    kc_count_fastq_files with 16 host threads -- host inflate, framing, PCIe
    and counting all inside the timed region, so it is never the headline.
    --gz-format bgzf: BGZF, whose member runs the threads inflate in
    parallel; gzip: one plain gzip member, whose deflate stream the threads
    decode in shares from guessed block starts, joined in order.  Checked against the same reads counted from HBM (equal
    histograms and k-mer totals).  The reference Tree binary counts a sample
    of the same format on the same host beside it."""
    import shutil

    import numpy as np

    from ktcountermemorygreedy_amd import kc
    from ktcountermemorygreedy_amd import workloads as W

    n_reads = args.reads or W.REPEAT_100MBP.n_reads
    threads = host_threads()
    fmt = args.gz_format
    write = W.write_bgzf if fmt == "bgzf" else W.write_gzip_parallel
    ws = W.REPEAT_100MBP.scaled(n_reads, "repeat_100mbp_%s_%d" % (fmt, n_reads))
    genome = ws.genome()
    t0 = time.time()
    tmp = tempfile.mkdtemp(prefix="kcbgzf_")
    path = os.path.join(tmp, "reads.fastq.gz")
    fq = ws.reads(fastq=True, genome=genome)
    fq_bytes = int(fq.nbytes)
    gz_bytes = write(path, fq, level=6, threads=threads)
    del fq
    prep_s = time.time() - t0
    print("prep: %d FASTQ bytes -> %d %s bytes in %.1f s" % (fq_bytes, gz_bytes, fmt, prep_s), flush=True)

    ctx = kc.Context(0)
    est_distinct = int(ws.genome_len * 1.1 + n_reads * 120 * 0.15)
    cap = 1 << max(20, int(np.ceil(np.log2(max(ws.genome_len * 1.1, est_distinct / 3.0) * 1.8))))

    def table():
        return kc.Table(ctx, 31, count_max=0xFFFFFFFF, capacity=cap, max_distinct=est_distinct, bloom=True,
                        low_capacity=est_distinct)

    t = table()
    runs = []
    for _ in range(max(1, args.steps)):
        t.clear()
        t1 = time.perf_counter()
        st = kc.count_fastq_files(t, [path], threads=threads, batch_bytes=256 << 20, queue_depth=3)
        runs.append((time.perf_counter() - t1, st))
        print("ingest %.3f s %s" % (runs[-1][0], st), flush=True)
    wall, st = min(runs, key=lambda r: r[0])
    stats = t.stats()
    assert stats["errors"] == 0, stats
    bins = t.histogram(1, 10000)
    t.close()
    # the same reads counted from HBM: equal k-mer totals and histograms
    framed = ws.reads(fastq=False, genome=genome)
    d = ctx.upload(framed)
    r = table()
    r.count_seq(d, framed.nbytes)
    ctx.sync()
    ref_stats, ref_bins = r.stats(), r.histogram(1, 10000)
    r.close()
    d.free()
    checks = {"kmers_equal_hbm_count": stats["kmers"] == ref_stats["kmers"],
              "histogram_equals_hbm_count": bool(np.array_equal(bins, ref_bins)),
              "framed_bytes_match": st["framed_bytes"] == framed.nbytes}
    del framed
    cpu = None
    ref = os.path.join(ROOT, "oracle", "_ref", "kmc_ref")
    if not args.no_cpu_baseline and os.path.exists(ref):
        # a bounded sample: the first reads of the workload as their own BGZF file
        sample_reads = min(n_reads, args.cpu_sample_reads or 4_000_000)
        spath = os.path.join(tmp, "sample.fastq.gz")
        write(spath, ws.scaled(sample_reads).reads(fastq=True, genome=genome), level=6, threads=threads)
        out_dir = os.path.join(tmp, "out")
        os.makedirs(out_dir)
        t1 = time.time()
        out = subprocess.run([ref, spath, "31", str(threads), "16", out_dir + "/", "1024", "2",
                              "4294967295", "4294967295"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                             text=True, timeout=900).stdout
        ref_wall = time.time() - t1
        km = re.search(r"Total read k-mer count: (\d+)", out)
        el = re.search(r"Total elapsed us: (\d+)", out)
        if km and el:
            kmers, us = int(km.group(1)), int(el.group(1))
            cpu = {"value": kmers / (us * 1e-6), "unit": "k-mers/s", "cores": threads, "kind": "reference",
                   "sample": "first %d reads as one %s file, Tree <file> 31 %d 16 <tmp> 1024 2 4294967295 "
                             "4294967295 (its reader inflates the file as one gzip stream); %d k-mers, elapsed "
                             "%.2f s (wall %.2f s)" % (sample_reads, fmt, threads, kmers, us * 1e-6, ref_wall),
                   "host": host_info()}
        else:
            cpu = {"value": None, "error": "reference run printed no totals", "output_tail": out[-2000:]}
    shutil.rmtree(tmp, ignore_errors=True)
    line = {
        "metric": "canonical 31-mers counted/s end to end from one gzip FASTQ file (host inflate + PCIe + count)",
        "value": stats["kmers"] / wall, "unit": "k-mers/s", "n_gpus": 1, "steps": len(runs), "warmup": 0,
        "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic",
        "config": {"workload": ws.name, "reads": n_reads, "files": 1,
                   "format": "BGZF" if fmt == "bgzf" else "gzip (one member)", "gzip_level": 6,
                   "fastq_bytes": fq_bytes, "gz_bytes": gz_bytes, "kmers": stats["kmers"], "host_threads": threads,
                   "batch_bytes": 256 << 20, "queue_depth": 3},
        "ingest": st, "roofline": None, "cpu_baseline": cpu, "prep_seconds": prep_s, "checks": checks,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def bench_ingest(args):
    """End-to-end from gzip FASTQ: the first reads of repeat_100mbp (default
    4M) as 16 .fastq.gz files (level 6), counted by kc_count_fastq_files
    (16 host threads inflate and frame in parallel into a kc_batch) -- host
    inflate included, so this is the PCIe- and zlib-inclusive rate, never the
    headline.  The reference Tree binary counts the same 16 files on the same
    host beside it."""
    import gzip
    import shutil
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np

    from ktcountermemorygreedy_amd import kc
    from ktcountermemorygreedy_amd import workloads as W

    n_reads = args.reads or 4_000_000
    n_files = 16
    ws = W.REPEAT_100MBP.scaled(n_reads, "repeat_100mbp_gz_%d" % n_reads)
    genome = ws.genome()
    t0 = time.time()
    fq = ws.reads(fastq=True, genome=genome)
    tmp = tempfile.mkdtemp(prefix="kcingest_")
    # cut at record boundaries (4 lines per record)
    nl = np.flatnonzero(fq == 10)
    rec_end = nl[3::4] + 1
    bounds = [0] + [int(rec_end[min(len(rec_end) - 1, (i + 1) * len(rec_end) // n_files - 1)])
                    for i in range(n_files)]
    paths = [os.path.join(tmp, "part%02d.fastq.gz" % i) for i in range(n_files)]

    def write(i):
        with open(paths[i], "wb") as f:
            f.write(gzip.compress(fq[bounds[i]:bounds[i + 1]].tobytes(), compresslevel=6))

    with ThreadPoolExecutor(n_files) as ex:
        list(ex.map(write, range(n_files)))
    gz_bytes = sum(os.path.getsize(p) for p in paths)
    prep_s = time.time() - t0
    framed_len = int(kc.fastq_frame(fq[:bounds[-1]].tobytes()).__len__())
    del fq

    ctx = kc.Context(0)
    est_distinct = int(ws.genome_len * 1.1 + n_reads * 120 * 0.15)
    cap = 1 << max(20, int(np.ceil(np.log2(max(ws.genome_len * 1.1, est_distinct / 3.0) * 1.8))))
    table = kc.Table(ctx, 31, count_max=0xFFFFFFFF, capacity=cap, max_distinct=est_distinct, bloom=True,
                     low_capacity=est_distinct)
    runs = []
    for _ in range(max(1, args.steps)):
        table.clear()
        t1 = time.perf_counter()
        st = kc.count_fastq_files(table, paths, threads=n_files, batch_bytes=256 << 20, queue_depth=3)
        runs.append((time.perf_counter() - t1, st))
    wall, st = min(runs, key=lambda r: r[0])
    stats = table.stats()
    assert stats["errors"] == 0, stats
    cpu = None
    ref = os.path.join(ROOT, "oracle", "_ref", "kmc_ref")
    if not args.no_cpu_baseline and os.path.exists(ref):
        out_dir = os.path.join(tmp, "out")
        os.makedirs(out_dir)
        t1 = time.time()
        out = subprocess.run([ref, ",".join(paths), "31", str(n_files), "16", out_dir + "/", "1024", "2",
                              "4294967295", "4294967295"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                             text=True, timeout=900).stdout
        ref_wall = time.time() - t1
        kmers = int(re.search(r"Total read k-mer count: (\d+)", out).group(1))
        us = int(re.search(r"Total elapsed us: (\d+)", out).group(1))
        cpu = {"value": kmers / (us * 1e-6), "unit": "k-mers/s", "cores": n_files, "kind": "reference",
               "sample": "the same %d .fastq.gz files, Tree <f1,...,f16> 31 %d 16 <tmp> 1024 2 4294967295 "
                         "4294967295; elapsed %.2f s (wall %.2f s); k-mers %s GPU" %
                         (n_files, n_files, us * 1e-6, ref_wall, "==" if kmers == stats["kmers"] else "!=")}
    shutil.rmtree(tmp, ignore_errors=True)
    line = {
        "metric": "canonical 31-mers counted/s end to end from gzip FASTQ (host inflate + PCIe + count)",
        "value": stats["kmers"] / wall, "unit": "k-mers/s", "n_gpus": 1, "steps": len(runs), "warmup": 0,
        "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic",
        "config": {"workload": ws.name, "reads": n_reads, "files": n_files, "gzip_level": 6, "gz_bytes": gz_bytes,
                   "framed_bytes": framed_len, "kmers": stats["kmers"], "host_threads": n_files,
                   "batch_bytes": 256 << 20, "queue_depth": 3},
        "ingest": st, "roofline": None, "cpu_baseline": cpu, "prep_seconds": prep_s,
        "checks": {"framed_bytes_match": st["framed_bytes"] == framed_len},
    }
    print(json.dumps(line), flush=True)
    table.close()
    ctx.close()


def bench_tools(args):
    """Device dump_tool / histogram_tool equivalents on count_5mbp (k=31,
    Bloom split, count_max 2^32-1): one step = kc_table_histogram (bins
    1..10000) + kc_table_dump_device (sorted "<bases>\t<count>\n" lines of
    every k-mer, into HBM).  Both outputs are checked against the reference
    goldens (sha256); the reference's dump_tool precise --sort and
    histogram_tool precise (built from source, 16 threads) run beside it on
    the same export."""
    import ctypes as C
    import hashlib
    import shutil

    import numpy as np

    from ktcountermemorygreedy_amd import kc
    from ktcountermemorygreedy_amd import workloads as W
    from oracle import oracle as O

    with open(os.path.join(ROOT, "tests", "golden", "count_5mbp.json")) as f:
        golden = json.load(f)
    ws = W.COUNT_5MBP
    seq = ws.reads()
    ctx = kc.Context(0)
    d_seq = ctx.upload(seq)
    kw, plan = kc.plan_table(ctx, d_seq, 31, seq.size, bloom=True)
    kw["max_distinct"] = max(kw["max_distinct"], golden["distinct"])
    table = kc.Table(ctx, 31, **dict(kw, capacity=max(kw["capacity"], 1 << 24)))
    table.count_seq(d_seq, seq.size)
    ctx.sync()
    n = C.c_uint64()
    kc.check(kc.lib().kc_table_dump_device(table.h, 1, 0xFFFFFFFF, 1, None, 0, C.byref(n)), "dump size")
    out = ctx.alloc(n.value + 4096)

    def step():
        bins = table.histogram(1, 10000)
        m = C.c_uint64()
        kc.check(kc.lib().kc_table_dump_device(table.h, 1, 0xFFFFFFFF, 1, kc._ptr(out), out.nbytes, C.byref(m)),
                 "kc_table_dump_device")
        ctx.sync()
        return bins, m.value

    for _ in range(args.warmup):
        step()
    times = []
    for _ in range(max(1, args.steps)):
        t0 = time.perf_counter()
        bins, nbytes = step()
        times.append(time.perf_counter() - t0)
    step_s = float(np.median(times))
    dump = out.numpy(np.uint8, nbytes).tobytes()
    hist = O.histogram_text(bins, 1, 10000)
    lines = dump.count(b"\n")
    checks = {"dump_sha256_golden": hashlib.sha256(dump).hexdigest() == golden["dump_sha256"],
              "hist_sha256_golden": hashlib.sha256(hist).hexdigest() == golden["hist_sha256"],
              "lines": lines}
    cpu = None
    d_ref = os.path.join(ROOT, "oracle", "_ref", "dump_tool_ref")
    h_ref = os.path.join(ROOT, "oracle", "_ref", "histogram_tool_ref")
    if not args.no_cpu_baseline and os.path.exists(d_ref) and os.path.exists(h_ref):
        tmp = tempfile.mkdtemp(prefix="kctools_")
        table.export_dir(tmp)
        threads = str(host_threads())
        t0 = time.time()
        subprocess.run([d_ref, "precise", tmp + "/", threads, "16", tmp + "/dump.txt", "1", "4294967295", "--sort"],
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, check=True, timeout=900)
        t1 = time.time()
        subprocess.run([h_ref, "precise", tmp + "/", threads, "16", tmp + "/hist.txt", "1", "10000"],
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, check=True, timeout=900)
        t2 = time.time()
        with open(tmp + "/dump.txt", "rb") as f:
            same_dump = f.read() == dump
        with open(tmp + "/hist.txt", "rb") as f:
            same_hist = f.read() == hist
        shutil.rmtree(tmp, ignore_errors=True)
        cpu = {"value": lines / (t2 - t0), "unit": "dump lines/s (dump + histogram)", "cores": host_threads(),
               "host": host_info(),
               "kind": "reference",
               "sample": "dump_tool precise <export> 16 16 dump.txt 1 4294967295 --sort (%.2f s) + histogram_tool "
                         "precise <export> 16 16 hist.txt 1 10000 (%.2f s) on the same export; outputs %s" %
                         (t1 - t0, t2 - t1, "equal" if same_dump and same_hist else "DIFFERENT")}
    line = {
        "metric": "sorted dump lines/s (dump_tool precise --sort + histogram_tool precise, on the device)",
        "value": lines / step_s, "unit": "lines/s", "n_gpus": 1, "steps": len(times), "warmup": args.warmup,
        "ms_per_step": step_s * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic",
        "config": {"workload": "tools_" + ws.name, "k": 31, "bloom": True, "distinct": lines,
                   "dump_bytes": nbytes, "step": "kc_table_histogram(1..10000) + kc_table_dump_device(sorted)"},
        "roofline": None, "cpu_baseline": cpu, "checks": checks,
    }
    print(json.dumps(line), flush=True)
    out.free()
    d_seq.free()
    table.close()
    ctx.close()


def serving_cpu_baseline(n_keys, n_batches, batch_keys=0):
    """The reference's point lookup (FlatConcurrentHashMap find_prepared with
    its query pool's prefetch loop) on this host: oracle/_ref/query_bench_ref,
    batches of log-uniform size in [1, 65536] or of batch_keys keys."""
    ref = os.path.join(ROOT, "oracle", "_ref", "query_bench_ref")
    if not os.path.exists(ref):
        return None
    threads = host_threads()
    args = [ref, str(n_keys), str(n_batches), str(threads), "5"] + ([str(batch_keys)] if batch_keys else [])
    out = subprocess.run(args, capture_output=True, text=True, timeout=600)
    if out.returncode != 0:
        return {"error": out.stderr[-500:]}
    r = json.loads(out.stdout.strip().splitlines()[-1])
    size = "%d keys" % batch_keys if batch_keys else "log-uniform size in [1, 65536]"
    return {"value": r["batch_p99_ms"], "unit": "ms", "cores": threads, "kind": "reference",
            "keys_per_s": r["keys_per_s"], "batch_p50_ms": r["batch_p50_ms"], "host": host_info(),
            "sample": "FlatConcurrentHashMap<1> of %d keys, %d batches of %s (half present), one batch per "
                      "thread, prefetch distance 8, no job running" % (n_keys, n_batches, size)}


def bench_serving(args):
    """Serving workload: Poisson count jobs (20/s, lognormal 1e4..1e7 reads
    from ten 20 Mbp genomes) and query batches (200/s, 1..65536 keys, half
    present) against device-resident tables through kc_server.  The schedule
    replays in real time; value = p99 latency of the query batches."""
    import numpy as np

    from ktcountermemorygreedy_amd import kc
    from ktcountermemorygreedy_amd.serving import ServingRun, ServingSpec, job_distinct_estimate

    spec = ServingSpec(duration_s=args.serve_seconds)
    ctx = kc.Context(0)
    t0 = time.time()
    run = ServingRun(ctx, spec)
    prep_s = time.time() - t0
    print("serving: %d jobs, %d query batches, pools and samples ready in %.1f s" %
          (run.job_t.size, run.q_t.size, prep_s), file=sys.stderr, flush=True)
    srv = kc.Server(0, spec.workers, spec.hbm_budget)
    summ = run.run(srv)
    srv.close()
    line = {
        "metric": "p99 latency of 65536-key query batches under a Poisson job + query stream (serving)",
        "value": summ["q64k_p99_ms"], "unit": "ms", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": summ["wall_s"] * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic",
        "config": {"workload": "serving", "job_rate_per_s": spec.job_rate, "query_rate_per_s": spec.query_rate,
                   "seconds": spec.duration_s, "genomes": spec.n_genomes, "genome_len": spec.genome_len,
                   "job_reads_lognormal": [spec.reads_mu, spec.reads_sigma, spec.min_reads, spec.max_reads],
                   "query_batch": [1, spec.max_query], "q64k_rate_per_s": spec.q64_rate,
                   "q64k_idle_batches": spec.idle_q64, "k": spec.k, "workers": spec.workers,
                   "hbm_budget_gb": spec.hbm_budget / 1e9},
        "serving": summ,
        "roofline": None,
        "cpu_baseline": None if args.no_cpu_baseline else serving_cpu_baseline(
            job_distinct_estimate(spec, int(np.median(run.job_reads))), 2000, spec.max_query),
        "cpu_baseline_mixed_sizes": None if args.no_cpu_baseline else serving_cpu_baseline(
            job_distinct_estimate(spec, int(np.median(run.job_reads))), 20000),
        "prep_seconds": prep_s,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def main():
    args = parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU; --gpus %d is not supported: the library counts on one GPU "
              "(sharding over several GPUs is designed, not implemented)" % args.gpus)
        sys.exit(2)
    if args.workload == "export_sort":
        return bench_sort(args)
    if args.workload == "serving":
        return bench_serving(args)
    if args.workload == "ingest_gz":
        return bench_ingest(args)
    if args.workload == "ingest_bgzf":
        return bench_ingest_bgzf(args)
    if args.workload == "tools_5mbp":
        return bench_tools(args)
    import numpy as np

    from ktcountermemorygreedy_amd import kc
    from ktcountermemorygreedy_amd import workloads as W

    ws = {"repeat_100mbp": W.REPEAT_100MBP, "count_5mbp": W.COUNT_5MBP,
          "uneven_50mbp": W.UNEVEN_50MBP}[args.workload]
    if args.reads:
        ws = ws.scaled(args.reads, ws.name)
    k = args.k
    nw = 1 if k <= 32 else 2
    bloom = args.workload in ("repeat_100mbp", "uneven_50mbp") and not args.no_bloom

    t0 = time.time()
    genome = ws.genome()
    seq_host = ws.reads(genome=genome)
    gen_s = time.time() - t0
    S = seq_host.size
    ctx = kc.Context(0)
    d_seq = ctx.upload(seq_host)

    # table sized from the input itself: one hash slice of its k-mers counted
    # exactly on the device and scaled (kc_plan_table), outside the timed region
    t0 = time.time()
    plan_kw, plan = kc.plan_table(ctx, d_seq, k, S, bloom=bloom, sample_slices=16)
    plan["seconds"] = time.time() - t0
    est_distinct = plan["distinct"]
    table = kc.Table(ctx, k, **plan_kw)

    # launch shape: autotuned on the workload itself (outside the timed region)
    from ktcountermemorygreedy_amd import autotune as AT
    tune = None
    if not args.no_autotune:
        t0 = time.time()
        tune = AT.tune_shape(table, d_seq, S, reps=3, use_cache=False)
        tune["seconds"] = time.time() - t0
        print("autotune: %s" % tune, file=sys.stderr, flush=True)
    launch = args.launch if args.launch != "auto" else ("graph" if args.workload == "uneven_50mbp" else "eager")
    batch = None
    if launch == "graph":
        table.trim()  # the graph owns its pipeline scratch
        batch = kc.Batch(table)

    bufs = {}

    def one_step(slot=None):
        table.clear()
        if slot is not None:
            ctx.record(slot)
        if batch is not None:
            batch.launch_device(d_seq, S)
        else:
            table.count_seq(d_seq, S)
        if slot is not None:
            ctx.record(slot + 1)
        if not bufs:
            hs, ls = table.export_sizes()
            bufs["h"] = ctx.alloc(int(hs * 1.05) + 4096)
            bufs["l"] = ctx.alloc(int(ls * 1.05) + 4096)
        return table.export_into(bufs["h"], bufs["l"])

    if args.copy_probe_bytes:
        a, b = ctx.alloc(args.copy_probe_bytes), ctx.alloc(args.copy_probe_bytes)
        kc.check(kc.lib().kc_copy_d2d(ctx.h, kc._ptr(b), kc._ptr(a), args.copy_probe_bytes), "copy probe")
        a.free()
        b.free()
    for _ in range(args.warmup):
        one_step()
    ctx.sync()

    t0 = time.perf_counter()
    for i in range(args.steps):
        hb, lb = one_step(2 * (i % 32))
    ctx.sync()
    elapsed = time.perf_counter() - t0
    count_ms = [ctx.elapsed_ms(2 * (i % 32), 2 * (i % 32) + 1) for i in range(max(0, args.steps - 32), args.steps)]

    st = table.stats()
    assert st["errors"] == 0, st
    K = st["kmers"]
    # size-independent checks: every window counted once; merged histogram
    # mass equals the number of windows (no count exceeds the bin range here)
    bins = table.histogram(1, 100000)
    mass = int((bins * np.arange(1, bins.size + 1)).sum())
    windows = int(np.sum(np.maximum(np.diff(np.flatnonzero(np.concatenate(
        [[True], seq_host == 10]))) - 1 - (k - 1), 0))) if ws.n_rate == 0 else K

    k_table = K - st["low"]
    alg_bytes = S + 2 * (8 * nw + 4) * k_table + (32 * K if bloom else 0)
    avg_count_s = float(np.mean(count_ms)) * 1e-3
    achieved = alg_bytes / avg_count_s / 1e9
    step_s = elapsed / args.steps
    value = K / step_s

    launch_info = {"launch": launch, "tune": tune}
    if batch is not None:
        launch_info["batch"] = batch.info()

    # host-batched line (PCIe-inclusive, never the headline): the reads as
    # uneven chunks of 1..64 reads packed by kc_batch from host memory (one
    # kc_batch_add_chunks call), with the staging queue (slot size x slots in
    # flight) tuned on this input; the same with one kc_batch_add per chunk
    host_batched = None
    if args.workload == "uneven_50mbp" and not args.no_host_batch:
        bounds = AT.chunk_bounds(seq_host, np.random.default_rng(7))
        q = AT.tune_queue(table, seq_host, bounds, use_cache=False)
        dt, info = AT.run_batches(table, seq_host, bounds, q["batch_bytes"], q["queue_depth"])
        st_h = table.stats()
        bins_h = table.histogram(1, 100000)
        same = bool(st_h["kmers"] == K and st_h["errors"] == 0 and np.array_equal(bins_h, bins))
        table.clear()
        dt1, _ = AT.run_batches(table, seq_host, bounds, q["batch_bytes"], q["queue_depth"], one_call=False)
        same = same and bool(np.array_equal(table.histogram(1, 100000), bins))
        table.clear()
        host_batched = {"kmers_per_s": st_h["kmers"] / dt, "seconds": dt, "chunks": int(bounds.size),
                        "batch_bytes": q["batch_bytes"], "queue_depth": q["queue_depth"],
                        "launches": info["launches"], "queue_trials": q["trials"],
                        "per_chunk_add_seconds": dt1, "same_as_device": same}

    multipass = None
    if args.passes > 1:
        multipass = bench_passes(ctx, args, table, d_seq, S, k, bloom, est_distinct, K, bins, tune)

    sample_reads = args.cpu_sample_reads or (300_000 if args.workload == "uneven_50mbp" else 4_000_000)
    cpu = None if args.no_cpu_baseline else cpu_baseline(ws, min(sample_reads, ws.n_reads), k, genome)
    line = {
        "metric": "canonical %d-mers counted/s on 1 MI355X vs host CPU" % k,
        "value": value, "unit": "k-mers/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": step_s * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic",
        "config": {"workload": ws.name, "reads": ws.n_reads, "read_len": ws.read_len, "k": k,
                   "bloom": bloom, "seq_bytes": S, "kmers": K, "table_keys": st["distinct"],
                   "first_occurrences": st["low"], "table_slots": st["capacity"],
                   "export_high_bytes": hb, "export_low_bytes": lb, "plan": plan,
                   "step": "clear + fused parse/canonical/Bloom/count + export packing (device)"},
        "roofline": {"bound": "hbm", "kernel": COUNT_KERNELS if nw <= 2 else DIRECT_KERNELS, "achieved": achieved,
                     "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic("count_pipeline", ws.name, k),
                     "alg_bytes_per_launch": alg_bytes, "avg_launch_ms": avg_count_s * 1e3},
        "cpu_baseline": cpu,
        "checks": {"windows_match": windows == K, "hist_mass_equals_kmers": mass == K},
        "launch": launch_info,
        "host_batched": host_batched,
        "multipass": multipass,
        "gen_seconds": gen_s,
    }
    print(json.dumps(line), flush=True)
    if batch is not None:
        batch.close()
    table.close()
    ctx.close()


def bench_passes(ctx, args, table, d_seq, S, k, bloom, est_distinct, K, bins, tune):
    """The out-of-HBM fallback on the same input: P passes (clear + count of
    hash slice p + export packing) into a table of 1/P the slots; checks
    that the slices' k-mers and summed histograms equal the one-pass run."""
    import numpy as np

    from ktcountermemorygreedy_amd import kc
    P = args.passes
    if P & (P - 1) or P > 256:
        raise SystemExit("--passes must be a power of two <= 256")
    one = table.stats()
    cap = max(1 << 20, one["capacity"] // P)
    md = est_distinct // P + (1 << 20)
    t = kc.Table(ctx, k, count_max=0xFFFFFFFF, capacity=cap, max_distinct=md, bloom=bloom,
                 low_capacity=md if bloom else 0)
    if tune:
        t.set_tune(**{f: tune[f] for f in ("front_blocks", "fine_slices", "path", "piece_bytes", "first_mode", "front_mode")
                      if f in tune})
    bufs = {}

    def passes(check=False):
        kmers, acc = 0, np.zeros_like(bins)
        for p in range(P):
            t.clear()
            t.set_pass(p, P)
            t.count_seq(d_seq, S)
            if not bufs:
                hs, ls = t.export_sizes()
                bufs["h"], bufs["l"] = ctx.alloc(int(hs * 1.25) + 4096), ctx.alloc(int(ls * 1.25) + 4096)
            t.export_into(bufs["h"], bufs["l"])
            if check:
                st = t.stats()
                assert st["errors"] == 0, st
                kmers += st["kmers"]
                acc += t.histogram(1, bins.size)
        return kmers, acc

    kmers, acc = passes(check=True)  # warm-up and checks
    ctx.sync()
    t0 = time.perf_counter()
    steps = max(1, args.steps)
    for _ in range(steps):
        passes()
    ctx.sync()
    step_s = (time.perf_counter() - t0) / steps
    for b in bufs.values():
        b.free()
    t.close()
    return {"passes": P, "table_slots": cap, "ms_per_step": step_s * 1e3, "kmers_per_s": K / step_s,
            "step": "P x (clear + count of one hash slice + export packing)",
            "checks": {"slice_kmers_sum": kmers == K, "summed_hist_equals_one_pass": bool(np.array_equal(acc, bins))}}


if __name__ == "__main__":
    main()
