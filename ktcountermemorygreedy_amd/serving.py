"""Serving workload: a Poisson stream of count jobs and point-query batches
against device-resident tables, driven through kc_server (include/kc_server.h).

The reference has no serving mode; its only point lookup is
FlatConcurrentHashMap::find_prepared (LowFrequencyQueryThreadPool.h:196-244).
This driver replays a seeded schedule:

  jobs     Poisson arrivals, `job_rate` per second.  Job size lognormal over
           reads (mu 12.664, sigma 1.485) clipped to [1e4, 1e7] reads of
           150 bp, from one of 10 uniform 20 Mbp genomes (seeds 101..110),
           0.5% substitutions, both strands.  Each genome has one pool of
           `pool_reads` framed reads in page-locked memory; a job counts a
           contiguous slice of its genome's pool into a new table (k = 31).
  queries  Poisson arrivals, `query_rate` per second, each a batch of
           log-uniform size in [1, 65536] against one of the 16 most recently
           completed tables: half of the keys are canonical 31-mers sampled
           from the job's reads (present), half are canonical 31-mers of
           random bases (absent with probability above 1 - 1e-9 each).
  q64k     a second Poisson stream, `q64_rate` per second, of batches of
           exactly 65536 keys (the batch size the latency target is quoted
           for), built the same way; after the schedule, `idle_q64` more such
           batches are sent one at a time with no job running (idle latency).

Every query answer is checked: present keys >= 1, absent keys 0.
"""
from __future__ import annotations

import math
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass

import numpy as np

from . import kc

_CODE = np.full(256, 255, dtype=np.uint8)
for _c, _v in zip(b"ACGTUacgtu", (0, 1, 2, 3, 3, 0, 1, 2, 3, 3)):
    _CODE[_c] = _v


def canonical_kmers(windows: np.ndarray, k: int) -> np.ndarray:
    """Canonical keys (k <= 32, one word) of an (n, k) array of base bytes.

    Position j of a window sits at bit 64 - 2k + 2j of the word and the
    reverse complement puts 3 - c at the mirrored position (kmer.h:11-119,
    GetKmer.h:58-82); the canonical key is the smaller number.  Windows
    must hold only ACGTU (either case)."""
    if k < 1 or k > 32 or windows.ndim != 2 or windows.shape[1] != k:
        raise ValueError("expected an (n, k) byte array with 1 <= k <= 32")
    codes = _CODE[windows].astype(np.uint64)
    if (codes == 255).any():
        raise ValueError("window holds a byte outside ACGTU")
    base0 = 64 - 2 * k
    sh = np.uint64(base0) + np.uint64(2) * np.arange(k, dtype=np.uint64)
    fwd = np.bitwise_or.reduce(codes << sh, axis=1)
    rev = np.bitwise_or.reduce((np.uint64(3) - codes) << sh[::-1], axis=1)
    return np.minimum(fwd, rev)


@dataclass
class ServingSpec:
    seed: int = 5
    job_rate: float = 20.0
    query_rate: float = 200.0
    duration_s: float = 5.0
    n_genomes: int = 10
    genome_len: int = 20_000_000
    genome_seed0: int = 101
    reads_mu: float = 12.664
    reads_sigma: float = 1.485
    min_reads: int = 10_000
    max_reads: int = 10_000_000
    read_len: int = 150
    sub_rate: float = 0.005
    max_query: int = 65536
    k: int = 31
    pool_reads: int = 10_000_000
    recent_tables: int = 16
    workers: int = 8
    hbm_budget: int = 160 << 30
    q64_rate: float = 10.0
    idle_q64: int = 64


def job_distinct_estimate(spec: ServingSpec, n_reads: int) -> int:
    """Upper estimate of distinct k-mers of a job: genome windows covered
    plus one new k-mer per window overlapping a substitution, +10%."""
    w = spec.read_len - spec.k + 1
    genomic = min(spec.genome_len, n_reads * w)
    errors = n_reads * spec.read_len * spec.sub_rate * spec.k
    return int(1.1 * (genomic + errors)) + 4096


@dataclass
class Schedule:
    job_t: np.ndarray       # arrival times (s)
    job_reads: np.ndarray   # reads per job
    job_genome: np.ndarray  # genome index per job
    job_off: np.ndarray     # first read of the job in its genome's pool
    q_t: np.ndarray         # query batch arrival times (s)
    q_n: np.ndarray         # keys per batch
    q_pick: np.ndarray      # picks the target among the recent tables
    q64_t: np.ndarray       # arrival times of the 65536-key batches
    q64_pick: np.ndarray


def make_schedule(spec: ServingSpec, rng: np.random.Generator) -> Schedule:
    if spec.max_reads > spec.pool_reads:
        raise ValueError("pool_reads must hold the largest job")
    n_jobs = max(1, int(rng.poisson(spec.job_rate * spec.duration_s)))
    job_t = np.sort(rng.uniform(0.0, spec.duration_s, n_jobs))
    sizes = np.exp(rng.normal(spec.reads_mu, spec.reads_sigma, n_jobs))
    job_reads = np.clip(np.rint(sizes), spec.min_reads, spec.max_reads).astype(np.int64)
    job_genome = rng.integers(0, spec.n_genomes, n_jobs)
    job_off = np.array([rng.integers(0, spec.pool_reads - n + 1) for n in job_reads], dtype=np.int64)
    n_q = int(rng.poisson(spec.query_rate * spec.duration_s))
    q_t = np.sort(rng.uniform(0.0, spec.duration_s, n_q))
    q_n = np.clip(np.rint(np.exp(rng.uniform(0.0, math.log(spec.max_query), n_q))), 1,
                  spec.max_query).astype(np.int64)
    q_pick = rng.integers(0, 1 << 30, n_q)
    n_q64 = int(rng.poisson(spec.q64_rate * spec.duration_s))
    q64_t = np.sort(rng.uniform(0.0, spec.duration_s, n_q64))
    q64_pick = rng.integers(0, 1 << 30, n_q64)
    return Schedule(job_t, job_reads, job_genome, job_off, q_t, q_n, q_pick, q64_t, q64_pick)


class ServingRun:
    """Builds the pools and the schedule (untimed), then replays it."""

    def __init__(self, ctx: kc.Context, spec: ServingSpec = ServingSpec()):
        self.ctx, self.spec = ctx, spec
        self.rng = np.random.default_rng(spec.seed)
        self.rec = spec.read_len + 1  # framed bytes per read
        sc = make_schedule(spec, self.rng)
        self.schedule = sc
        self.job_t, self.job_reads, self.job_genome, self.job_off = sc.job_t, sc.job_reads, sc.job_genome, sc.job_off
        self.q_t, self.q_n, self.q_pick = sc.q_t, sc.q_n, sc.q_pick
        self.q64_t, self.q64_pick = sc.q64_t, sc.q64_pick
        n_jobs = self.job_t.size
        # pools: only the genomes the schedule uses
        self.pools = {}
        for g in sorted(set(int(x) for x in self.job_genome)):
            genome = kc.synth_genome(spec.genome_seed0 + g, spec.genome_len)
            pool = ctx.pinned(spec.pool_reads * self.rec)
            kc.synth_reads(genome, 1000 + spec.genome_seed0 + g, spec.pool_reads, spec.read_len, True,
                           spec.sub_rate, out=pool.array)
            self.pools[g] = pool
        # per job: present-key sample drawn from its slice of the pool
        self.present = [self._sample_present(j) for j in range(n_jobs)]
        self.absent = self._random_kmers(spec.max_query)

    def job_bytes(self, j):
        pool = self.pools[int(self.job_genome[j])].array
        a = int(self.job_off[j]) * self.rec
        return pool[a:a + int(self.job_reads[j]) * self.rec]

    def _sample_present(self, j):
        spec = self.spec
        n = min(spec.max_query // 2 + 1, int(self.job_reads[j]))
        seq = self.job_bytes(j)
        rd = self.rng.integers(0, int(self.job_reads[j]), n)
        pos = self.rng.integers(0, spec.read_len - spec.k + 1, n)
        idx = (rd * self.rec + pos)[:, None] + np.arange(spec.k)[None, :]
        return canonical_kmers(seq[idx], spec.k)

    def _random_kmers(self, n):
        bases = np.frombuffer(b"ACGT", dtype=np.uint8)[self.rng.integers(0, 4, (n, self.spec.k))]
        return canonical_kmers(bases, self.spec.k)

    def run(self, server: kc.Server, table_base=1):
        spec = self.spec
        lock = threading.Lock()
        completed = []     # job indices in completion order
        jobs, queries, q64 = [], [], []
        events = sorted([(t, 0, j) for j, t in enumerate(self.job_t)] + [(t, 1, q) for q, t in enumerate(self.q_t)] +
                        [(t, 2, q) for q, t in enumerate(self.q64_t)])
        ex = ThreadPoolExecutor(max_workers=64)

        def wait_job(j, rid):
            res, _ = server.wait(rid)
            with lock:
                jobs.append((j, res))
                if res["status"] == kc.KC_OK:
                    completed.append(j)

        def wait_query(j, n_present, rid, big):
            res, counts = server.wait(rid)
            ok = res["status"] == kc.KC_OK
            bad = 0
            if ok:
                bad = int((counts[:n_present] == 0).sum() + (counts[n_present:] != 0).sum())
            with lock:
                (q64 if big else queries).append((j, int(counts.size) if counts is not None else 0, res, bad))

        futs = []
        t0 = time.monotonic()
        skipped = 0
        for t, kind, i in events:
            delay = t0 + t - time.monotonic()
            if delay > 0:
                time.sleep(delay)
            if kind == 0:
                seq = self.job_bytes(i)
                rid = server.submit_count(table_base + i, seq, k=spec.k,
                                          max_distinct=job_distinct_estimate(spec, int(self.job_reads[i])))
                futs.append(ex.submit(wait_job, i, rid))
            else:
                with lock:
                    recent = completed[-spec.recent_tables:]
                if not recent:
                    skipped += 1
                    continue
                big = kind == 2
                j = recent[int((self.q64_pick if big else self.q_pick)[i]) % len(recent)]
                n = spec.max_query if big else int(self.q_n[i])
                keys, n_present = self.query_keys(j, n)
                rid = server.submit_query(table_base + j, keys)
                futs.append(ex.submit(wait_query, j, n_present, rid, big))
        for f in futs:
            f.result()
        wall = time.monotonic() - t0
        ex.shutdown()
        idle = self.run_idle(server, table_base, completed[-spec.recent_tables:])
        return self._summary(jobs, queries, q64, idle, skipped, wall, server)

    def query_keys(self, j, n):
        """n keys for job j's table: half present, half absent."""
        n_present = min(n // 2, self.present[j].size)
        return np.concatenate([self.present[j][:n_present], self.absent[:n - n_present]]), n_present

    def run_idle(self, server, table_base, recent):
        """65536-key batches one at a time with no job running."""
        out = []
        if not recent:
            return out
        for i in range(self.spec.idle_q64):
            j = recent[i % len(recent)]
            keys, n_present = self.query_keys(j, self.spec.max_query)
            res, counts = server.wait(server.submit_query(table_base + j, keys))
            bad = int((counts[:n_present] == 0).sum() + (counts[n_present:] != 0).sum()) \
                if res["status"] == kc.KC_OK else 0
            out.append((j, keys.size, res, bad))
        return out

    def _summary(self, jobs, queries, q64, idle, skipped, wall, server):
        def pct(v, p):
            return float(np.percentile(v, p)) if len(v) else None

        def lat(qs):
            return [r["latency_ms"] for _, _, r, _ in qs if r["status"] == kc.KC_OK]

        jl = [r["latency_ms"] for _, r in jobs if r["status"] == kc.KC_OK]
        ql = [r["latency_ms"] for _, _, r, _ in queries if r["status"] == kc.KC_OK]
        kmers = sum(r["kmers"] for _, r in jobs if r["status"] == kc.KC_OK)
        resident, resident_bytes = server.resident()
        return {
            "jobs": len(jobs),
            "jobs_failed": sum(1 for _, r in jobs if r["status"] != kc.KC_OK),
            "job_status": sorted(set(r["status"] for _, r in jobs)),
            "job_p50_ms": pct(jl, 50), "job_p99_ms": pct(jl, 99),
            "job_queue_p99_ms": pct([r["queue_ms"] for _, r in jobs], 99),
            "kmers": int(kmers), "wall_s": wall, "kmers_per_s": kmers / wall,
            "queries": len(queries), "queries_skipped": skipped,
            "queries_failed": sum(1 for _, _, r, _ in queries if r["status"] != kc.KC_OK),
            "query_keys": int(sum(n for _, n, _, _ in queries)),
            "query_p50_ms": pct(ql, 50), "query_p99_ms": pct(ql, 99),
            "query_wrong": int(sum(b for _, _, _, b in queries)),
            "query_merged_share": float(np.mean([r["batched"] > 1 for _, _, r, _ in queries])) if queries else None,
            "q64k": len(q64), "q64k_failed": sum(1 for _, _, r, _ in q64 if r["status"] != kc.KC_OK),
            "q64k_wrong": int(sum(b for _, _, _, b in q64)),
            "q64k_p50_ms": pct(lat(q64), 50), "q64k_p99_ms": pct(lat(q64), 99),
            "q64k_max_ms": max(lat(q64)) if q64 else None,
            # of which waiting in the server's queue (submit to start)
            "q64k_queue_p50_ms": pct([r["queue_ms"] for _, _, r, _ in q64 if r["status"] == kc.KC_OK], 50),
            "q64k_queue_p99_ms": pct([r["queue_ms"] for _, _, r, _ in q64 if r["status"] == kc.KC_OK], 99),
            "q64k_idle": len(idle), "q64k_idle_wrong": int(sum(b for _, _, _, b in idle)),
            "q64k_idle_failed": sum(1 for _, _, r, _ in idle if r["status"] != kc.KC_OK),
            "q64k_idle_p50_ms": pct(lat(idle), 50), "q64k_idle_p99_ms": pct(lat(idle), 99),
            "resident_tables": resident, "resident_gb": resident_bytes / 1e9,
        }
