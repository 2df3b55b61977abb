"""Autotuner front end.

Two searches, both timed on the device they tune for:

  shape   kc_autotune (C++, in libkc.so): the counting call over a
          device-resident sample for each launch shape (partitioned path:
          persistent front blocks per CU x fine slices per coarse region;
          direct path) -> the fastest shape, stored in the table's tune.
  queue   host batching: kc_batch over host-resident chunks for each staging
          slot size (batch_bytes) x slots in flight (queue_depth) -> the
          highest end-to-end rate (host packing + H2D copy + counting).

Results are cached in a JSON file ($KC_TUNE_CACHE, default
~/.cache/ktcountermemorygreedy_amd/tune.json) under a key made of k, the
Bloom flag, the table size and the input size class, so a later run on the
same kind of input reuses them.
"""
from __future__ import annotations

import json
import math
import os
import time

import numpy as np

from . import kc


def cache_path():
    return os.environ.get("KC_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "ktcountermemorygreedy_amd", "tune.json"))


def _load():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(key, value):
    data = _load()
    data[key] = value
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home only loses the cache


def tune_key(table: kc.Table, nbytes: int, what: str):
    st = table.stats()
    size_class = int(math.ceil(math.log2(max(nbytes, 2))))
    return "%s|k%d|bloom%d|slots%d|bytes2^%d" % (what, table.k, int(table.bloom), st["capacity"], size_class)


def tune_shape(table: kc.Table, d_sample, nbytes=None, reps=1, use_cache=True):
    """Launch shape for counting calls of `table` (left empty); returns the
    tune dict, with 'ms' (best time on the sample) and 'cached'."""
    n = getattr(d_sample, "valid", d_sample.nbytes) if nbytes is None else nbytes
    key = tune_key(table, n, "shape")
    if use_cache:
        hit = _load().get(key)
        if hit:
            table.set_tune(**{f: hit[f] for f in ("front_blocks", "fine_slices", "batch_bytes", "queue_depth", "path",
                                                  "piece_bytes", "first_mode", "front_mode") if f in hit})
            return dict(hit, cached=True)
    tune, ms = table.autotune(d_sample, n, reps=reps)
    rec = dict(tune, ms=ms)
    if use_cache:
        _store(key, rec)
    return dict(rec, cached=False)


def chunk_bounds(seq: np.ndarray, rng, max_reads=64):
    """Cuts a framed stream into chunks of 1..max_reads whole reads (uneven
    byte lengths); returns the end offset of every chunk."""
    ends = np.flatnonzero(seq == 10) + 1
    cuts, i = [], 0
    while i < ends.size:
        i = min(ends.size, i + int(rng.integers(1, max_reads + 1)))
        cuts.append(int(ends[i - 1]))
    return np.asarray(cuts, dtype=np.int64)


def run_batches(table: kc.Table, seq: np.ndarray, bounds, batch_bytes, queue_depth, one_call=True):
    """Counts every chunk through one kc_batch, all in one
    kc_batch_add_chunks call (or one kc_batch_add per chunk); returns
    (seconds, info)."""
    b = kc.Batch(table, batch_bytes=batch_bytes, queue_depth=queue_depth)
    try:
        t0 = time.perf_counter()
        if one_call:
            b.add_chunks(seq, bounds)
        else:
            start = 0
            for end in bounds:
                b.add(seq[start:end])
                start = int(end)
        b.flush()
        b.sync()
        dt = time.perf_counter() - t0
        return dt, b.info()
    finally:
        b.close()


def tune_queue(table: kc.Table, seq: np.ndarray, bounds, candidates=None, use_cache=True):
    """Staging slot size x slots in flight for host batches of `seq` (the
    table is cleared before each trial and left empty).  Returns the best
    {'batch_bytes', 'queue_depth', 'seconds'} plus every trial."""
    key = tune_key(table, int(seq.size), "queue")
    if use_cache:
        hit = _load().get(key)
        if hit:
            return dict(hit, cached=True)
    if candidates is None:
        candidates = [(bb, qd) for bb in (64 << 20, 256 << 20, 1 << 30) for qd in (2, 3, 4)
                      if bb <= max(64 << 20, 2 * seq.size)]
    trials = []
    for bb, qd in candidates:
        table.clear()
        dt, _ = run_batches(table, seq, bounds, bb, qd)
        trials.append({"batch_bytes": bb, "queue_depth": qd, "seconds": dt})
    table.clear()
    best = min(trials, key=lambda r: r["seconds"])
    rec = dict(best, trials=trials)
    if use_cache:
        _store(key, rec)
    return dict(rec, cached=False)
