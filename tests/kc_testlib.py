"""Shared helpers: run the HIP engine through its C ABI and derive the
merged views with the oracle's tool semantics (the oracle is the checker)."""
from __future__ import annotations

import json
import os

import numpy as np

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_json(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def read_golden(name, mode="rb"):
    with open(os.path.join(GOLDEN, name), mode) as f:
        return f.read()


def views_from_export(high: bytes, low: bytes, k: int, count_max: int, hist_max=10000):
    """dump --sort text and histogram text of exported high/low bytes,
    computed with the reference tools' merge rules (oracle)."""
    nw = O.words_for_k(k)
    hk, hc = O.parse_high(high, k, count_max)
    lk = O.unpack(low, k) if low else np.zeros((0, nw), np.uint64)
    mk, mc = O.merge(hk, hc, lk, nw, count_max)
    dump = O.dump_text(mk, mc, k)
    bins = O.histogram(hk, hc, lk, nw, count_max, 1, hist_max)
    return dump, O.histogram_text(bins, 1, hist_max)


def gpu_count(ctx, seq: bytes, k: int, count_max: int, bloom=False, host_path=False, max_distinct=None,
              **kw):
    from ktcountermemorygreedy_amd import kc
    md = max_distinct or max(1024, len(seq))
    t = kc.Table(ctx, k, count_max=count_max, max_distinct=md, bloom=bloom, **kw)
    if host_path:
        t.count_seq_host(seq)
    else:
        d = ctx.upload(np.frombuffer(seq, dtype=np.uint8))
        t.count_seq(d, len(seq))
        ctx.sync()
        d.free()
    return t


def gpu_views(ctx, seq: bytes, k: int, count_max: int, bloom=False, hist_max=10000, host_path=False, **kw):
    t = gpu_count(ctx, seq, k, count_max, bloom=bloom, host_path=host_path, **kw)
    st = t.stats()
    assert st["errors"] == 0, st
    high, low = t.export_device()
    dump, hist_from_files = views_from_export(high, low, k, count_max, hist_max)
    dev_hist = O.histogram_text(t.histogram(1, hist_max), 1, hist_max)
    t.close()
    return dump, dev_hist, hist_from_files, st
