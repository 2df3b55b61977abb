"""Device export-record sorts against the reference sorts' outputs (KATs
from the reference headers) and the oracle at larger sizes."""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.kc_testlib import GOLDEN

pytestmark = pytest.mark.gpu


def pack_records(keys, counts, nw, record_bytes):
    n = keys.shape[0]
    rw = record_bytes // 4
    a = np.zeros((n, rw), dtype=np.uint32)
    a[:, :2 * nw] = keys.view(np.uint32).reshape(n, 2 * nw)
    a[:, 2 * nw] = counts
    return a


def unpack_records(a, nw):
    n = a.shape[0]
    keys = np.ascontiguousarray(a[:, :2 * nw]).view(np.uint64).reshape(n, nw)
    return keys, a[:, 2 * nw].copy()


def device_sort(ctx, keys, counts, nw, k, bits, record_bytes, cpb=10):
    a = pack_records(keys, counts, nw, record_bytes)
    d = ctx.upload(a)
    s = ctx.alloc(a.nbytes + 16)
    in_scratch = ctx.sort_records(d, s, keys.shape[0], nw, k, cpb, bits, record_bytes)
    out = (s if in_scratch else d).numpy(np.uint32, a.size).reshape(a.shape)
    return unpack_records(out, nw)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "sort_kat_*.npz"))),
                         ids=lambda p: p.split("/")[-1])
@pytest.mark.parametrize("packed", [True, False], ids=["packed", "padded"])
def test_sort_matches_reference_kat(ctx, path, packed):
    d = np.load(path)
    nw, k, bits = int(d["nw"]), int(d["k"]), int(d["bits"])
    rb = 8 * nw + (4 if packed else 8)
    sk, sc = device_sort(ctx, d["keys"], d["counts"], nw, k, bits, rb)
    assert np.array_equal(sk, d["sorted_keys"])
    assert np.array_equal(sc, d["sorted_counts"])


@pytest.mark.parametrize("nw,k,bits", [(1, 31, 12), (1, 31, 8), (2, 63, 12), (2, 51, 8), (1, 25, 12)])
def test_sort_large_vs_oracle(ctx, nw, k, bits):
    rng = np.random.default_rng(k * 100 + bits)
    n = 300_000
    keys = rng.integers(0, 2**64 - 1, size=(n, nw), dtype=np.uint64)
    pad = 64 * nw - 2 * k
    w = nw - 1
    while pad > 0:
        b = min(pad, 64)
        keys[:, w] &= np.uint64((~((1 << b) - 1)) & 0xFFFFFFFFFFFFFFFF)
        pad -= b
        w -= 1
    keys[n // 3: n // 3 + 20000] = keys[:20000]  # equal keys: stability
    counts = np.arange(n, dtype=np.uint32)
    want = O.sort_records12(keys, counts, k, 10) if bits == 12 else O.sort_records8(keys, counts, k)
    got = device_sort(ctx, keys, counts, nw, k, bits, 8 * nw + 4)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])


def test_sort_edges(ctx):
    from ktcountermemorygreedy_amd import kc
    for n in (0, 1, 2, 4095, 4096, 4097):
        keys = (np.arange(n, 0, -1, dtype=np.uint64) << np.uint64(30)).reshape(-1, 1)
        counts = np.arange(n, dtype=np.uint32)
        got = device_sort(ctx, keys, counts, 1, 31, 12, 12)
        want = O.sort_records12(keys, counts, 31, 10) if n else (keys, counts)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    d = ctx.alloc(64)
    s = ctx.alloc(64)
    with pytest.raises(kc.KCError):
        ctx.sort_records(d, s, 2, 1, 9, 10, 12, 12)  # k below the common prefix
    with pytest.raises(kc.KCError):
        ctx.sort_records(d, d, 2, 1, 31, 10, 12, 12)  # same buffer


def test_synth_records_sort_and_histogram_properties(ctx):
    """Seeded Zipf record workload: sortedness on bits [2,44), multiset kept,
    histogram equal to numpy's over the counts."""
    n = 1 << 22
    d = ctx.alloc(12 * n)
    s = ctx.alloc(12 * n)
    ctx.synth_records(d, n, 12, seed=4, zipf_s=1.2)
    before = d.numpy(np.uint32, 3 * n).reshape(n, 3)
    bins = ctx.records_histogram(d, n, 1, 12, 1, 10000)
    c = before[:, 2]
    want = np.bincount(c[(c >= 1) & (c <= 10000)], minlength=10001)[1:]
    assert np.array_equal(bins, want)
    in_s = ctx.sort_records(d, s, n, 1, 31, 10, 12, 12)
    after = (s if in_s else d).numpy(np.uint32, 3 * n).reshape(n, 3)
    k0 = before[:, 0].astype(np.uint64) | (before[:, 1].astype(np.uint64) << np.uint64(32))
    k1 = after[:, 0].astype(np.uint64) | (after[:, 1].astype(np.uint64) << np.uint64(32))
    assert len(np.unique(k0)) == n  # distinct keys (bijection)
    sk = (k1 >> np.uint64(2)) & np.uint64((1 << 42) - 1)
    assert np.all(sk[1:] >= sk[:-1])
    assert np.array_equal(np.sort(k0), np.sort(k1))
    pos = {int(x): i for i, x in enumerate(k0[:1000])}
    m = {int(x): int(cc) for x, cc in zip(k1, after[:, 2]) if int(x) in pos}
    assert all(m[int(x)] == int(before[i, 2]) for x, i in pos.items())


def _sorted_stably(keys, counts, out_keys, out_counts, k, cpb=10):
    """The output holds the input's (key, count) pairs, ordered by sort-key
    bits [2(32-k), 2(32-cpb)) of word 0 (k <= 32), equal sort keys in input
    order (counts are the input positions)."""
    lo, hi = 64 - 2 * k, 64 - 2 * cpb
    sk = (out_keys[:, 0] >> np.uint64(lo)) & np.uint64((1 << (hi - lo)) - 1)
    assert np.all(sk[1:] >= sk[:-1])
    eq = sk[1:] == sk[:-1]
    assert np.all(out_counts[1:][eq] > out_counts[:-1][eq])
    assert np.array_equal(keys[out_counts, 0], out_keys[:, 0])


def test_sort_keys_concentrated_on_few_prefixes(ctx):
    """Half of 300k keys on eight values of the top sort-key bits (digits
    with ~19k records in every tile's run) and the rest spread, with equal
    keys: equal to the oracle."""
    rng = np.random.default_rng(77)
    n = 300_000
    keys = rng.integers(0, 2**62, size=(n, 1), dtype=np.uint64) << np.uint64(2)
    hot = rng.random(n) < 0.5
    keys[hot, 0] = (keys[hot, 0] & np.uint64((1 << 40) - 1)) | np.uint64(0x5A5 << 52)
    keys[n // 2: n // 2 + 5000] = keys[:5000]  # equal keys: stability
    counts = np.arange(n, dtype=np.uint32)
    want = O.sort_records12(keys, counts, 31, 10)
    got = device_sort(ctx, keys, counts, 1, 31, 12, 12)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])


def test_sort_large_skewed_prefixes(ctx):
    """2^24 keys whose sort-key bits [32, 44) take 1100 of their 4096 values:
    sorted, stable, the same (key, count) pairs."""
    rng = np.random.default_rng(78)
    n = 1 << 24
    # sort-key bits [32, 44) on 1100 of their 4096 values (~15k records
    # each); bits [2, 32) and [44, 64) random
    pv = rng.choice(4096, size=1100, replace=False).astype(np.uint64)
    pre = pv[rng.integers(0, 1100, size=n)] << np.uint64(32)
    low = rng.integers(0, 2**30, size=n, dtype=np.uint64) << np.uint64(2)
    top = rng.integers(0, 2**20, size=n, dtype=np.uint64) << np.uint64(44)
    keys = (top | pre | low).reshape(n, 1)
    counts = np.arange(n, dtype=np.uint32)
    got = device_sort(ctx, keys, counts, 1, 31, 12, 12)
    _sorted_stably(keys, counts, got[0], got[1], 31)
