"""Device hashes and Bloom probes, bit-exact with the reference headers."""
import numpy as np
import pytest

from tests.kc_testlib import read_golden

pytestmark = pytest.mark.gpu


def _rows():
    rows = {1: [], 2: [], 4: []}
    for line in read_golden("hash_kat.txt", "r").splitlines():
        if line.startswith("#") or not line.strip():
            continue
        f = line.split()
        nw = int(f[0])
        rows[nw].append(([int(x, 16) for x in f[1:1 + nw]], [int(x, 16) for x in f[1 + nw:]]))
    return rows


@pytest.mark.parametrize("nw", [1, 2, 4])
def test_device_hashes_match_kat(ctx, nw):
    from ktcountermemorygreedy_amd import kc
    rows = _rows()[nw]
    keys = np.array([w for w, _ in rows], dtype=np.uint64)
    d = keys.reshape(-1)
    kinds = [kc.HASH_XXH3_64, kc.HASH_XXH3_128_LO, kc.HASH_XXH3_128_HI, kc.HASH_RAPIDHASH,
             kc.HASH_KOMIHASH, kc.HASH_CONCURRENT_MAP, kc.HASH_FUNC]
    for col, kind in enumerate(kinds):
        got = ctx.hash_keys(kind, d, nw)
        want = np.array([v[col] for _, v in rows], dtype=np.uint64)
        assert np.array_equal(got, want), (nw, kind)
    blk, msk = ctx.bloom_probe(d, nw, 1 << 20)
    assert np.array_equal(blk, np.array([v[7] for _, v in rows], np.uint64))
    assert np.array_equal(msk, np.array([v[8] for _, v in rows], np.uint64))
