"""Host side of the autotuner: chunk cutting for host batches and the tune
cache (no device calls)."""
import numpy as np

from ktcountermemorygreedy_amd import autotune as AT


def test_chunk_bounds_cover_stream_at_read_ends():
    rng = np.random.default_rng(1)
    lens = rng.integers(20, 9000, 5000)
    seq = np.concatenate([np.concatenate([np.full(n, ord("A"), np.uint8), [10]]) for n in lens]).astype(np.uint8)
    b = AT.chunk_bounds(seq, np.random.default_rng(2), max_reads=64)
    assert b[-1] == seq.size and np.all(np.diff(b) > 0)
    assert np.all(seq[b - 1] == 10)  # every chunk ends with a newline
    ends = np.flatnonzero(seq == 10) + 1
    reads_per_chunk = np.diff(np.searchsorted(ends, np.concatenate([[0], b]), side="left"))
    assert reads_per_chunk.min() >= 1 and reads_per_chunk.max() <= 64


def test_tune_cache_roundtrip(tmp_path, monkeypatch):
    path = tmp_path / "tune.json"
    monkeypatch.setenv("KC_TUNE_CACHE", str(path))
    assert AT._load() == {}
    AT._store("shape|k31|bloom1|slots1024|bytes2^20", {"front_blocks": 512, "fine_slices": 4})
    AT._store("queue|k31|bloom1|slots1024|bytes2^20", {"batch_bytes": 1 << 26, "queue_depth": 3})
    data = AT._load()
    assert data["shape|k31|bloom1|slots1024|bytes2^20"]["front_blocks"] == 512
    assert data["queue|k31|bloom1|slots1024|bytes2^20"]["queue_depth"] == 3
