"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the cpu_baseline leg of ``bench.py`` as the checker.  The product package
never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    P = C.c_void_p
    u32, u64, i32 = C.c_uint32, C.c_uint64, C.c_int
    sig = {
        "kco_words_for_k": (u32, [u32]),
        "kco_fastq_frame": (u64, [P, u64, P]),
        "kco_extract": (u64, [P, u64, u32, P, u64]),
        "kco_encode_window": (i32, [P, u32, P]),
        "kco_count": (u64, [P, u64, u32, P, P]),
        "kco_packed_kmer_bytes": (u32, [u32]),
        "kco_count_bytes": (u32, [u32]),
        "kco_pack_kmer": (None, [P, u32, P]),
        "kco_unpack_kmer": (None, [P, u32, P]),
        "kco_merge": (u64, [P, P, u64, P, u64, u32, u32, u32, u32, P, P]),
        "kco_histogram": (None, [P, P, u64, P, u64, u32, u32, u32, u32, P]),
        "kco_dump_text": (u64, [P, P, u64, u32, P]),
        "kco_histogram_text": (u64, [P, u32, u32, P]),
        "kco_sort_records8": (None, [P, P, u64, u32, u32]),
        "kco_sort_records12": (i32, [P, P, u64, u32, u32, u32]),
        "kco_xxh3_64": (u64, [P, u64]),
        "kco_xxh3_128": (None, [P, u64, P, P]),
        "kco_rapidhash": (u64, [P, u64]),
        "kco_komihash": (u64, [P, u64, u64]),
        "kco_mix_hash": (u64, [u64, u64]),
        "kco_hash_func": (u64, [P, u32]),
        "kco_concurrent_map_hash": (u64, [P, u32]),
        "kco_bloom_probe": (None, [P, u32, u64, P, P]),
        "kco_extract_slice": (u64, [P, u64, u32, u32, u32, P, u64]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def words_for_k(k):
    return lib().kco_words_for_k(k)


def fastq_frame(fq: bytes) -> bytes:
    src = np.frombuffer(fq, dtype=np.uint8)
    out = np.empty(max(1, len(fq)), dtype=np.uint8)
    n = lib().kco_fastq_frame(_p(src), len(fq), _p(out))
    return out[:n].tobytes()


def extract(seq: bytes, k: int) -> np.ndarray:
    """All canonical k-mers in stream order, shape (n, nw) uint64."""
    nw = words_for_k(k)
    src = np.frombuffer(seq, dtype=np.uint8)
    cap = max(1, len(seq))
    out = np.empty((cap, nw), dtype=np.uint64)
    n = lib().kco_extract(_p(src), len(seq), k, _p(out), cap)
    return out[:n].copy()


def extract_slice(seq, k: int, pass_bits: int, pass_id: int, threads: int = 1) -> np.ndarray:
    """Canonical k-mers of hash slice pass_id of 2^pass_bits (the engine's
    multi-pass filter), in stream order.  seq: bytes or a uint8 array; with
    threads > 1 the input is cut at newlines and the parts run in parallel
    (ctypes releases the GIL)."""
    nw = words_for_k(k)
    src = np.frombuffer(seq, dtype=np.uint8) if isinstance(seq, (bytes, bytearray)) else \
        np.ascontiguousarray(seq, dtype=np.uint8)
    cuts = [0]
    for i in range(1, max(1, threads)):
        c = int(src.size * i // threads)
        nl = np.flatnonzero(src[c:c + (1 << 20)] == 10)
        if nl.size:
            cuts.append(max(cuts[-1], c + int(nl[0]) + 1))
    cuts.append(src.size)

    def run(a, b):
        part = src[a:b]
        cap = max(1, part.size // (1 << pass_bits) // 16 + 4096)
        while True:
            out = np.empty((cap, nw), dtype=np.uint64)
            n = lib().kco_extract_slice(C.c_void_p(part.ctypes.data), part.size, k, pass_bits, pass_id,
                                        _p(out), cap)
            if n <= cap:
                return out[:n].copy()
            cap = int(n)

    if len(cuts) == 2:
        return run(0, src.size)
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(len(cuts) - 1) as ex:
        parts = list(ex.map(lambda ab: run(*ab), zip(cuts[:-1], cuts[1:])))
    return np.concatenate(parts) if parts else np.zeros((0, nw), np.uint64)


def encode_window(s: bytes, k: int):
    nw = words_for_k(k)
    out = np.zeros(nw, dtype=np.uint64)
    src = np.frombuffer(s, dtype=np.uint8)
    ok = lib().kco_encode_window(_p(src), k, _p(out))
    return out if ok else None


def count(keys: np.ndarray, nw: int):
    keys = np.ascontiguousarray(keys, dtype=np.uint64).reshape(-1, nw).copy()
    n = keys.shape[0]
    uniq = np.empty((max(1, n), nw), dtype=np.uint64)
    cnt = np.empty(max(1, n), dtype=np.uint64)
    u = lib().kco_count(_p(keys), n, nw, _p(uniq), _p(cnt))
    return uniq[:u].copy(), cnt[:u].copy()


def packed_kmer_bytes(k):
    return lib().kco_packed_kmer_bytes(k)


def count_bytes(count_max):
    return lib().kco_count_bytes(count_max)


def pack(keys: np.ndarray, k: int) -> bytes:
    nw = words_for_k(k)
    keys = np.ascontiguousarray(keys, dtype=np.uint64).reshape(-1, nw)
    pb = packed_kmer_bytes(k)
    out = np.zeros(keys.shape[0] * pb, dtype=np.uint8)
    for i in range(keys.shape[0]):
        lib().kco_pack_kmer(_p(keys[i]), k, C.c_void_p(out.ctypes.data + i * pb))
    return out.tobytes()


def unpack(buf: bytes, k: int) -> np.ndarray:
    nw = words_for_k(k)
    pb = packed_kmer_bytes(k)
    n = len(buf) // pb
    src = np.frombuffer(buf, dtype=np.uint8)
    out = np.zeros((n, nw), dtype=np.uint64)
    for i in range(n):
        lib().kco_unpack_kmer(C.c_void_p(src.ctypes.data + i * pb), k, _p(out[i]))
    return out


def split_export(uniq: np.ndarray, counts: np.ndarray, k: int, count_max: int):
    """Simplest valid export split: singletons -> low, rest -> high with
    min(count, count_max).  Returns (infos, high, low) file bytes."""
    nw = words_for_k(k)
    uniq = uniq.reshape(-1, nw)
    cb = count_bytes(count_max)
    single = counts == 1
    low = pack(uniq[single], k)
    hk = uniq[~single]
    hc = np.minimum(counts[~single], count_max).astype(np.uint32)
    pb = packed_kmer_bytes(k)
    kb = np.frombuffer(pack(hk, k), dtype=np.uint8).reshape(-1, pb)
    cbytes = hc.view(np.uint8).reshape(-1, 4)[:, :cb]
    high = np.concatenate([kb, cbytes], axis=1).tobytes() if len(hk) else b""
    infos = np.array([k, count_max], dtype=np.uint32).tobytes()
    return infos, high, low


def parse_high(buf: bytes, k: int, count_max: int):
    nw = words_for_k(k)
    pb = packed_kmer_bytes(k)
    cb = count_bytes(count_max)
    rec = pb + cb
    n = len(buf) // rec
    arr = np.frombuffer(buf, dtype=np.uint8).reshape(n, rec)
    keys = unpack(arr[:, :pb].tobytes(), k) if n else np.zeros((0, nw), np.uint64)
    cpad = np.zeros((n, 4), dtype=np.uint8)
    cpad[:, :cb] = arr[:, pb:]
    return keys, cpad.view(np.uint32).reshape(-1).copy()


def merge(high_keys, high_counts, low_keys, nw, count_max, min_freq=1, max_freq=0xFFFFFFFF):
    hk = np.ascontiguousarray(high_keys, dtype=np.uint64).reshape(-1, nw)
    hc = np.ascontiguousarray(high_counts, dtype=np.uint32)
    lk = np.ascontiguousarray(low_keys, dtype=np.uint64).reshape(-1, nw)
    m = hk.shape[0] + lk.shape[0]
    ko = np.empty((max(1, m), nw), dtype=np.uint64)
    co = np.empty(max(1, m), dtype=np.uint32)
    r = lib().kco_merge(_p(hk), _p(hc), hk.shape[0], _p(lk), lk.shape[0], nw, count_max,
                        min_freq, max_freq, _p(ko), _p(co))
    return ko[:r].copy(), co[:r].copy()


def histogram(high_keys, high_counts, low_keys, nw, count_max, min_freq=1, max_freq=10000):
    hk = np.ascontiguousarray(high_keys, dtype=np.uint64).reshape(-1, nw)
    hc = np.ascontiguousarray(high_counts, dtype=np.uint32)
    lk = np.ascontiguousarray(low_keys, dtype=np.uint64).reshape(-1, nw)
    bins = np.zeros(max_freq - min_freq + 1, dtype=np.int64)
    lib().kco_histogram(_p(hk), _p(hc), hk.shape[0], _p(lk), lk.shape[0], nw, count_max,
                        min_freq, max_freq, _p(bins))
    return bins


def dump_text(keys, counts, k) -> bytes:
    nw = words_for_k(k)
    keys = np.ascontiguousarray(keys, dtype=np.uint64).reshape(-1, nw)
    counts = np.ascontiguousarray(counts, dtype=np.uint32)
    n = keys.shape[0]
    size = lib().kco_dump_text(_p(keys), _p(counts), n, k, None)
    out = np.empty(max(1, size), dtype=np.uint8)
    lib().kco_dump_text(_p(keys), _p(counts), n, k, _p(out))
    return out[:size].tobytes()


def histogram_text(bins, min_freq, max_freq) -> bytes:
    bins = np.ascontiguousarray(bins, dtype=np.int64)
    size = lib().kco_histogram_text(_p(bins), min_freq, max_freq, None)
    out = np.empty(max(1, size), dtype=np.uint8)
    lib().kco_histogram_text(_p(bins), min_freq, max_freq, _p(out))
    return out[:size].tobytes()


def sort_records8(keys, counts, k):
    nw = words_for_k(k)
    keys = np.ascontiguousarray(keys, dtype=np.uint64).reshape(-1, nw).copy()
    counts = np.ascontiguousarray(counts, dtype=np.uint32).copy()
    lib().kco_sort_records8(_p(keys), _p(counts), keys.shape[0], nw, k)
    return keys, counts


def sort_records12(keys, counts, k, cpb=10):
    nw = words_for_k(k)
    keys = np.ascontiguousarray(keys, dtype=np.uint64).reshape(-1, nw).copy()
    counts = np.ascontiguousarray(counts, dtype=np.uint32).copy()
    rc = lib().kco_sort_records12(_p(keys), _p(counts), keys.shape[0], nw, k, cpb)
    if rc != 0:
        raise ValueError("sort_records12 rejected arguments (rc=%d)" % rc)
    return keys, counts


def _kb(words):
    a = np.ascontiguousarray(words, dtype=np.uint64)
    return a, a.nbytes


def xxh3_64(words):
    a, n = _kb(words)
    return lib().kco_xxh3_64(_p(a), n)


def xxh3_128(words):
    a, n = _kb(words)
    lo, hi = C.c_uint64(), C.c_uint64()
    lib().kco_xxh3_128(_p(a), n, C.byref(lo), C.byref(hi))
    return lo.value, hi.value


def rapidhash(words):
    a, n = _kb(words)
    return lib().kco_rapidhash(_p(a), n)


def komihash(words, seed=0):
    a, n = _kb(words)
    return lib().kco_komihash(_p(a), n, seed)


def hash_func(words):
    a, _ = _kb(words)
    return lib().kco_hash_func(_p(a), a.size)


def concurrent_map_hash(words):
    a, _ = _kb(words)
    return lib().kco_concurrent_map_hash(_p(a), a.size)


def bloom_probe(words, capacity):
    a, _ = _kb(words)
    b, m = C.c_uint64(), C.c_uint64()
    lib().kco_bloom_probe(_p(a), a.size, capacity, C.byref(b), C.byref(m))
    return b.value, m.value


def count_fastq_views(fq: bytes, k: int, count_max: int, hist_min=1, hist_max=10000):
    """Oracle merged views of a FASTQ input: (dump --sort text, histogram text)."""
    seq = fastq_frame(fq)
    nw = words_for_k(k)
    keys = extract(seq, k)
    uniq, cnt = count(keys, nw)
    clamped = np.minimum(cnt, count_max).astype(np.uint32)
    dump = dump_text(uniq, clamped, k)
    bins = np.zeros(hist_max - hist_min + 1, dtype=np.int64)
    vals, freq = np.unique(clamped, return_counts=True)
    for v, f in zip(vals, freq):
        if hist_min <= v <= hist_max:
            bins[v - hist_min] += f
    return dump, histogram_text(bins, hist_min, hist_max), keys.shape[0]
