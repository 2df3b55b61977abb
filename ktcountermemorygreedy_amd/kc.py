"""Python binding of libkc.so (ctypes over the C ABI in include/kc_api.h).

Device buffers are DeviceBuffer objects: HBM allocated through the library's
own HIP runtime (kc_dev_alloc), so that only one HIP runtime is ever loaded
in the process.  There is no CPU fallback: using this module on a machine
without the built library raises, and every call checks the status code.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libkc.so")

KC_OK, KC_EINVAL, KC_ENOMEM, KC_EFULL, KC_EDEVICE, KC_EIO, KC_ENOTSUP = 0, -1, -2, -3, -4, -5, -6
HASH_XXH3_64, HASH_XXH3_128_LO, HASH_XXH3_128_HI, HASH_RAPIDHASH, HASH_KOMIHASH, \
    HASH_CONCURRENT_MAP, HASH_FUNC = range(7)


class KCError(RuntimeError):
    def __init__(self, status, what):
        super().__init__("%s failed: %s (%d)" % (what, lib().kc_strerror(status).decode(), status))
        self.status = status


class TableConfig(C.Structure):
    _fields_ = [
        ("k", C.c_uint32), ("count_max", C.c_uint32), ("filter_min", C.c_uint32),
        ("filter_max", C.c_uint32), ("capacity", C.c_uint64), ("max_distinct", C.c_uint64),
        ("bloom", C.c_uint32), ("count_path", C.c_uint32), ("bloom_blocks", C.c_uint64),
        ("low_capacity", C.c_uint64),
    ]


class Stats(C.Structure):
    _fields_ = [("kmers", C.c_uint64), ("distinct", C.c_uint64), ("low", C.c_uint64),
                ("capacity", C.c_uint64), ("errors", C.c_uint64)]


class GenomeSpec(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("length", C.c_uint64), ("n_segments", C.c_uint32),
                ("segment_len", C.c_uint32), ("repeat_fraction", C.c_double),
                ("segment_divergence", C.c_double)]


class ReadSpec(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("n_reads", C.c_uint64), ("read_len", C.c_uint32),
                ("both_strands", C.c_uint32), ("sub_rate", C.c_double), ("n_rate", C.c_double),
                ("long_fraction", C.c_double), ("long_mu", C.c_double), ("long_sigma", C.c_double),
                ("long_sub_rate", C.c_double), ("min_len", C.c_uint32), ("threads", C.c_uint32)]


class Tune(C.Structure):
    _fields_ = [("front_blocks", C.c_uint32), ("fine_slices", C.c_uint32), ("batch_bytes", C.c_uint64),
                ("queue_depth", C.c_uint32), ("path", C.c_uint32), ("piece_bytes", C.c_uint64),
                ("first_mode", C.c_uint32), ("front_mode", C.c_uint32)]


class BatchInfo(C.Structure):
    _fields_ = [("path", C.c_uint32), ("queue_depth", C.c_uint32), ("capacity", C.c_uint64),
                ("front_blocks", C.c_uint32), ("fine_slices", C.c_uint32), ("launches", C.c_uint64),
                ("graphs_captured", C.c_uint64), ("chunks", C.c_uint64), ("pieces", C.c_uint64),
                ("bytes", C.c_uint64)]


class IngestStats(C.Structure):
    _fields_ = [("files", C.c_uint64), ("threads", C.c_uint64), ("input_bytes", C.c_uint64),
                ("framed_bytes", C.c_uint64), ("chunks", C.c_uint64), ("launches", C.c_uint64),
                ("inflate_seconds", C.c_double), ("wall_seconds", C.c_double)]


class PassPlan(C.Structure):
    _fields_ = [("n_passes", C.c_uint32), ("export_dir", C.c_char_p), ("dump_path", C.c_char_p),
                ("dump_min", C.c_uint32), ("dump_max", C.c_uint32), ("dump_sorted", C.c_int),
                ("hist_min", C.c_uint32), ("hist_max", C.c_uint32), ("hist_bins", C.c_void_p)]


class Plan(C.Structure):
    _fields_ = [("kmers", C.c_uint64), ("distinct", C.c_uint64), ("repeated", C.c_uint64),
                ("device_bytes", C.c_uint64), ("passes", C.c_uint32), ("sample_slices", C.c_uint32)]


class RequestResult(C.Structure):
    _fields_ = [("status", C.c_int), ("batched", C.c_int), ("latency_ms", C.c_double),
                ("queue_ms", C.c_double), ("kmers", C.c_uint64), ("distinct", C.c_uint64)]


_lib = None

# every symbol include/kc_api.h and include/kc_host.h declare
EXPORTS = {
    "kc_strerror": (C.c_char_p, [C.c_int]),
    "kc_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "kc_ctx_create_leaving_cus": (C.c_int, [C.c_int, C.c_uint32, C.POINTER(C.c_void_p)]),
    "kc_ctx_destroy": (None, [C.c_void_p]),
    "kc_ctx_sync": (C.c_int, [C.c_void_p]),
    "kc_ctx_stream": (C.c_void_p, [C.c_void_p]),
    "kc_dev_alloc": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    "kc_dev_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "kc_host_alloc_pinned": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    "kc_host_free_pinned": (C.c_int, [C.c_void_p, C.c_void_p]),
    "kc_host_device_ptr": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "kc_copy_h2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    "kc_copy_d2h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    "kc_copy_d2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    "kc_event_record": (C.c_int, [C.c_void_p, C.c_int]),
    "kc_event_elapsed_ms": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_float)]),
    "kc_device_mem_info": (C.c_int, [C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "kc_table_create": (C.c_int, [C.c_void_p, C.POINTER(TableConfig), C.POINTER(C.c_void_p)]),
    "kc_table_destroy": (None, [C.c_void_p]),
    "kc_table_clear": (C.c_int, [C.c_void_p]),
    "kc_table_trim": (C.c_int, [C.c_void_p]),
    "kc_table_count_seq": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "kc_table_count_seq_host": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "kc_table_increment": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "kc_table_query": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "kc_table_query_on": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "kc_table_stats": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    "kc_table_front_redos": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "kc_table_set_pass": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32]),
    "kc_table_export_dir": (C.c_int, [C.c_void_p, C.c_char_p]),
    "kc_table_export_dir_append": (C.c_int, [C.c_void_p, C.c_char_p]),
    "kc_table_export_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64),
                                         C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "kc_table_histogram": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]),
    "kc_table_histogram_noalloc": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]),
    "kc_table_dump_device": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_int, C.c_void_p, C.c_uint64,
                                       C.POINTER(C.c_uint64)]),
    "kc_table_dump_file": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_int, C.c_char_p]),
    "kc_export_low": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p]),
    "kc_export_high": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32,
                                 C.c_uint32, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "kc_table_load_dir": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.POINTER(C.c_void_p)]),
    "kc_histogram_dir": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.c_uint32, C.c_uint32, C.c_void_p]),
    "kc_dump_dir": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.c_uint32, C.c_uint32, C.c_int, C.c_char_p]),
    "kc_plan_table": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint64, C.POINTER(TableConfig),
                                C.POINTER(Plan)]),
    "kc_plan_sizes": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(TableConfig), C.POINTER(Plan)]),
    "kc_table_set_tune": (C.c_int, [C.c_void_p, C.POINTER(Tune)]),
    "kc_table_get_tune": (C.c_int, [C.c_void_p, C.POINTER(Tune)]),
    "kc_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.POINTER(Tune),
                              C.POINTER(C.c_float)]),
    "kc_batch_create": (C.c_int, [C.c_void_p, C.POINTER(Tune), C.POINTER(C.c_void_p)]),
    "kc_batch_destroy": (None, [C.c_void_p]),
    "kc_batch_add": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "kc_batch_add_chunks": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    "kc_batch_flush": (C.c_int, [C.c_void_p]),
    "kc_batch_sync": (C.c_int, [C.c_void_p]),
    "kc_batch_launch_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "kc_batch_get_info": (C.c_int, [C.c_void_p, C.POINTER(BatchInfo)]),
    "kc_sort_records": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32,
                                  C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_int)]),
    "kc_records_histogram": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32,
                                       C.c_uint32, C.c_uint32, C.c_void_p]),
    "kc_synth_records": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint64, C.c_double]),
    "kc_hash_keys": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p]),
    "kc_bloom_probe_keys": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint64,
                                      C.c_void_p, C.c_void_p]),
    "kc_server_create": (C.c_int, [C.c_int, C.c_uint32, C.c_uint64, C.POINTER(C.c_void_p)]),
    "kc_server_destroy": (None, [C.c_void_p]),
    "kc_server_submit_count": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p, C.c_uint64, C.c_uint64,
                                         C.c_int, C.POINTER(C.c_uint64)]),
    "kc_server_submit_query": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p,
                                         C.POINTER(C.c_uint64)]),
    "kc_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(RequestResult)]),
    "kc_server_drop_table": (C.c_int, [C.c_void_p, C.c_uint64]),
    "kc_server_table_histogram": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_void_p]),
    "kc_server_resident": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "kc_fastq_frame": (C.c_uint64, [C.c_void_p, C.c_uint64, C.c_void_p]),
    "kc_fastq_frame_file": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p), C.POINTER(C.c_uint64)]),
    "kc_fastq_frame_files": (C.c_int, [C.POINTER(C.c_char_p), C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p),
                                       C.POINTER(C.c_uint64), C.c_void_p]),
    "kc_host_free": (None, [C.c_void_p]),
    "kc_count_fastq_files": (C.c_int, [C.c_void_p, C.POINTER(C.c_char_p), C.c_uint32, C.c_uint32,
                                       C.POINTER(Tune), C.POINTER(IngestStats)]),
    "kc_count_fastq_files_passes": (C.c_int, [C.c_void_p, C.POINTER(C.c_char_p), C.c_uint32, C.c_uint32,
                                              C.POINTER(Tune), C.POINTER(PassPlan), C.POINTER(IngestStats)]),
    "kc_merge_sorted_dumps": (C.c_int, [C.POINTER(C.c_char_p), C.c_uint32, C.c_char_p]),
    "kc_synth_genome": (C.c_int, [C.POINTER(GenomeSpec), C.c_void_p]),
    "kc_synth_reads_size": (C.c_int64, [C.c_void_p, C.c_uint64, C.POINTER(ReadSpec), C.c_int]),
    "kc_synth_reads": (C.c_int64, [C.c_void_p, C.c_uint64, C.POINTER(ReadSpec), C.c_int,
                                   C.c_void_p, C.c_uint64]),
}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libkc.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status, what):
    if status != KC_OK:
        raise KCError(status, what)


def words_for_k(k):
    """Words of a k-mer key (kmer<N>, main.cpp:617-633)."""
    return 1 if k <= 32 else 2 if k <= 64 else 4


def _ptr(t):
    if t is None:
        return None
    if isinstance(t, np.ndarray):
        if not t.flags["C_CONTIGUOUS"]:
            raise ValueError("expected a contiguous array")
        return t.ctypes.data_as(C.c_void_p)
    if isinstance(t, DeviceBuffer):
        return C.c_void_p(t.ptr)
    raise TypeError("expected a numpy array (host) or a DeviceBuffer (device)")


class DeviceBuffer:
    """A block of HBM owned by a Context (kc_dev_alloc / kc_dev_free)."""

    def __init__(self, ctx, nbytes, ptr=None, owner=None):
        self.ctx, self.nbytes = ctx, int(nbytes)
        self._owner = owner
        if ptr is None:
            p = C.c_void_p()
            check(lib().kc_dev_alloc(ctx.h, max(1, self.nbytes), C.byref(p)), "kc_dev_alloc")
            self.ptr = p.value
            self._own = True
            ctx._buffers.add(self)
        else:
            self.ptr = ptr
            self._own = False

    @classmethod
    def from_numpy(cls, ctx, arr, pad=16):
        arr = np.ascontiguousarray(arr)
        b = cls(ctx, arr.nbytes + pad)
        if arr.nbytes:
            check(lib().kc_copy_h2d(ctx.h, C.c_void_p(b.ptr), arr.ctypes.data_as(C.c_void_p), arr.nbytes),
                  "kc_copy_h2d")
        b.valid = arr.nbytes
        return b

    def upload(self, arr, offset=0):
        arr = np.ascontiguousarray(arr)
        if offset + arr.nbytes > self.nbytes:
            raise ValueError("upload exceeds the buffer")
        check(lib().kc_copy_h2d(self.ctx.h, C.c_void_p(self.ptr + offset), arr.ctypes.data_as(C.c_void_p),
                                arr.nbytes), "kc_copy_h2d")

    def numpy(self, dtype=np.uint8, count=None, offset=0):
        dtype = np.dtype(dtype)
        if count is None:
            count = (self.nbytes - offset) // dtype.itemsize
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            check(lib().kc_copy_d2h(self.ctx.h, out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr + offset),
                                    out.nbytes), "kc_copy_d2h")
        return out

    def view(self, offset, nbytes):
        return DeviceBuffer(self.ctx, nbytes, ptr=self.ptr + offset, owner=self)

    def free(self):
        if self._own and self.ptr:
            if self.ctx.h:  # a closed context has freed its buffers
                lib().kc_dev_free(self.ctx.h, C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedBuffer:
    """Page-locked host memory (kc_host_alloc_pinned); .array is a uint8 view."""

    def __init__(self, ctx, nbytes):
        self.ctx, self.nbytes = ctx, int(nbytes)
        p = C.c_void_p()
        check(lib().kc_host_alloc_pinned(ctx.h, max(1, self.nbytes), C.byref(p)), "kc_host_alloc_pinned")
        self.ptr = p.value
        self.array = np.ctypeslib.as_array((C.c_uint8 * max(1, self.nbytes)).from_address(self.ptr))
        ctx._buffers.add(self)

    def free(self):
        if self.ptr and self.ctx.h:  # a closed context has freed its buffers
            self.array = None
            lib().kc_host_free_pinned(self.ctx.h, C.c_void_p(self.ptr))
        self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class _Native:
    """A native handle and the handles that must be destroyed before it (a
    table's batches, a context's tables).  Parents hold these records
    strongly and the Python objects weakly-by-id, so a parent destroys its
    children first even when the garbage collector finalizes a cycle of
    them in any order (weak references are cleared before finalizers run);
    a destroyed record is a no-op afterwards."""

    def __init__(self, h, destroy, parent=None):
        self.h, self._destroy, self.children, self._parent = h, destroy, {}, parent
        if parent is not None:
            parent.children[id(self)] = self

    def destroy(self):
        if self.h:
            for c in list(self.children.values()):
                c.destroy()
            self.children.clear()
            self._destroy(self.h)
            self.h = None
        if self._parent is not None:
            self._parent.children.pop(id(self), None)
            self._parent = None


class Context:
    """kc_ctx: one device and one HIP stream.  left_cus > 0 keeps the
    stream's kernels off compute units 0 .. left_cus-1 (a CU mask), leaving
    them to other streams."""

    def __init__(self, device=0, left_cus=0):
        h = C.c_void_p()
        if left_cus:
            check(lib().kc_ctx_create_leaving_cus(device, left_cus, C.byref(h)), "kc_ctx_create_leaving_cus")
        else:
            check(lib().kc_ctx_create(device, C.byref(h)), "kc_ctx_create")
        self.h = h
        self.device = device
        self._tables = _Native(h, lambda _: None)  # tables (and their batches) destroyed before the context
        self._buffers = weakref.WeakSet()  # device and pinned buffers it allocated, freed on close

    def sync(self):
        check(lib().kc_ctx_sync(self.h), "kc_ctx_sync")

    @property
    def stream_ptr(self):
        return lib().kc_ctx_stream(self.h)

    def close(self):
        if self.h:
            self._tables.destroy()
            for b in list(self._buffers):
                b.free()
            lib().kc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def upload(self, arr, pad=16):
        return DeviceBuffer.from_numpy(self, arr, pad)

    def pinned(self, nbytes):
        return PinnedBuffer(self, nbytes)

    def record(self, slot):
        check(lib().kc_event_record(self.h, slot), "kc_event_record")

    def elapsed_ms(self, a, b):
        ms = C.c_float()
        check(lib().kc_event_elapsed_ms(self.h, a, b, C.byref(ms)), "kc_event_elapsed_ms")
        return ms.value

    # -- stateless device ops (keys: uint64 numpy array, n x nw) --
    def hash_keys(self, kind, keys: np.ndarray, nw):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        n = keys.size // nw
        dk = self.upload(keys)
        out = self.alloc(8 * n)
        check(lib().kc_hash_keys(self.h, kind, _ptr(dk), n, nw, _ptr(out)), "kc_hash_keys")
        return out.numpy(np.uint64, n)

    def bloom_probe(self, keys: np.ndarray, nw, capacity):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        n = keys.size // nw
        dk = self.upload(keys)
        blk, msk = self.alloc(8 * n), self.alloc(8 * n)
        check(lib().kc_bloom_probe_keys(self.h, _ptr(dk), n, nw, capacity, _ptr(blk), _ptr(msk)),
              "kc_bloom_probe_keys")
        return blk.numpy(np.uint64, n), msk.numpy(np.uint64, n)

    def sort_records(self, recs, scratch, n, nw, k, cpb, digit_bits, record_bytes):
        """Stable sort of device records; returns True when the result is in scratch."""
        flag = C.c_int(0)
        check(lib().kc_sort_records(self.h, _ptr(recs), _ptr(scratch), n, nw, k, cpb, digit_bits,
                                    record_bytes, C.byref(flag)), "kc_sort_records")
        return bool(flag.value)

    def records_histogram(self, recs, n, nw, record_bytes, min_freq=1, max_freq=10000):
        bins = np.zeros(max_freq - min_freq + 1, dtype=np.int64)
        check(lib().kc_records_histogram(self.h, _ptr(recs), n, nw, record_bytes, min_freq, max_freq,
                                         _ptr(bins)), "kc_records_histogram")
        return bins

    def export_low(self, keys: np.ndarray, k) -> bytes:
        """low.bin records of host keys (n x nw uint64), packed on the device."""
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        n = keys.size // words_for_k(k)
        kb = (k // 32) * 8 + (2 * (k % 32) + 7) // 8
        dk, out = self.upload(keys), self.alloc(max(16, n * kb))
        check(lib().kc_export_low(self.h, _ptr(dk), n, k, _ptr(out)), "kc_export_low")
        self.sync()
        return out.numpy(np.uint8, n * kb).tobytes()

    def export_high(self, keys: np.ndarray, counts: np.ndarray, k, count_max=0xFFFFFFFF, filter_min=0,
                    filter_max=0xFFFFFFFF) -> bytes:
        """high.bin records of host (key, count) pairs, filtered as the counter does."""
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        counts = np.ascontiguousarray(counts, dtype=np.uint32)
        n = counts.size
        dk, dc = self.upload(keys), self.upload(counts)
        size = C.c_uint64()
        check(lib().kc_export_high(self.h, _ptr(dk), _ptr(dc), n, k, count_max, filter_min, filter_max, None, 0,
                                   C.byref(size)), "kc_export_high(size)")
        out = self.alloc(max(16, size.value))
        check(lib().kc_export_high(self.h, _ptr(dk), _ptr(dc), n, k, count_max, filter_min, filter_max, _ptr(out),
                                   out.nbytes, C.byref(size)), "kc_export_high")
        return out.numpy(np.uint8, size.value).tobytes()

    def histogram_dir(self, path, min_freq=1, max_freq=10000, precise=True):
        """histogram_tool over a temp_dir, on the device."""
        bins = np.zeros(max_freq - min_freq + 1, dtype=np.int64)
        check(lib().kc_histogram_dir(self.h, os.fsencode(path), 1 if precise else 0, min_freq, max_freq, _ptr(bins)),
              "kc_histogram_dir")
        return bins

    def dump_dir(self, path, out_path, min_freq=1, max_freq=0xFFFFFFFF, sorted=True, precise=True):
        """dump_tool over a temp_dir into out_path, on the device."""
        check(lib().kc_dump_dir(self.h, os.fsencode(path), 1 if precise else 0, min_freq, max_freq,
                                1 if sorted else 0, os.fsencode(out_path)), "kc_dump_dir")

    def synth_records(self, recs, n, record_bytes=12, seed=4, zipf_s=1.2):
        check(lib().kc_synth_records(self.h, _ptr(recs), n, record_bytes, seed, zipf_s), "kc_synth_records")


class Table:
    """kc_table: a device-resident counting table (ConcurrentCountingHashMap<N>)."""

    def __init__(self, ctx: Context, k, count_max=0xFFFFFFFF, capacity=0, max_distinct=0,
                 bloom=False, bloom_blocks=0, low_capacity=0, filter_min=0, filter_max=0xFFFFFFFF,
                 path="auto", _handle=None):
        if _handle is None:
            paths = {"auto": 0, "direct": 1, "partitioned": 2}
            cfg = TableConfig(k=k, count_max=count_max, filter_min=filter_min, filter_max=filter_max,
                              capacity=capacity, max_distinct=max_distinct, bloom=1 if bloom else 0,
                              count_path=paths[path], bloom_blocks=bloom_blocks, low_capacity=low_capacity)
            h = C.c_void_p()
            check(lib().kc_table_create(ctx.h, C.byref(cfg), C.byref(h)), "kc_table_create")
        else:
            h = _handle
        self._native = _Native(h, lambda x: lib().kc_table_destroy(x), ctx._tables)
        self.ctx, self.k, self.nw, self.count_max = ctx, k, words_for_k(k), count_max
        self.bloom = bool(bloom)

    @property
    def h(self):
        return self._native.h

    @classmethod
    def load_dir(cls, ctx: Context, path, with_low=True):
        """A temp_dir (infos.bin, high.bin, low.bin) loaded into a table whose
        histogram / dump / lookups are the tools' merges of the files."""
        h = C.c_void_p()
        check(lib().kc_table_load_dir(ctx.h, os.fsencode(path), 1 if with_low else 0, C.byref(h)),
              "kc_table_load_dir")
        with open(os.path.join(path, "infos.bin"), "rb") as f:
            k, count_max = np.frombuffer(f.read(8), dtype=np.uint32)
        return cls(ctx, int(k), int(count_max), bloom=with_low, _handle=h)

    def close(self):
        self._native.destroy()  # its batches first

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clear(self):
        check(lib().kc_table_clear(self.h), "kc_table_clear")

    def trim(self):
        """Frees the counting scratch kept for reuse (kc_table_trim)."""
        check(lib().kc_table_trim(self.h), "kc_table_trim")

    def count_seq(self, d_seq, nbytes=None):
        n = getattr(d_seq, "valid", d_seq.nbytes) if nbytes is None else nbytes
        check(lib().kc_table_count_seq(self.h, _ptr(d_seq), n), "kc_table_count_seq")

    def count_seq_host(self, seq: bytes):
        buf = np.frombuffer(seq, dtype=np.uint8)
        check(lib().kc_table_count_seq_host(self.h, _ptr(buf), len(seq)), "kc_table_count_seq_host")
        self.ctx.sync()

    def increment(self, keys: np.ndarray, inc: np.ndarray = None):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        n = keys.size // self.nw
        dk = self.ctx.upload(keys)
        di = self.ctx.upload(np.ascontiguousarray(inc, dtype=np.uint32)) if inc is not None else None
        check(lib().kc_table_increment(self.h, _ptr(dk), n, _ptr(di)), "kc_table_increment")
        self.ctx.sync()

    def query(self, keys: np.ndarray, ctx: "Context" = None) -> np.ndarray:
        """Point lookups; with ctx (another context on the same device) the
        lookup runs on that context's stream (kc_table_query_on)."""
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        n = keys.size // self.nw
        c = ctx or self.ctx
        dk = c.upload(keys)
        out = c.alloc(4 * n)
        if ctx is None:
            check(lib().kc_table_query(self.h, _ptr(dk), n, _ptr(out)), "kc_table_query")
        else:
            check(lib().kc_table_query_on(self.h, ctx.h, _ptr(dk), n, _ptr(out)), "kc_table_query_on")
        return out.numpy(np.uint32, n)

    def stats(self):
        s = Stats()
        check(lib().kc_table_stats(self.h, C.byref(s)), "kc_table_stats")
        return {f: getattr(s, f) for f, _ in Stats._fields_}

    def front_redos(self):
        """Partitioned passes since the last clear whose speculative coarse
        regions overflowed and were redone exactly."""
        v = C.c_uint64()
        check(lib().kc_table_front_redos(self.h, C.byref(v)), "kc_table_front_redos")
        return v.value

    def export_dir(self, path, append=False):
        if append:
            check(lib().kc_table_export_dir_append(self.h, path.encode()), "kc_table_export_dir_append")
        else:
            check(lib().kc_table_export_dir(self.h, path.encode()), "kc_table_export_dir")

    def set_pass(self, pass_id, n_passes):
        """Later counting calls count only the k-mers of hash slice pass_id of
        n_passes (a power of two <= 256); set_pass(0, 1) restores one pass."""
        check(lib().kc_table_set_pass(self.h, pass_id, n_passes), "kc_table_set_pass")

    def export_sizes(self):
        hb, lb = C.c_uint64(), C.c_uint64()
        check(lib().kc_table_export_device(self.h, None, 0, C.byref(hb), None, 0, C.byref(lb)),
              "kc_table_export_device(size)")
        return hb.value, lb.value

    def export_into(self, dh, dl):
        hb, lb = C.c_uint64(), C.c_uint64()
        check(lib().kc_table_export_device(self.h, _ptr(dh), dh.nbytes, C.byref(hb), _ptr(dl), dl.nbytes,
                                           C.byref(lb)), "kc_table_export_device")
        return hb.value, lb.value

    def export_device(self):
        """Packed (high, low) export records as host bytes."""
        hs, ls = self.export_sizes()
        dh, dl = self.ctx.alloc(max(hs, 16)), self.ctx.alloc(max(ls, 16))
        hb, lb = self.export_into(dh, dl)
        return dh.numpy(np.uint8, hb).tobytes(), dl.numpy(np.uint8, lb).tobytes()

    def dump(self, min_freq=1, max_freq=0xFFFFFFFF, sorted=True) -> bytes:
        """dump_tool precise text built on the device."""
        n = C.c_uint64()
        check(lib().kc_table_dump_device(self.h, min_freq, max_freq, 1 if sorted else 0, None, 0, C.byref(n)),
              "kc_table_dump_device(size)")
        buf = self.ctx.alloc(max(n.value, 16))
        check(lib().kc_table_dump_device(self.h, min_freq, max_freq, 1 if sorted else 0, _ptr(buf), buf.nbytes,
                                         C.byref(n)), "kc_table_dump_device")
        return buf.numpy(np.uint8, n.value).tobytes()

    def dump_file(self, path, min_freq=1, max_freq=0xFFFFFFFF, sorted=True):
        check(lib().kc_table_dump_file(self.h, min_freq, max_freq, 1 if sorted else 0, path.encode()),
              "kc_table_dump_file")

    def histogram(self, min_freq=1, max_freq=10000):
        bins = np.zeros(max_freq - min_freq + 1, dtype=np.int64)
        check(lib().kc_table_histogram(self.h, min_freq, max_freq, _ptr(bins)), "kc_table_histogram")
        return bins

    def histogram_noalloc(self, min_freq=1, max_freq=10000):
        """kc_table_histogram_noalloc: the same bins, no device allocation."""
        bins = np.zeros(max_freq - min_freq + 1, dtype=np.int64)
        check(lib().kc_table_histogram_noalloc(self.h, min_freq, max_freq, _ptr(bins)),
              "kc_table_histogram_noalloc")
        return bins

    def set_tune(self, **kw):
        t = Tune(**{k: v for k, v in kw.items() if v is not None})
        check(lib().kc_table_set_tune(self.h, C.byref(t)), "kc_table_set_tune")

    def tune(self):
        t = Tune()
        check(lib().kc_table_get_tune(self.h, C.byref(t)), "kc_table_get_tune")
        return {f: getattr(t, f) for f, _ in Tune._fields_}

    def autotune(self, d_sample, nbytes=None, reps=2):
        """Launch-shape search over a device-resident sample (leaves the
        table empty); returns (tune dict, best ms)."""
        n = getattr(d_sample, "valid", d_sample.nbytes) if nbytes is None else nbytes
        t, ms = Tune(), C.c_float()
        check(lib().kc_autotune(self.h, _ptr(d_sample), n, reps, C.byref(t), C.byref(ms)), "kc_autotune")
        return {f: getattr(t, f) for f, _ in Tune._fields_}, ms.value


class Batch:
    """kc_batch: framed read chunks packed into staging slots, each slot
    counted into `table` by one captured-graph launch."""

    def __init__(self, table: Table, batch_bytes=0, queue_depth=0, path=None, front_blocks=0, fine_slices=0):
        paths = {None: None, "auto": 0, "direct": 1, "partitioned": 2}
        h = C.c_void_p()
        if batch_bytes or queue_depth or path is not None or front_blocks or fine_slices:
            base = table.tune()
            t = Tune(front_blocks=front_blocks or base["front_blocks"], fine_slices=fine_slices or base["fine_slices"],
                     batch_bytes=batch_bytes or base["batch_bytes"], queue_depth=queue_depth or base["queue_depth"],
                     path=base["path"] if path is None else paths[path], piece_bytes=base["piece_bytes"],
                     first_mode=base["first_mode"], front_mode=base["front_mode"])
            check(lib().kc_batch_create(table.h, C.byref(t), C.byref(h)), "kc_batch_create")
        else:
            check(lib().kc_batch_create(table.h, None, C.byref(h)), "kc_batch_create")
        self._native = _Native(h, lambda x: lib().kc_batch_destroy(x), table._native)
        self.table = table

    @property
    def h(self):
        return self._native.h

    def add(self, chunk):
        buf = np.frombuffer(chunk, dtype=np.uint8) if isinstance(chunk, (bytes, bytearray)) else \
            np.ascontiguousarray(chunk, dtype=np.uint8)
        check(lib().kc_batch_add(self.h, _ptr(buf), buf.size), "kc_batch_add")

    def add_chunks(self, buf, ends):
        """Chunks buf[ends[i-1]:ends[i]] of one host buffer in one call
        (the batches of add() per chunk)."""
        buf = np.frombuffer(buf, dtype=np.uint8) if isinstance(buf, (bytes, bytearray)) else \
            np.ascontiguousarray(buf, dtype=np.uint8)
        e = np.ascontiguousarray(ends, dtype=np.uint64)
        if e.size and int(e[-1]) > buf.size:
            raise ValueError("chunk ends past the buffer")
        check(lib().kc_batch_add_chunks(self.h, _ptr(buf), _ptr(e), e.size), "kc_batch_add_chunks")

    def flush(self):
        check(lib().kc_batch_flush(self.h), "kc_batch_flush")

    def sync(self):
        check(lib().kc_batch_sync(self.h), "kc_batch_sync")

    def launch_device(self, d_seq, nbytes=None):
        n = getattr(d_seq, "valid", d_seq.nbytes) if nbytes is None else nbytes
        check(lib().kc_batch_launch_device(self.h, _ptr(d_seq), n), "kc_batch_launch_device")

    def info(self):
        i = BatchInfo()
        check(lib().kc_batch_get_info(self.h, C.byref(i)), "kc_batch_get_info")
        return {f: getattr(i, f) for f, _ in BatchInfo._fields_}

    def close(self):
        self._native.destroy()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- host helpers ----------------------------------------------------------

def device_mem_info(device=0):
    """(free, total) HBM bytes of a device as the HIP runtime reports them."""
    f, t = C.c_uint64(), C.c_uint64()
    check(lib().kc_device_mem_info(device, C.byref(f), C.byref(t)), "kc_device_mem_info")
    return f.value, t.value


def fastq_frame(fq: bytes) -> bytes:
    src = np.frombuffer(fq, dtype=np.uint8)
    out = np.empty(max(1, len(fq)), dtype=np.uint8)
    n = lib().kc_fastq_frame(_ptr(src), len(fq), _ptr(out))
    return out[:n].tobytes()


def fastq_frame_file(path) -> np.ndarray:
    p, n = C.c_void_p(), C.c_uint64()
    check(lib().kc_fastq_frame_file(path.encode(), C.byref(p), C.byref(n)), "kc_fastq_frame_file")
    try:
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), shape=(n.value,)).copy() \
            if n.value else np.zeros(0, np.uint8)
    finally:
        lib().kc_host_free(p)


def fastq_frame_files(paths, threads=4):
    """The ingest reader without a table (kc_fastq_frame_files): the framed
    reads of the files and the ingest statistics."""
    arr = (C.c_char_p * max(1, len(paths)))(*[os.fsencode(p) for p in paths])
    p, n, st = C.c_void_p(), C.c_uint64(), IngestStats()
    check(lib().kc_fastq_frame_files(arr, len(paths), threads, C.byref(p), C.byref(n), C.byref(st)),
          "kc_fastq_frame_files")
    try:
        data = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), shape=(n.value,)).copy() \
            if n.value else np.zeros(0, np.uint8)
    finally:
        lib().kc_host_free(p)
    return data, {f: getattr(st, f) for f, _ in IngestStats._fields_}


def count_fastq_files(table: "Table", paths, threads=4, batch_bytes=0, queue_depth=0):
    """Streaming ingest of FASTQ(.gz) files into table (host threads inflate
    and frame, a kc_batch counts); returns the ingest statistics."""
    arr = (C.c_char_p * max(1, len(paths)))(*[os.fsencode(p) for p in paths])
    st = IngestStats()
    tune = None
    if batch_bytes or queue_depth:
        base = table.tune()
        tune = Tune(front_blocks=base["front_blocks"], fine_slices=base["fine_slices"],
                    batch_bytes=batch_bytes or base["batch_bytes"], queue_depth=queue_depth or base["queue_depth"],
                    path=base["path"], piece_bytes=base["piece_bytes"], first_mode=base["first_mode"],
                    front_mode=base["front_mode"])
    check(lib().kc_count_fastq_files(table.h, arr, len(paths), threads, C.byref(tune) if tune else None,
                                     C.byref(st)), "kc_count_fastq_files")
    return {f: getattr(st, f) for f, _ in IngestStats._fields_}


def count_fastq_files_passes(table: "Table", paths, n_passes, threads=4, export_dir=None, dump_path=None,
                             dump_range=(1, 0xFFFFFFFF), dump_sorted=True, hist_range=None):
    """Out-of-HBM counting: n_passes passes over the files, pass p counting
    hash slice p in the cleared table, exports appended to export_dir, dumps
    merged into dump_path, histograms summed.  Returns (bins or None, stats)."""
    arr = (C.c_char_p * max(1, len(paths)))(*[os.fsencode(p) for p in paths])
    bins = None
    plan = PassPlan(n_passes=n_passes, export_dir=os.fsencode(export_dir) if export_dir else None,
                    dump_path=os.fsencode(dump_path) if dump_path else None,
                    dump_min=dump_range[0], dump_max=dump_range[1], dump_sorted=int(bool(dump_sorted)))
    if hist_range is not None:
        bins = np.zeros(hist_range[1] - hist_range[0] + 1, dtype=np.int64)
        plan.hist_min, plan.hist_max, plan.hist_bins = hist_range[0], hist_range[1], bins.ctypes.data
    st = IngestStats()
    check(lib().kc_count_fastq_files_passes(table.h, arr, len(paths), threads, None, C.byref(plan), C.byref(st)),
          "kc_count_fastq_files_passes")
    return bins, {f: getattr(st, f) for f, _ in IngestStats._fields_}


def plan_table(ctx: "Context", d_seq, k, nbytes=None, bloom=False, count_max=0xFFFFFFFF, sample_slices=16,
               hbm_budget=0):
    """Table sizing from one exactly counted hash slice of the input on the
    device (kc_plan_table).  Returns (table kwargs for Table(ctx, k, **kw),
    plan dict with the scaled estimates and the pass count)."""
    n = getattr(d_seq, "valid", d_seq.nbytes) if nbytes is None else nbytes
    cfg = TableConfig(k=k, count_max=count_max, filter_min=0, filter_max=0xFFFFFFFF, bloom=1 if bloom else 0)
    p = Plan()
    check(lib().kc_plan_table(ctx.h, _ptr(d_seq), n, sample_slices, hbm_budget, C.byref(cfg), C.byref(p)),
          "kc_plan_table")
    kw = {"count_max": count_max, "capacity": cfg.capacity, "max_distinct": cfg.max_distinct, "bloom": bool(bloom),
          "bloom_blocks": cfg.bloom_blocks, "low_capacity": cfg.low_capacity}
    return kw, {f: getattr(p, f) for f, _ in Plan._fields_}


def plan_sizes(k, distinct, repeated, bloom=False, count_max=0xFFFFFFFF, hbm_budget=0):
    """kc_plan_table's sizing rule from known totals (host only): returns
    (table kwargs for Table(ctx, k, **kw), plan dict)."""
    cfg = TableConfig(k=k, count_max=count_max, filter_min=0, filter_max=0xFFFFFFFF, bloom=1 if bloom else 0)
    p = Plan()
    check(lib().kc_plan_sizes(distinct, repeated, hbm_budget, C.byref(cfg), C.byref(p)), "kc_plan_sizes")
    kw = {"count_max": count_max, "capacity": cfg.capacity, "max_distinct": cfg.max_distinct, "bloom": bool(bloom),
          "bloom_blocks": cfg.bloom_blocks, "low_capacity": cfg.low_capacity}
    return kw, {f: getattr(p, f) for f, _ in Plan._fields_}


def merge_sorted_dumps(paths, out_path):
    """k-way merge of key-sorted dump texts (host)."""
    arr = (C.c_char_p * max(1, len(paths)))(*[os.fsencode(p) for p in paths])
    check(lib().kc_merge_sorted_dumps(arr, len(paths), os.fsencode(out_path)), "kc_merge_sorted_dumps")


def synth_genome(seed, length, n_segments=0, segment_len=0, repeat_fraction=0.0,
                 segment_divergence=0.0) -> np.ndarray:
    spec = GenomeSpec(seed=seed, length=length, n_segments=n_segments, segment_len=segment_len,
                      repeat_fraction=repeat_fraction, segment_divergence=segment_divergence)
    out = np.empty(length, dtype=np.uint8)
    check(lib().kc_synth_genome(C.byref(spec), _ptr(out)), "kc_synth_genome")
    return out


def synth_reads(genome: np.ndarray, seed, n_reads, read_len=150, both_strands=True, sub_rate=0.005,
                n_rate=0.0, long_fraction=0.0, long_mu=0.0, long_sigma=0.0, long_sub_rate=0.0,
                min_len=0, fastq=False, threads=0, out=None) -> np.ndarray:
    spec = ReadSpec(seed=seed, n_reads=n_reads, read_len=read_len, both_strands=1 if both_strands else 0,
                    sub_rate=sub_rate, n_rate=n_rate, long_fraction=long_fraction, long_mu=long_mu,
                    long_sigma=long_sigma, long_sub_rate=long_sub_rate, min_len=min_len, threads=threads)
    size = lib().kc_synth_reads_size(_ptr(genome), genome.size, C.byref(spec), 1 if fastq else 0)
    if size < 0:
        raise KCError(int(size), "kc_synth_reads_size")
    if out is None:
        out = np.empty(size, dtype=np.uint8)
    wrote = lib().kc_synth_reads(_ptr(genome), genome.size, C.byref(spec), 1 if fastq else 0,
                                 _ptr(out), out.size)
    if wrote < 0:
        raise KCError(int(wrote), "kc_synth_reads")
    return out[:wrote]


class Server:
    """kc_server: count jobs and query batches against resident tables.

    Host arrays handed to submit_* are kept alive here until wait()."""

    def __init__(self, device=0, workers=4, hbm_budget=64 << 30):
        h = C.c_void_p()
        check(lib().kc_server_create(device, workers, hbm_budget, C.byref(h)), "kc_server_create")
        self.h = h
        self._hold = {}

    def close(self):
        if self.h:
            lib().kc_server_destroy(self.h)
            self.h = None
            self._hold.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit_count(self, table_id, seq: np.ndarray, k=31, max_distinct=0, bloom=False):
        seq = np.ascontiguousarray(seq, dtype=np.uint8)
        rid = C.c_uint64()
        check(lib().kc_server_submit_count(self.h, table_id, k, _ptr(seq), seq.size, max_distinct,
                                           1 if bloom else 0, C.byref(rid)), "kc_server_submit_count")
        self._hold[rid.value] = (seq, None)
        return rid.value

    def submit_query(self, table_id, keys: np.ndarray, nw=1):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        n = keys.size // nw
        out = np.zeros(max(n, 1), dtype=np.uint32)
        rid = C.c_uint64()
        check(lib().kc_server_submit_query(self.h, table_id, _ptr(keys), n, _ptr(out), C.byref(rid)),
              "kc_server_submit_query")
        self._hold[rid.value] = (keys, out[:n])
        return rid.value

    def wait(self, rid):
        """Returns (result dict, query counts or None)."""
        r = RequestResult()
        check(lib().kc_server_wait(self.h, rid, C.byref(r)), "kc_server_wait")
        _, out = self._hold.pop(rid)
        return {f: getattr(r, f) for f, _ in RequestResult._fields_}, out

    def histogram(self, table_id, min_freq=1, max_freq=10000):
        bins = np.zeros(max_freq - min_freq + 1, dtype=np.int64)
        check(lib().kc_server_table_histogram(self.h, table_id, min_freq, max_freq, _ptr(bins)),
              "kc_server_table_histogram")
        return bins

    def drop(self, table_id):
        check(lib().kc_server_drop_table(self.h, table_id), "kc_server_drop_table")

    def resident(self):
        n, b = C.c_uint64(), C.c_uint64()
        check(lib().kc_server_resident(self.h, C.byref(n), C.byref(b)), "kc_server_resident")
        return n.value, b.value
