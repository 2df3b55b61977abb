/*
 * kc_api.h — C ABI of the MI355X k-mer counting engine (libkc.so).
 *
 * The engine is a drop-in for KTCounterMemoryGreedy's counting path.  The
 * reference is header-only C++ with no FFI; each entry point below names the
 * reference interface it replaces (file:line).  Plain pointers and sizes
 * only.  Pointers documented as "device" must point to HBM (hipMalloc or a
 * torch CUDA tensor); "host" pointers are ordinary process memory.
 *
 * Status: every call returns an int, KC_OK (0) or a negative KC_E* code.
 * Nothing calls exit().  Calls on one kc_ctx are issued in order on that
 * context's HIP stream; a kc_ctx must be used by one host thread at a time.
 *
 * Key layout (identical to kmer<N>, kmer.h:11-119 and GetKmer.h:71-77):
 * nw = 1 (k<=32), 2 (k<=64) or 4 (k<=128) little-endian u64 words, word 0
 * most significant, the k-mer in words 0..(k-1)/32 with string position j
 * at bit 64*m-2k+2j of those m words; padding bits and words zero.
 */
#ifndef KC_API_H
#define KC_API_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KC_OK 0
#define KC_EINVAL (-1)   /* bad argument */
#define KC_ENOMEM (-2)   /* device or host allocation failed */
#define KC_EFULL (-3)    /* table overflowed (a full FIRST-list segment spills into the table) */
#define KC_EDEVICE (-4)  /* HIP runtime error */
#define KC_EIO (-5)      /* file error */
#define KC_ENOTSUP (-6)  /* configuration the library does not support */

typedef struct kc_ctx kc_ctx;
typedef struct kc_table kc_table;

/* Table configuration.  Mirrors the Tree CLI knobs (main.cpp:484-554):
 * k, filter_min, filter_max, count_max; capacity replaces map_capacity. */
/* Counting paths: partitioned (hash-partition scatter, then each table
 * partition counted in LDS) or direct (global atomics per k-mer).  Both fill
 * the same table layout; AUTO picks partitioned for large batches. */
#define KC_PATH_AUTO 0
#define KC_PATH_DIRECT 1
#define KC_PATH_PARTITIONED 2

typedef struct kc_table_config {
    uint32_t k;            /* 1..128 */
    uint32_t count_max;    /* saturation value (main.cpp:599-615), >= 1 */
    uint32_t filter_min;   /* high.bin filter (FinalDrainWriter.h:101-104) */
    uint32_t filter_max;
    uint64_t capacity;     /* table slots, power of two; 0 = derive from max_distinct */
    uint64_t max_distinct; /* sizing hint when capacity == 0 */
    uint32_t bloom;        /* 1 = Bloom singleton split (BloomFilter.h:149-298) */
    uint32_t count_path;   /* KC_PATH_AUTO / _DIRECT / _PARTITIONED */
    uint64_t bloom_blocks; /* 64-bit Bloom blocks, power of two; 0 = derive */
    uint64_t low_capacity; /* FIRST-occurrence list capacity; 0 = derive */
} kc_table_config;

typedef struct kc_stats {
    uint64_t kmers;        /* windows counted (main.cpp:418 "Total read k-mer count") */
    uint64_t distinct;     /* keys stored in the table */
    uint64_t low;          /* FIRST occurrences routed to low.bin (Bloom on) */
    uint64_t capacity;     /* table slots */
    uint64_t errors;       /* nonzero: overflow or failed wait; results invalid */
} kc_stats;

/* Launch shape of the counting pipeline and of host batches, as chosen by
 * kc_autotune.  No reference counterpart: the reference sizes its thread
 * pipeline from the CLI and a pre-read sample (main.cpp:248-256, 336-347).
 * Zero fields take the built-in defaults. */
typedef struct kc_tune {
    uint32_t front_blocks;  /* persistent blocks of the parse / scatter front end */
    uint32_t fine_slices;   /* blocks per coarse region in the second scatter level */
    uint64_t batch_bytes;   /* capacity of one host batch (one staging queue slot) */
    uint32_t queue_depth;   /* batches in flight (staging slots), 2..8 */
    uint32_t path;          /* KC_PATH_* of batch launches; AUTO picks by batch size */
    uint64_t piece_bytes;   /* largest input one partitioned pass counts (0 = 4 GB); larger
                               inputs go in pieces, bounding the pass's two key arrays to
                               2 x 8N x piece_bytes bytes */
    uint32_t first_mode;    /* FIRST occurrences in the partition pass: 0 = by key width (staged
                               for k <= 32, appended above), 1 = staged per wave and moved to the
                               list at the end, 2 = appended per wave step */
    uint32_t front_mode;    /* how the partitioned pass places keys in its coarse regions:
                               0 = auto (speculative on inputs of 32 MB or more), 1 = exact (a
                               count pass over the input sizes every region first), 2 = speculative
                               whenever the pass has two scatter levels (regions of a fixed
                               capacity; a region that overflows makes the same launch sequence
                               redo the front end exactly) */
} kc_tune;

const char *kc_strerror(int status);

/* ---- context ---------------------------------------------------------- */
int kc_ctx_create(int device, kc_ctx **out);
/* A context whose stream leaves compute units 0 .. left_cus-1 to other
 * streams (a CU mask, hipExtStreamCreateWithCUMask); a plain context where the
 * runtime has no CU masks or the device has no more than 2*left_cus CUs.  No
 * reference counterpart: the serving loop runs its count jobs on such
 * contexts so that lookups never wait for a count kernel's blocks. */
int kc_ctx_create_leaving_cus(int device, uint32_t left_cus, kc_ctx **out);
void kc_ctx_destroy(kc_ctx *ctx);
int kc_ctx_sync(kc_ctx *ctx);
/* hipStream_t of the context, for callers that record events on it. */
void *kc_ctx_stream(kc_ctx *ctx);

/* ---- device memory and timing on the context's device (no reference
 *      counterpart: the reference is host-only).  Copies are synchronous
 *      with respect to the host.  Events: a pool of KC_EVENT_SLOTS timing
 *      events recorded on the context stream. ---- */
#define KC_EVENT_SLOTS 64
int kc_dev_alloc(kc_ctx *ctx, uint64_t bytes, void **out);
int kc_dev_free(kc_ctx *ctx, void *d);
int kc_host_alloc_pinned(kc_ctx *ctx, uint64_t bytes, void **out);
int kc_host_free_pinned(kc_ctx *ctx, void *h);
/* The device address of page-locked memory from kc_host_alloc_pinned (it is
 * mapped into the device's address space), for kernels that read or write
 * it in place instead of copying it.  No reference counterpart. */
int kc_host_device_ptr(kc_ctx *ctx, void *h, void **d);
int kc_copy_h2d(kc_ctx *ctx, void *d, const void *h, uint64_t bytes);
int kc_copy_d2h(kc_ctx *ctx, void *h, const void *d, uint64_t bytes);
int kc_copy_d2d(kc_ctx *ctx, void *dst, const void *src, uint64_t bytes);
int kc_event_record(kc_ctx *ctx, int slot);
int kc_event_elapsed_ms(kc_ctx *ctx, int slot_start, int slot_end, float *ms);
/* Free and total HBM of a device, as the HIP runtime reports them now
 * (serving admits count jobs by it). */
int kc_device_mem_info(int device, uint64_t *free_bytes, uint64_t *total_bytes);

/* ---- counting table: replaces ConcurrentCountingHashMap<N>
 *      (ConcurrentCountingHashMap.h:211-312) and, on the production path,
 *      ConcurrentMap<N> (ConcurrentMap.h:231-300) fed by KmerTree. ---- */
int kc_table_create(kc_ctx *ctx, const kc_table_config *cfg, kc_table **out);
void kc_table_destroy(kc_table *t);
/* Empties the table, Bloom filter, FIRST list and counters (device fill). */
int kc_table_clear(kc_table *t);
/* Frees the staging and partition scratch that counting keeps for reuse;
 * the table stays countable and queryable.  No counterpart in the
 * reference, whose buffers live until the process ends; resident serving
 * tables need it. */
int kc_table_trim(kc_table *t);

/* Fused parse + canonical extraction + (Bloom split) + count over a
 * framed byte stream in device memory: sequence lines each followed by
 * '\n' (FastqReader.h:657-687), any byte outside ACGTUacgtu resetting the
 * window (GetKmer.h:58-82, FastqParser.h:245-518).  Asynchronous. */
int kc_table_count_seq(kc_table *t, const uint8_t *d_seq, uint64_t nbytes);
/* Same from host memory: staged through pinned buffers, H2D overlapped
 * with counting; chunks cut with a k-1 byte halo (FastqReader.h:542-557).
 * A page-locked h_seq (kc_host_alloc_pinned) is copied from directly.
 * Returns once h_seq may be reused; counting itself is asynchronous. */
int kc_table_count_seq_host(kc_table *t, const uint8_t *h_seq, uint64_t nbytes);
/* increment(key, inc) for n keys (ConcurrentCountingHashMap.h:238);
 * d_inc may be NULL for +1 each.  Keys: n*nw device words. */
int kc_table_increment(kc_table *t, const uint64_t *d_keys, uint64_t n,
                       const uint32_t *d_inc);
/* Point lookup: merged count min(c, count_max) of each key, 0 if absent
 * (FlatConcurrentHashMap::find_prepared, FlatConcurrentHashMap.h:142-205,
 * with the tools' merge of a low.bin key, LowFrequencyQueryThreadPool.h:
 * 213-232).  Read-only: the table, its FIRST list and its counters are not
 * changed, so counting may go on afterwards.  With the Bloom split on, a
 * lookup after the counts changed first builds a hash index of the FIRST
 * list (2 to 4 slots of 8N bytes per FIRST occurrence in device memory,
 * kept with the table); n = 0 builds it without looking anything up.
 * Asynchronous, except for that index build. */
int kc_table_query(kc_table *t, const uint64_t *d_keys, uint64_t n,
                   uint32_t *d_counts);
/* The same lookup launched on ctx's stream instead of the table's (same
 * device), for lookups that must not queue behind other work on the table's
 * stream: a serving loop's lookup stream next to job streams that several
 * tables share.  The caller makes sure no count of this table is still in
 * flight (kc_ctx_sync of the table's context after its last count): the
 * lookup does not wait for the table's stream, which may be busy with other
 * tables.  Later work enqueued on the table's stream (counts,
 * kc_table_destroy) waits for the lookup.  No reference counterpart beyond
 * kc_table_query's. */
int kc_table_query_on(kc_table *t, kc_ctx *ctx, const uint64_t *d_keys, uint64_t n,
                      uint32_t *d_counts);
int kc_table_stats(kc_table *t, kc_stats *out);
/* How many partitioned counting passes since the table was created or last
 * cleared had a speculative coarse region overflow and redid their front end
 * exactly (kc_tune.front_mode).  No reference counterpart: the reference's
 * tree grows its nodes on demand (NewKmerTree.h:600-614). */
int kc_table_front_redos(kc_table *t, uint64_t *out);

/* Multi-pass counting for tables larger than HBM: the k-mers are split into
 * n_passes hash slices (n_passes a power of two <= 256; slice = top
 * log2(n_passes) bits of XXH3_64 of the key) and later counting calls
 * (count_seq, count_seq_host, increment, batch launches: a batch captures its
 * graphs again after a change, so flush and sync it before changing the
 * pass) count only the k-mers of slice `pass`; kc_stats.kmers counts those.  A table
 * then needs room for one slice's distinct k-mers.  Counting every slice,
 * each into a cleared table, and concatenating the exports gives the
 * one-pass result.  The out-of-memory fallback of the reference partitions
 * the exported files by key prefix instead (dump_tool.cpp:1043-1134,
 * Partition.h:135-346).  (1, 1) restores one pass. */
int kc_table_set_pass(kc_table *t, uint32_t pass, uint32_t n_passes);

/* Export in the reference's on-disk format: infos.bin (main.cpp:590-597),
 * high.bin (FinalDrainWriter.h:71-146), low.bin (ExportWriter.h:296-313). */
int kc_table_export_dir(kc_table *t, const char *dir);
/* Same, appending the records to existing high.bin / low.bin (created when
 * absent): the exports of the passes of a multi-pass count concatenate into
 * one export, as the reference's partition files do (Partition.h:135-346). */
int kc_table_export_dir_append(kc_table *t, const char *dir);
/* Device-side export: packed high/low records into device buffers.
 * Either buffer may be NULL to query the required sizes. */
int kc_table_export_device(kc_table *t, uint8_t *d_high, uint64_t high_cap,
                           uint64_t *high_bytes, uint8_t *d_low,
                           uint64_t low_cap, uint64_t *low_bytes);
/* histogram_tool precise over the table's exported view
 * (histogram_tool.cpp:734-790): bins[f-min] for f=min..max, host int64. */
int kc_table_histogram(kc_table *t, uint32_t min_freq, uint32_t max_freq,
                       int64_t *h_bins);

/* kc_table_histogram without allocating: the device bins live in a static
 * device array of KC_NOALLOC_HIST_BINS bins per device (512 KiB, held until
 * the process ends), cleared at the start of each call.  No counterpart in
 * the reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid allocation).
 * KC_EINVAL when max - min + 1 > KC_NOALLOC_HIST_BINS.  Calls that overlap
 * on one device race on the shared array. */
#define KC_NOALLOC_HIST_BINS 65536
int kc_table_histogram_noalloc(kc_table *t, uint32_t min_freq, uint32_t max_freq,
                               int64_t *h_bins);

/* dump_tool precise over the exported view (dump_tool.cpp:1202-1270):
 * "<bases>\t<count>\n" lines of merged counts in [min, max], sorted
 * lexicographically when sorted != 0 (the --sort order), built on the
 * device.  d_out may be NULL to size; the file variant writes path. */
int kc_table_dump_device(kc_table *t, uint32_t min_freq, uint32_t max_freq, int sorted,
                         uint8_t *d_out, uint64_t out_cap, uint64_t *out_bytes);
int kc_table_dump_file(kc_table *t, uint32_t min_freq, uint32_t max_freq, int sorted,
                       const char *path);

/* ---- the on-disk format for caller keys: ExportWriter<N>
 *      (ExportWriter.h:70-313) and FinalDrainWriter<N>
 *      (FinalDrainWriter.h:43-146).  Device keys in, device bytes out,
 *      asynchronous on ctx's stream except where a size is returned. ---- */
/* low.bin records of n keys (ExportWriter::byte_pack_to_buffer,
 * ExportWriter.h:296-313): k/32 full words as LE u64, then the top
 * ceil(2(k%32)/8) bytes of the masked tail word.  d_out holds
 * n * packed-k-mer bytes. */
int kc_export_low(kc_ctx *ctx, const uint64_t *d_keys, uint64_t n, uint32_t k, uint8_t *d_out);
/* high.bin records of n (key, count) pairs (write_kmer_record /
 * write_map_record, FinalDrainWriter.h:71-146): the packed key, then the
 * low count_bytes(count_max) bytes of the count (main.cpp:599-615; counts
 * are written as given, not clamped).  A pair is left out when
 * count + 1 < filter_min (in u32, so 2^32-1 wraps to 0) or count >
 * filter_max (NewKmerTree.h:1303-1311, FinalDrainWriter.h:95-104); pass 0
 * and 0xFFFFFFFF to keep every pair.  Kept records stay in input order.
 * d_out NULL: *out_bytes only.  Synchronous. */
int kc_export_high(kc_ctx *ctx, const uint64_t *d_keys, const uint32_t *d_counts, uint64_t n, uint32_t k,
                   uint32_t count_max, uint32_t filter_min, uint32_t filter_max, uint8_t *d_out,
                   uint64_t out_cap, uint64_t *out_bytes);

/* ---- the reference's tools over a temp_dir (infos.bin, high.bin, low.bin
 *      as written by the reference's counter or by kc_table_export_dir) ---- */
/* Loads a temp_dir into a new table on ctx: k and count_max from infos.bin
 * (main.cpp:590-597), high.bin records as table counts and, with with_low,
 * low.bin keys as the table's FIRST list.  The table's histogram, dump and
 * lookups are then histogram_tool's / dump_tool's precise merges of the
 * files (histogram_tool.cpp:734-816, dump_tool.cpp:1202-1270); without
 * low.bin, their approximate modes (high.bin alone: histogram_tool.cpp:
 * 840-880, dump_tool.cpp:1294-1400).  Empty files are an empty export (the
 * reference tools refuse an empty high.bin). */
int kc_table_load_dir(kc_ctx *ctx, const char *dir, int with_low, kc_table **out);
/* histogram_tool <precise|approximate> <dir> ... <min> <max>: bins[f-min]
 * for f = min..max, host int64. */
int kc_histogram_dir(kc_ctx *ctx, const char *dir, int precise, uint32_t min_freq, uint32_t max_freq,
                     int64_t *h_bins);
/* dump_tool <precise|approximate> <dir> ... <path> <min> <max> [--sort]:
 * "<bases>\t<count>\n" lines into path, sorted lexicographically when
 * sorted != 0. */
int kc_dump_dir(kc_ctx *ctx, const char *dir, int precise, uint32_t min_freq, uint32_t max_freq, int sorted,
                const char *path);

/* ---- table sizing from a device sample.  The reference plans its Bloom
 *      filters and maps from the file size, a pre-read quality average and a
 *      prefix sample (main.cpp:41-217, FastqPrefixCounter.h:28-133); here one
 *      hash slice of the input's own k-mers is counted exactly and scaled. ---- */
typedef struct kc_plan {
    uint64_t kmers;         /* estimated windows of the input */
    uint64_t distinct;      /* estimated distinct k-mers (the FIRST occurrences when the Bloom split is on) */
    uint64_t repeated;      /* estimated distinct k-mers seen at least twice */
    uint64_t device_bytes;  /* table + Bloom filter + FIRST list of the planned table (one pass) */
    uint32_t passes;        /* hash-slice passes (kc_table_set_pass) that keep device_bytes <= hbm_budget */
    uint32_t sample_slices; /* the estimate is slice 0 of this many, scaled up */
} kc_plan;
/* Counts hash slice 0 of sample_slices (a power of two <= 256; raised for
 * large inputs so the probe table stays below 2^26 keys) of the framed input
 * d_seq exactly in a scratch table, scales its distinct and repeated keys,
 * and fills cfg (k, count_max, filters, bloom, count_path as given) with
 * capacity (table keys at <= ~62% load), bloom_blocks (3.7-7.4 distinct keys per
 * 64-bit block), low_capacity and max_distinct for ONE of plan->passes passes
 * (hbm_budget = 0: no limit, one pass).  Synchronous; the scratch table is
 * freed before returning. */
int kc_plan_table(kc_ctx *ctx, const uint8_t *d_seq, uint64_t nbytes, uint32_t sample_slices, uint64_t hbm_budget,
                  kc_table_config *cfg, kc_plan *plan);
/* The sizing rule of kc_plan_table from known totals (distinct k-mers and
 * distinct k-mers seen at least twice), e.g. a caller's own estimate; host
 * only.  Fills cfg's sizes and plan's device_bytes / passes (kmers and
 * sample_slices are left 0).  KC_ENOMEM when even 256 passes do not fit. */
int kc_plan_sizes(uint64_t distinct, uint64_t repeated, uint64_t hbm_budget, kc_table_config *cfg, kc_plan *plan);

/* ---- launch shape and batched counting.  No reference counterpart: the
 *      reference feeds 32 KiB ring blocks with a k-1 base overlap to parser
 *      threads (FastqReader.h:542-557, definition.h:53-54). ---- */

/* Shape used by later counting calls of this table (kc_autotune's result,
 * or the caller's).  tune->path applies when the table's count_path is
 * AUTO. */
int kc_table_set_tune(kc_table *t, const kc_tune *tune);
int kc_table_get_tune(kc_table *t, kc_tune *out);

/* Times kc_table_count_seq over a device-resident sample (reps timed runs
 * per candidate after one warm-up) for each launch shape: partitioned with
 * {2,4,8,16} persistent front blocks per CU x {2,4,8,16} fine slices, and
 * direct; stores the fastest in the table's tune and returns it (with
 * front_mode exact when that shape's speculative front end had to be redone
 * on the sample).  Every trial starts from an empty table and the table is
 * left empty. */
int kc_autotune(kc_table *t, const uint8_t *d_sample, uint64_t nbytes, uint32_t reps, kc_tune *best,
                float *best_ms);

/* Batches have no counterpart in the reference, whose reader threads cut
 * the input into fixed 32 KiB ring blocks with a k-1 base overlap
 * (FastqReader.h:542-557, definition.h:53-54) and hand them to parser
 * threads one at a time.
 *
 * A batch packs framed read chunks ("seq\n" streams of any length) into
 * page-locked staging slots of tune->batch_bytes; each full slot (or
 * kc_batch_flush) is copied to HBM on a copy stream and counted into the
 * table by ONE launch of a captured HIP graph of the persistent pipeline,
 * on the table's stream.  tune->queue_depth slots rotate, so packing, copy
 * and counting of consecutive batches overlap.  A chunk not ending in '\n'
 * gets one; a chunk larger than a slot is counted in pieces with a 64-byte
 * halo.  tune NULL: the table's tune.  Counting is asynchronous:
 * kc_batch_sync waits for it.  Same results as kc_table_count_seq over the
 * concatenated chunks. */
typedef struct kc_batch kc_batch;
typedef struct kc_batch_info {
    uint32_t path;            /* KC_PATH_DIRECT or KC_PATH_PARTITIONED */
    uint32_t queue_depth;
    uint64_t capacity;        /* bytes per batch */
    uint32_t front_blocks;    /* persistent blocks of the captured launch */
    uint32_t fine_slices;
    uint64_t launches;        /* graph launches */
    uint64_t graphs_captured;
    uint64_t chunks;          /* chunks added */
    uint64_t pieces;          /* slot-sized pieces of oversized chunks */
    uint64_t bytes;           /* bytes counted */
} kc_batch_info;
int kc_batch_create(kc_table *t, const kc_tune *tune, kc_batch **out);
void kc_batch_destroy(kc_batch *b);
/* Copies the chunk into the current slot; returns when h_chunk may be reused. */
int kc_batch_add(kc_batch *b, const uint8_t *h_chunk, uint64_t len);
/* n_chunks consecutive chunks of one host buffer, chunk i being
 * h_base[ends[i-1], ends[i]) (ends[-1] = 0, ends non-decreasing): the same
 * batches as kc_batch_add per chunk, packed in multithreaded span copies
 * with one call.  Returns when h_base may be reused. */
int kc_batch_add_chunks(kc_batch *b, const uint8_t *h_base, const uint64_t *ends, uint64_t n_chunks);
int kc_batch_flush(kc_batch *b);
int kc_batch_sync(kc_batch *b);
/* One graph launch over a packed batch already in HBM (16-byte aligned);
 * graphs are cached per input pointer. */
int kc_batch_launch_device(kc_batch *b, const uint8_t *d_seq, uint64_t nbytes);
int kc_batch_get_info(kc_batch *b, kc_batch_info *out);

/* ---- sorts: ExportRecordRadixSort (ExportRecordRadixSort.h:256-269) and
 *      radix_sort_export_records_12bit (ExportRecordRadixSort12.h:639-646).
 * Records are {nw u64 key words, u32 count} at record_bytes stride
 * (8*nw+4 packed, or 16/24/40 as ExportRecord<N> in memory).  digit_bits 8 gives
 * the 8-bit sort's key range, 12 the 12-bit sort's (common_prefix_bases
 * used only there).  The result lands in d_recs or d_scratch;
 * *result_in_scratch says which.  Stable, like the reference. ---- */
int kc_sort_records(kc_ctx *ctx, void *d_recs, void *d_scratch, uint64_t n,
                    uint32_t nw, uint32_t k, uint32_t common_prefix_bases,
                    uint32_t digit_bits, uint32_t record_bytes,
                    int *result_in_scratch);

/* Abundance histogram of a record array's counts: bins[f-min] for f in
 * [min, max], host int64 (histogram_tool's bins over exact counts). */
int kc_records_histogram(kc_ctx *ctx, const void *d_recs, uint64_t n, uint32_t nw,
                         uint32_t record_bytes, uint32_t min_freq, uint32_t max_freq,
                         int64_t *h_bins);
/* Seeded record workload (no reference counterpart): key i is a 62-bit
 * bijection of i in the k=31 layout, counts Zipf-like with exponent s,
 * P(c >= x) = x^-(s-1), clipped to [1, 2^32-1].  record_bytes 12 or 16. */
int kc_synth_records(kc_ctx *ctx, void *d_recs, uint64_t n, uint32_t record_bytes,
                     uint64_t seed, double zipf_s);

/* ---- hashes of the raw key bytes, bit-exact with the vendored headers ---- */
#define KC_HASH_XXH3_64 0      /* XXH3_64bits (xxhash.h:6428) */
#define KC_HASH_XXH3_128_LO 1  /* XXH3_128bits low64 (xxhash.h:7249) */
#define KC_HASH_XXH3_128_HI 2  /* XXH3_128bits high64 */
#define KC_HASH_RAPIDHASH 3    /* rapidhash (rapidhash.h:500) */
#define KC_HASH_KOMIHASH 4     /* komihash seed 0 (komihash.h:763) */
#define KC_HASH_CONCURRENT_MAP 5 /* ConcurrentMap::hash_func (ConcurrentMap.h:141-148) */
#define KC_HASH_FUNC 6         /* hash_func<N> (HashFunction.h:44-63) */
int kc_hash_keys(kc_ctx *ctx, int kind, const uint64_t *d_keys, uint64_t n,
                 uint32_t nw, uint64_t *d_out);
/* ConcurrentBloomFilter::prepare_insert (BloomFilter.h:215-222). */
int kc_bloom_probe_keys(kc_ctx *ctx, const uint64_t *d_keys, uint64_t n,
                        uint32_t nw, uint64_t capacity, uint64_t *d_block,
                        uint64_t *d_mask);

#ifdef __cplusplus
}
#endif
#endif
