// Internal declarations shared by the engine's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/kc_api.h"

#define KC_HIP(expr)                                                      \
    do {                                                                  \
        hipError_t kc_e_ = (expr);                                        \
        if (kc_e_ != hipSuccess) {                                        \
            kc_set_error(#expr, kc_e_);                                   \
            return KC_EDEVICE;                                            \
        }                                                                 \
    } while (0)

void kc_set_error(const char *what, hipError_t e);

struct kc_ctx {
    int device;
    hipStream_t stream;
    uint32_t cu_count;
    hipEvent_t events[KC_EVENT_SLOTS];
};

// Slot layout in HBM.
//  nw = 1: 16 B  {u64 key, u64 count}            key EMPTY = ~0
//  nw = 2: 32 B  {u64 w0, u64 w1, u64 count, pad} w0/w1 EMPTY = ~0
// ~0 is never a canonical key: for k <= 31 / 33..63 the padding bits are
// zero, and for k == 32 all-T's reverse complement (all-A's) is smaller.
constexpr uint64_t KC_EMPTY = ~0ULL;
// Bit 63 of a count word marks nothing in the stored data; counts stay
// below 2^63 by construction (at most 2^64 bytes of input).

// Device counters.  Statistics are spread over KC_LANES lanes, one 128-byte
// line each, so that blocks update different lines (block b uses lane
// b % KC_LANES); readers sum the lanes.  The FIRST-occurrence list is cut
// into KC_LOW_SEGS segments with one cursor line each (an append goes to a
// segment mixed from its block, wave and step), for the same reason.
enum kc_stat_idx : int {
    KC_STAT_KMERS = 0,
    KC_STAT_DISTINCT = 1,
    KC_STAT_ERRORS = 2,
    KC_STAT_FRONT_REDOS = 3,  // speculative front ends redone exactly (kc_table_front_redos)
    KC_STAT_N = 4,
};
constexpr int KC_LANES = 64;
constexpr int KC_LINE_WORDS = 16;
constexpr int KC_LOW_SEGS = 1024;
constexpr uint64_t KC_STAT_WORDS = (uint64_t)KC_STAT_N * KC_LANES * KC_LINE_WORDS;
constexpr uint64_t KC_CURSOR_BASE = KC_STAT_WORDS;
constexpr uint64_t KC_COUNTER_WORDS = KC_CURSOR_BASE + (uint64_t)KC_LOW_SEGS * KC_LINE_WORDS;

// Launch shape of one partitioned counting pipeline.  Every size is fixed
// before the first kernel runs, so the pass is enqueued without host
// synchronisation (and can be captured in a graph): the key arrays hold
// key_cap keys, an upper bound on the windows of the input.
struct kc_pipe_shape {
    uint32_t front_blocks;  // persistent blocks of the parse / scatter front end
    uint32_t fine_slices;   // blocks per coarse region in the second scatter level
    uint64_t key_cap;
    uint32_t first_mode;    // partition pass FIRST hand-over: 1 staged, 2 per wave step, 3 per wave group of steps
    uint32_t front_mode;    // kc_tune.front_mode: 0 auto, 1 exact, 2 speculative
};

// Device scratch of the partitioned counting pass, sized for a shape.
struct kc_part_scratch {
    uint64_t *keys_a = nullptr, *keys_b = nullptr;  // front_cap / key_cap keys of nw words
    uint32_t *m1 = nullptr, *m2 = nullptr;          // digit counts per block / slice
    uint64_t *o1 = nullptr, *o2 = nullptr;          // their exclusive offsets
    uint64_t *sums = nullptr, *starts = nullptr, *part_off = nullptr;
    uint8_t *dig = nullptr;                         // each key's fine digit (u8, or u16 above 256 digits)
    uint64_t *gcur = nullptr;                       // front scatter: [coarse][group] cursors, region ends,
                                                    // region starts, then the speculative overflow flag
    uint64_t key_cap = 0;
    uint64_t front_cap = 0;  // keys of keys_a (and digits): key_cap plus the speculative regions' slack
    uint64_t m1_n = 0, m2_n = 0, sums_n = 0, part_n = 0, dig_bytes = 0;
};

struct kc_table {
    kc_ctx *ctx;
    kc_table_config cfg;
    uint32_t nw;
    uint64_t epoch;         // bumped by every call that changes the counts
    uint64_t cap;           // slots
    uint32_t part_bits;     // log2(partitions)
    uint32_t part_log2;     // log2(slots per partition)
    uint32_t bloom_part_log2; // log2(Bloom blocks per partition)
    uint64_t *slots;        // cap * (nw == 1 ? 2 : 4) words
    uint64_t bloom_blocks;  // 0 when Bloom is off
    uint64_t *bloom;        // bloom_blocks words: one 64-bit block each
    uint64_t low_cap;       // KC_LOW_SEGS * seg_cap
    uint64_t seg_cap;       // keys per FIRST-list segment
    uint64_t *low;          // low_cap * nw words, segment s at s * seg_cap
    uint64_t *seg_off;      // KC_LOW_SEGS + 1 exclusive offsets (device)
    uint64_t *counters;     // KC_COUNTER_WORDS device u64
    uint64_t *tile_counts;  // export scan scratch (export_scan_words(n_tiles) words)
    uint64_t n_tiles;
    // host staging for kc_table_count_seq_host
    uint8_t *pinned[2];
    uint8_t *dstage[2];
    uint64_t stage_bytes;
    hipEvent_t stage_done[2];
    hipEvent_t copy_done;   // last copy out of a page-locked caller buffer
    kc_part_scratch ps;
    kc_tune tune;           // launch shape of counting calls (zeros: defaults)
    uint32_t pass_bits;     // multi-pass counting: log2(passes) (kc_table_set_pass)
    uint32_t pass;
    // Read-only index of the FIRST list for point lookups: an open-addressing
    // set of its keys (idx_cap slots of nw words, then one error word), built
    // when a lookup finds it older than the table's counts (idx_epoch != epoch).
    uint64_t *low_idx;
    uint64_t low_idx_cap;   // allocated slots
    uint64_t low_idx_mask;  // slots in use - 1 (0: no index, the list is empty)
    uint64_t low_idx_epoch;
    // lookups launched on another context (kc_table_query_on): the table's
    // stream hands over to that stream and waits for the lookup after it
    hipEvent_t query_ready, query_done;
    // host copy of the device counters (read_counters), sized at creation
    std::vector<uint64_t> hcounters, hseg_off;
    // test-only fault injection (KC_TEST_SKEW_COUNTS): skews one digit count
    // of the partitioned pass so that the scatter guards must report it
    uint32_t fault_skew;
    // lazy clear: the table is logically empty but its slots and Bloom filter
    // have not been written; the next partitioned pass initialises them, and
    // any other access writes them first (kc_table_materialize)
    bool fresh;
};

// Every call that changes a table's counts marks the table (read-only
// indexes built before are then rebuilt on their next use).
inline void kc_table_touch(kc_table *t) { ++t->epoch; }
// Writes the empty slots and Bloom filter of a lazily cleared table (no-op
// otherwise); every access to the table other than a fresh partitioned pass
// comes after it.
int kc_table_materialize(kc_table *t);

// Slots / Bloom blocks of one partition that the LDS counting pass holds.
constexpr uint32_t kc_lds_part_slots(uint32_t nw, bool bloom) { return (nw == 1 ? 4096u : 2048u) << (bloom ? 0 : 1); }
constexpr uint32_t kc_lds_part_bloom(uint32_t) { return 4096u; }

// Grid of a FIRST-list walk (see low_seg_key): enough chunks of `per_block`
// keys for the fullest segment, at most 64.
inline unsigned kc_low_chunks(uint64_t most, uint64_t per_block)
{
    const uint64_t c = (most + per_block - 1) / per_block;
    return (unsigned)(c < 1 ? 1 : (c > 64 ? 64 : c));
}

// Launch helpers implemented in kc_count.hip / kc_sort.hip / kc_part.hip
int kc_launch_fill(kc_ctx *ctx, uint64_t *p, uint64_t n_words, uint64_t value);
uint64_t kc_scan_tmp_words(uint64_t n);
int kc_scan_u32(kc_ctx *ctx, const uint32_t *in, uint64_t n, uint64_t *out, uint64_t *sums);
// gate non-null: the scan runs only when *gate is nonzero (device flag)
int kc_scan_u32_on(hipStream_t st, const uint32_t *in, uint64_t n, uint64_t *out, uint64_t *sums,
                   const unsigned long long *gate = nullptr);
int kc_sort_range(kc_ctx *ctx, uint32_t *a, uint32_t *b, uint64_t n, uint32_t rw, uint32_t lo, uint32_t hi,
                  int *in_scratch);
// histogram_tool precise over a table's exported view with caller-provided
// device bins (kc_count.hip); allocates nothing.
int kc_histogram_into(kc_table *t, uint32_t min_freq, uint32_t max_freq, int64_t *h_bins, unsigned long long *gb);
