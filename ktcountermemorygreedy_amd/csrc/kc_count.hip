// MI355X (gfx950) k-mer counting engine: context, counting table, fused
// parse/canonical/Bloom/count kernel, point queries, export packing and the
// histogram merge.  All integer work: no MFMA.  Bound: HBM (random 16-32 B
// read-modify-writes on the table and Bloom filter, streaming input bytes).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "kc_dev_common.h"
#include "kc_guard.h"
#include "kc_pack_dev.h"

namespace {
thread_local char g_last_error[256];
constexpr int kBlock = 256;
constexpr int kTile = kc::kParseTile;           // bytes per block
}  // namespace

void kc_set_error(const char *what, hipError_t e)
{
    snprintf(g_last_error, sizeof g_last_error, "%s: %s", what, hipGetErrorString(e));
}

extern "C" const char *kc_strerror(int s)
{
    switch (s) {
    case KC_OK: return "ok";
    case KC_EINVAL: return "invalid argument";
    case KC_ENOMEM: return "out of memory";
    case KC_EFULL: return "table full";
    case KC_EDEVICE: return g_last_error[0] ? g_last_error : "HIP error";
    case KC_EIO: return "I/O error";
    case KC_ENOTSUP: return "not supported";
    default: return "unknown status";
    }
}

namespace kc {

// ---- fused count kernel ------------------------------------------------
// A block stages kTile input bytes (+ the previous segment as halo) in LDS,
// transposed so that lane t reading its own dword j hits consecutive banks.
// Thread t then rolls over halo + its 64-byte segment and handles every
// window whose last byte lies in the segment (and at or after emit_from,
// which excludes windows owned by an earlier chunk).  Block b takes tiles
// b, b + grid, ...: one block per tile, or a persistent grid over a batch
// whose size is read on the device.
template <int NW, bool BLOOM>
__global__ __launch_bounds__(kBlock) void k_count_seq(SeqSrc src, TableView tv)
{
    __shared__ ParseLds lds;
    uint64_t n, emit_from;
    seq_bounds(src, n, emit_from);
    uint32_t n_kmers = 0, n_new = 0, n_err = 0;
    const uint64_t ntiles = (n + kTile - 1) / kTile;
    uint32_t step = 0;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x, ++step) {
        stage_tile<halo_segs<NW>()>(src.seq, n, tile, lds);
        roll_segment<NW>(lds, tile, n, emit_from, tv.k, [&](const uint64_t *key, uint32_t e) {
            if (!in_pass<NW>(key, tv.pass_bits, tv.pass, tv.rw)) return;
            ++n_kmers;
            const KeyHash kh = key_hash<NW>(tv, key);
            bool to_table = true;
            if constexpr (BLOOM) {
                const bool first = bloom_first(tv, kh);
                // each byte position of the tile appends to its own segment,
                // so that even a grid of a few blocks fills every segment
                const bool kept = first_append<NW>(tv, first, key, step * kParseSeg + e);
                to_table = !first || !kept;  // a FIRST the list has no room for is counted here
            }
            if (to_table) table_add<NW>(tv, key, kh, 1, n_new, n_err);
        });
        __syncthreads();  // the next tile overwrites lds
    }
    stat_add(tv.counters, KC_STAT_KMERS, n_kmers);
    stat_add(tv.counters, KC_STAT_DISTINCT, n_new);
    stat_add(tv.counters, KC_STAT_ERRORS, n_err);
}

// increment(key, inc) for a batch of keys.
template <int NW>
__global__ __launch_bounds__(kBlock) void k_increment(const uint64_t *__restrict__ keys, uint64_t n,
                                                      const uint32_t *__restrict__ inc, TableView tv)
{
    uint32_t n_new = 0, n_err = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
        uint64_t key[NW];
        caller_key<NW>(keys, i, tv.rw, key);
        if (!in_pass<NW>(key, tv.pass_bits, tv.pass, tv.rw)) continue;
        table_add<NW>(tv, key, key_hash<NW>(tv, key), inc ? inc[i] : 1, n_new, n_err);
    }
    stat_add(tv.counters, KC_STAT_DISTINCT, n_new);
    stat_add(tv.counters, KC_STAT_ERRORS, n_err);
}

// ---- read-only FIRST-list index ---------------------------------------
// A point lookup of a Bloom-split table answers the table count plus one
// when the key is on the FIRST list (the merge rule of
// LowFrequencyQueryThreadPool.h:213-232 for one key).  The list is an array
// of unique keys (one FIRST per distinct key), so an open-addressing set of
// them at load <= 1/2 is built once per change of the counts and probed by
// lookups: a lookup reads the table and the index and writes neither.
template <int NW>
__device__ __forceinline__ uint64_t idx_hash(const uint64_t *key)
{
    uint64_t x = key[0];
#pragma unroll
    for (int w = 1; w < NW; ++w) x ^= ((key[w] >> (29 - 4 * w)) | (key[w] << (35 + 4 * w))) * 0x94D049BB133111EBULL;
    x ^= x >> 31;
    x *= 0xBF58476D1CE4E5B9ULL;
    x ^= x >> 29;
    x *= 0x94D049BB133111EBULL;
    return x ^ (x >> 32);
}

// Index slots: the key (one or two words, claimed with a CAS on word 0 from
// KC_EMPTY), or four key words and a fifth word claimed from 0 to 1 (a
// four-word key can have an all-ones word 0).
template <int NW> constexpr int idx_words() { return NW == 4 ? 5 : NW; }

// Inserts every FIRST-list key (block (chunk, segment) walks one segment).
// idx[(imask + 1) * idx_words] counts keys that found no slot (never at load 1/2).
template <int NW>
__global__ __launch_bounds__(kBlock) void k_low_index_build(TableView tv, uint64_t *__restrict__ idx, uint64_t imask)
{
    uint32_t n_err = 0;
    const uint32_t seg = blockIdx.y;
    const uint64_t n = low_seg_fill(tv, seg);
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
        const uint64_t *src = low_seg_key(tv, seg, i, NW);
        uint64_t key[NW];
#pragma unroll
        for (int w = 0; w < NW; ++w) key[w] = src[w];
        uint64_t s = idx_hash<NW>(key) & imask;
        bool done = false;
        for (uint64_t probe = 0; probe <= imask && !done; ++probe, s = (s + 1) & imask) {
            uint64_t *slot = idx + s * idx_words<NW>();
            // the list's keys are unique: an occupied slot holds another key
            if constexpr (NW == 4) {
                if (atomicCAS((unsigned long long *)(slot + 4), 0ull, 1ull) == 0) {
#pragma unroll
                    for (int w = 0; w < 4; ++w) slot[w] = key[w];
                    done = true;
                }
            } else if (atomicCAS((unsigned long long *)slot, (unsigned long long)KC_EMPTY, (unsigned long long)key[0]) ==
                       KC_EMPTY) {
                if constexpr (NW == 2) slot[1] = key[1];
                done = true;
            }
        }
        n_err += !done;
    }
    uint64_t e = n_err;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o, 64);
    if (__lane_id() == 0 && e) atomicAdd((unsigned long long *)(idx + (imask + 1) * idx_words<NW>()), (unsigned long long)e);
}

template <int NW>
__device__ __forceinline__ bool idx_contains(const uint64_t *__restrict__ idx, uint64_t imask, const uint64_t *key)
{
    uint64_t s = idx_hash<NW>(key) & imask;
    for (uint64_t probe = 0; probe <= imask; ++probe, s = (s + 1) & imask) {
        const uint64_t *slot = idx + s * idx_words<NW>();
        if constexpr (NW == 4) {
            if (slot[4] == 0) return false;
            if (slot[0] == key[0] && slot[1] == key[1] && slot[2] == key[2] && slot[3] == key[3]) return true;
        } else {
            const uint64_t w0 = slot[0];
            if (w0 == KC_EMPTY) return false;
            if (w0 == key[0] && (NW == 1 || slot[NW - 1] == key[NW - 1])) return true;
        }
    }
    return false;
}

// Merged count of each key: table count (+1 when on the FIRST list, IDX),
// clamped at count_max.  Reads only.
template <int NW, bool IDX>
__global__ __launch_bounds__(kBlock) void k_query(const uint64_t *__restrict__ keys, uint64_t n,
                                                  uint32_t *__restrict__ out, TableView tv,
                                                  const uint64_t *__restrict__ idx, uint64_t imask)
{
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
        uint64_t key[NW];
        caller_key<NW>(keys, i, tv.rw, key);
        uint64_t c = table_find<NW>(tv, key);
        if constexpr (IDX) c += idx_contains<NW>(idx, imask, key);
        out[i] = (uint32_t)(c < tv.count_max ? c : tv.count_max);
    }
}

template <int NW>
__global__ void k_init_slots(uint64_t *slots, uint64_t cap)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cap; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t *s = slots + i * slot_words<NW>();
        if constexpr (NW == 4) {
#pragma unroll
            for (int w = 0; w < 8; ++w) s[w] = 0;  // state word 0: empty
        } else {
            s[0] = KC_EMPTY;
            s[1] = NW == 1 ? 0 : KC_EMPTY;
            if constexpr (NW == 2) {
                s[2] = 0;
                s[3] = 0;
            }
        }
    }
}

__global__ void k_fill(uint64_t *p, uint64_t n, uint64_t v)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---- export (packing helpers: kc_pack_dev.h) ---------------------------
// A slot's words in registers: key words, count, and for four-word keys the
// state word (count_word / kStateWord index them as in the slot).
template <int NW> constexpr int kSlotRegs = NW == 4 ? 6 : NW + 1;

template <int NW>
__device__ __forceinline__ void load_slot_regs(const TableView &tv, uint64_t i, uint64_t *r)
{
    const uint64_t *s = tv.slots + i * slot_words<NW>();
    if constexpr (NW == 1) {
        const ulonglong2 v = *reinterpret_cast<const ulonglong2 *>(s);
        r[0] = v.x;
        r[1] = v.y;
    } else if constexpr (NW == 2) {
        const ulonglong2 a = *reinterpret_cast<const ulonglong2 *>(s);
        r[0] = a.x;
        r[1] = a.y;
        r[2] = s[2];
    } else {
#pragma unroll
        for (int w = 0; w < kSlotRegs<NW>; ++w) r[w] = s[w];
    }
}

// Classifies a slot: 1 = high record (count in cc), 2 = low record
// (singleton without an active FIRST list), 0 = nothing.
template <int NW>
__device__ __forceinline__ int classify_regs(const TableView &tv, const uint64_t *r, bool singles_to_low,
                                             uint32_t &cc)
{
    const bool used = NW == 4 ? r[kStateWord] == kSlotReady : r[0] != KC_EMPTY;
    if (!used) return 0;
    const uint64_t c = r[count_word<NW>()];
    if (singles_to_low && c == 1) return 2;
    if (c == 0) return 0;
    cc = (uint32_t)(c < tv.count_max ? c : tv.count_max);
    return export_filtered(cc, tv) ? 0 : 1;
}

// Exclusive scan of one value per thread of a kBlock-thread block: a shuffle
// scan inside each wave, then the wave totals (two barriers).  The export
// scans a tile's high and low record counts at once, packed as h | l << 16
// (each at most kExportTile).
__device__ __forceinline__ uint32_t block_scan_u32(uint32_t v, uint32_t *wsum, uint32_t &total)
{
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    uint32_t base = 0;
    total = 0;
#pragma unroll
    for (uint32_t w = 0; w < kBlock / 64; ++w) {
        const uint32_t sw = wsum[w];
        base += w < wave ? sw : 0;
        total += sw;
    }
    __syncthreads();
    return base + x - v;
}

// Export scan scratch (one allocation of export_scan_words(n_tiles) words):
// exclusive offsets of the 2 x n_tiles record counts (high counts of every
// tile, then low counts), then the high and low totals; the u32 counts; the
// scan's block sums.
inline uint64_t export_scan_words(uint64_t n_tiles)
{
    return 2 * n_tiles + 2 + n_tiles + kc_scan_tmp_words(2 * n_tiles);
}
struct ExportScan {
    uint64_t *off;    // [2 n_tiles + 2]
    uint32_t *cnt;    // [2 n_tiles]
    uint64_t *sums;
};
inline ExportScan export_scan_of(uint64_t *base, uint64_t n_tiles)
{
    ExportScan e;
    e.off = base;
    e.cnt = reinterpret_cast<uint32_t *>(base + 2 * n_tiles + 2);
    e.sums = base + 2 * n_tiles + 2 + n_tiles;
    return e;
}

// Pass 1: per-tile record counts -> cnt[t] (high), cnt[n_tiles + t] (low).
template <int NW>
__global__ __launch_bounds__(kBlock) void k_export_count(TableView tv, bool singles_to_low, uint64_t n_tiles,
                                                         uint32_t *cnt)
{
    __shared__ uint32_t wsum[kBlock / 64];
    const uint64_t cap = tv.mask + 1;
    for (uint64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        uint32_t hl = 0;
        uint64_t r[kExportPer][kSlotRegs<NW>];
#pragma unroll
        for (int j = 0; j < kExportPer; ++j) {  // every slot load issued before the first use
            const uint64_t i = tile * kExportTile + (uint64_t)j * kBlock + threadIdx.x;
            if (i < cap) load_slot_regs<NW>(tv, i, r[j]);
        }
#pragma unroll
        for (int j = 0; j < kExportPer; ++j) {
            const uint64_t i = tile * kExportTile + (uint64_t)j * kBlock + threadIdx.x;
            uint32_t cc;
            const int kind = i < cap ? classify_regs<NW>(tv, r[j], singles_to_low, cc) : 0;
            hl += kind == 1 ? 1u : kind == 2 ? 1u << 16 : 0u;
        }
        uint32_t total;
        block_scan_u32(hl, wsum, total);
        if (threadIdx.x == 0) {
            cnt[tile] = total & 0xFFFF;
            cnt[n_tiles + tile] = total >> 16;
        }
    }
}

// Totals after the scan of the counts: high = the offset of the first low
// count, low = the last low offset + count - high.
__global__ void k_export_totals(uint64_t *off, const uint32_t *cnt, uint64_t n_tiles)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    off[2 * n_tiles] = off[n_tiles];
    off[2 * n_tiles + 1] = off[2 * n_tiles - 1] + cnt[2 * n_tiles - 1] - off[n_tiles];
}

// Pass 2: pack records at the scanned offsets.  Low records from the table
// (singletons) follow the n_list FIRST-list records in the low buffer.  Each
// tile's high records, then its low records, are packed into LDS and copied
// out as one contiguous byte range.  One- and two-word slots of the next tile
// are loaded while this one is packed and written.
template <int NW>
__global__ __launch_bounds__(kBlock) void k_export_write(TableView tv, PackSpec ps, bool singles_to_low,
                                                         uint64_t n_tiles, const uint64_t *off,
                                                         uint8_t *out_high, uint8_t *out_low, uint64_t n_list)
{
    constexpr bool kPrefetch = NW <= 2;
    __shared__ __align__(16) uint8_t stage[export_stage<NW>()];
    __shared__ uint32_t wsum[kBlock / 64];
    const uint64_t cap = tv.mask + 1;
    const uint32_t hrec = ps.kmer_bytes + ps.count_bytes;
    const uint64_t low0 = off[n_tiles];
    uint64_t r[kExportPer][kSlotRegs<NW>];
    auto load_tile = [&](uint64_t tile) {
#pragma unroll
        for (int j = 0; j < kExportPer; ++j) {
            const uint64_t i = tile * kExportTile + (uint64_t)j * kBlock + threadIdx.x;
            if (i < cap) load_slot_regs<NW>(tv, i, r[j]);
        }
    };
    if (kPrefetch && blockIdx.x < n_tiles) load_tile(blockIdx.x);
    for (uint64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        if (!kPrefetch) load_tile(tile);
        uint64_t key[kExportPer][NW];
        uint32_t cc[kExportPer];
        int kind[kExportPer];
        uint32_t hl = 0;
#pragma unroll
        for (int j = 0; j < kExportPer; ++j) {
            const uint64_t i = tile * kExportTile + (uint64_t)j * kBlock + threadIdx.x;
            kind[j] = i < cap ? classify_regs<NW>(tv, r[j], singles_to_low, cc[j]) : 0;
#pragma unroll
            for (int w = 0; w < NW; ++w) key[j][w] = r[j][w];
            hl += kind[j] == 1 ? 1u : kind[j] == 2 ? 1u << 16 : 0u;
        }
        if (kPrefetch && tile + gridDim.x < n_tiles) load_tile(tile + gridDim.x);
        uint32_t total;
        const uint32_t at0 = block_scan_u32(hl, wsum, total);
        const uint32_t th = total & 0xFFFF, tl = total >> 16;
        // high records of the tile
        uint8_t *dh = out_high + off[tile] * hrec;
        uint8_t *sh = stage + stage_skew(dh);
        uint32_t at = at0 & 0xFFFF;
#pragma unroll
        for (int j = 0; j < kExportPer; ++j) {
            if (kind[j] == 1) {
                uint8_t *d = sh + at * hrec;
                pack_key<NW>(d, key[j], ps);
                put_bytes(d + ps.kmer_bytes, cc[j], ps.count_bytes, 0);
                ++at;
            }
        }
        __syncthreads();
        stage_copy(dh, sh, th * hrec);
        __syncthreads();
        if (tl) {
            uint8_t *dl = out_low + (n_list + off[n_tiles + tile] - low0) * ps.kmer_bytes;
            uint8_t *sl = stage + stage_skew(dl);
            at = at0 >> 16;
#pragma unroll
            for (int j = 0; j < kExportPer; ++j) {
                if (kind[j] == 2) {
                    pack_key<NW>(sl + at * ps.kmer_bytes, key[j], ps);
                    ++at;
                }
            }
            __syncthreads();
            stage_copy(dl, sl, tl * ps.kmer_bytes);
            __syncthreads();
        }
    }
}

// FIRST-list records: block (chunk, segment) packs rows [chunk * tile,
// (chunk + 1) * tile) of one segment (they are consecutive in the
// concatenated list) into LDS and copies them out as one byte range.
template <int NW>
__global__ __launch_bounds__(kBlock) void k_export_low_list(TableView tv, PackSpec ps, uint8_t *out_low)
{
    __shared__ __align__(16) uint8_t stage[export_stage<NW>()];
    const uint32_t seg = blockIdx.y;
    const uint64_t s0 = tv.seg_off[seg], cnt = tv.seg_off[seg + 1] - s0;
    for (uint64_t r0 = blockIdx.x * kExportTile; r0 < cnt; r0 += (uint64_t)gridDim.x * kExportTile) {
        const uint32_t rn = (uint32_t)(cnt - r0 < kExportTile ? cnt - r0 : kExportTile);
        uint8_t *d = out_low + (s0 + r0) * ps.kmer_bytes;
        uint8_t *sd = stage + stage_skew(d);
        uint64_t key[kExportPer][NW];
#pragma unroll
        for (int j = 0; j < kExportPer; ++j) {  // every row load issued before the first use
            const uint32_t r = (uint32_t)j * kBlock + threadIdx.x;
            if (r < rn) {
                const uint64_t *src = tv.low + ((uint64_t)seg * tv.seg_cap + r0 + r) * NW;
#pragma unroll
                for (int w = 0; w < NW; ++w) key[j][w] = src[w];
            }
        }
#pragma unroll
        for (int j = 0; j < kExportPer; ++j) {
            const uint32_t r = (uint32_t)j * kBlock + threadIdx.x;
            if (r < rn) pack_key<NW>(sd + r * ps.kmer_bytes, key[j], ps);
        }
        __syncthreads();
        stage_copy(d, sd, rn * ps.kmer_bytes);
        __syncthreads();
    }
}

// ---- histogram (histogram_tool precise over the exported view) ----------
// High pass (HighFrequencyInsertThreadPool.h:150-177): for each exported
// high record c: skip if c+1 < min; hist[c]++ when c in [min,max].
// Low pass (LowFrequencyQueryThreadPool.h:186-245): found -> hist[c]--,
// hist[min(c+1,count_max)]++; missing -> hist[1]++.
constexpr int kLdsBins = 4096;

struct HistSpec {
    uint32_t min_freq, max_freq;
    bool singles_low;  // singletons are low records (no active FIRST list)
};

__device__ __forceinline__ void hist_add(int *lds_bins, unsigned long long *gbins, const HistSpec &h,
                                         uint64_t f, int delta)
{
    if (f < h.min_freq || f > h.max_freq) return;
    const uint64_t b = f - h.min_freq;
    if (b < kLdsBins) atomicAdd(lds_bins + b, delta);
    else atomicAdd(gbins + b, (unsigned long long)(long long)delta);
}

__device__ __forceinline__ void hist_flush(int *lds_bins, unsigned long long *gbins, uint64_t nb)
{
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < kLdsBins && b < nb; b += kBlock)
        if (lds_bins[b]) atomicAdd(gbins + b, (unsigned long long)(long long)lds_bins[b]);
}

// exported high count of a slot, or -1 when the slot yields no high record
template <int NW>
__device__ __forceinline__ int64_t high_view(const TableView &tv, uint64_t c, bool singles_low)
{
    if (c == 0) return -1;
    if (singles_low && c == 1) return -1;
    const uint32_t cc = (uint32_t)(c < tv.count_max ? c : tv.count_max);
    if (export_filtered(cc, tv)) return -1;
    return (int64_t)cc;
}

template <int NW>
__global__ __launch_bounds__(kBlock) void k_hist_table(TableView tv, HistSpec h, unsigned long long *gbins)
{
    __shared__ int lds_bins[kLdsBins];
    for (uint32_t b = threadIdx.x; b < kLdsBins; b += kBlock) lds_bins[b] = 0;
    __syncthreads();
    const uint64_t cap = tv.mask + 1;
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < cap; i += (uint64_t)gridDim.x * kBlock) {
        const uint64_t *s = tv.slots + i * slot_words<NW>();
        if (!slot_used<NW>(s)) continue;
        const uint64_t c = s[count_word<NW>()];
        if (h.singles_low && c == 1) {
            // low record whose lookup misses the high set
            hist_add(lds_bins, gbins, h, 1, 1);
            continue;
        }
        const int64_t cc = high_view<NW>(tv, c, h.singles_low);
        if (cc < 0) continue;
        if ((uint64_t)cc + 1 < h.min_freq) continue;
        hist_add(lds_bins, gbins, h, (uint64_t)cc, 1);
    }
    hist_flush(lds_bins, gbins, (uint64_t)h.max_freq - h.min_freq + 1);
}

template <int NW>
__global__ __launch_bounds__(kBlock) void k_hist_low_list(TableView tv, HistSpec h, unsigned long long *gbins)
{
    __shared__ int lds_bins[kLdsBins];
    for (uint32_t b = threadIdx.x; b < kLdsBins; b += kBlock) lds_bins[b] = 0;
    __syncthreads();
    const uint32_t seg = blockIdx.y;
    const uint64_t n = low_seg_fill(tv, seg);
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
        const uint64_t *key = low_seg_key(tv, seg, i, NW);
        const uint64_t c = table_find<NW>(tv, key);
        const int64_t cc = high_view<NW>(tv, c, false);
        if (cc >= 0 && (uint64_t)cc + 1 >= h.min_freq) {
            hist_add(lds_bins, gbins, h, (uint64_t)cc, -1);
            uint64_t merged = (uint64_t)cc + 1;
            if (merged > tv.count_max) merged = tv.count_max;
            hist_add(lds_bins, gbins, h, merged, 1);
        } else {
            hist_add(lds_bins, gbins, h, 1, 1);
        }
    }
    hist_flush(lds_bins, gbins, (uint64_t)h.max_freq - h.min_freq + 1);
}

// ---- hash KAT kernels ----------------------------------------------------
template <int NW>
__global__ void k_hash(int kind, const uint64_t *keys, uint64_t n, uint64_t *out)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t w[NW];
#pragma unroll
        for (int j = 0; j < NW; ++j) w[j] = keys[i * NW + j];
        uint64_t r = 0, lo, hi;
        switch (kind) {
        case KC_HASH_XXH3_64: r = xxh3_64<NW>(w); break;
        case KC_HASH_XXH3_128_LO: xxh3_128<NW>(w, lo, hi); r = lo; break;
        case KC_HASH_XXH3_128_HI: xxh3_128<NW>(w, lo, hi); r = hi; break;
        case KC_HASH_RAPIDHASH: r = rapidhash<NW>(w); break;
        case KC_HASH_KOMIHASH: r = komihash<NW>(w); break;
        case KC_HASH_CONCURRENT_MAP: r = mix_hash(komihash<NW>(w), rapidhash<NW>(w)); break;
        case KC_HASH_FUNC: r = hash_func<NW>(w); break;
        }
        out[i] = r;
    }
}

template <int NW>
__global__ void k_bloom_probe(const uint64_t *keys, uint64_t n, uint64_t cap_mask, uint64_t *blk, uint64_t *msk)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t w[NW];
#pragma unroll
        for (int j = 0; j < NW; ++j) w[j] = keys[i * NW + j];
        bloom_probe<NW>(w, cap_mask, blk[i], msk[i]);
    }
}

}  // namespace kc

using namespace kc;

// ======================================================================
// host side
// ======================================================================
namespace {

uint64_t next_pow2(uint64_t v)
{
    uint64_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

bool is_pow2(uint64_t v) { return v && !(v & (v - 1)); }

uint32_t ilog2(uint64_t v)
{
    uint32_t r = 0;
    while ((1ULL << r) < v) ++r;
    return r;
}

uint32_t grid_for(uint64_t n, uint32_t cu)
{
    const uint64_t want = (n + kBlock - 1) / kBlock;
    const uint64_t cap = (uint64_t)cu * 16;
    return (uint32_t)std::max<uint64_t>(1, std::min(want, cap));
}

TableView view_of(kc_table *t)
{
    TableView v;
    v.slots = t->slots;
    v.mask = t->cap - 1;
    v.part_bits = t->part_bits;
    v.part_log2 = t->part_log2;
    v.bloom = t->bloom;
    v.bloom_mask = t->bloom_blocks ? t->bloom_blocks - 1 : 0;
    v.bloom_part_log2 = t->bloom_part_log2;
    v.low = t->low;
    v.seg_cap = t->seg_cap;
    v.seg_off = t->seg_off;
    v.counters = reinterpret_cast<unsigned long long *>(t->counters);
    v.k = t->cfg.k;
    v.count_max = t->cfg.count_max;
    v.filter_min = t->cfg.filter_min;
    v.filter_max = t->cfg.filter_max;
    v.pass_bits = t->pass_bits;
    v.pass = t->pass;
    v.rw = t->cfg.k <= 32 ? 1 : t->cfg.k <= 64 ? 2 : 4;
    return v;
}

PackSpec pack_spec(const kc_table *t) { return kc_pack_spec(t->cfg.k, t->cfg.count_max); }

bool low_list_active(const kc_table *t) { return t->bloom != nullptr; }

struct HostCounters {
    uint64_t stat[KC_STAT_N];
    uint64_t seg_fill[KC_LOW_SEGS];  // FIRST keys stored per segment
    uint64_t low;                    // sum of seg_fill
};

// Reads the device counters, sums the statistic lanes and, when the FIRST
// list is active, uploads its segment offsets for the kernels that read it.
// The host buffers belong to the table (sized at creation): nothing is
// allocated here.
int read_counters(kc_table *t, HostCounters &hc)
{
    std::vector<uint64_t> &h = t->hcounters, &off = t->hseg_off;
    KC_HIP(hipMemcpyAsync(h.data(), t->counters, KC_COUNTER_WORDS * 8, hipMemcpyDeviceToHost, t->ctx->stream));
    KC_HIP(hipStreamSynchronize(t->ctx->stream));
    for (int st = 0; st < KC_STAT_N; ++st) {
        hc.stat[st] = 0;
        for (int l = 0; l < KC_LANES; ++l) hc.stat[st] += h[((uint64_t)st * KC_LANES + l) * KC_LINE_WORDS];
    }
    hc.low = 0;
    for (int sgi = 0; sgi < KC_LOW_SEGS; ++sgi) {
        const uint64_t c = h[KC_CURSOR_BASE + (uint64_t)sgi * KC_LINE_WORDS];
        // a segment's cursor runs past its capacity when it overflowed; the
        // FIRSTs it had no room for were counted in the table instead
        hc.seg_fill[sgi] = std::min(c, t->seg_cap);
        off[sgi] = hc.low;
        hc.low += hc.seg_fill[sgi];
    }
    off[KC_LOW_SEGS] = hc.low;
    if (t->seg_off) {
        KC_HIP(hipMemcpyAsync(t->seg_off, off.data(), off.size() * 8, hipMemcpyHostToDevice, t->ctx->stream));
        KC_HIP(hipStreamSynchronize(t->ctx->stream));
    }
    return KC_OK;
}

}  // namespace

int kc_launch_fill(kc_ctx *ctx, uint64_t *p, uint64_t n, uint64_t v)
{
    if (!n) return KC_OK;
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n, ctx->cu_count)), dim3(kBlock), 0, ctx->stream, p, n, v);
    KC_HIP(hipGetLastError());
    return KC_OK;
}

extern "C" int kc_ctx_create(int device, kc_ctx **out)
{
    if (!out) return KC_EINVAL;
    *out = nullptr;
    int ndev = 0;
    KC_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return KC_EINVAL;
    KC_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    KC_HIP(hipGetDeviceProperties(&prop, device));
    kc_ctx *c = new (std::nothrow) kc_ctx;
    if (!c) return KC_ENOMEM;
    c->device = device;
    c->cu_count = (uint32_t)prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return KC_EDEVICE;
    }
    for (int i = 0; i < KC_EVENT_SLOTS; ++i) c->events[i] = nullptr;
    *out = c;
    return KC_OK;
}

extern "C" int kc_ctx_create_leaving_cus(int device, uint32_t left_cus, kc_ctx **out)
{
    if (!out) return KC_EINVAL;
    int rc = kc_ctx_create(device, out);
    kc_ctx *c = *out;
    // a mask only where enough CUs stay on the stream to be worth it
    if (rc != KC_OK || !left_cus || c->cu_count <= 2 * left_cus) return rc;
    std::vector<uint32_t> mask((c->cu_count + 31) / 32, 0u);
    for (uint32_t cu = left_cus; cu < c->cu_count; ++cu) mask[cu / 32] |= 1u << (cu % 32);
    hipStream_t ms = nullptr;
    if (hipExtStreamCreateWithCUMask(&ms, (uint32_t)mask.size(), mask.data()) == hipSuccess) {
        // the masked stream replaces the plain one; a runtime without CU masks keeps it
        (void)hipStreamDestroy(c->stream);
        c->stream = ms;
    } else {
        (void)hipGetLastError();
    }
    return KC_OK;
}

extern "C" int kc_dev_alloc(kc_ctx *ctx, uint64_t bytes, void **out)
{
    if (!ctx || !out) return KC_EINVAL;
    *out = nullptr;
    KC_HIP(hipSetDevice(ctx->device));
    if (hipMalloc(out, bytes ? bytes : 1) != hipSuccess) return KC_ENOMEM;
    return KC_OK;
}

extern "C" int kc_dev_free(kc_ctx *ctx, void *d)
{
    if (!ctx) return KC_EINVAL;
    if (!d) return KC_OK;
    KC_HIP(hipStreamSynchronize(ctx->stream));
    KC_HIP(hipFree(d));
    return KC_OK;
}

extern "C" int kc_host_alloc_pinned(kc_ctx *ctx, uint64_t bytes, void **out)
{
    if (!ctx || !out) return KC_EINVAL;
    *out = nullptr;
    if (hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocMapped) != hipSuccess) return KC_ENOMEM;
    return KC_OK;
}

extern "C" int kc_host_device_ptr(kc_ctx *ctx, void *h, void **d)
{
    if (!ctx || !h || !d) return KC_EINVAL;
    *d = nullptr;
    KC_HIP(hipSetDevice(ctx->device));
    KC_HIP(hipHostGetDevicePointer(d, h, 0));
    return KC_OK;
}

extern "C" int kc_host_free_pinned(kc_ctx *ctx, void *h)
{
    if (!ctx) return KC_EINVAL;
    if (!h) return KC_OK;
    KC_HIP(hipStreamSynchronize(ctx->stream));
    KC_HIP(hipHostFree(h));
    return KC_OK;
}

static int copy_sync(kc_ctx *ctx, void *dst, const void *src, uint64_t bytes, hipMemcpyKind kind)
{
    if (!ctx || (bytes && (!dst || !src))) return KC_EINVAL;
    if (!bytes) return KC_OK;
    KC_HIP(hipMemcpyAsync(dst, src, bytes, kind, ctx->stream));
    KC_HIP(hipStreamSynchronize(ctx->stream));
    return KC_OK;
}

extern "C" int kc_copy_h2d(kc_ctx *ctx, void *d, const void *h, uint64_t bytes)
{
    return copy_sync(ctx, d, h, bytes, hipMemcpyHostToDevice);
}

extern "C" int kc_copy_d2h(kc_ctx *ctx, void *h, const void *d, uint64_t bytes)
{
    return copy_sync(ctx, h, d, bytes, hipMemcpyDeviceToHost);
}

extern "C" int kc_copy_d2d(kc_ctx *ctx, void *dst, const void *src, uint64_t bytes)
{
    return copy_sync(ctx, dst, src, bytes, hipMemcpyDeviceToDevice);
}

extern "C" int kc_event_record(kc_ctx *ctx, int slot)
{
    if (!ctx || slot < 0 || slot >= KC_EVENT_SLOTS) return KC_EINVAL;
    if (!ctx->events[slot]) KC_HIP(hipEventCreate(&ctx->events[slot]));
    KC_HIP(hipEventRecord(ctx->events[slot], ctx->stream));
    return KC_OK;
}

extern "C" int kc_event_elapsed_ms(kc_ctx *ctx, int a, int b, float *ms)
{
    if (!ctx || !ms || a < 0 || b < 0 || a >= KC_EVENT_SLOTS || b >= KC_EVENT_SLOTS) return KC_EINVAL;
    if (!ctx->events[a] || !ctx->events[b]) return KC_EINVAL;
    KC_HIP(hipEventSynchronize(ctx->events[b]));
    KC_HIP(hipEventElapsedTime(ms, ctx->events[a], ctx->events[b]));
    return KC_OK;
}

extern "C" int kc_device_mem_info(int device, uint64_t *free_bytes, uint64_t *total_bytes)
{
    if (!free_bytes || !total_bytes) return KC_EINVAL;
    int ndev = 0;
    KC_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return KC_EINVAL;
    KC_HIP(hipSetDevice(device));
    size_t f = 0, t = 0;
    KC_HIP(hipMemGetInfo(&f, &t));
    *free_bytes = f;
    *total_bytes = t;
    return KC_OK;
}

extern "C" void kc_ctx_destroy(kc_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (int i = 0; i < KC_EVENT_SLOTS; ++i)
        if (ctx->events[i]) (void)hipEventDestroy(ctx->events[i]);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" int kc_ctx_sync(kc_ctx *ctx)
{
    if (!ctx) return KC_EINVAL;
    KC_HIP(hipStreamSynchronize(ctx->stream));
    return KC_OK;
}

extern "C" void *kc_ctx_stream(kc_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

// Empty slots and a zero Bloom filter.
static int init_slots_bloom(kc_table *t)
{
    kc_ctx *c = t->ctx;
    if (t->nw == 1)
        hipLaunchKernelGGL(k_init_slots<1>, dim3(grid_for(t->cap, c->cu_count)), dim3(kBlock), 0, c->stream, t->slots, t->cap);
    else if (t->nw == 2)
        hipLaunchKernelGGL(k_init_slots<2>, dim3(grid_for(t->cap, c->cu_count)), dim3(kBlock), 0, c->stream, t->slots, t->cap);
    else
        hipLaunchKernelGGL(k_init_slots<4>, dim3(grid_for(t->cap, c->cu_count)), dim3(kBlock), 0, c->stream, t->slots, t->cap);
    KC_HIP(hipGetLastError());
    if (t->bloom) return kc_launch_fill(c, t->bloom, t->bloom_blocks, 0);
    return KC_OK;
}

int kc_table_materialize(kc_table *t)
{
    if (!t->fresh) return KC_OK;
    const int rc = init_slots_bloom(t);
    if (rc == KC_OK) t->fresh = false;
    return rc;
}

static int table_init_device(kc_table *t)
{
    kc_ctx *c = t->ctx;
    if (t->nw == 1)
        hipLaunchKernelGGL(k_init_slots<1>, dim3(grid_for(t->cap, c->cu_count)), dim3(kBlock), 0, c->stream, t->slots, t->cap);
    else if (t->nw == 2)
        hipLaunchKernelGGL(k_init_slots<2>, dim3(grid_for(t->cap, c->cu_count)), dim3(kBlock), 0, c->stream, t->slots, t->cap);
    else
        hipLaunchKernelGGL(k_init_slots<4>, dim3(grid_for(t->cap, c->cu_count)), dim3(kBlock), 0, c->stream, t->slots, t->cap);
    KC_HIP(hipGetLastError());
    int rc;
    if (t->bloom && (rc = kc_launch_fill(c, t->bloom, t->bloom_blocks, 0)) != KC_OK) return rc;
    if ((rc = kc_launch_fill(c, t->counters, KC_COUNTER_WORDS, 0)) != KC_OK) return rc;
    t->fresh = false;
    kc_table_touch(t);
    return KC_OK;
}

// Smallest table partition the Bloom-driven split above makes.
constexpr uint64_t kMinPartSlots = 256;

extern "C" int kc_table_create(kc_ctx *ctx, const kc_table_config *cfg, kc_table **out)
{
    if (!ctx || !cfg || !out) return KC_EINVAL;
    *out = nullptr;
    if (cfg->k < 1 || cfg->k > 128 || cfg->count_max == 0 || cfg->filter_max < cfg->filter_min) return KC_EINVAL;
    if (cfg->count_path > KC_PATH_PARTITIONED) return KC_EINVAL;
    if (cfg->capacity && !is_pow2(cfg->capacity)) return KC_EINVAL;
    if (cfg->bloom_blocks && !is_pow2(cfg->bloom_blocks)) return KC_EINVAL;
    if (!cfg->capacity && !cfg->max_distinct) return KC_EINVAL;
    KC_HIP(hipSetDevice(ctx->device));

    kc_table *t = new (std::nothrow) kc_table();
    if (!t) return KC_ENOMEM;
    try {
        t->hcounters.resize(KC_COUNTER_WORDS);
        t->hseg_off.resize(KC_LOW_SEGS + 1);
    } catch (...) {
        delete t;
        return KC_ENOMEM;
    }
    t->epoch = 1;
    t->low_idx_epoch = 0;
    // test-only fault injection, see kc_internal.h
    if (const char *f = std::getenv("KC_TEST_SKEW_COUNTS")) t->fault_skew = (uint32_t)std::atoi(f);
    t->ctx = ctx;
    t->cfg = *cfg;
    // key words inside the table: k = 64 keys (kmer<2>) take the four-word slots
    t->nw = cfg->k <= 32 ? 1 : cfg->k <= 63 ? 2 : 4;
    t->cap = cfg->capacity ? cfg->capacity : std::max<uint64_t>(1024, next_pow2(cfg->max_distinct + cfg->max_distinct / 2));
    const uint64_t hint = cfg->max_distinct ? cfg->max_distinct : t->cap / 2;
    // partitions of the size the LDS counting pass holds; a Bloom filter
    // larger than the table's partitions carry (error-rich input: many more
    // distinct keys than keys seen twice) splits the table into more, smaller
    // partitions instead, so that each Bloom part still fits the LDS copy
    const uint64_t lds_slots = kc_lds_part_slots(t->nw, cfg->bloom != 0);
    uint64_t n_parts = t->cap > lds_slots ? t->cap / lds_slots : 1;
    if (cfg->bloom && cfg->bloom_blocks) {
        const uint64_t want = cfg->bloom_blocks / kc_lds_part_bloom(t->nw);
        n_parts = std::max(n_parts, std::min(want, std::min<uint64_t>(t->cap / kMinPartSlots, 1ull << 18)));
        n_parts = std::max<uint64_t>(n_parts, 1);
    }
    t->part_bits = ilog2(n_parts);
    t->part_log2 = ilog2(t->cap) - t->part_bits;
    if (cfg->bloom) {
        t->bloom_blocks = cfg->bloom_blocks ? cfg->bloom_blocks : n_parts * kc_lds_part_bloom(t->nw);
        t->bloom_blocks = std::max<uint64_t>(t->bloom_blocks, n_parts);
        t->bloom_part_log2 = ilog2(t->bloom_blocks) - t->part_bits;
        // Appends go to segments hashed from block, wave and step
        // (first_segment), so the segments fill evenly up to a few appends:
        // one append moves at most 64 FIRSTs (1024 in the partition pass).
        // 1/8 slack plus 4096 keys; a FIRST that still finds its segment full
        // is counted in the table instead.
        const uint64_t want = cfg->low_capacity ? cfg->low_capacity : hint;
        t->seg_cap = (want + want / 8 + KC_LOW_SEGS - 1) / KC_LOW_SEGS + 4096;
        t->low_cap = t->seg_cap * KC_LOW_SEGS;
    }
    t->n_tiles = (t->cap + kExportTile - 1) / kExportTile;
    const uint64_t slot_bytes = 8ull * (t->nw == 1 ? 2 : t->nw == 2 ? 4 : 8);
    const uint64_t need = t->cap * slot_bytes + 8 * t->bloom_blocks + 8ull * t->nw * t->low_cap +
                          8 * KC_COUNTER_WORDS + 8 * export_scan_words(t->n_tiles);
    size_t free_b = 0, total_b = 0;
    KC_HIP(hipMemGetInfo(&free_b, &total_b));
    if (need > free_b) {
        delete t;
        return KC_ENOMEM;
    }
    bool ok = hipMalloc((void **)&t->slots, t->cap * slot_bytes) == hipSuccess &&
              hipMalloc((void **)&t->counters, 8 * KC_COUNTER_WORDS) == hipSuccess &&
              hipMalloc((void **)&t->tile_counts, 8 * export_scan_words(t->n_tiles)) == hipSuccess &&
              hipMalloc((void **)&t->seg_off, 8 * (KC_LOW_SEGS + 1)) == hipSuccess;
    if (ok && t->bloom_blocks) ok = hipMalloc((void **)&t->bloom, 8 * t->bloom_blocks) == hipSuccess;
    if (ok && t->low_cap) ok = hipMalloc((void **)&t->low, 8ull * t->nw * t->low_cap) == hipSuccess;
    if (!ok) {
        kc_table_destroy(t);
        return KC_ENOMEM;
    }
    int rc = table_init_device(t);
    if (rc != KC_OK) {
        kc_table_destroy(t);
        return rc;
    }
    *out = t;
    return KC_OK;
}

extern "C" void kc_table_destroy(kc_table *t)
{
    if (!t) return;
    (void)hipSetDevice(t->ctx->device);
    (void)hipStreamSynchronize(t->ctx->stream);
    for (void *p : {(void *)t->slots, (void *)t->bloom, (void *)t->low, (void *)t->counters, (void *)t->seg_off,
                    (void *)t->tile_counts, (void *)t->low_idx})
        if (p) (void)hipFree(p);
    kc_part_free(t);
    for (int i = 0; i < 2; ++i) {
        if (t->pinned[i]) (void)hipHostFree(t->pinned[i]);
        if (t->dstage[i]) (void)hipFree(t->dstage[i]);
        if (t->stage_done[i]) (void)hipEventDestroy(t->stage_done[i]);
    }
    if (t->copy_done) (void)hipEventDestroy(t->copy_done);
    if (t->query_ready) (void)hipEventDestroy(t->query_ready);
    if (t->query_done) (void)hipEventDestroy(t->query_done);
    delete t;
}

// Staging and partition scratch are kept between counting calls for reuse;
// a table that is only queried from now on gives them back.
extern "C" int kc_table_trim(kc_table *t)
{
    if (!t) return KC_EINVAL;
    KC_HIP(hipSetDevice(t->ctx->device));
    KC_HIP(hipStreamSynchronize(t->ctx->stream));
    kc_part_free(t);
    for (int i = 0; i < 2; ++i) {
        if (t->pinned[i]) KC_HIP(hipHostFree(t->pinned[i]));
        if (t->dstage[i]) KC_HIP(hipFree(t->dstage[i]));
        if (t->stage_done[i]) KC_HIP(hipEventDestroy(t->stage_done[i]));
        t->pinned[i] = nullptr;
        t->dstage[i] = nullptr;
        t->stage_done[i] = nullptr;
    }
    if (t->copy_done) KC_HIP(hipEventDestroy(t->copy_done));
    t->copy_done = nullptr;
    t->stage_bytes = 0;
    return KC_OK;
}

// A table the partitioned pass can count is cleared lazily: only the
// counters are zeroed now; the next partitioned pass starts its partitions
// empty in LDS and writes every slot and Bloom word (no read of the old
// ones), and any other access writes the empty table first
// (kc_table_materialize).
extern "C" int kc_table_clear(kc_table *t)
{
    if (!t) return KC_EINVAL;
    KC_HIP(hipSetDevice(t->ctx->device));
    if (!kc_part_eligible(t)) return table_init_device(t);
    const int rc = kc_launch_fill(t->ctx, t->counters, KC_COUNTER_WORDS, 0);
    if (rc != KC_OK) return rc;
    t->fresh = true;
    kc_table_touch(t);
    return KC_OK;
}

// Batches of at least this many bytes / keys take the partitioned path
// under KC_PATH_AUTO.
constexpr uint64_t kPartMinBytes = 32ull << 20;
constexpr uint64_t kPartMinKeys = 4ull << 20;

static bool use_partitioned(const kc_table *t, uint64_t size, uint64_t min_size)
{
    const uint32_t p = t->cfg.count_path != KC_PATH_AUTO ? t->cfg.count_path : t->tune.path;
    if (p == KC_PATH_DIRECT || !kc_part_eligible(t)) return false;
    return p == KC_PATH_PARTITIONED || size >= min_size;
}

// Direct path: the fused count kernel over `src` with `grid` blocks on st.
int kc_direct_enqueue(const kc_table *t, const SeqSrc &src, uint32_t grid, hipStream_t st)
{
    const TableView v = view_of(const_cast<kc_table *>(t));
    const bool bloom = low_list_active(t);
    const dim3 g(grid);
    if (t->nw == 1) {
        if (bloom) hipLaunchKernelGGL((k_count_seq<1, true>), g, dim3(kBlock), 0, st, src, v);
        else hipLaunchKernelGGL((k_count_seq<1, false>), g, dim3(kBlock), 0, st, src, v);
    } else if (t->nw == 2) {
        if (bloom) hipLaunchKernelGGL((k_count_seq<2, true>), g, dim3(kBlock), 0, st, src, v);
        else hipLaunchKernelGGL((k_count_seq<2, false>), g, dim3(kBlock), 0, st, src, v);
    } else {
        if (bloom) hipLaunchKernelGGL((k_count_seq<4, true>), g, dim3(kBlock), 0, st, src, v);
        else hipLaunchKernelGGL((k_count_seq<4, false>), g, dim3(kBlock), 0, st, src, v);
    }
    KC_HIP(hipGetLastError());
    return KC_OK;
}

static int launch_count(kc_table *t, const uint8_t *d_seq, uint64_t nbytes, uint64_t emit_from)
{
    if (!nbytes) return KC_OK;
    kc_table_touch(t);
    const TableView v = view_of(t);
    if (use_partitioned(t, nbytes, kPartMinBytes)) return kc_part_count(t, d_seq, nbytes, emit_from, nullptr, 0, v);
    int rc = kc_table_materialize(t);
    if (rc != KC_OK) return rc;
    const uint64_t blocks = (nbytes + kTile - 1) / kTile;
    return kc_direct_enqueue(t, SeqSrc{d_seq, nbytes, emit_from, nullptr}, (uint32_t)std::min<uint64_t>(blocks, 1u << 30),
                             t->ctx->stream);
}

extern "C" int kc_table_count_seq(kc_table *t, const uint8_t *d_seq, uint64_t nbytes)
{
    if (!t || (!d_seq && nbytes)) return KC_EINVAL;
    if (reinterpret_cast<uintptr_t>(d_seq) % 16) return KC_EINVAL;
    return launch_count(t, d_seq, nbytes, 0);
}

extern "C" int kc_table_count_seq_host(kc_table *t, const uint8_t *h_seq, uint64_t nbytes)
{
    if (!t || (!h_seq && nbytes)) return KC_EINVAL;
    if (!nbytes) return KC_OK;
    constexpr uint64_t kStage = 256ull << 20;
    constexpr uint64_t kHalo = 128;  // >= k-1 for k <= 128, multiple of 16
    hipStream_t s = t->ctx->stream;
    KC_HIP(hipSetDevice(t->ctx->device));
    // Page-locked sources are copied to the device directly; pageable ones
    // go through two page-locked staging buffers.
    hipPointerAttribute_t attr{};
    const bool direct = hipPointerGetAttributes(&attr, h_seq) == hipSuccess && attr.type == hipMemoryTypeHost;
    (void)hipGetLastError();  // a pageable pointer leaves hipErrorInvalidValue behind
    // sized for the largest call so far (small jobs keep small buffers), at
    // most kStage; a larger call grows them
    const uint64_t want = std::min<uint64_t>(kStage, (nbytes + 15) / 16 * 16) + kHalo;
    if (!t->dstage[0] || (!direct && !t->pinned[0]) || want > t->stage_bytes) {
        if (t->dstage[0] && want > t->stage_bytes) {
            KC_HIP(hipStreamSynchronize(s));
            for (int i = 0; i < 2; ++i) {
                if (t->pinned[i]) KC_HIP(hipHostFree(t->pinned[i]));
                KC_HIP(hipFree(t->dstage[i]));
                t->pinned[i] = t->dstage[i] = nullptr;
            }
        }
        if (!t->dstage[0]) t->stage_bytes = want;
        for (int i = 0; i < 2; ++i) {
            if (!direct && !t->pinned[i])
                KC_HIP(hipHostMalloc((void **)&t->pinned[i], t->stage_bytes, hipHostMallocDefault));
            if (!t->dstage[i]) KC_HIP(hipMalloc((void **)&t->dstage[i], t->stage_bytes));
            if (!t->stage_done[i]) KC_HIP(hipEventCreateWithFlags(&t->stage_done[i], hipEventDisableTiming));
        }
        if (!t->copy_done) KC_HIP(hipEventCreateWithFlags(&t->copy_done, hipEventDisableTiming));
    }
    uint64_t pos = 0;
    int buf = 0;
    const uint64_t stage_data = t->stage_bytes - kHalo;
    while (pos < nbytes) {
        const uint64_t chunk = std::min(stage_data, nbytes - pos);
        const uint64_t halo = std::min(kHalo, pos);
        // the staging buffers are reused two chunks later: wait for their use
        KC_HIP(hipEventSynchronize(t->stage_done[buf]));
        if (direct) {
            KC_HIP(hipMemcpyAsync(t->dstage[buf], h_seq + pos - halo, chunk + halo, hipMemcpyHostToDevice, s));
            if (pos + chunk == nbytes) KC_HIP(hipEventRecord(t->copy_done, s));
        } else {
            std::memcpy(t->pinned[buf], h_seq + pos - halo, chunk + halo);
            KC_HIP(hipMemcpyAsync(t->dstage[buf], t->pinned[buf], chunk + halo, hipMemcpyHostToDevice, s));
        }
        int rc = launch_count(t, t->dstage[buf], chunk + halo, halo);
        if (rc != KC_OK) return rc;
        KC_HIP(hipEventRecord(t->stage_done[buf], s));
        pos += chunk;
        buf ^= 1;
    }
    // The caller may reuse h_seq on return: wait for the copies out of it.
    if (direct) KC_HIP(hipEventSynchronize(t->copy_done));
    return KC_OK;
}

extern "C" int kc_table_increment(kc_table *t, const uint64_t *d_keys, uint64_t n, const uint32_t *d_inc)
{
    if (!t || (!d_keys && n)) return KC_EINVAL;
    if (!n) return KC_OK;
    kc_table_touch(t);
    const TableView v = view_of(t);
    // increments add to the table directly (no Bloom split), as increment()
    // of the reference's maps does
    if (!d_inc && !t->bloom && use_partitioned(t, n, kPartMinKeys))
        return kc_part_count(t, nullptr, 0, 0, d_keys, n, v);
    int rc = kc_table_materialize(t);
    if (rc != KC_OK) return rc;
    const uint32_t g = grid_for(n, t->ctx->cu_count);
    if (t->nw == 1) hipLaunchKernelGGL(k_increment<1>, dim3(g), dim3(kBlock), 0, t->ctx->stream, d_keys, n, d_inc, v);
    else if (t->nw == 2) hipLaunchKernelGGL(k_increment<2>, dim3(g), dim3(kBlock), 0, t->ctx->stream, d_keys, n, d_inc, v);
    else hipLaunchKernelGGL(k_increment<4>, dim3(g), dim3(kBlock), 0, t->ctx->stream, d_keys, n, d_inc, v);
    KC_HIP(hipGetLastError());
    return KC_OK;
}

// Brings the FIRST-list index up to date with the table's counts (nothing
// to do when the counts have not changed since it was built).  *mask is the
// index's slot mask, 0 when lookups need no index (no Bloom split, or an
// empty list).  Writes only the index: the table, its list and its counters
// are left as they are.
static int ensure_low_index(kc_table *t, uint64_t *mask)
{
    *mask = 0;
    if (!low_list_active(t)) return KC_OK;
    if (t->low_idx_epoch == t->epoch) {
        *mask = t->low_idx_mask;
        return KC_OK;
    }
    HostCounters hc;
    int rc = read_counters(t, hc);
    if (rc != KC_OK) return rc;
    if (hc.stat[KC_STAT_ERRORS]) return KC_EFULL;
    t->low_idx_mask = 0;
    if (hc.low) {
        const uint64_t cap = std::max<uint64_t>(1024, next_pow2(2 * hc.low));
        const uint64_t iw = t->nw == 4 ? 5 : t->nw;  // index words per key (idx_words)
        const uint64_t words = cap * iw + 1;  // + the error word
        if (cap > t->low_idx_cap) {
            KC_HIP(hipStreamSynchronize(t->ctx->stream));
            if (t->low_idx) KC_HIP(hipFree(t->low_idx));
            t->low_idx = nullptr;
            t->low_idx_cap = 0;
            if (hipMalloc((void **)&t->low_idx, words * 8) != hipSuccess) {
                t->low_idx = nullptr;
                return KC_ENOMEM;
            }
            t->low_idx_cap = cap;
        }
        if ((rc = kc_launch_fill(t->ctx, t->low_idx, cap * iw, t->nw == 4 ? 0 : KC_EMPTY)) != KC_OK) return rc;
        if ((rc = kc_launch_fill(t->ctx, t->low_idx + cap * iw, 1, 0)) != KC_OK) return rc;
        uint64_t most = 0;
        for (int sgi = 0; sgi < KC_LOW_SEGS; ++sgi) most = std::max(most, hc.seg_fill[sgi]);
        const TableView v = view_of(t);
        const dim3 g(kc_low_chunks(most, 4 * kBlock), KC_LOW_SEGS);
        if (t->nw == 1) hipLaunchKernelGGL(k_low_index_build<1>, g, dim3(kBlock), 0, t->ctx->stream, v, t->low_idx, cap - 1);
        else if (t->nw == 2) hipLaunchKernelGGL(k_low_index_build<2>, g, dim3(kBlock), 0, t->ctx->stream, v, t->low_idx, cap - 1);
        else hipLaunchKernelGGL(k_low_index_build<4>, g, dim3(kBlock), 0, t->ctx->stream, v, t->low_idx, cap - 1);
        KC_HIP(hipGetLastError());
        uint64_t err = 0;
        KC_HIP(hipMemcpyAsync(&err, t->low_idx + cap * iw, 8, hipMemcpyDeviceToHost, t->ctx->stream));
        KC_HIP(hipStreamSynchronize(t->ctx->stream));
        if (err) return KC_EDEVICE;  // a set at load 1/2 always has room: a broken list
        t->low_idx_mask = cap - 1;
    }
    t->low_idx_epoch = t->epoch;
    *mask = t->low_idx_mask;
    return KC_OK;
}

extern "C" int kc_table_query_on(kc_table *t, kc_ctx *ctx, const uint64_t *d_keys, uint64_t n, uint32_t *d_counts)
{
    if (!t || !ctx || ctx->device != t->ctx->device || (n && (!d_keys || !d_counts))) return KC_EINVAL;
    KC_HIP(hipSetDevice(t->ctx->device));
    const bool other = ctx != t->ctx;
    if (other && !t->query_done) KC_HIP(hipEventCreateWithFlags(&t->query_done, hipEventDisableTiming));
    // a lazily cleared table is written empty on its own stream first, and
    // only then does the lookup stream wait for that stream
    const bool was_fresh = t->fresh;
    uint64_t imask = 0;
    int rc = kc_table_materialize(t);
    if (rc == KC_OK) rc = ensure_low_index(t, &imask);  // synchronous when it builds the index
    if (rc != KC_OK) return rc;
    if (!n) return KC_OK;
    hipStream_t st = ctx->stream;
    if (other && was_fresh) {
        if (!t->query_ready) KC_HIP(hipEventCreateWithFlags(&t->query_ready, hipEventDisableTiming));
        KC_HIP(hipEventRecord(t->query_ready, t->ctx->stream));
        KC_HIP(hipStreamWaitEvent(st, t->query_ready, 0));
    }
    const TableView v = view_of(t);
    const uint32_t g = grid_for(n, ctx->cu_count);
    const uint64_t *idx = t->low_idx;
    if (t->nw == 1) {
        if (imask) hipLaunchKernelGGL((k_query<1, true>), dim3(g), dim3(kBlock), 0, st, d_keys, n, d_counts, v, idx, imask);
        else hipLaunchKernelGGL((k_query<1, false>), dim3(g), dim3(kBlock), 0, st, d_keys, n, d_counts, v, idx, imask);
    } else if (t->nw == 2) {
        if (imask) hipLaunchKernelGGL((k_query<2, true>), dim3(g), dim3(kBlock), 0, st, d_keys, n, d_counts, v, idx, imask);
        else hipLaunchKernelGGL((k_query<2, false>), dim3(g), dim3(kBlock), 0, st, d_keys, n, d_counts, v, idx, imask);
    } else {
        if (imask) hipLaunchKernelGGL((k_query<4, true>), dim3(g), dim3(kBlock), 0, st, d_keys, n, d_counts, v, idx, imask);
        else hipLaunchKernelGGL((k_query<4, false>), dim3(g), dim3(kBlock), 0, st, d_keys, n, d_counts, v, idx, imask);
    }
    KC_HIP(hipGetLastError());
    if (other) {  // later work on the table (a count, its destruction) waits for the lookup
        KC_HIP(hipEventRecord(t->query_done, st));
        KC_HIP(hipStreamWaitEvent(t->ctx->stream, t->query_done, 0));
    }
    return KC_OK;
}

extern "C" int kc_table_query(kc_table *t, const uint64_t *d_keys, uint64_t n, uint32_t *d_counts)
{
    if (!t) return KC_EINVAL;
    return kc_table_query_on(t, t->ctx, d_keys, n, d_counts);
}

extern "C" int kc_table_stats(kc_table *t, kc_stats *out)
{
    if (!t || !out) return KC_EINVAL;
    HostCounters hc;
    int rc = read_counters(t, hc);
    if (rc != KC_OK) return rc;
    out->kmers = hc.stat[KC_STAT_KMERS];
    out->distinct = hc.stat[KC_STAT_DISTINCT];
    out->low = hc.low;
    out->capacity = t->cap;
    out->errors = hc.stat[KC_STAT_ERRORS];
    return KC_OK;
}

extern "C" int kc_table_front_redos(kc_table *t, uint64_t *out)
{
    if (!t || !out) return KC_EINVAL;
    HostCounters hc;
    int rc = read_counters(t, hc);
    if (rc != KC_OK) return rc;
    *out = hc.stat[KC_STAT_FRONT_REDOS];
    return KC_OK;
}

extern "C" int kc_table_export_device(kc_table *t, uint8_t *d_high, uint64_t high_cap, uint64_t *high_bytes,
                                      uint8_t *d_low, uint64_t low_cap, uint64_t *low_bytes)
{
    if (!t || !high_bytes || !low_bytes) return KC_EINVAL;
    kc_ctx *c = t->ctx;
    HostCounters hc;
    int rc = kc_table_materialize(t);
    if (rc == KC_OK) rc = read_counters(t, hc);
    if (rc != KC_OK) return rc;
    if (hc.stat[KC_STAT_ERRORS]) return KC_EFULL;
    const PackSpec ps = pack_spec(t);
    const uint64_t hrec = ps.kmer_bytes + ps.count_bytes;
    const bool list = low_list_active(t);
    const uint64_t n_list = list ? hc.low : 0;
    const TableView v = view_of(t);
    const uint32_t g = (uint32_t)std::min<uint64_t>(t->n_tiles, (uint64_t)c->cu_count * 8);
    const uint64_t N = t->n_tiles;
    const ExportScan es = export_scan_of(t->tile_counts, N);
    if (t->nw == 1) hipLaunchKernelGGL(k_export_count<1>, dim3(g), dim3(kBlock), 0, c->stream, v, !list, N, es.cnt);
    else if (t->nw == 2) hipLaunchKernelGGL(k_export_count<2>, dim3(g), dim3(kBlock), 0, c->stream, v, !list, N, es.cnt);
    else hipLaunchKernelGGL(k_export_count<4>, dim3(g), dim3(kBlock), 0, c->stream, v, !list, N, es.cnt);
    KC_HIP(hipGetLastError());
    if ((rc = kc_scan_u32_on(c->stream, es.cnt, 2 * N, es.off, es.sums)) != KC_OK) return rc;
    hipLaunchKernelGGL(k_export_totals, dim3(1), dim3(64), 0, c->stream, es.off, es.cnt, N);
    KC_HIP(hipGetLastError());
    uint64_t tot[2];
    KC_HIP(hipMemcpyAsync(tot, es.off + 2 * N, 16, hipMemcpyDeviceToHost, c->stream));
    KC_HIP(hipStreamSynchronize(c->stream));
    *high_bytes = tot[0] * hrec;
    *low_bytes = (tot[1] + n_list) * ps.kmer_bytes;
    if (!d_high || !d_low) return KC_OK;
    if (*high_bytes > high_cap || *low_bytes > low_cap) return KC_EINVAL;
    if (t->nw == 1) hipLaunchKernelGGL(k_export_write<1>, dim3(g), dim3(kBlock), 0, c->stream, v, ps, !list, N, es.off, d_high, d_low, n_list);
    else if (t->nw == 2) hipLaunchKernelGGL(k_export_write<2>, dim3(g), dim3(kBlock), 0, c->stream, v, ps, !list, N, es.off, d_high, d_low, n_list);
    else hipLaunchKernelGGL(k_export_write<4>, dim3(g), dim3(kBlock), 0, c->stream, v, ps, !list, N, es.off, d_high, d_low, n_list);
    KC_HIP(hipGetLastError());
    if (n_list) {
        // chunks per segment: enough blocks for the fullest segment, capped
        uint64_t most = 0;
        for (int sgi = 0; sgi < KC_LOW_SEGS; ++sgi) most = std::max(most, hc.seg_fill[sgi]);
        const dim3 g2((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((most + kExportTile - 1) / kExportTile, 64)),
                      KC_LOW_SEGS);
        if (t->nw == 1) hipLaunchKernelGGL(k_export_low_list<1>, g2, dim3(kBlock), 0, c->stream, v, ps, d_low);
        else if (t->nw == 2) hipLaunchKernelGGL(k_export_low_list<2>, g2, dim3(kBlock), 0, c->stream, v, ps, d_low);
        else hipLaunchKernelGGL(k_export_low_list<4>, g2, dim3(kBlock), 0, c->stream, v, ps, d_low);
        KC_HIP(hipGetLastError());
    }
    return KC_OK;
}

static int write_file(const std::string &path, const void *data, uint64_t n, bool append = false)
{
    FILE *f = std::fopen(path.c_str(), append ? "ab" : "wb");
    if (!f) return KC_EIO;
    const bool ok = n == 0 || std::fwrite(data, 1, n, f) == n;
    return (std::fclose(f) == 0 && ok) ? KC_OK : KC_EIO;
}

static int export_dir(kc_table *t, const char *dir, bool append)
{
    if (!t || !dir) return KC_EINVAL;
    std::string d(dir);
    if (!d.empty() && d.back() != '/') d += '/';
    uint64_t hb = 0, lb = 0;
    int rc = kc_table_export_device(t, nullptr, 0, &hb, nullptr, 0, &lb);
    if (rc != KC_OK) return rc;
    uint8_t *dh = nullptr, *dl = nullptr;
    KC_HIP(hipMalloc((void **)&dh, std::max<uint64_t>(hb, 16)));
    if (hipMalloc((void **)&dl, std::max<uint64_t>(lb, 16)) != hipSuccess) {
        (void)hipFree(dh);
        return KC_ENOMEM;
    }
    uint64_t hb2 = 0, lb2 = 0;
    rc = kc_table_export_device(t, dh, hb, &hb2, dl, lb, &lb2);
    std::vector<uint8_t> hh, hl;
    if (rc == KC_OK) {
        try {
            hh.resize(hb2);
            hl.resize(lb2);
        } catch (...) {
            rc = KC_ENOMEM;
        }
    }
    // on the table's stream: the context's stream does not order with the
    // null stream, so a plain hipMemcpy could read the records before the
    // export kernels have written them
    hipStream_t st = t->ctx->stream;
    if (rc == KC_OK && (hipMemcpyAsync(hh.data(), dh, hb2, hipMemcpyDeviceToHost, st) != hipSuccess ||
                        hipMemcpyAsync(hl.data(), dl, lb2, hipMemcpyDeviceToHost, st) != hipSuccess ||
                        hipStreamSynchronize(st) != hipSuccess))
        rc = KC_EDEVICE;
    (void)hipFree(dh);
    (void)hipFree(dl);
    if (rc != KC_OK) return rc;
    const uint32_t infos[2] = {t->cfg.k, t->cfg.count_max};
    if ((rc = write_file(d + "infos.bin", infos, sizeof infos)) != KC_OK) return rc;
    if ((rc = write_file(d + "high.bin", hh.data(), hh.size(), append)) != KC_OK) return rc;
    return write_file(d + "low.bin", hl.data(), hl.size(), append);
}

extern "C" int kc_table_export_dir(kc_table *t, const char *dir)
{
    return kc_guard([&] { return export_dir(t, dir, false); });
}

// high.bin and low.bin are record streams, so the exports of several passes
// (or tables) concatenate into one valid export.
extern "C" int kc_table_export_dir_append(kc_table *t, const char *dir)
{
    return kc_guard([&] { return export_dir(t, dir, true); });
}

extern "C" int kc_table_set_pass(kc_table *t, uint32_t pass, uint32_t n_passes)
{
    if (!t || n_passes == 0 || n_passes > 256 || !is_pow2(n_passes) || pass >= n_passes) return KC_EINVAL;
    t->pass_bits = ilog2(n_passes);
    t->pass = pass;
    return KC_OK;
}

// histogram_tool precise over the table's exported view into host bins,
// with the device bins in gb (at least max - min + 1 words; cleared here by
// a fill kernel).  Allocates nothing: the host counters are the table's own.
int kc_histogram_into(kc_table *t, uint32_t min_freq, uint32_t max_freq, int64_t *h_bins, unsigned long long *gb)
{
    kc_ctx *c = t->ctx;
    HostCounters hc;
    int rc = kc_table_materialize(t);
    if (rc == KC_OK) rc = read_counters(t, hc);
    if (rc != KC_OK) return rc;
    if (hc.stat[KC_STAT_ERRORS]) return KC_EFULL;
    const uint64_t nb = (uint64_t)max_freq - min_freq + 1;
    if ((rc = kc_launch_fill(c, reinterpret_cast<uint64_t *>(gb), nb, 0)) != KC_OK) return rc;
    const bool list = low_list_active(t);
    HistSpec hs{min_freq, max_freq, !list};
    const TableView v = view_of(t);
    const uint32_t g = grid_for(t->cap, c->cu_count);
    if (t->nw == 1) hipLaunchKernelGGL(k_hist_table<1>, dim3(g), dim3(kBlock), 0, c->stream, v, hs, gb);
    else if (t->nw == 2) hipLaunchKernelGGL(k_hist_table<2>, dim3(g), dim3(kBlock), 0, c->stream, v, hs, gb);
    else hipLaunchKernelGGL(k_hist_table<4>, dim3(g), dim3(kBlock), 0, c->stream, v, hs, gb);
    if (list && hc.low) {
        uint64_t most = 0;
        for (int sgi = 0; sgi < KC_LOW_SEGS; ++sgi) most = std::max(most, hc.seg_fill[sgi]);
        const dim3 g2(kc_low_chunks(most, 4 * kBlock), KC_LOW_SEGS);
        if (t->nw == 1) hipLaunchKernelGGL(k_hist_low_list<1>, g2, dim3(kBlock), 0, c->stream, v, hs, gb);
        else if (t->nw == 2) hipLaunchKernelGGL(k_hist_low_list<2>, g2, dim3(kBlock), 0, c->stream, v, hs, gb);
        else hipLaunchKernelGGL(k_hist_low_list<4>, g2, dim3(kBlock), 0, c->stream, v, hs, gb);
    }
    KC_HIP(hipGetLastError());
    KC_HIP(hipMemcpyAsync(h_bins, gb, nb * 8, hipMemcpyDeviceToHost, c->stream));
    KC_HIP(hipStreamSynchronize(c->stream));
    return KC_OK;
}

extern "C" int kc_table_histogram(kc_table *t, uint32_t min_freq, uint32_t max_freq, int64_t *h_bins)
{
    if (!t || !h_bins || max_freq < min_freq) return KC_EINVAL;
    KC_HIP(hipSetDevice(t->ctx->device));
    const uint64_t nb = (uint64_t)max_freq - min_freq + 1;
    unsigned long long *gb = nullptr;
    if (hipMalloc((void **)&gb, nb * 8) != hipSuccess) return KC_ENOMEM;
    const int rc = kc_histogram_into(t, min_freq, max_freq, h_bins, gb);
    (void)hipStreamSynchronize(t->ctx->stream);
    (void)hipFree(gb);
    return rc;
}

extern "C" int kc_hash_keys(kc_ctx *ctx, int kind, const uint64_t *d_keys, uint64_t n, uint32_t nw, uint64_t *d_out)
{
    if (!ctx || (n && (!d_keys || !d_out)) || kind < 0 || kind > KC_HASH_FUNC) return KC_EINVAL;
    if (!n) return KC_OK;
    const uint32_t g = grid_for(n, ctx->cu_count);
    switch (nw) {
    case 1: hipLaunchKernelGGL(k_hash<1>, dim3(g), dim3(kBlock), 0, ctx->stream, kind, d_keys, n, d_out); break;
    case 2: hipLaunchKernelGGL(k_hash<2>, dim3(g), dim3(kBlock), 0, ctx->stream, kind, d_keys, n, d_out); break;
    case 4: hipLaunchKernelGGL(k_hash<4>, dim3(g), dim3(kBlock), 0, ctx->stream, kind, d_keys, n, d_out); break;
    default: return KC_EINVAL;
    }
    KC_HIP(hipGetLastError());
    return KC_OK;
}

extern "C" int kc_bloom_probe_keys(kc_ctx *ctx, const uint64_t *d_keys, uint64_t n, uint32_t nw, uint64_t capacity,
                                   uint64_t *d_block, uint64_t *d_mask)
{
    if (!ctx || !is_pow2(capacity) || (n && (!d_keys || !d_block || !d_mask))) return KC_EINVAL;
    if (!n) return KC_OK;
    const uint32_t g = grid_for(n, ctx->cu_count);
    switch (nw) {
    case 1: hipLaunchKernelGGL(k_bloom_probe<1>, dim3(g), dim3(kBlock), 0, ctx->stream, d_keys, n, capacity - 1, d_block, d_mask); break;
    case 2: hipLaunchKernelGGL(k_bloom_probe<2>, dim3(g), dim3(kBlock), 0, ctx->stream, d_keys, n, capacity - 1, d_block, d_mask); break;
    case 4: hipLaunchKernelGGL(k_bloom_probe<4>, dim3(g), dim3(kBlock), 0, ctx->stream, d_keys, n, capacity - 1, d_block, d_mask); break;
    default: return KC_EINVAL;
    }
    KC_HIP(hipGetLastError());
    return KC_OK;
}

kc::TableView kc_table_view(kc_table *t) { return view_of(t); }

// FIRST-list size (uploads the segment offsets) and whether the list is
// active; KC_EFULL when counting overflowed.
int kc_table_low_count(kc_table *t, uint64_t *n_low, bool *list_active)
{
    HostCounters hc;
    int rc = kc_table_materialize(t);
    if (rc == KC_OK) rc = read_counters(t, hc);
    if (rc != KC_OK) return rc;
    if (hc.stat[KC_STAT_ERRORS]) return KC_EFULL;
    *list_active = low_list_active(t);
    *n_low = *list_active ? hc.low : 0;
    return KC_OK;
}
