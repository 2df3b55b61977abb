// Device-side pieces shared by the counting kernels: the table view and its
// partitioned slot indexing, the global-atomic insert/lookup, the Bloom
// singleton split, the rolling canonical encoder and counter helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kc_hash_dev.h"
#include "kc_internal.h"

namespace kc {

// The table is cut into 2^part_bits partitions of part_slots slots (and
// the Bloom filter into partitions of bloom_part 64-bit blocks).  A key's
// partition is the top part_bits of coarse_mix; inside the partition the
// slot probe start, the Bloom block and the Bloom mask come from slot_mix
// (KeyHash below), and the probe wraps within the partition.
struct TableView {
    uint64_t *slots;
    uint64_t mask;               // total slots - 1
    uint32_t part_bits;          // log2(partitions)
    uint32_t part_log2;          // log2(part_slots)
    uint64_t *bloom;             // nullptr when the Bloom split is off
    uint64_t bloom_mask;         // total blocks - 1
    uint32_t bloom_part_log2;    // log2(blocks per partition)
    uint64_t *low;
    uint64_t seg_cap;            // keys per FIRST-list segment
    const uint64_t *seg_off;     // KC_LOW_SEGS + 1 exclusive offsets (after counting)
    unsigned long long *counters;
    uint32_t k;
    uint32_t count_max;
    uint32_t filter_min;
    uint32_t filter_max;
    uint32_t pass_bits;          // multi-pass counting: log2(passes), 0 = one pass
    uint32_t pass;               // the pass whose k-mers are counted
    uint32_t rw;                 // words of a caller's key (kmer<N>: 1, 2 or 4; 2 at k = 64)
};

// Key i of a caller's array of rw-word keys (kmer<N>), widened to the
// table's NW words (k = 64 keys are kmer<2> outside and four words inside).
template <int NW>
__device__ __forceinline__ void caller_key(const uint64_t *keys, uint64_t i, uint32_t rw, uint64_t *key)
{
#pragma unroll
    for (int w = 0; w < NW; ++w) key[w] = (uint32_t)w < rw ? keys[i * rw + w] : 0;
}

// Slot layout by key words: {key, count} (k <= 32), {word 0, word 1, count,
// 0} (k = 33..63) and {words 0..3, count, state, 0, 0} (k = 64..128).  One-
// and two-word keys claim a slot with a CAS on word 0, whose all-ones value
// (KC_EMPTY) no canonical key of k <= 63 has.  Four-word keys can have an
// all-ones word 0 (the palindrome T^32 A^32 at k = 64, for one), so their
// slots carry a state word instead: 0 empty, 1 claimed, 2 key published.
template <int NW> constexpr int slot_words() { return NW == 1 ? 2 : NW == 2 ? 4 : 8; }
template <int NW> constexpr int count_word() { return NW == 1 ? 1 : NW == 2 ? 2 : 4; }
constexpr int kStateWord = 5;
constexpr uint64_t kSlotClaimed = 1, kSlotReady = 2;

// Whether a slot holds a key (after counting has finished).
template <int NW>
__device__ __forceinline__ bool slot_used(const uint64_t *s)
{
    if constexpr (NW == 4) return s[kStateWord] == kSlotReady;
    else return s[0] != KC_EMPTY;
}

// The counter's export filter (NewKmerTree.h:1303-1311, FinalDrainWriter.h:
// 95-104): a record with count c is left out of high.bin when c + 1 <
// filter_min or c > filter_max, with c + 1 computed in u32 as the reference
// does, so a count saturated at 2^32 - 1 wraps to 0 and is left out whenever
// filter_min >= 1.  (The tools' own range test is u64,
// HighFrequencyInsertThreadPool.h:161.)
__device__ __forceinline__ bool export_filtered(uint32_t c, const TableView &tv)
{
    return (uint32_t)(c + 1u) < tv.filter_min || c > tv.filter_max;
}

// A table of 2^L partitions splits L into a coarse digit (the top
// min(L, 8) bits) and a fine digit (the other L - min(L, 8) bits); the two
// scatter levels sort by them in turn.  Both come from one cheap mix of the
// key (32 bits; L <= 18): the coarse digit is its top bits, the fine digit
// the bits below.  The parse passes compute it for every window, so it is
// kept to one 32-bit multiply (two for two-word keys): each key word is
// folded to 32 bits (low half xor the high half rotated by 13), the fold is
// xor-shifted by 16 so that its high bits reach the product's top bits, and
// multiplied by an odd constant.  On repeat-rich and error-rich k = 21, 31
// and 63 keys its top 16 and 18 bits load the partitions as a Poisson law
// does (standard deviation / sqrt(mean) within 1% of 1, as for the 64-bit
// multiply-shift it replaced).
constexpr uint32_t kCoarseMaxBits = 8;

__device__ __forceinline__ uint32_t fold32(uint64_t w)
{
    const uint32_t hi = (uint32_t)(w >> 32);
    return (uint32_t)w ^ ((hi << 13) | (hi >> 19));
}

template <int NW>
__device__ __forceinline__ uint32_t coarse_mix(const uint64_t *key)
{
    uint32_t y = fold32(key[0]);
#pragma unroll
    for (int w = 1; w < NW; ++w) y = (y * 0x85EBCA77u) ^ fold32(key[w]);
    y ^= y >> 16;
    return y * 0x9E3779B1u;
}

__device__ __forceinline__ uint32_t coarse_bits(uint32_t part_bits)
{
    return part_bits < kCoarseMaxBits ? part_bits : kCoarseMaxBits;
}

__device__ __forceinline__ uint32_t coarse_digit(uint32_t mix, uint32_t l1) { return l1 ? mix >> (32 - l1) : 0; }

// The fine digit: the l2 bits of the mix below the coarse digit's l1.
__device__ __forceinline__ uint32_t fine_digit(uint32_t mix, uint32_t l1, uint32_t l2)
{
    return l2 ? (mix >> (32 - l1 - l2)) & ((1u << l2) - 1) : 0;
}

// Multi-pass counting (tables larger than HBM are counted in 2^pass_bits
// passes over the input, each holding one hash slice of the k-mers): a key
// belongs to pass (XXH3_64 >> (64 - pass_bits)), a hash independent of the
// partition mix, so every slice spreads over all partitions.  The hash is
// over the caller's key words (rw: a k = 64 key hashes as the 16 bytes of a
// kmer<2>, though the table keeps it in four words).
template <int NW>
__device__ __forceinline__ bool in_pass(const uint64_t *key, uint32_t pass_bits, uint32_t pass, uint32_t rw = NW)
{
    if (pass_bits == 0) return true;
    uint64_t h;
    if constexpr (NW == 4) h = rw == 2 ? xxh3_64<2>(key) : xxh3_64<4>(key);
    else h = xxh3_64<NW>(key);
    return (uint32_t)(h >> (64 - pass_bits)) == pass;
}

template <int NW>
__device__ __forceinline__ uint64_t part_of_key(const uint64_t *key, uint32_t part_bits)
{
    const uint32_t l1 = coarse_bits(part_bits), l2 = part_bits - l1;
    const uint32_t mix = coarse_mix<NW>(key);
    const uint64_t c = coarse_digit(mix, l1);
    return l2 ? (c << l2) | fine_digit(mix, l1, l2) : c;
}

// Where a key lives.  Its partition is (coarse digit, fine digit) as
// above.  Inside the partition, one slot_mix gives the Bloom block and mask
// (bloom_in_part) and the linear probe start (bits 44..63, which the Bloom
// probe does not use).
struct KeyHash {
    uint64_t part;
    uint64_t h;
};

template <int NW>
__device__ __forceinline__ KeyHash key_hash(const TableView &tv, const uint64_t *key)
{
    KeyHash kh;
    kh.part = part_of_key<NW>(key, tv.part_bits);
    kh.h = slot_mix<NW>(key);
    return kh;
}

__device__ __forceinline__ uint32_t probe_start(uint64_t h) { return (uint32_t)(h >> 44); }

// Slot index of probe step `probe` of a key.
__device__ __forceinline__ uint64_t slot_index(const TableView &tv, const KeyHash &kh, uint64_t probe)
{
    const uint64_t pmask = (1ULL << tv.part_log2) - 1;
    return (kh.part << tv.part_log2) | ((probe_start(kh.h) + probe) & pmask);
}

__device__ __forceinline__ uint64_t lanemask_lt()
{
    return (1ULL << __lane_id()) - 1ULL;
}

__device__ __forceinline__ uint64_t ld_agent(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void add_agent(uint64_t *p, uint64_t v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wave sum of a 32-bit statistic, then one atomic per wave into this
// block's lane of the statistic (lanes live on separate 128-byte lines).
__device__ __forceinline__ void stat_add(unsigned long long *counters, int stat, uint32_t v)
{
    uint64_t s = v;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (__lane_id() == 0 && s)
        atomicAdd(counters + ((uint64_t)stat * KC_LANES + (blockIdx.x % KC_LANES)) * KC_LINE_WORDS,
                  (unsigned long long)s);
}

// FIRST-list walks run on a grid (chunks, KC_LOW_SEGS): block (x, s) takes
// keys x*B + t, (x + gridDim.x)*B + t, ... of segment s, so every key is
// read from its segment directly (no search over the segment offsets).
__device__ __forceinline__ uint64_t low_seg_fill(const TableView &tv, uint32_t seg)
{
    return tv.seg_off[seg + 1] - tv.seg_off[seg];
}

__device__ __forceinline__ const uint64_t *low_seg_key(const TableView &tv, uint32_t seg, uint64_t r, int nw)
{
    return tv.low + ((uint64_t)seg * tv.seg_cap + r) * nw;
}

// ---- table insert / add ------------------------------------------------
// Open addressing, linear probing.  A slot is claimed with a 64-bit CAS on
// its first key word; counts are 64-bit atomic adds (saturation is applied
// when results are read, so min(true count, count_max) is exact).
template <int NW>
__device__ __forceinline__ bool table_add(const TableView &tv, const uint64_t *key, const KeyHash &kh, uint64_t inc,
                                          uint32_t &n_new, uint32_t &n_err)
{
    const uint64_t psize = 1ULL << tv.part_log2;
    if constexpr (NW == 4) {
        // Four-word key: the state word's CAS claims a slot, the owner writes
        // the key and publishes it with a release store; others wait for the
        // publication (bounded) and compare the key.
        for (uint64_t probe = 0; probe < psize; ++probe) {
            uint64_t *s = tv.slots + slot_index(tv, kh, probe) * slot_words<NW>();
            uint64_t st = __hip_atomic_load(s + kStateWord, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
            bool owner = false;
            if (st == 0) {
                st = atomicCAS((unsigned long long *)(s + kStateWord), 0ull, (unsigned long long)kSlotClaimed);
                owner = st == 0;
            }
            if (owner) {
#pragma unroll
                for (int w = 0; w < 4; ++w) s[w] = key[w];
                __hip_atomic_store(s + kStateWord, kSlotReady, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
            __builtin_amdgcn_wave_barrier();
            if (owner) {
                ++n_new;
                add_agent(s + count_word<NW>(), inc);
                return true;
            }
            for (int spin = 0; st == kSlotClaimed && spin < (1 << 16); ++spin) {
                __builtin_amdgcn_s_sleep(1);
                st = __hip_atomic_load(s + kStateWord, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
            }
            if (st != kSlotReady) {
                ++n_err;
                return false;
            }
            if (s[0] == key[0] && s[1] == key[1] && s[2] == key[2] && s[3] == key[3]) {
                add_agent(s + count_word<NW>(), inc);
                return true;
            }
        }
        ++n_err;
        return false;
    }
    for (uint64_t probe = 0; probe < psize; ++probe) {
        uint64_t *s = tv.slots + slot_index(tv, kh, probe) * slot_words<NW>();
        // A plain read settles the common case (key present: one atomic add).
        // A stale EMPTY is harmless: the CAS below returns the real word.
        uint64_t w0 = s[0];
        bool owner = false;
        if (w0 == KC_EMPTY) {
            w0 = atomicCAS((unsigned long long *)s, (unsigned long long)KC_EMPTY, (unsigned long long)key[0]);
            owner = w0 == KC_EMPTY;
        }
        if constexpr (NW == 1) {
            if (owner || w0 == key[0]) {
                n_new += owner;
                add_agent(s + 1, inc);
                return true;
            }
        } else {
            // Two-word key: the CAS winner publishes word 1; a thread that
            // finds word 0 equal waits for word 1 (bounded) and compares.
            if (owner) __hip_atomic_store(s + 1, key[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __builtin_amdgcn_wave_barrier();
            if (owner) {
                ++n_new;
                add_agent(s + 2, inc);
                return true;
            }
            if (w0 == key[0]) {
                uint64_t w1 = ld_agent(s + 1);
                for (int spin = 0; w1 == KC_EMPTY && spin < (1 << 16); ++spin) {
                    __builtin_amdgcn_s_sleep(1);
                    w1 = ld_agent(s + 1);
                }
                if (w1 == KC_EMPTY) {
                    ++n_err;
                    return false;
                }
                if (w1 == key[1]) {
                    add_agent(s + 2, inc);
                    return true;
                }
            }
        }
    }
    ++n_err;
    return false;
}

// The slot holding key, or nullptr when absent.
template <int NW>
__device__ __forceinline__ uint64_t *table_locate(const TableView &tv, const uint64_t *key)
{
    const KeyHash kh = key_hash<NW>(tv, key);
    const uint64_t psize = 1ULL << tv.part_log2;
    for (uint64_t probe = 0; probe < psize; ++probe) {
        uint64_t *s = tv.slots + slot_index(tv, kh, probe) * slot_words<NW>();
        if constexpr (NW == 4) {
            if (s[kStateWord] == 0) return nullptr;
            if (s[0] == key[0] && s[1] == key[1] && s[2] == key[2] && s[3] == key[3]) return s;
        } else {
            const uint64_t w0 = s[0];
            if (w0 == KC_EMPTY) return nullptr;
            if (w0 == key[0] && (NW == 1 || s[1] == key[NW - 1])) return s;
        }
    }
    return nullptr;
}

// Lookup: returns the raw stored count (0 when absent).
template <int NW>
__device__ __forceinline__ uint64_t table_find(const TableView &tv, const uint64_t *key)
{
    const uint64_t *s = table_locate<NW>(tv, key);
    return s ? s[count_word<NW>()] : 0;
}

// Bloom singleton split (BloomFilter.h:233-273, FastqClassifier.h:641-650):
// true when this occurrence is the key's FIRST.  The reference keeps two
// words per block, BF2 (seen once) and BF1 (seen twice), and classifies
// FIRST / SECOND / THIRD_PLUS; its counter sends SECOND and THIRD_PLUS down
// the same path, and BF1 only ever receives bits BF2 holds, so the FIRST
// decision depends on BF2 alone: FIRST exactly when the fetch_or of the
// key's mask into BF2 finds a bit of the mask missing.  One word per block
// is kept here.  Only atomic results decide a FIRST, so each distinct key
// gets at most one FIRST; a plain read settles keys already seen.
__device__ __forceinline__ bool bloom_first(const TableView &tv, const KeyHash &kh)
{
    uint64_t blk, m;
    bloom_in_part(kh.h, (1ULL << tv.bloom_part_log2) - 1, blk, m);
    blk |= kh.part << tv.bloom_part_log2;
    uint64_t *cell = tv.bloom + blk;
    if ((ld_agent(cell) & m) == m) return false;
    const uint64_t b = __hip_atomic_fetch_or(cell, m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return (b & m) != m;
}

// ---- rolling canonical state (GetKmer.h:58-82) -------------------------
// Four input bytes at once (FastqParser.h:280-305 semantics, GetKmer.h
// table: A/a=0 C/c=1 G/g=2 T/t/U/u=3, anything else invalid): `codes` gets
// the 2-bit code of every byte in that byte's low bits; the result has bit 7
// set in every invalid byte.  A byte is valid when it equals (case folded)
// the base its code names (A C G T) or is U, which shares T's code.
__device__ __forceinline__ uint32_t invalid_bytes(uint32_t w, uint32_t &codes)
{
    codes = ((w >> 1) ^ (w >> 2)) & 0x03030303u;
    const uint32_t c = w & 0xDFDFDFDFu;
    const uint32_t named = __builtin_amdgcn_perm(0u, 0x54474341u, codes);  // byte i = "ACGT"[code i]
    const auto nonzero = [](uint32_t z) { return (((z & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | z) & 0x80808080u; };
    return nonzero(c ^ named) & nonzero(c ^ 0x55555555u);
}

// The codes of invalid_bytes packed to 8 bits, byte 0's code lowest.
__device__ __forceinline__ uint32_t pack_codes(uint32_t codes)
{
    const uint32_t u = codes | (codes >> 6);
    return (u | (u >> 12)) & 0xFFu;
}

// Bit 7 of each byte of invalid_bytes gathered to bits 0..3.
__device__ __forceinline__ uint32_t gather_invalid(uint32_t iv)
{
    return (((iv >> 7) * 0x00204081u) >> 21) & 0xFu;
}

// Reverses the order of the 32 two-bit fields of x.
__device__ __forceinline__ uint64_t pair_reverse(uint64_t x)
{
    x = __builtin_bitreverse64(x);
    return ((x >> 1) & 0x5555555555555555ULL) | ((x & 0x5555555555555555ULL) << 1);
}

// History bytes h[0..4*HW) (byte 0 oldest) packed 4 words per u64 (byte
// j's code at bits 2j of its word) with their invalid flags (bit j).
template <int HW>
__device__ __forceinline__ void pack_history(const uint32_t *h, uint64_t *packed, uint64_t &invalid)
{
    invalid = 0;
#pragma unroll
    for (int q = 0; q < HW; ++q) {
        uint32_t codes;
        const uint32_t iv = invalid_bytes(h[q], codes);
        if (q % 8 == 0) packed[q / 8] = 0;
        packed[q / 8] |= (uint64_t)pack_codes(codes) << (8 * (q % 8));
        invalid |= (uint64_t)gather_invalid(iv) << (4 * q);
    }
}

// The codes, complements and invalid flags of four input bytes
// (invalid_bytes): byte b's code at bits 8b..8b+1 of codes, code ^ 3 at the
// same bits of comps, bit 8b+7 of iv set when the byte is not a base.
struct Word4 {
    uint32_t codes, comps, iv;
};

__device__ __forceinline__ Word4 word4(uint32_t v)
{
    Word4 x;
    x.iv = invalid_bytes(v, x.codes);
    x.comps = x.codes ^ 0x03030303u;
    return x;
}

template <int NW> struct Roll;

// The rolling state of GetKmer.h:58-82, kept so that a push is four 32-bit
// operations per key word: the forward codes with the newest at the top and
// the older ones below it, not masked (canon() keeps the top 2k bits), and
// the complements bottom-aligned, newest at bits 0-1 (canon() shifts them up
// by rev_shift, which drops the bases older than the window).  Shifts across
// words are funnel shifts (v_alignbit).  canon() returns the key layout of
// kmer<N>: the forward word(s) masked and the reverse complement, whichever
// is numerically smaller.
template <> struct Roll<1> {
    uint32_t f0 = 0, f1 = 0;  // forward, f1:f0
    uint32_t r0 = 0, r1 = 0;  // complement, r1:r0, bottom-aligned
    uint64_t back_mask;
    uint32_t rev_shift;
    __device__ Roll(uint32_t k)
    {
        back_mask = (k % 32) ? (((1ULL << (2 * (k % 32))) - 1ULL) << (64 - 2 * (k % 32))) : ~0ULL;
        rev_shift = (k % 32) ? 2 * (32 - k % 32) : 0;
    }
    // code: the new base's code in bits 0-1 (higher bits ignored); comp:
    // code ^ 3, exactly two bits.
    __device__ __forceinline__ void push(uint32_t code, uint32_t comp)
    {
        f0 = __builtin_amdgcn_alignbit(f1, f0, 2);
        f1 = __builtin_amdgcn_alignbit(code, f1, 2);
        r1 = __builtin_amdgcn_alignbit(r1, r0, 30);
        r0 = (r0 << 2) | comp;
    }
    __device__ __forceinline__ void canon(uint64_t *out) const
    {
        const uint64_t f = (((uint64_t)f1 << 32) | f0) & back_mask;
        const uint64_t r = (((uint64_t)r1 << 32) | r0) << rev_shift;
        out[0] = f < r ? f : r;
    }
    // The state pushes of the 32 history bytes h[0..8) (byte 0 first) leave,
    // without rolling them: the forward word is the packed history, the
    // complement its reverse.  `have` = the trailing run of valid bytes.
    __device__ __forceinline__ void seed(const uint32_t *h, uint32_t &have)
    {
        uint64_t p[1], inv;
        pack_history<8>(h, p, inv);
        f0 = (uint32_t)p[0];
        f1 = (uint32_t)(p[0] >> 32);
        const uint64_t y = pair_reverse(~p[0]);
        r0 = (uint32_t)y;
        r1 = (uint32_t)(y >> 32);
        have = (uint32_t)__clz((uint32_t)inv);
    }
};

template <> struct Roll<2> {
    uint32_t f[4] = {0, 0, 0, 0};  // forward: f[3]:f[2] = key word 0 (newest codes), f[1]:f[0] = word 1
    uint32_t r[4] = {0, 0, 0, 0};  // complement, bottom-aligned 128 bits
    uint64_t back_mask;
    uint32_t rev_shift;
    __device__ Roll(uint32_t k)
    {
        back_mask = (k % 32) ? (((1ULL << (2 * (k % 32))) - 1ULL) << (64 - 2 * (k % 32))) : ~0ULL;
        rev_shift = (k % 32) ? 2 * (32 - k % 32) : 0;
    }
    __device__ __forceinline__ void push(uint32_t code, uint32_t comp)
    {
        f[0] = __builtin_amdgcn_alignbit(f[1], f[0], 2);
        f[1] = __builtin_amdgcn_alignbit(f[2], f[1], 2);
        f[2] = __builtin_amdgcn_alignbit(f[3], f[2], 2);
        f[3] = __builtin_amdgcn_alignbit(code, f[3], 2);
        r[3] = __builtin_amdgcn_alignbit(r[3], r[2], 30);
        r[2] = __builtin_amdgcn_alignbit(r[2], r[1], 30);
        r[1] = __builtin_amdgcn_alignbit(r[1], r[0], 30);
        r[0] = (r[0] << 2) | comp;
    }
    __device__ __forceinline__ void canon(uint64_t *out) const
    {
        // word 0 most significant; the tail word (word 1) is masked
        const uint64_t f0 = ((uint64_t)f[3] << 32) | f[2];
        const uint64_t f1 = (((uint64_t)f[1] << 32) | f[0]) & back_mask;
        const uint64_t ylo = ((uint64_t)r[1] << 32) | r[0], yhi = ((uint64_t)r[3] << 32) | r[2];
        const uint64_t r1 = ylo << rev_shift;
        const uint64_t r0 = rev_shift ? (yhi << rev_shift) | (ylo >> (64 - rev_shift)) : yhi;
        const bool fwd = (f0 < r0) || (f0 == r0 && f1 < r1);
        out[0] = fwd ? f0 : r0;
        out[1] = fwd ? f1 : r1;
    }
    // As Roll<1>::seed over the 64 history bytes h[0..16): the newer 32
    // codes are the head word, the older 32 the tail word.
    __device__ __forceinline__ void seed(const uint32_t *h, uint32_t &have)
    {
        uint64_t p[2], inv;
        pack_history<16>(h, p, inv);
        f[0] = (uint32_t)p[0];
        f[1] = (uint32_t)(p[0] >> 32);
        f[2] = (uint32_t)p[1];
        f[3] = (uint32_t)(p[1] >> 32);
        const uint64_t ylo = pair_reverse(~p[1]), yhi = pair_reverse(~p[0]);
        r[0] = (uint32_t)ylo;
        r[1] = (uint32_t)(ylo >> 32);
        r[2] = (uint32_t)yhi;
        r[3] = (uint32_t)(yhi >> 32);
        have = (uint32_t)__clzll((long long)inv);
    }
};

// Four-word keys (k = 64..128; k = 64 is kept in four words here and
// written out as kmer<2>): the k-mer spans words 0..last, last = (k-1)/32,
// word `last` masked to its top 2 * (k - 32 * last) bits and the words after
// it zero, as the oracle and kmer<N> lay them out.  The forward codes are a
// 256-bit number with the newest code at the top (f[7] holds the top 32
// bits), the complements a 256-bit number with the newest at the bottom
// (r[0] the lowest 32 bits).
template <> struct Roll<4> {
    uint32_t f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t r[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint64_t back_mask;
    uint32_t last, pad;  // key words in use - 1; zero bits below the reverse complement
    __device__ Roll(uint32_t k)
    {
        last = (k - 1) / 32;
        const uint32_t rem = k - 32 * last;  // 1..32 bases in word `last`
        back_mask = rem < 32 ? (((1ULL << (2 * rem)) - 1ULL) << (64 - 2 * rem)) : ~0ULL;
        pad = 64 * (last + 1) - 2 * k;
    }
    __device__ __forceinline__ void push(uint32_t code, uint32_t comp)
    {
#pragma unroll
        for (int i = 0; i < 7; ++i) f[i] = __builtin_amdgcn_alignbit(f[i + 1], f[i], 2);
        f[7] = __builtin_amdgcn_alignbit(code, f[7], 2);
#pragma unroll
        for (int i = 7; i > 0; --i) r[i] = __builtin_amdgcn_alignbit(r[i], r[i - 1], 30);
        r[0] = (r[0] << 2) | comp;
    }
    __device__ __forceinline__ void canon(uint64_t *out) const
    {
        uint64_t fk[4], rw[4], rs[4], rk[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            fk[j] = ((uint64_t)f[7 - 2 * j] << 32) | f[6 - 2 * j];  // word j from the top
            rw[j] = ((uint64_t)r[2 * j + 1] << 32) | r[2 * j];      // word j from the bottom
        }
        // the reverse complement: the low 2k bits of rw, shifted up by pad
#pragma unroll
        for (int j = 0; j < 4; ++j) rs[j] = pad ? (rw[j] << pad) | (j ? rw[j - 1] >> (64 - pad) : 0) : rw[j];
        // key word j of it is rs[last - j]
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t src = last - j;  // wraps for j > last
            rk[j] = src == 0 ? rs[0] : src == 1 ? rs[1] : src == 2 ? rs[2] : src == 3 ? rs[3] : 0;
            if ((uint32_t)j > last) fk[j] = 0;
            if ((uint32_t)j == last) fk[j] &= back_mask;
        }
        // numeric (lexicographic by word) minimum
        bool fwd = false, decided = false;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (!decided && fk[j] != rk[j]) {
                fwd = fk[j] < rk[j];
                decided = true;
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) out[j] = (fwd || !decided) ? fk[j] : rk[j];
    }
    // The state after the 128 history bytes h[0..32) (byte 0 first).
    __device__ __forceinline__ void seed(const uint32_t *h, uint32_t &have)
    {
        uint64_t lo[2], hi[2], inv_lo, inv_hi;
        pack_history<16>(h, lo, inv_lo);       // bytes 0..63 (older)
        pack_history<16>(h + 16, hi, inv_hi);  // bytes 64..127 (newer)
        const uint64_t p[4] = {lo[0], lo[1], hi[0], hi[1]};  // p[3] the newest 32 codes
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            f[2 * j] = (uint32_t)p[j];
            f[2 * j + 1] = (uint32_t)(p[j] >> 32);
            const uint64_t y = pair_reverse(~p[3 - j]);
            r[2 * j] = (uint32_t)y;
            r[2 * j + 1] = (uint32_t)(y >> 32);
        }
        have = inv_hi ? (uint32_t)__clzll((long long)inv_hi) : 64 + (inv_lo ? (uint32_t)__clzll((long long)inv_lo) : 64);
    }
};

// Pushes byte b of a Word4 and updates the trailing run of valid bytes
// (any invalid byte resets the window, FastqParser.h:280-305 / GetKmer.h).
template <int NW>
__device__ __forceinline__ void push_byte(Roll<NW> &roll, const Word4 &x, int b, uint32_t &have)
{
    roll.push(x.codes >> (8 * b), (x.comps >> (8 * b)) & 3u);
    const bool valid = !((x.iv >> (8 * b + 7)) & 1u);
    have = valid ? have + 1 : 0;
}

// FIRST-list segment of an append by (block, wave, step): a mix of the three,
// so that every segment fills alike whatever the grid (a few blocks of one
// tile each, or 2^18 partitions) and appends of one wave do not collide.
__device__ __forceinline__ uint32_t first_segment(uint32_t block, uint32_t wave, uint32_t step)
{
    uint32_t h = block * 0x9E3779B1u ^ (wave + 1) * 0x85EBCA77u ^ step * 0xC2B2AE3Du;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    return h % KC_LOW_SEGS;
}

// Appends the FIRST occurrences of this wave step to a FIRST-list segment
// (one cursor atomic per wave step).  Every lane of the wave must call it;
// `first` selects the lanes that append.  Returns false for a FIRST the
// segment has no room for: the caller counts that occurrence in the table
// instead (+1), which gives the same merged count (table count + 1 if the key
// is on the FIRST list), so a full segment never loses a k-mer.
template <int NW>
__device__ __forceinline__ bool first_append(const TableView &tv, bool first, const uint64_t *key, uint32_t rot)
{
    const uint64_t firsts = __ballot(first);
    if (!firsts) return true;
    const int leader = __ffsll((unsigned long long)firsts) - 1;
    const uint32_t seg = first_segment(blockIdx.x, threadIdx.x >> 6, rot);
    unsigned long long b0 = 0;
    if ((int)__lane_id() == leader)
        b0 = atomicAdd(tv.counters + KC_CURSOR_BASE + (uint64_t)seg * KC_LINE_WORDS, (unsigned long long)__popcll(firsts));
    b0 = __shfl(b0, leader, 64);
    if (!first) return true;
    const uint64_t at = b0 + __popcll(firsts & lanemask_lt());
    if (at >= tv.seg_cap) return false;
    uint64_t *dst = tv.low + ((uint64_t)seg * tv.seg_cap + at) * NW;
#pragma unroll
    for (int w = 0; w < NW; ++w) dst[w] = key[w];
    return true;
}

// Framed input of one counting call.  When hdr is set, the byte count and
// the first emitted position are read from the device words {hdr[0],
// hdr[1]} (the count clamped to n, the capacity the launch was sized for),
// so that one captured graph serves batches of any size; otherwise they are
// n and emit_from.
struct SeqSrc {
    const uint8_t *seq;
    uint64_t n;
    uint64_t emit_from;
    const uint64_t *hdr;
};

__device__ __forceinline__ void seq_bounds(const SeqSrc &s, uint64_t &n, uint64_t &emit_from)
{
    if (s.hdr) {
        const uint64_t h0 = s.hdr[0];
        n = h0 < s.n ? h0 : s.n;
        emit_from = s.hdr[1];
    } else {
        n = s.n;
        emit_from = s.emit_from;
    }
}

// ---- parse tile: 256 threads x 64 bytes -------------------------------
constexpr int kParseBlock = 256;
constexpr int kParseSeg = 64;
constexpr int kParseTile = kParseBlock * kParseSeg;

// LDS image of one tile (+ the previous H segments as halo: H = 1 for
// k <= 64, 2 for up to 127 bytes of history), transposed so that lane t
// reading its own dword j touches consecutive banks; segment s of the tile
// is column s + H.
struct ParseLds {
    uint32_t w[16][kParseBlock + 2];
};

template <int NW> constexpr int halo_segs() { return NW == 4 ? 2 : 1; }

// Stages tile `tile` of seq[0, n) into L (all threads; ends with a barrier).
// A whole tile issues its four 16-byte loads per thread back to back (no
// wait in between); only the last, partial tile takes the byte-wise path.
template <int H = 1>
__device__ __forceinline__ void stage_tile(const uint8_t *__restrict__ seq, uint64_t n, uint64_t tile, ParseLds &L)
{
    const uint32_t t = threadIdx.x;
    const uint64_t base = tile * kParseTile;
    uint4 v[4];
    if (base + kParseTile <= n) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
            v[r] = *reinterpret_cast<const uint4 *>(seq + base + 16ull * (t + kParseBlock * r));
    } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint64_t off = base + 16ull * (t + kParseBlock * r);
            v[r] = make_uint4(0, 0, 0, 0);
            if (off + 16 <= n) {
                v[r] = *reinterpret_cast<const uint4 *>(seq + off);
            } else if (off < n) {
                uint32_t w[4] = {0, 0, 0, 0};
                for (uint64_t b = off; b < n; ++b) w[(b - off) >> 2] |= (uint32_t)seq[b] << (8 * ((b - off) & 3));
                v[r] = make_uint4(w[0], w[1], w[2], w[3]);
            }
        }
    }
    uint32_t hv = 0;
    if (t < 16 * H && base >= (uint64_t)kParseSeg * H) {
        const uint64_t off = base - (uint64_t)kParseSeg * H + 4ull * t;
        if (off + 4 <= n) hv = *reinterpret_cast<const uint32_t *>(seq + off);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t q = t + kParseBlock * r;  // uint4 index inside the tile
        const uint32_t seg = q >> 2, j0 = (q & 3) * 4;
        L.w[j0 + 0][seg + H] = v[r].x;
        L.w[j0 + 1][seg + H] = v[r].y;
        L.w[j0 + 2][seg + H] = v[r].z;
        L.w[j0 + 3][seg + H] = v[r].w;
    }
    if (t < 16 * H) L.w[t & 15][t >> 4] = hv;
    __syncthreads();
}

// Loads the 64-byte segment at `start` straight into registers with HW
// words of history before it: w[0..HW-1] = the 4*HW bytes before the
// segment (32 bytes cover k <= 33, 64 bytes k <= 63), w[HW..HW+15] = the
// segment; bytes outside [0, n) read as 0 (invalid).  The history words are
// the previous lane's segment tail, so they hit in L1.
template <int HW>
__device__ __forceinline__ void load_segment(const uint8_t *__restrict__ seq, uint64_t n, uint64_t start,
                                             uint32_t (&w)[HW + 16])
{
    static_assert(HW % 4 == 0, "history is whole 16-byte groups");
    constexpr int G = HW / 4 + 4;
    constexpr uint64_t H = 4 * HW;
    if (start >= H && start + kParseSeg <= n) {
        const uint4 *p = reinterpret_cast<const uint4 *>(seq + start - H);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const uint4 v = p[g];
            w[4 * g] = v.x;
            w[4 * g + 1] = v.y;
            w[4 * g + 2] = v.z;
            w[4 * g + 3] = v.w;
        }
        return;
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
        uint4 v = make_uint4(0, 0, 0, 0);
        if (start + 16 * g >= H) {
            const uint64_t off = start + 16 * g - H;
            if (off + 16 <= n) {
                v = *reinterpret_cast<const uint4 *>(seq + off);
            } else if (off < n) {
                uint32_t x[4] = {0, 0, 0, 0};
                for (uint64_t b = off; b < n; ++b) x[(b - off) >> 2] |= (uint32_t)seq[b] << (8 * ((b - off) & 3));
                v = make_uint4(x[0], x[1], x[2], x[3]);
            }
        }
        w[4 * g] = v.x;
        w[4 * g + 1] = v.y;
        w[4 * g + 2] = v.z;
        w[4 * g + 3] = v.w;
    }
}

// Rolls thread t's segment of a staged tile and calls emit(key, e) for every
// window whose last byte lies in the segment, at or after emit_from, e being
// that byte's index in the segment (the same for every lane of one call)
// (GetKmer.h:58-82 semantics; any invalid byte resets the window).
template <int NW, typename F>
__device__ __forceinline__ void roll_segment(const ParseLds &L, uint64_t tile, uint64_t n, uint64_t emit_from,
                                             uint32_t k, F &&emit)
{
    constexpr int kHaloWords = NW == 1 ? 8 : NW == 2 ? 16 : 32;  // >= k-1 bytes of history
    constexpr int H = halo_segs<NW>();
    const uint32_t t = threadIdx.x;
    const uint64_t start = tile * kParseTile + (uint64_t)t * kParseSeg;
    if (start >= n) return;
    Roll<NW> roll(k);
    uint32_t have, hist[kHaloWords];
    // the last kHaloWords words before the segment (column t + H)
#pragma unroll
    for (int j = 0; j < kHaloWords; ++j) {
        const int back = kHaloWords - j;  // words before the segment
        hist[j] = L.w[16 - 1 - ((back - 1) & 15)][t + H - 1 - (back - 1) / 16];
    }
    roll.seed(hist, have);
    // bytes of this segment before emit_from emit nothing
    const uint32_t skip = emit_from > start ? (uint32_t)(emit_from - start < 64 ? emit_from - start : 64) : 0;
    for (int j = 0; j < 16; ++j) {
        const Word4 x = word4(L.w[j][t + H]);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            push_byte<NW>(roll, x, b, have);
            if (have >= k && (uint32_t)(4 * j + b) >= skip) {
                uint64_t key[NW];
                roll.canon(key);
                emit(key, (uint32_t)(4 * j + b));
            }
        }
    }
}

}  // namespace kc

// Partitioned counting pass (kc_part.hip).
bool kc_part_eligible(const kc_table *t);
void kc_part_free(kc_table *t);
void kc_part_scratch_free(kc_part_scratch &s);
kc_pipe_shape kc_part_shape(const kc_table *t, const kc_tune &tune, uint64_t key_cap);
int kc_part_reserve(const kc_table *t, kc_part_scratch &s, const kc_pipe_shape &sh);
// Enqueues the whole pass on st without host synchronisation: over the
// framed input `seq` (when d_keys is null) or the n_keys keys of d_keys.
// fresh: the table is lazily cleared (kc_table::fresh) and this pass writes
// every slot and Bloom word.
int kc_part_enqueue(const kc_table *t, kc_part_scratch &s, const kc_pipe_shape &sh, const kc::SeqSrc &seq,
                    const uint64_t *d_keys, uint64_t n_keys, const kc::TableView &tv, hipStream_t st,
                    bool fresh = false);
int kc_part_count(kc_table *t, const uint8_t *d_seq, uint64_t nbytes, uint64_t emit_from, const uint64_t *d_keys,
                  uint64_t n_keys, const kc::TableView &tv);
// Direct path (kc_count.hip): the fused count kernel with `grid` blocks.
int kc_direct_enqueue(const kc_table *t, const kc::SeqSrc &src, uint32_t grid, hipStream_t st);
kc::TableView kc_table_view(kc_table *t);
