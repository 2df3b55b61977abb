// Device hashes of k-mer key words, bit-exact with the reference's vendored
// headers for the key sizes on the path (8, 16 and 32 bytes):
//   XXH3_64bits / XXH3_128bits  include/xxhash.h 0.8.3 (:4696-4830, :6900-7070)
//   rapidhash V3                include/rapidhash.h (:243-330)
//   komihash 5.29               include/komihash.h (:565-890)
//   hash_func / mix_hash        src/HashFunction.h (:9-63)
// Keys arrive as words (little-endian u64 as stored in kmer<N>), so every
// 8-byte read of the message is a whole word; 4-byte reads are its halves.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kc {

// XXH3 default secret, read as little-endian u64 at the offsets the
// short-input paths use (xxhash.h:4398).
struct xxh3_secret_words {
    static constexpr uint64_t s0 = 0xbe4ba423396cfeb8ULL, s8 = 0x1cad21f72c81017cULL,
                              s16 = 0xdb979083e96dd4deULL, s24 = 0x1f67b3b7a4a44072ULL,
                              s32 = 0x78e5c0cc4ee679cbULL, s40 = 0x2172ffcc7dd05a82ULL,
                              s48 = 0x8e2443f7744608b8ULL, s56 = 0x4c263a81e69035e0ULL;
};

constexpr uint64_t XP1 = 0x9E3779B185EBCA87ULL;
constexpr uint64_t XP2 = 0xC2B2AE3D27D4EB4FULL;
constexpr uint64_t XP4 = 0x85EBCA77C2B2AE63ULL;
constexpr uint32_t XP32_2 = 0x85EBCA77U;
constexpr uint64_t MX1 = 0x165667919E3779F9ULL;
constexpr uint64_t MX2 = 0x9FB21C651E98DF25ULL;

__device__ __forceinline__ uint64_t rotl(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
__device__ __forceinline__ uint64_t mulhi(uint64_t a, uint64_t b) { return __umul64hi(a, b); }
__device__ __forceinline__ uint64_t fold(uint64_t a, uint64_t b) { return (a * b) ^ mulhi(a, b); }
__device__ __forceinline__ uint64_t xsh(uint64_t v, int s) { return v ^ (v >> s); }
__device__ __forceinline__ uint64_t av3(uint64_t h) { h = xsh(h, 37); h *= MX1; return xsh(h, 32); }
__device__ __forceinline__ uint64_t bswap(uint64_t x) { return __builtin_bswap64(x); }

// XXH3_64bits over NW words (8*NW bytes, seed 0).
template <int NW>
__device__ __forceinline__ uint64_t xxh3_64(const uint64_t *w)
{
    using S = xxh3_secret_words;
    if constexpr (NW == 1) {
        // len 4..8: input64 = readLE32(p+len-4) + (readLE32(p) << 32)
        const uint64_t in64 = (w[0] >> 32) | (w[0] << 32);
        uint64_t h = in64 ^ (S::s8 ^ S::s16);
        h ^= rotl(h, 49) ^ rotl(h, 24);
        h *= MX2;
        h ^= (h >> 35) + 8;
        h *= MX2;
        return xsh(h, 28);
    } else if constexpr (NW == 2) {
        const uint64_t lo = w[0] ^ (S::s24 ^ S::s32);
        const uint64_t hi = w[1] ^ (S::s40 ^ S::s48);
        return av3(16 + bswap(lo) + hi + fold(lo, hi));
    } else {
        uint64_t acc = 32 * XP1;
        acc += fold(w[0] ^ S::s0, w[1] ^ S::s8);
        acc += fold(w[2] ^ S::s16, w[3] ^ S::s24);
        return av3(acc);
    }
}

// XXH3_128bits over NW words (seed 0).
template <int NW>
__device__ __forceinline__ void xxh3_128(const uint64_t *w, uint64_t &olo, uint64_t &ohi)
{
    using S = xxh3_secret_words;
    if constexpr (NW == 1) {
        // len 4..8: input_64 = readLE32(p) + (readLE32(p+len-4) << 32) = word
        const uint64_t keyed = w[0] ^ (S::s16 ^ S::s24);
        const uint64_t m = XP1 + (8u << 2);
        uint64_t lo = keyed * m, hi = mulhi(keyed, m);
        hi += lo << 1;
        lo ^= hi >> 3;
        lo = xsh(lo, 35);
        lo *= MX2;
        lo = xsh(lo, 28);
        olo = lo;
        ohi = av3(hi);
    } else if constexpr (NW == 2) {
        const uint64_t bl = S::s32 ^ S::s40, bh = S::s48 ^ S::s56;
        const uint64_t a = w[0] ^ w[1] ^ bl;
        uint64_t mlo = a * XP1, mhi = mulhi(a, XP1);
        mlo += (uint64_t)(16 - 1) << 54;
        const uint64_t ihi = w[1] ^ bh;
        mhi += ihi + (uint64_t)(uint32_t)ihi * (uint64_t)(XP32_2 - 1);
        mlo ^= bswap(mhi);
        const uint64_t hlo = mlo * XP2;
        const uint64_t hhi = mulhi(mlo, XP2) + mhi * XP2;
        olo = av3(hlo);
        ohi = av3(hhi);
    } else {
        uint64_t alo = 32 * XP1, ahi = 0;
        alo += fold(w[0] ^ S::s0, w[1] ^ S::s8);
        alo ^= w[2] + w[3];
        ahi += fold(w[2] ^ S::s16, w[3] ^ S::s24);
        ahi ^= w[0] + w[1];
        const uint64_t hlo = alo + ahi;
        const uint64_t hhi = alo * XP1 + ahi * XP4 + 32 * XP2;
        olo = av3(hlo);
        ohi = 0 - av3(hhi);
    }
}

// rapidhash (seed 0, default secret) over NW words.
__device__ __forceinline__ void mum(uint64_t &a, uint64_t &b)
{
    const uint64_t lo = a * b, hi = mulhi(a, b);
    a = lo;
    b = hi;
}
__device__ __forceinline__ uint64_t rmix(uint64_t a, uint64_t b) { mum(a, b); return a ^ b; }

template <int NW>
__device__ __forceinline__ uint64_t rapidhash(const uint64_t *w)
{
    constexpr uint64_t r0 = 0x2d358dccaa6c78a5ULL, r1 = 0x8bb84b93962eacc9ULL,
                       r2 = 0x4b33a62ed433d4a3ULL, r7 = 0xaaaaaaaaaaaaaaaaULL;
    uint64_t seed = rmix(r2, r1);
    (void)r0;
    uint64_t a, b;
    constexpr uint64_t len = 8 * NW;
    if constexpr (NW == 1) {
        seed ^= len;
        a = w[0];
        b = w[0];
    } else if constexpr (NW == 2) {
        seed ^= len;
        a = w[0];
        b = w[1];
    } else {
        seed = rmix(w[0] ^ r2, w[1] ^ seed);
        a = w[2] ^ len;
        b = w[3];
    }
    a ^= r1;
    b ^= seed;
    mum(a, b);
    return rmix(a ^ r7, b ^ r1 ^ len);
}

// komihash (seed 0) over NW words; the 128-bit multiply adds its high half
// into the second lane (komihash.h:565-600).
__device__ __forceinline__ void khm(uint64_t u, uint64_t v, uint64_t &rl, uint64_t &rha)
{
    const uint64_t lo = u * v, hi = mulhi(u, v);
    rl = lo;
    rha += hi;
}

template <int NW>
__device__ __forceinline__ uint64_t komihash(const uint64_t *w)
{
    uint64_t s1 = 0x243F6A8885A308D3ULL, s5 = 0x452821E638D01377ULL;
    uint64_t r1, r2;
    khm(s1, s5, s1, s5);
    s1 ^= s5;
    if constexpr (NW == 1) {
        r1 = s1 ^ w[0];
        r2 = s5 ^ 1;
    } else {
        khm(w[0] ^ s1, w[1] ^ s5, s1, s5);
        s1 ^= s5;
        if constexpr (NW == 4) {
            khm(w[2] ^ s1, w[3] ^ s5, s1, s5);
            s1 ^= s5;
        }
        r1 = 1 ^ s1;
        r2 = s5;
    }
    khm(r1, r2, s1, s5);
    s1 ^= s5;
    khm(s1, s5, s1, s5);
    s1 ^= s5;
    return s1;
}

__device__ __forceinline__ uint64_t mix_hash(uint64_t h1, uint64_t h2)
{
    constexpr uint64_t F = 0x9ddfea08eb382d69ULL;
    uint64_t a = (h1 ^ h2) * F;
    a ^= a >> 47;
    uint64_t b = (h2 ^ a) * F;
    b ^= b >> 47;
    return b * F;
}

template <int NW>
__device__ __forceinline__ uint64_t hash_func(const uint64_t *w)
{
    uint64_t h = 0x9e3779b97f4a7c15ULL * NW + 0x85ebca77c2b2ae63ULL;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        uint64_t x = w[i];
        x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
        x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
        x ^= x >> 31;
        h ^= x;
        h = rotl(h, 27);
        h = h * 0x9e3779b97f4a7c15ULL + 0x165667b19e3779f9ULL;
    }
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdULL;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ULL;
    h ^= h >> 33;
    return h;
}

// ConcurrentBloomFilter::prepare_insert + calculate_insert_num
// (BloomFilter.h:215-222, :287-297).
__device__ __forceinline__ void bloom_from(uint64_t lo, uint64_t hi, uint64_t cap_mask, uint64_t &block,
                                           uint64_t &mask)
{
    const uint64_t h2 = hi | 1ULL;
    mask = (1ULL << (lo & 63)) | (1ULL << ((lo + h2) & 63)) | (1ULL << ((lo + 2 * h2) & 63));
    block = lo & cap_mask;
}

template <int NW>
__device__ __forceinline__ void bloom_probe(const uint64_t *w, uint64_t cap_mask,
                                            uint64_t &block, uint64_t &mask)
{
    uint64_t lo, hi;
    xxh3_128<NW>(w, lo, hi);
    bloom_from(lo, hi, cap_mask, block, mask);
}

// ---- the table's own key hash (device layout, not a reference hash) ------
// Where a key sits inside its partition: its Bloom block (low bits), the
// three Bloom bits (bits 32..43) and the slot probe start (bits 44..63), all
// from one 64-bit mix of the key words: a 32-bit xor-shift (a word swap on a
// 32-bit ALU), one 64-bit multiply and a final 32-bit xor-shift.  (A second
// multiply round left the block, mask and probe-start bits of repeat-rich
// k = 21 and 31 keys inside each partition no more uniform: both give
// Poisson counts.)  The reference's probe (bloom_from) takes block and mask
// from XXH3_128, whose first mask bit is the block's low bits, so a key's
// mask in its block depends on 6 bits only; three positions a, a+b, a+2b
// (b odd, so distinct) from 12 free bits give ~2^11 masks per block and
// about a third of the false positives at the same size, for a fraction of
// the multiplies.
constexpr uint64_t kSlotMixC = 0xD6E8FEB86659FD93ULL;
constexpr uint64_t kSlotFoldC = 0x9E3779B97F4A7C15ULL;

template <int NW>
__device__ __forceinline__ uint64_t slot_mix(const uint64_t *w)
{
    uint64_t x = w[0];
    if constexpr (NW >= 2) x ^= w[1] * kSlotFoldC;
    if constexpr (NW == 4) x ^= (w[2] ^ ((w[3] >> 32) | (w[3] << 32))) * kSlotMixC;
    x = (x ^ (x >> 32)) * kSlotMixC;
    return x ^ (x >> 32);
}

// Bloom block (in a part of cap_mask + 1 blocks) and 3-bit mask of a key.
__device__ __forceinline__ void bloom_in_part(uint64_t h, uint64_t cap_mask, uint64_t &block, uint64_t &mask)
{
    const uint32_t a = (uint32_t)(h >> 32) & 63u, b = ((uint32_t)(h >> 38) & 63u) | 1u;
    mask = (1ULL << a) | (1ULL << ((a + b) & 63u)) | (1ULL << ((a + 2u * b) & 63u));
    block = h & cap_mask;
}

} // namespace kc
