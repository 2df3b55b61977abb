// The on-disk record format on the device: packed k-mer bytes
// (ExportWriter.h:296-313: full words as LE u64, then the top tail_bytes of
// the masked tail word; main.cpp:599-615 count bytes) and the LDS-staged
// copy that writes a tile's packed records as one contiguous byte range.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kc {

struct PackSpec {
    uint32_t full_words, tail_bytes, kmer_bytes, count_bytes;
    uint64_t tail_mask;
};

// Packing of a k (1..128) and count_max: high records are kmer_bytes +
// count_bytes long (FinalDrainWriter.h:43-52), low records kmer_bytes.
inline PackSpec kc_pack_spec(uint32_t k, uint32_t count_max)
{
    PackSpec p;
    p.full_words = k / 32;
    const uint32_t tail_bits = 2 * (k % 32);
    p.tail_bytes = (tail_bits + 7) / 8;
    p.kmer_bytes = p.full_words * 8 + p.tail_bytes;
    p.count_bytes = count_max <= 0xFF ? 1 : count_max <= 0xFFFF ? 2 : count_max <= 0xFFFFFF ? 3 : 4;
    p.tail_mask = tail_bits ? (~0ULL << (64 - tail_bits)) : 0;
    return p;
}

__device__ __forceinline__ void put_bytes(uint8_t *dst, uint64_t v, uint32_t nbytes, uint32_t skip)
{
    for (uint32_t b = 0; b < nbytes; ++b) dst[b] = (uint8_t)(v >> (8 * (skip + b)));
}

template <int NW>
__device__ __forceinline__ void pack_key(uint8_t *dst, const uint64_t *key, const PackSpec &ps)
{
    for (uint32_t w = 0; w < ps.full_words; ++w) put_bytes(dst + 8 * w, key[w], 8, 0);
    if (ps.tail_bytes) {
        const uint64_t tail = key[ps.full_words] & ps.tail_mask;
        put_bytes(dst + 8 * ps.full_words, tail, ps.tail_bytes, 8 - ps.tail_bytes);
    }
}

// The inverse (ExportReader / FinalDrainReader unpacking): the key words,
// tail bytes back at the top of the tail word, padding zero.
template <int NW>
__device__ __forceinline__ void unpack_key(const uint8_t *src, const PackSpec &ps, uint64_t *key)
{
#pragma unroll
    for (int w = 0; w < NW; ++w) key[w] = 0;
    for (uint32_t w = 0; w < ps.full_words; ++w) {
        uint64_t v = 0;
        for (uint32_t b = 0; b < 8; ++b) v |= (uint64_t)src[8 * w + b] << (8 * b);
        key[w] = v;
    }
    if (ps.tail_bytes) {
        uint64_t v = 0;
        for (uint32_t b = 0; b < ps.tail_bytes; ++b) v |= (uint64_t)src[8 * ps.full_words + b] << (8 * (8 - ps.tail_bytes + b));
        key[ps.full_words] = v;
    }
}

// Export compaction is deterministic: a counting pass writes per-tile
// record counts, a scan gives offsets, and the writing pass places each
// tile's records at its offset.  Records are packed into an LDS staging
// buffer first and the tile's contiguous output bytes are then written with
// coalesced stores.  Blocks of kPackBlock threads, kExportPer records each.
constexpr int kPackBlock = 256;
constexpr int kExportPer = 8;
constexpr uint64_t kExportTile = (uint64_t)kPackBlock * kExportPer;
// LDS staging of a tile's records by key words: 20 bytes hold a k <= 64 high
// record (16 key bytes + 4 count bytes), 36 a k <= 128 one; 16 more for the
// skew below.  Declare the staging buffer __align__(16).
template <int NW> constexpr uint32_t export_stage() { return (uint32_t)kExportTile * (NW == 4 ? 36 : 20) + 16; }

// Records for dst are packed into the staging buffer from byte stage_skew(dst)
// on, so that LDS and destination bytes agree modulo 16 and stage_copy moves
// aligned 16-byte words.
__device__ __forceinline__ uint32_t stage_skew(const uint8_t *dst)
{
    return (uint32_t)(reinterpret_cast<uintptr_t>(dst) & 15);
}

// Copies nbytes of LDS staging at lds to dst.  When lds and dst agree modulo
// 16 (lds = buffer + stage_skew(dst)): the head bytes up to a 16-byte
// boundary, then whole 16-byte words (consecutive threads take consecutive
// words, so every wave store is one contiguous 1 KB segment), then the tail
// bytes.  Otherwise bytes up to a 4-byte boundary, then 4-byte stores.
__device__ __forceinline__ void stage_copy(uint8_t *dst, const uint8_t *lds, uint32_t nbytes)
{
    const uintptr_t a = reinterpret_cast<uintptr_t>(dst) & 15;
    if (a == (reinterpret_cast<uintptr_t>(lds) & 15)) {
        const uint32_t end = (uint32_t)a + nbytes;
        const uint32_t q0 = ((uint32_t)a + 15) / 16, q1 = end / 16;  // whole words [q0, q1) from the aligned base
        if (q0 > q1) {  // inside one word
            if (threadIdx.x < nbytes) dst[threadIdx.x] = lds[threadIdx.x];
            return;
        }
        const uint32_t head = q0 * 16 - (uint32_t)a, tail = end - q1 * 16;
        if (threadIdx.x < head) dst[threadIdx.x] = lds[threadIdx.x];
        const uint32_t t0 = q1 * 16 - (uint32_t)a;  // first tail byte
        if (threadIdx.x < tail) dst[t0 + threadIdx.x] = lds[t0 + threadIdx.x];
        uint4 *dw = reinterpret_cast<uint4 *>(dst - a);
        const uint4 *sw = reinterpret_cast<const uint4 *>(lds - a);
        for (uint32_t q = q0 + threadIdx.x; q < q1; q += kPackBlock) dw[q] = sw[q];
        return;
    }
    const uint32_t head = (uint32_t)((4 - (reinterpret_cast<uintptr_t>(dst) & 3)) & 3);
    const uint32_t h = head < nbytes ? head : nbytes;
    if (threadIdx.x < h) dst[threadIdx.x] = lds[threadIdx.x];
    const uint32_t words = (nbytes - h) / 4;
    uint32_t *dw = reinterpret_cast<uint32_t *>(dst + h);
    for (uint32_t w = threadIdx.x; w < words; w += kPackBlock) {
        const uint8_t *b = lds + h + 4 * w;
        dw[w] = (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24);
    }
    const uint32_t done = h + 4 * words;
    if (threadIdx.x < nbytes - done) dst[done + threadIdx.x] = lds[done + threadIdx.x];
}

}  // namespace kc
