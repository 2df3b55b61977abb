// Partitioned counting: the MI355X form of KmerTree's prefix partitioning
// (NewKmerTree.h:329-407, 600-614, 983-1137) feeding per-node maps
// (NewKmerTree.h:526-570).  Where the reference moves 4 KiB blocks down a
// 256 x 16 x 16 x 16 tree by the top bits of the key, this pass sends every
// k-mer to one of 2^L table partitions by hash digits (a coarse digit of up
// to 8 bits and a fine digit from the bits below it, of one cheap mix of
// the key), in two exact scatter levels, and then counts each
// partition in an LDS-resident copy of its sub-table (and Bloom part):
//
//   front count    G persistent blocks; block b parses tiles b, b+G, ... and
//                  builds an LDS histogram of the coarse digit (top L1 <= 8
//                  bits) -> counts matrix [coarse][block]
//   scan           exclusive scan: exact coarse offsets per block
//   front scatter  the same blocks parse the same tiles again and write each
//                  key to its coarse region (write-combined through LDS)
//   fine count / scan / scatter   S blocks per coarse region, each over one
//                  contiguous slice of the region, on the remaining L2 bits
//   part count     one 1024-thread block per partition: load the sub-table
//                  (+ Bloom part) into LDS, run the Bloom split and the
//                  table inserts with LDS atomics, write back what changed.
//
// Every grid and buffer size is fixed by the launch shape before the first
// kernel runs (the input size is read on the device), so the pass is
// enqueued without a host round trip and can be captured in a graph.
//
// HBM traffic per k-mer: 8N bytes written and read twice by the scatter
// levels, read once by the part pass; the random accesses all hit LDS.  The
// FIRST list and the statistics use the same helpers as the direct path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "kc_dev_common.h"

namespace kcp {
using namespace kc;

constexpr int kKeyBlock = 256;
constexpr int kKeysPerThread = 16;
constexpr uint64_t kKeyTile = (uint64_t)kKeyBlock * kKeysPerThread;  // keys per front-end tile
constexpr int kPartThreads = 1024;
// How the partition pass hands FIRST occurrences to the FIRST list by
// default (kc_tune.first_mode 0): appended per wave every few steps from
// registers (kFirstGroup below); the autotuner also tries the other modes.
constexpr uint32_t kFirstModeDefault = 3;
// Steps of keys the partition pass keeps in flight ahead of the one counted.
constexpr int kPartPrefetch = 1;
// Front scatter blocks of one-word keys: 256 threads (the count pass's
// tiles), 32-byte chunks (8192-key chunks, ~32-key runs, two blocks per CU).
// 1024-thread blocks with 16-byte chunks (longer runs, one block per CU)
// measured slower on this ALU-heavy pass.
constexpr int kFrontBlock1 = kParseBlock;
constexpr int kFrontChunk1 = 32;
// Two-word keys: 16-byte chunks (4096 keys of 16 bytes = 64 KB of stage).
constexpr int kFrontBlock2 = kParseBlock;
constexpr int kFrontChunk2 = 16;
// The front scatter's blocks form kFrontGroups groups (block b in group
// b % kFrontGroups: the XCD the hardware dispatches it to); each group has
// one region per coarse digit and its blocks append their chunk runs to it
// through a cursor.  A block's open (partly written) lines then continue
// where another block of its XCD stopped, so the L2 of the XCD completes
// them; with a region per block, 64 blocks x 512 open runs filled the 4 MB
// L2 of an XCD and partial lines reached memory.
constexpr uint32_t kFrontGroups = 8;
// Group words: [coarse][group] cursors, region ends, region starts, then the
// overflow flag of a speculative scatter.
constexpr uint64_t kGcurEnd = 256ull * kFrontGroups;
constexpr uint64_t kGcurBase = 2ull * 256 * kFrontGroups;
constexpr uint64_t kGcurFlag = 3ull * 256 * kFrontGroups;
constexpr uint64_t kGroupWords = kGcurFlag + 1;
// Speculative front end (kc_tune.front_mode): every (coarse digit, group)
// region gets the same capacity, a share of key arrays kSpecSlackDiv-th
// larger than the input's window bound, and the scatter runs without a count
// pass; a run that does not fit raises the overflow flag, and the exact front
// end (count pass, scan, region placement, scatter) that follows it in the
// same launch sequence runs only then.  Hash digits load the regions evenly
// (Poisson: ~0.1% spread at 1M keys per region), so the redo is for skewed
// input (a few keys repeated all over).  Auto mode speculates from
// kSpecMinRegion keys per region (inputs of 32 MB or more).
constexpr uint64_t kSpecSlackDiv = 8;  // capacity slack: key_cap / 8
constexpr uint64_t kSpecAlign = 128;   // region capacity in whole lines of keys and digits
constexpr uint64_t kSpecMinRegion = 1u << 14;

struct Digits {
    uint32_t part_bits;  // L
    uint32_t l1;         // coarse bits = min(L, 8)
    uint32_t l2;         // fine bits = L - l1
    uint32_t pass_bits;  // multi-pass counting: only keys of `pass` enter
    uint32_t pass;
    uint32_t dig16;      // the front scatter writes fine digits as u16 (else u8)
};

template <int NW>
__device__ __forceinline__ bool in_pass_of(const uint64_t *key, const Digits &d)
{
    return in_pass<NW>(key, d.pass_bits, d.pass);
}

template <int NW>
__device__ __forceinline__ uint32_t coarse_of(const uint64_t *key, const Digits &d)
{
    return coarse_digit(coarse_mix<NW>(key), d.l1);
}

template <int NW>
__device__ __forceinline__ uint32_t fine_of(const uint64_t *key, const Digits &d)
{
    return fine_digit(coarse_mix<NW>(key), d.l1, d.l2);
}

// The front scatters write each key's fine digit beside it (u8, or u16
// above 256 fine digits), so that the fine count reads 1-2 bytes per key
// instead of the key.  dig is null when there is no fine level.
__device__ __forceinline__ void put_digit(uint8_t *dig, uint64_t at, uint32_t f, const Digits &d)
{
    if (!d.dig16) dig[at] = (uint8_t)f;
    else reinterpret_cast<uint16_t *>(dig)[at] = (uint16_t)f;
}

// A digit whose chunk does not fit its region (scatter guard): no staged
// key of it is written.  Real offsets differences stay far from it.
constexpr uint64_t kNoDelta = 0x8000000000000000ull;

// Exclusive scan of one value per thread of a 256-thread block: a shuffle
// scan inside each wave, then the four wave totals (two barriers).
__device__ __forceinline__ uint32_t scan256(uint32_t v, uint32_t *wsum, uint32_t &total)
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
    uint32_t wbase = 0;
    total = 0;
#pragma unroll
    for (uint32_t w = 0; w < 4; ++w) {
        const uint32_t sw = wsum[w];
        wbase += w < wave ? sw : 0;
        total += sw;
    }
    __syncthreads();
    return wbase + x - v;
}

// Exclusive scan of one value per thread of a TB-thread block: a shuffle
// scan inside each wave, then the wave totals (two barriers).
template <int TB>
__device__ __forceinline__ uint32_t scan_block(uint32_t v, uint32_t *wsum, uint32_t &total)
{
    constexpr uint32_t W = TB / 64;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    uint32_t wbase = 0;
    total = 0;
#pragma unroll
    for (uint32_t w = 0; w < W; ++w) {
        const uint32_t sw = wsum[w];
        wbase += w < wave ? sw : 0;
        total += sw;
    }
    __syncthreads();
    return wbase + x - v;
}

// The exact front end after a speculative scatter runs only when that
// scatter overflowed (its flag, `gate`, is nonzero); a null gate always runs.
// The flag is written by an earlier kernel of the stream and not changed
// while these run, so every thread of a launch decides alike.
__device__ __forceinline__ bool gated_off(const unsigned long long *gate) { return gate && *gate == 0; }

// ---- front end over framed bytes (persistent: block b owns tiles b + iG) --
template <int NW, bool PASS>
__global__ __launch_bounds__(kParseBlock) void k_front_count_seq(SeqSrc src, uint32_t k, Digits d,
                                                                 uint32_t *__restrict__ m1,
                                                                 const unsigned long long *gate)
{
    __shared__ ParseLds lds;
    __shared__ uint32_t hist[256];
    if (gated_off(gate)) return;
    uint64_t n, emit_from;
    seq_bounds(src, n, emit_from);
    hist[threadIdx.x] = 0;
    const uint64_t ntiles = (n + kParseTile - 1) / kParseTile;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        stage_tile(src.seq, n, tile, lds);
        roll_segment<NW>(lds, tile, n, emit_from, k, [&](const uint64_t *key, uint32_t) {
            if (!PASS || in_pass_of<NW>(key, d)) atomicAdd(&hist[coarse_of<NW>(key, d)], 1u);
        });
        __syncthreads();  // the next tile overwrites lds
    }
    __syncthreads();
    // counts in [coarse][group][block of the group] order: the scan lays each
    // group's blocks out as one region per coarse digit
    const uint32_t J = gridDim.x / kFrontGroups;
    const uint32_t col = (blockIdx.x % kFrontGroups) * J + blockIdx.x / kFrontGroups;
    if (threadIdx.x < (1u << d.l1)) m1[(uint64_t)threadIdx.x * gridDim.x + col] = hist[threadIdx.x];
}

// Scatter guard.  A scatter writes each key at its digit region's offset
// plus its rank, the region's size coming from a separate count pass.  Each
// block keeps how many keys its region of every digit has left: when a chunk
// holds more keys of a digit than that, none of them is written (they would
// run into the next region) and an error is counted, and so is a region left
// with holes at the end.  Either makes the counting call report KC_EFULL
// instead of returning wrong counts.
__device__ __forceinline__ void guard_report(unsigned long long *counters, uint32_t n_err)
{
    stat_add(counters, KC_STAT_ERRORS, n_err);
}

// Front scatter with write combining: each thread loads its 64-byte
// segment (and k-1 <= 4*HW bytes of history) straight into registers and
// rolls it in chunks of CB bytes; the block's keys of one chunk (up to
// TB x CB) are placed in LDS grouped by coarse digit and written out as
// contiguous runs (TB x CB / 256 keys per digit per chunk on average), each
// run at a place taken from its group's cursor for the digit (one atomic per
// digit and chunk).  A tile is TB x 64 bytes, the tiles of
// k_front_count_seq at TB = 256, so the group regions of the count pass hold
// exactly the keys of their blocks.
// Speculative (overflow non-null): the regions have a fixed capacity; a run
// that does not fit raises *overflow instead of counting an error, and the
// exact front end that follows redoes the placement.  gate: see gated_off.
template <int NW, int TB, int CB, bool PASS>
__global__ __launch_bounds__(TB) void k_front_scatter_seq_staged(SeqSrc src, uint32_t k, Digits d,
                                                                 unsigned long long *gcur,
                                                                 uint64_t *__restrict__ out, uint8_t *__restrict__ dig,
                                                                 unsigned long long *counters,
                                                                 const unsigned long long *gate,
                                                                 unsigned long long *overflow)
{
    constexpr int HW = NW == 1 ? 8 : 16;
    constexpr int kChunks = kParseSeg / CB;
    constexpr int kWords = CB / 4;
    constexpr int kStage = TB * CB;
    static_assert(TB == 256, "thread t takes the cursor of coarse digit t");
    __shared__ __align__(16) uint64_t stage[kStage * NW];
    __shared__ uint8_t sbin[kStage];
    __shared__ uint32_t hist[256], loff[256];
    __shared__ uint64_t delta[256];
    uint32_t *wsum = loff + 256 - TB / 64;  // the scan's wave sums: loff is written only after the scan
    if (gated_off(gate)) return;
    const uint32_t t = threadIdx.x;
    const uint32_t nbins = 1u << d.l1;
    uint32_t n_err = 0;
    uint64_t n, emit_from;
    seq_bounds(src, n, emit_from);
    // thread t's digit: its group's cursor and region end
    const uint32_t gslot = t * kFrontGroups + blockIdx.x % kFrontGroups;
    const uint64_t my_end = t < nbins ? gcur[kGcurEnd + gslot] : 0;
    hist[t] = 0;
    __syncthreads();
    constexpr uint64_t kTileBytes = (uint64_t)TB * kParseSeg;
    const uint64_t ntiles = (n + kTileBytes - 1) / kTileBytes;
    // this thread's segment of the tile; the next tile's is loaded once the
    // last chunk is parsed, so it is in flight during that chunk's write-out
    uint32_t w[HW + 16];
    if (blockIdx.x < ntiles) load_segment<HW>(src.seq, n, (uint64_t)blockIdx.x * kTileBytes + (uint64_t)t * kParseSeg, w);
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const uint64_t start = tile * kTileBytes + (uint64_t)t * kParseSeg;
        // segment bytes that emit nothing: before emit_from, or past the end
        const uint32_t skip = start >= n ? 64u
                            : emit_from > start ? (uint32_t)(emit_from - start < 64 ? emit_from - start : 64) : 0u;
        Roll<NW> roll(k);
        uint32_t have;
        roll.seed(w, have);  // the history: the previous segment's last 4*HW bytes
#pragma unroll
        for (int q = 0; q < kChunks; ++q) {
            uint64_t key[CB * NW];
            uint32_t rb[CB];       // rank << 8 | coarse digit
            uint32_t emitted = 0;  // bit e: byte e of the chunk emitted a key
#pragma unroll
            for (int jj = 0; jj < kWords; ++jj) {
                const Word4 x = word4(w[HW + q * kWords + jj]);
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const int e = jj * 4 + b;
                    push_byte<NW>(roll, x, b, have);
                    const bool emit = have >= k && (uint32_t)(4 * (q * kWords + jj) + b) >= skip;
                    if (emit) {
                        roll.canon(&key[e * NW]);
                        if (!PASS || in_pass_of<NW>(&key[e * NW], d)) {
                            const uint32_t c = coarse_of<NW>(&key[e * NW], d);
                            rb[e] = (atomicAdd(&hist[c], 1u) << 8) | c;
                            emitted |= 1u << e;
                        }
                    }
                }
            }
            if (q == kChunks - 1 && tile + gridDim.x < ntiles)
                load_segment<HW>(src.seq, n, start + (uint64_t)gridDim.x * kTileBytes, w);
            __syncthreads();
            uint32_t total;
            const uint32_t h = hist[t];
            const uint32_t off = scan_block<TB>(h, wsum, total);
            loff[t] = off;
            // staged key i of digit t goes to i + delta[t]; a run that would
            // pass the end of its group's region is not written
            uint64_t pos = 0;
            if (h) pos = atomicAdd(gcur + gslot, (unsigned long long)h);
            const bool fits = pos + h <= my_end;
            n_err += !fits;
            delta[t] = fits ? pos - off : kNoDelta;
            __syncthreads();
#pragma unroll
            for (int e = 0; e < CB; ++e) {
                if (emitted & (1u << e)) {
                    const uint32_t p = loff[rb[e] & 0xFF] + (rb[e] >> 8);
#pragma unroll
                    for (int x = 0; x < NW; ++x) stage[p * NW + x] = key[e * NW + x];
                    sbin[p] = (uint8_t)rb[e];
                }
            }
            __syncthreads();
            for (uint32_t i = t; i < total; i += TB) {
                const uint64_t dl = delta[sbin[i]];
                if (dl == kNoDelta) continue;
                const uint64_t at = i + dl;
                if constexpr (NW == 1) {
                    out[at] = stage[i];
                } else {
                    *reinterpret_cast<ulonglong2 *>(out + at * 2) = *reinterpret_cast<const ulonglong2 *>(stage + i * 2);
                }
                if (dig) put_digit(dig, at, fine_of<NW>(stage + i * NW, d), d);
            }
            __syncthreads();
            hist[t] = 0;
            __syncthreads();
        }
    }
    if (!overflow) guard_report(counters, n_err);
    else if (n_err) atomicOr(overflow, 1ull);
}

// Group cursors of the front scatter: for coarse digit c and group x the
// region [o1[c][x][0], o1[c][x + 1][0]) of the scan (the next group's start,
// or the next digit's, or the total): cursors, region ends, region starts.
// Under a gate (the redo of a speculative front end) it counts the redo.
__global__ void k_group_cursors(const uint64_t *o1, const uint32_t *m1, uint32_t G, uint32_t n_coarse,
                                unsigned long long *gcur, const unsigned long long *gate,
                                unsigned long long *counters)
{
    if (gated_off(gate)) return;
    const uint32_t c = threadIdx.x;
    if (gate) stat_add(counters, KC_STAT_FRONT_REDOS, c == 0 ? 1u : 0u);  // whole waves: stat_add reduces over the wave
    if (c >= n_coarse) return;
    const uint32_t J = G / kFrontGroups;
    const uint64_t last = (uint64_t)n_coarse * G - 1;
    for (uint32_t x = 0; x < kFrontGroups; ++x) {
        const uint64_t at = (uint64_t)c * G + (uint64_t)x * J, next = at + J;
        gcur[c * kFrontGroups + x] = o1[at];
        gcur[kGcurBase + c * kFrontGroups + x] = o1[at];
        gcur[kGcurEnd + c * kFrontGroups + x] = next <= last ? o1[next] : o1[last] + m1[last];
    }
}

// Regions of a speculative front scatter: region i = (coarse c, group x),
// i = c * kFrontGroups + x, holds [i * cap, (i + 1) * cap); the overflow flag
// is cleared.
__global__ void k_spec_regions(unsigned long long *gcur, uint32_t n_coarse, uint64_t cap)
{
    for (uint32_t i = threadIdx.x; i < n_coarse * kFrontGroups; i += blockDim.x) {
        const uint64_t b = (uint64_t)i * cap;
        gcur[i] = b;
        gcur[kGcurBase + i] = b;
        gcur[kGcurEnd + i] = b + cap;
    }
    if (threadIdx.x == 0) gcur[kGcurFlag] = 0;
}

// Regions of the key-input front end (one contiguous region per coarse
// digit, from the coarse starts) in the group words the fine level reads:
// group 0 holds the region, the other groups are empty.
__global__ void k_group_from_starts(const uint64_t *starts, uint32_t n_coarse, unsigned long long *gcur)
{
    const uint32_t c = threadIdx.x;
    if (c >= n_coarse) return;
    for (uint32_t x = 0; x < kFrontGroups; ++x) {
        const uint32_t i = c * kFrontGroups + x;
        gcur[kGcurBase + i] = x ? starts[c + 1] : starts[c];
        gcur[i] = starts[c + 1];
        gcur[kGcurEnd + i] = starts[c + 1];
    }
}

// Scatter guard of the grouped front scatter: a region whose cursor did not
// reach its end holds keys the count pass saw and the scatter did not.
__global__ void k_group_check(const unsigned long long *gcur, uint32_t n_coarse, unsigned long long *counters,
                              const unsigned long long *gate)
{
    if (gated_off(gate)) return;
    uint32_t n_err = 0;
    for (uint32_t i = threadIdx.x; i < n_coarse * kFrontGroups; i += blockDim.x)
        n_err += gcur[i] != gcur[kGcurEnd + i];
    guard_report(counters, n_err);
}

template <int NW>
__global__ __launch_bounds__(kKeyBlock) void k_front_count_keys(const uint64_t *__restrict__ keys, uint64_t n,
                                                                Digits d, uint32_t *__restrict__ m1)
{
    __shared__ uint32_t hist[256];
    hist[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t ntiles = (n + kKeyTile - 1) / kKeyTile;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        for (int j = 0; j < kKeysPerThread; ++j) {
            const uint64_t i = tile * kKeyTile + (uint64_t)j * kKeyBlock + threadIdx.x;
            if (i < n && in_pass_of<NW>(keys + i * NW, d)) atomicAdd(&hist[coarse_of<NW>(keys + i * NW, d)], 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x < (1u << d.l1)) m1[(uint64_t)threadIdx.x * gridDim.x + blockIdx.x] = hist[threadIdx.x];
}

template <int NW>
__global__ __launch_bounds__(kKeyBlock) void k_front_scatter_keys(const uint64_t *__restrict__ keys, uint64_t n,
                                                                  Digits d, const uint64_t *__restrict__ o1,
                                                                  const uint32_t *__restrict__ m1,
                                                                  uint64_t *__restrict__ out, uint8_t *__restrict__ dig,
                                                                  unsigned long long *counters)
{
    __shared__ unsigned long long cur[256], lend[256];
    const uint32_t nbins = 1u << d.l1;
    if (threadIdx.x < nbins) {
        const uint64_t j = (uint64_t)threadIdx.x * gridDim.x + blockIdx.x;
        cur[threadIdx.x] = o1[j];
        lend[threadIdx.x] = o1[j] + m1[j];
    }
    __syncthreads();
    uint32_t n_err = 0;
    const uint64_t ntiles = (n + kKeyTile - 1) / kKeyTile;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        for (int j = 0; j < kKeysPerThread; ++j) {
            const uint64_t i = tile * kKeyTile + (uint64_t)j * kKeyBlock + threadIdx.x;
            if (i < n && in_pass_of<NW>(keys + i * NW, d)) {
                const uint32_t c = coarse_of<NW>(keys + i * NW, d);
                const uint64_t pos = atomicAdd(&cur[c], 1ull);
                if (pos >= lend[c]) {  // past this block's region (scatter guard)
                    ++n_err;
                    continue;
                }
#pragma unroll
                for (int w = 0; w < NW; ++w) out[pos * NW + w] = keys[i * NW + w];
                if (dig) put_digit(dig, pos, fine_of<NW>(keys + i * NW, d), d);
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < nbins) n_err += cur[threadIdx.x] < lend[threadIdx.x];
    guard_report(counters, n_err);
}

// Test-only fault injection (KC_TEST_SKEW_COUNTS): adds delta (+1 or -1) to
// the first nonzero count of a count matrix, so that the scatter that uses
// it has one region too large or too small.  Every write stays inside the
// key arrays: a total one larger is still below key_cap (a pass has fewer
// windows than input bytes).
__global__ void k_skew_count(uint32_t *m, uint64_t n, int delta)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (uint64_t i = 0; i < n; ++i)
        if (m[i]) {
            m[i] = (uint32_t)((int64_t)m[i] + delta);
            return;
        }
}

// Coarse region starts: o1[c * G] for every coarse digit, plus the total.
__global__ void k_coarse_starts(const uint64_t *o1, const uint32_t *m1, uint32_t G, uint32_t n_coarse,
                                uint64_t *starts, const unsigned long long *gate)
{
    if (gated_off(gate)) return;
    const uint32_t c = threadIdx.x;
    if (c < n_coarse) starts[c] = o1[(uint64_t)c * G];
    if (c == 0) {
        const uint64_t last = (uint64_t)n_coarse * G - 1;
        starts[n_coarse] = o1[last] + m1[last];
    }
}

// ---- second level: fine digit inside each coarse region -----------------
// The keys of coarse region c are its group regions [start, min(cursor, end))
// in group order: back to back when the count pass placed them, a fixed
// capacity apart when the scatter speculated.  Block b works on slice
// s = b % S of coarse region c = b / S: the part [L*s/S, L*(s+1)/S) of their
// concatenation (L keys in all), returned as one piece [rb[x], re[x]) per
// group region in LDS (empty pieces are [0, 0)).  Its counts go to
// m2[((c << L2) + f) * S + s], so that the scan lays partition (c, f) out as
// slices 0..S-1 in order.
__device__ __forceinline__ void slice_pieces(const unsigned long long *gcur, uint32_t S, uint32_t &c, uint32_t &s,
                                             uint64_t *rb, uint64_t *re)
{
    c = blockIdx.x / S;
    s = blockIdx.x % S;
    if (threadIdx.x == 0) {
        uint64_t b[kFrontGroups], l[kFrontGroups], L = 0;
#pragma unroll
        for (uint32_t x = 0; x < kFrontGroups; ++x) {
            const uint64_t i = (uint64_t)c * kFrontGroups + x;
            const uint64_t e = gcur[i] < gcur[kGcurEnd + i] ? gcur[i] : gcur[kGcurEnd + i];
            b[x] = gcur[kGcurBase + i];
            l[x] = e > b[x] ? e - b[x] : 0;
            L += l[x];
        }
        const uint64_t lo = L * s / S, hi = L * (s + 1) / S;
        uint64_t acc = 0;
#pragma unroll
        for (uint32_t x = 0; x < kFrontGroups; ++x) {
            const uint64_t a = lo > acc ? lo : acc, z = hi < acc + l[x] ? hi : acc + l[x];
            rb[x] = a < z ? b[x] + (a - acc) : 0;
            re[x] = a < z ? b[x] + (z - acc) : 0;
            acc += l[x];
        }
    }
    __syncthreads();
}

// Fine count over the digits the front scatter wrote: 16-byte loads (16 u8
// or 8 u16 digits), U of them per thread in flight, LDS histogram.
template <typename DT>
__global__ __launch_bounds__(kKeyBlock) void k_fine_count_dig(const DT *__restrict__ dig,
                                                              const unsigned long long *__restrict__ gcur, Digits d,
                                                              uint32_t S, uint32_t *__restrict__ m2)
{
    __shared__ uint32_t hist[1024];
    __shared__ uint64_t rb[kFrontGroups], re[kFrontGroups];
    uint32_t c, s;
    slice_pieces(gcur, S, c, s, rb, re);
    const uint32_t nf = 1u << d.l2, fmask = nf - 1;  // the mask keeps any stored value inside hist
    for (uint32_t f = threadIdx.x; f < nf; f += kKeyBlock) hist[f] = 0;
    __syncthreads();
    constexpr uint32_t PER = 16 / sizeof(DT);  // digits per 16-byte load
    for (uint32_t g = 0; g < kFrontGroups; ++g) {
        const uint64_t beg = rb[g], end = re[g];
        if (beg >= end) continue;
        const uint64_t a0 = (beg + PER - 1) / PER * PER < end ? (beg + PER - 1) / PER * PER : end;
        const uint64_t a1 = end / PER * PER > a0 ? end / PER * PER : a0;
        for (uint64_t i = beg + threadIdx.x; i < a0; i += kKeyBlock) atomicAdd(&hist[dig[i] & fmask], 1u);
        for (uint64_t i = a1 + threadIdx.x; i < end; i += kKeyBlock) atomicAdd(&hist[dig[i] & fmask], 1u);
        const uint4 *v = reinterpret_cast<const uint4 *>(dig);
        const uint64_t q0 = a0 / PER, q1 = a1 / PER;
        constexpr uint32_t U = 4;
        for (uint64_t qb = q0; qb < q1; qb += (uint64_t)U * kKeyBlock) {
            uint4 x[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const uint64_t q = qb + u * kKeyBlock + threadIdx.x;
                x[u] = v[q < q1 ? q : q0];
            }
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                if (qb + u * kKeyBlock + threadIdx.x >= q1) continue;
                const uint32_t w[4] = {x[u].x, x[u].y, x[u].z, x[u].w};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
#pragma unroll
                    for (uint32_t b = 0; b < 4 / sizeof(DT); ++b)
                        atomicAdd(&hist[(w[j] >> (8 * sizeof(DT) * b)) & fmask], 1u);
                }
            }
        }
    }
    __syncthreads();
    const uint64_t row = (uint64_t)c << d.l2;
    for (uint32_t f = threadIdx.x; f < nf; f += kKeyBlock) m2[(row + f) * S + s] = hist[f];
}

// Fine scatter with write combining: a slice is handled in chunks of
// TB x FPT keys held in registers; each chunk's keys are ranked per fine
// digit, placed in LDS grouped by digit and written out as runs.  Longer
// runs leave fewer partly written lines: a chunk is 16384 one-word or 8192
// two-word keys (one 1024-thread block per CU; ~64-key runs at 256 fine
// digits, ~16 or ~8 at 1024).  The next chunk's keys are loaded while this
// one is written out, so one block per CU still keeps reads in flight.
// SB: keep each staged key's digit in LDS (sbin); without it the digit is
// recomputed from the staged key at write-out, which leaves the LDS to
// larger chunks.
template <int NW, int NF, int TB, int FPT, bool SB = true>
__global__ __launch_bounds__(TB) void k_fine_scatter_staged(const uint64_t *__restrict__ keys,
                                                            const unsigned long long *__restrict__ gcur, Digits d,
                                                            uint32_t S, const uint64_t *__restrict__ o2,
                                                            const uint32_t *__restrict__ m2,
                                                            uint64_t *__restrict__ out,
                                                            unsigned long long *counters)
{
    using BinT = typename std::conditional<NF <= 256, uint8_t, uint16_t>::type;
    constexpr uint32_t CH = (uint32_t)TB * FPT;
    __shared__ __align__(16) uint64_t stage[CH * NW];
    static_assert(NF <= TB, "thread t keeps the region of fine digit t");
    __shared__ BinT sbin[SB ? CH : 1];
    __shared__ uint32_t hist[NF], loff[NF];
    __shared__ uint64_t delta[NF];
    __shared__ uint64_t rb[kFrontGroups], re[kFrontGroups];
    uint32_t *tmp = loff + NF - TB / 64;  // the scan's wave sums: loff is rewritten only after the scan
    uint32_t c, s;
    slice_pieces(gcur, S, c, s, rb, re);
    const uint32_t nf = 1u << d.l2;
    const uint32_t t = threadIdx.x;
    const uint64_t row = (uint64_t)c << d.l2;
    uint32_t n_err = 0;
    // thread t's digit: where its region continues and how many keys it has left
    const bool mine = t < nf;
    uint64_t my_base = mine ? o2[(row + t) * S + s] : 0;
    uint32_t my_left = mine ? m2[(row + t) * S + s] : 0;
    if (mine) hist[t] = 0;
    __syncthreads();
    // a chunk's FPT keys per thread (index clamped into the chunk)
    uint64_t key[FPT][NW];
    auto load_chunk = [&](uint64_t cs, uint64_t end) {
        const uint32_t cn = (uint32_t)(end - cs < CH ? end - cs : CH);
#pragma unroll
        for (uint32_t j = 0; j < (uint32_t)FPT; ++j) {
            const uint32_t i = j * TB + t;
            const uint64_t *src = keys + (cs + (i < cn ? i : cn - 1)) * NW;
            if constexpr (NW == 2) {
                const ulonglong2 v = *reinterpret_cast<const ulonglong2 *>(src);
                key[j][0] = v.x;
                key[j][1] = v.y;
            } else {
                key[j][0] = src[0];
            }
        }
    };
    // chunks of the slice's pieces in order: [cs, end) of piece x
    uint32_t x = 0;
    while (x < kFrontGroups && rb[x] >= re[x]) ++x;
    uint64_t cs = x < kFrontGroups ? rb[x] : 0, end = x < kFrontGroups ? re[x] : 0;
    if (x < kFrontGroups) load_chunk(cs, end);
    while (x < kFrontGroups) {
        const uint32_t cn = (uint32_t)(end - cs < CH ? end - cs : CH);
        uint32_t bin[FPT], rank[FPT];
#pragma unroll
        for (uint32_t j = 0; j < (uint32_t)FPT; ++j) {
            const uint32_t i = j * TB + t;
            if (i < cn) {
                bin[j] = fine_of<NW>(key[j], d);
                rank[j] = atomicAdd(&hist[bin[j]], 1u);
            }
        }
        __syncthreads();
        const uint32_t h = mine ? hist[t] : 0;
        uint32_t total;
        const uint32_t off = scan_block<TB>(h, tmp, total);
        // staged key p of digit t goes to p + delta[t]; a chunk with more
        // keys of the digit than its region has left is not written
        const bool fits = h <= my_left;
        if (mine) {
            loff[t] = off;
            n_err += !fits;
            delta[t] = fits ? my_base - off : kNoDelta;
        }
        __syncthreads();
#pragma unroll
        for (uint32_t j = 0; j < (uint32_t)FPT; ++j) {
            const uint32_t i = j * TB + t;
            if (i < cn) {
                const uint32_t p = loff[bin[j]] + rank[j];
#pragma unroll
                for (int w = 0; w < NW; ++w) stage[p * NW + w] = key[j][w];
                if constexpr (SB) sbin[p] = (BinT)bin[j];
            }
        }
        // the next chunk (the rest of this piece, else the next piece's
        // first): its loads stay in flight through the write-out below (the
        // barriers wait for LDS operations only)
        uint32_t nx = x;
        uint64_t ncs = cs + CH, nend = end;
        if (ncs >= end) {
            ++nx;
            while (nx < kFrontGroups && rb[nx] >= re[nx]) ++nx;
            ncs = nx < kFrontGroups ? rb[nx] : 0;
            nend = nx < kFrontGroups ? re[nx] : 0;
        }
        if (nx < kFrontGroups) load_chunk(ncs, nend);
        __syncthreads();
        for (uint32_t p = t; p < cn; p += TB) {
            uint32_t f;
            if constexpr (SB) f = sbin[p];
            else f = fine_of<NW>(stage + p * NW, d);
            const uint64_t dl = delta[f];
            if (dl == kNoDelta) continue;
            const uint64_t g = p + dl;
            if constexpr (NW == 2) {
                *reinterpret_cast<ulonglong2 *>(out + g * 2) = *reinterpret_cast<const ulonglong2 *>(stage + p * 2);
            } else {
                out[g] = stage[p];
            }
        }
        __syncthreads();
        if (mine) {
            my_base += h;
            my_left = fits ? my_left - h : 0;
            hist[t] = 0;
        }
        __syncthreads();
        x = nx;
        cs = ncs;
        end = nend;
    }
    n_err += mine && my_left != 0;  // keys the count pass saw and this one did not
    guard_report(counters, n_err);
}

// Partition p starts at o2[p * S] (two levels) or at the coarse start p (one
// level); part_off[2^L] = total keys (the fine scan's total, or the last
// coarse start).
__global__ void k_part_offsets(const uint64_t *o2, const uint32_t *m2, const uint64_t *starts, Digits d, uint32_t S,
                               uint64_t *part_off)
{
    const uint64_t n_parts = 1ull << d.part_bits;
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p <= n_parts;
         p += (uint64_t)gridDim.x * blockDim.x) {
        if (p == n_parts) part_off[p] = d.l2 ? o2[n_parts * S - 1] + m2[n_parts * S - 1] : starts[1u << d.l1];
        else part_off[p] = d.l2 ? o2[p * S] : starts[p];
    }
}

// ---- per-partition counting in LDS ---------------------------------------
template <int NW, bool BLOOM>
struct PartLds {
    static constexpr uint32_t PS = kc_lds_part_slots(NW, BLOOM);
    static constexpr uint32_t BS = BLOOM ? kc_lds_part_bloom(NW) : 1;
    unsigned long long key[PS * NW];  // word 0 at [s], word 1 at [PS + s]
    uint32_t cnt[PS];                 // occurrences added in this pass
    unsigned long long bloom[BS];     // one 64-bit Bloom block each
};

// The Bloom singleton split on the LDS copy (block and mask from
// bloom_in_part): FIRST exactly when the fetch_or of the key's mask into its
// block finds a bit missing (bloom_first in kc_dev_common.h: the reference's
// BF1 word does not change that decision and is not kept).  (A plain read
// first, with the atomic only for keys that miss a bit, measured slower: a
// wave of 64 keys nearly always holds one that needs the atomic, so the
// wave issues both.)
template <int NW, bool BLOOM>
__device__ __forceinline__ bool lds_bloom_first(PartLds<NW, BLOOM> &S, uint64_t h, uint32_t bmask)
{
    uint64_t blk, m;
    bloom_in_part(h, bmask, blk, m);
    const uint64_t b = atomicOr(&S.bloom[blk], (unsigned long long)m);
    return (b & m) != m;
}

template <int NW, bool BLOOM>
__device__ __forceinline__ bool lds_insert(PartLds<NW, BLOOM> &S, const uint64_t *key, uint64_t h, uint32_t pmask,
                                           uint32_t &n_new, uint32_t &n_err)
{
    constexpr uint32_t PS = PartLds<NW, BLOOM>::PS;
    uint32_t s = probe_start(h) & pmask;
    if constexpr (NW == 1) {
        // One exit condition for the whole probe loop (claimed, found, or
        // every slot tried) keeps its exec-mask bookkeeping short; the count
        // is added once the loop has settled.  (A plain read of the first
        // probe slot before the CAS measured slower, as for the Bloom word.)
        uint64_t w0;
        uint32_t left = pmask + 1;
        bool go;
        do {
            w0 = atomicCAS(&S.key[s], (unsigned long long)KC_EMPTY, (unsigned long long)key[0]);
            const bool settled = (w0 == KC_EMPTY) | (w0 == key[0]);
            --left;
            go = !settled & (left != 0);
            s = go ? (s + 1) & pmask : s;
        } while (go);
        if ((w0 == KC_EMPTY) | (w0 == key[0])) {
            n_new += w0 == KC_EMPTY;
            atomicAdd(&S.cnt[s], 1u);
            return true;
        }
        ++n_err;
        return false;
    } else {
        for (uint32_t probe = 0; probe <= pmask; ++probe, s = (s + 1) & pmask) {
            // one LDS CAS per probe: it claims an empty slot or returns the
            // slot's key (a present key is the common case and is not written)
            const uint64_t w0 = atomicCAS(&S.key[s], (unsigned long long)KC_EMPTY, (unsigned long long)key[0]);
            const bool owner = w0 == KC_EMPTY;
            if (owner)
                __hip_atomic_store(&S.key[PS + s], (unsigned long long)key[1], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            __builtin_amdgcn_wave_barrier();
            if (owner) {
                ++n_new;
                atomicAdd(&S.cnt[s], 1u);
                return true;
            }
            if (w0 == key[0]) {
                uint64_t w1 = __hip_atomic_load(&S.key[PS + s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                for (int spin = 0; w1 == KC_EMPTY && spin < (1 << 16); ++spin) {
                    __builtin_amdgcn_s_sleep(1);
                    w1 = __hip_atomic_load(&S.key[PS + s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
                if (w1 == KC_EMPTY) {
                    ++n_err;
                    return false;
                }
                if (w1 == key[1]) {
                    atomicAdd(&S.cnt[s], 1u);
                    return true;
                }
            }
        }
        ++n_err;
        return false;
    }
}

// Position of the n-th key counted by wave `wave` of a partition block (the
// block takes 1024 consecutive keys per step, wave w the w-th 64 of them).
__device__ __forceinline__ uint64_t wave_pos(uint64_t beg, uint32_t wave, uint32_t n)
{
    return beg + (uint64_t)(n >> 6) * kPartThreads + wave * 64u + (n & 63u);
}

// FIRST occurrences (the first occurrence of a key, for the FIRST list):
//   FM = kFirstStaged: staged in `first_out` (the pipeline's other key
//     array, idle during this pass) at the positions of the wave's own keys:
//     the wave's n-th FIRST goes where its n-th key was, so runs of different
//     waves and partitions never overlap and no atomic is needed while
//     counting.  After the loop each wave moves its run to the FIRST list in
//     pieces of up to 1024 keys, one cursor atomic per piece.
//   FM = kFirstStep: each wave step appends its FIRSTs at once (one cursor
//     atomic per wave step).
//   FM = kFirstGroup: each lane keeps its FIRSTs of kGroup consecutive
//     steps in registers and the wave appends them together (one cursor
//     atomic and one contiguous run per kGroup steps): the FIRST bytes move
//     once, with few atomics and no LDS.
// Segments rotate with the partition, the wave and the step or piece.
// FRESH: the table is logically empty (a lazy clear): partitions start from
// empty slots and zero Bloom words in LDS instead of loading them, and every
// slot and Bloom word is written back, also for partitions with no keys.
constexpr int kFirstStaged = 1, kFirstStep = 2, kFirstGroup = 3;
template <int NW> constexpr int kGroup = 4;  // 4 steps: <= 64 VGPRs, two 1024-thread blocks per CU

template <int NW, bool BLOOM, int FM, bool FRESH>
__global__ __launch_bounds__(kPartThreads) void k_count_parts(const uint64_t *__restrict__ keys,
                                                              const uint64_t *__restrict__ part_off, TableView tv,
                                                              uint64_t *__restrict__ first_out)
{
    constexpr uint32_t PS = PartLds<NW, BLOOM>::PS;
    constexpr int SW = slot_words<NW>();
    __shared__ PartLds<NW, BLOOM> S;
    const uint64_t part = blockIdx.x;
    const uint64_t beg = part_off[part], end = part_off[part + 1];
    if (!FRESH && beg == end) return;  // nothing to count: the partition stays as it is
    const uint32_t psz = 1u << tv.part_log2;
    const uint64_t sbase = part << tv.part_log2;
    for (uint32_t s = threadIdx.x; s < psz; s += kPartThreads) {
        if constexpr (FRESH) {
            S.key[s] = KC_EMPTY;
            if constexpr (NW == 2) S.key[PS + s] = KC_EMPTY;
        } else {
            const uint64_t *g = tv.slots + (sbase + s) * SW;
            S.key[s] = g[0];
            if constexpr (NW == 2) S.key[PS + s] = g[1];
        }
        S.cnt[s] = 0;
    }
    uint32_t bsz = 0;
    uint64_t bbase = 0;
    if constexpr (BLOOM) {
        bsz = 1u << tv.bloom_part_log2;
        bbase = part << tv.bloom_part_log2;
        for (uint32_t b = threadIdx.x; b < bsz; b += kPartThreads) S.bloom[b] = FRESH ? 0 : tv.bloom[bbase + b];
    }
    __syncthreads();

    uint32_t n_k = 0, n_new = 0, n_err = 0;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint32_t fcur = 0;  // FIRSTs staged by this wave (wave-uniform)
    // kFirstGroup: this lane's FIRSTs of the group's steps and the waves' ballots
    constexpr int G = FM == kFirstGroup ? kGroup<NW> : 1;
    uint64_t fk[G][NW];
    uint64_t fbs[G];
    uint32_t held = 0, group = 0;  // steps held (wave-uniform), groups appended
    auto append_group = [&]() {
        uint32_t tot = 0;
#pragma unroll
        for (int j = 0; j < G; ++j)
            if ((uint32_t)j < held) tot += (uint32_t)__popcll(fbs[j]);
        if (tot) {
            const uint32_t seg = first_segment((uint32_t)part, wave, group);
            unsigned long long at0 = 0;
            if (lane == 0) at0 = atomicAdd(tv.counters + KC_CURSOR_BASE + (uint64_t)seg * KC_LINE_WORDS, (unsigned long long)tot);
            at0 = __shfl(at0, 0, 64);
            uint32_t pre = 0;
#pragma unroll
            for (int j = 0; j < G; ++j) {
                if ((uint32_t)j >= held) break;
                if ((fbs[j] >> lane) & 1) {
                    const uint64_t at = at0 + pre + (uint32_t)__popcll(fbs[j] & lanemask_lt());
                    if (at < tv.seg_cap) {
                        uint64_t *dst = tv.low + ((uint64_t)seg * tv.seg_cap + at) * NW;
#pragma unroll
                        for (int w = 0; w < NW; ++w) dst[w] = fk[j][w];
                    } else {
                        // no room on the list: counted in the partition's table instead
                        lds_insert<NW, BLOOM>(S, fk[j], slot_mix<NW>(fk[j]), psz - 1, n_new, n_err);
                    }
                }
                pre += (uint32_t)__popcll(fbs[j]);
            }
        }
        ++group;
        held = 0;
    };
    // the keys of the next PF steps are in flight while this step's key is
    // counted (indexes past the end read the partition's first key)
    constexpr int PF = kPartPrefetch;
    uint64_t nxt[PF][NW];
#pragma unroll
    for (int p = 0; p < PF; ++p) {
        const uint64_t i = beg + (uint64_t)p * kPartThreads + threadIdx.x;
#pragma unroll
        for (int w = 0; w < NW; ++w) nxt[p][w] = keys[(i < end ? i : beg) * NW + w];
    }
    for (uint64_t gbase = beg; gbase < end; gbase += (uint64_t)G * kPartThreads) {
#pragma unroll
        for (int j = 0; j < G; ++j) {
            const uint64_t base = gbase + (uint64_t)j * kPartThreads;
            if (base >= end) break;
            const uint64_t i = base + threadIdx.x;
            const bool valid = i < end;
            uint64_t key[NW];
#pragma unroll
            for (int w = 0; w < NW; ++w) key[w] = nxt[0][w];
#pragma unroll
            for (int p = 0; p + 1 < PF; ++p)
#pragma unroll
                for (int w = 0; w < NW; ++w) nxt[p][w] = nxt[p + 1][w];
            if (base + (uint64_t)PF * kPartThreads < end) {
                const uint64_t jj = i + (uint64_t)PF * kPartThreads;
#pragma unroll
                for (int w = 0; w < NW; ++w) nxt[PF - 1][w] = keys[(jj < end ? jj : beg) * NW + w];
            }
            n_k += valid;
            const uint64_t h = slot_mix<NW>(key);  // Bloom probe and slot probe start
            bool to_table = valid;
            if constexpr (BLOOM) {
                const bool first = valid && lds_bloom_first<NW, BLOOM>(S, h, bsz - 1);
                if constexpr (FM == kFirstStaged) {
                    const uint64_t fb = __ballot(first);
                    if (first) {
                        uint64_t *dst = first_out + wave_pos(beg, wave, fcur + (uint32_t)__popcll(fb & lanemask_lt())) * NW;
#pragma unroll
                        for (int w = 0; w < NW; ++w) dst[w] = key[w];
                    }
                    fcur += (uint32_t)__popcll(fb);
                    to_table = valid && !first;
                } else if constexpr (FM == kFirstStep) {
                    const bool kept = first_append<NW>(tv, first, key, (uint32_t)((base - beg) >> 10));
                    to_table = valid && (!first || !kept);  // no room on the list: counted in the table
                } else {
                    fbs[j] = __ballot(first);
#pragma unroll
                    for (int w = 0; w < NW; ++w) fk[j][w] = key[w];
                    held = (uint32_t)j + 1;
                    to_table = valid && !first;
                }
            }
            if (to_table) lds_insert<NW, BLOOM>(S, key, h, psz - 1, n_new, n_err);
        }
        if constexpr (BLOOM && FM == kFirstGroup) append_group();
    }
    if constexpr (BLOOM && FM == kFirstStaged) {
        __threadfence_block();  // the wave's staged FIRSTs are read back by other lanes
        for (uint32_t p0 = 0; p0 < fcur; p0 += 1024) {
            const uint32_t cnt = fcur - p0 < 1024 ? fcur - p0 : 1024;
            const uint32_t seg = first_segment((uint32_t)part, wave, p0 / 1024);
            unsigned long long at = 0;
            if (lane == 0)
                at = atomicAdd(tv.counters + KC_CURSOR_BASE + (uint64_t)seg * KC_LINE_WORDS, (unsigned long long)cnt);
            at = __shfl(at, 0, 64);
            for (uint32_t n = lane; n < cnt; n += 64) {
                const uint64_t *src = first_out + wave_pos(beg, wave, p0 + n) * NW;
                if (at + n < tv.seg_cap) {
                    uint64_t *dst = tv.low + ((uint64_t)seg * tv.seg_cap + at + n) * NW;
#pragma unroll
                    for (int w = 0; w < NW; ++w) dst[w] = src[w];
                } else {
                    // no room on the list: the occurrence is counted in the
                    // partition's table (still in LDS) instead
                    uint64_t key[NW];
#pragma unroll
                    for (int w = 0; w < NW; ++w) key[w] = src[w];
                    lds_insert<NW, BLOOM>(S, key, slot_mix<NW>(key), psz - 1, n_new, n_err);
                }
            }
        }
    }
    __syncthreads();
    for (uint32_t s = threadIdx.x; s < psz; s += kPartThreads) {
        const uint32_t c = S.cnt[s];
        uint64_t *g = tv.slots + (sbase + s) * SW;
        if constexpr (FRESH) {  // every slot: the memory held no table before
            g[0] = S.key[s];
            if constexpr (NW == 2) {
                g[1] = S.key[PS + s];
                g[3] = 0;
            }
            g[NW == 1 ? 1 : 2] = c;
        } else {
            if (!c) continue;
            g[0] = S.key[s];
            if constexpr (NW == 2) g[1] = S.key[PS + s];
            g[NW == 1 ? 1 : 2] += c;
        }
    }
    if constexpr (BLOOM) {
        for (uint32_t b = threadIdx.x; b < bsz; b += kPartThreads) tv.bloom[bbase + b] = S.bloom[b];
    }
    stat_add(tv.counters, KC_STAT_KMERS, n_k);
    stat_add(tv.counters, KC_STAT_DISTINCT, n_new);
    stat_add(tv.counters, KC_STAT_ERRORS, n_err);
}

}  // namespace kcp

using namespace kcp;

namespace {

Digits digits_of(const kc_table *t)
{
    Digits d;
    d.part_bits = t->part_bits;
    d.l1 = std::min<uint32_t>(d.part_bits, kCoarseMaxBits);
    d.l2 = d.part_bits - d.l1;
    d.pass_bits = t->pass_bits;
    d.pass = t->pass;
    d.dig16 = d.l2 > 8;
    return d;
}

template <typename T>
bool dev_alloc(T *&p, uint64_t n)
{
    return hipMalloc((void **)&p, std::max<uint64_t>(n, 1) * sizeof(T)) == hipSuccess;
}

}  // namespace

// Whether the LDS partition pass can hold this table's partitions.
bool kc_part_eligible(const kc_table *t)
{
    if (t->nw > 2) return false;  // four-word keys (k >= 64) count on the direct path
    const bool bloom = t->bloom != nullptr;
    if ((1ull << t->part_log2) > kc_lds_part_slots(t->nw, bloom)) return false;
    if (bloom && (1ull << t->bloom_part_log2) > kc_lds_part_bloom(t->nw)) return false;
    return t->part_bits <= 18;
}

void kc_part_scratch_free(kc_part_scratch &s)
{
    for (void *p : {(void *)s.keys_a, (void *)s.keys_b, (void *)s.m1, (void *)s.o1, (void *)s.m2, (void *)s.o2,
                    (void *)s.sums, (void *)s.part_off, (void *)s.starts, (void *)s.dig, (void *)s.gcur})
        if (p) (void)hipFree(p);
    s = kc_part_scratch();
}

void kc_part_free(kc_table *t) { kc_part_scratch_free(t->ps); }

// Launch shape for up to key_cap keys: the tuned values, else
// 8 front blocks per CU (rounded to a multiple of 64) and 8 slices per
// coarse region.
kc_pipe_shape kc_part_shape(const kc_table *t, const kc_tune &tune, uint64_t key_cap)
{
    kc_pipe_shape sh;
    uint32_t g = tune.front_blocks ? tune.front_blocks : t->ctx->cu_count * 8;
    g = std::max<uint32_t>(64, std::min<uint32_t>(g, 1u << 16));
    sh.front_blocks = (g + 63) / 64 * 64;
    sh.fine_slices = std::max<uint32_t>(1, std::min<uint32_t>(tune.fine_slices ? tune.fine_slices : 8, 64));
    sh.key_cap = key_cap;
    sh.first_mode = tune.first_mode >= 1 && tune.first_mode <= 3 ? tune.first_mode : kFirstModeDefault;
    sh.front_mode = tune.front_mode <= 2 ? tune.front_mode : 0;
    return sh;
}

// Keys the first key array (and the digit array) holds: key_cap plus the
// speculative regions' slack and one alignment unit per region.
static uint64_t front_cap_for(uint64_t key_cap)
{
    const uint64_t kc = std::max<uint64_t>(key_cap, 1);
    return kc + kc / kSpecSlackDiv + 256 * kFrontGroups * kSpecAlign;
}

// Capacity of each speculative region for a front key array of front_cap
// keys (whole lines), or 0 when this pass places its keys exactly: one
// scatter level, the count-skew test hook, front_mode 1, or (auto) regions
// smaller than kSpecMinRegion keys.
static uint64_t spec_region_cap(const kc_table *t, const kc_pipe_shape &sh, const Digits &d, uint64_t front_cap)
{
    if (!d.l2 || sh.front_mode == 1 || t->fault_skew == 1 || t->fault_skew == 2) return 0;
    const uint64_t cap = front_cap / ((uint64_t)kFrontGroups << d.l1) / kSpecAlign * kSpecAlign;
    if (cap == 0 || (sh.front_mode == 0 && cap < kSpecMinRegion)) return 0;
    return cap;
}

// Sizes the scratch for a shape (reallocating what is too small).
int kc_part_reserve(const kc_table *t, kc_part_scratch &s, const kc_pipe_shape &sh)
{
    const Digits d = digits_of(t);
    const uint64_t n_coarse = 1ull << d.l1, n_parts = 1ull << d.part_bits;
    const uint64_t m1n = n_coarse * sh.front_blocks;
    const uint64_t m2n = d.l2 ? n_parts * sh.fine_slices : 0;
    const uint64_t sumsn = kc_scan_tmp_words(std::max(m1n, m2n));
    const uint64_t nw = t->nw;
    auto grow = [](auto *&p, uint64_t &have, uint64_t need) {
        if (p && have >= need) return true;
        if (p) (void)hipFree(p);
        p = nullptr;
        have = 0;
        if (!dev_alloc(p, need)) return false;
        have = need;
        return true;
    };
    uint64_t kc_a = s.front_cap, kc_b = s.key_cap, m1c = s.m1_n, o1c = s.m1_n, m2c = s.m2_n, o2c = s.m2_n;
    uint64_t sc = s.sums_n, pc = s.part_n, stc = s.starts ? 257 : 0, dgc = s.dig_bytes, gcc = s.gcur ? kGroupWords : 0;
    const uint64_t keys_words = std::max<uint64_t>(sh.key_cap, 1) * nw;
    const uint64_t front_words = front_cap_for(sh.key_cap) * nw;
    // the digit array (+16: the fine count's last 16-byte load may reach past the final digit)
    const uint64_t dig_need = d.l2 ? front_cap_for(sh.key_cap) * (d.dig16 ? 2 : 1) + 16 : 0;
    kc_a *= nw;
    kc_b *= nw;
    const bool ok = grow(s.keys_a, kc_a, front_words) && grow(s.keys_b, kc_b, keys_words) &&
                    grow(s.m1, m1c, m1n) && grow(s.o1, o1c, m1n) && (d.l2 == 0 || grow(s.m2, m2c, m2n)) &&
                    (d.l2 == 0 || grow(s.o2, o2c, m2n)) && grow(s.sums, sc, sumsn) &&
                    grow(s.part_off, pc, n_parts + 1) && grow(s.starts, stc, 257) && grow(s.gcur, gcc, kGroupWords) &&
                    (d.l2 == 0 || grow(s.dig, dgc, dig_need));
    if (!ok) {
        kc_part_scratch_free(s);
        return KC_ENOMEM;
    }
    s.key_cap = std::min(kc_a, kc_b) / nw;
    s.front_cap = kc_a / nw;
    s.m1_n = std::min(m1c, o1c);
    s.m2_n = d.l2 ? std::min(m2c, o2c) : s.m2_n;
    s.sums_n = sc;
    s.part_n = pc;
    s.dig_bytes = dgc;
    return KC_OK;
}

int kc_part_enqueue(const kc_table *t, kc_part_scratch &s, const kc_pipe_shape &sh, const SeqSrc &seq,
                    const uint64_t *d_keys, uint64_t n_keys, const TableView &tv, hipStream_t st, bool fresh)
{
    const int nw = (int)t->nw;
    const Digits d = digits_of(t);
    const uint32_t n_coarse = 1u << d.l1;
    const uint64_t n_parts = 1ull << d.part_bits;
    const uint32_t G = sh.front_blocks, S = sh.fine_slices;
    const bool from_seq = d_keys == nullptr;
    if (G % kFrontGroups) return KC_EINVAL;
    // the shape the scratch was sized for must cover this launch
    const uint64_t m1n = (uint64_t)n_coarse * G, m2n = d.l2 ? n_parts * S : 0;
    const uint64_t need_keys = from_seq ? seq.n - std::min(seq.n, seq.hdr ? 0 : seq.emit_from) : n_keys;
    const uint64_t dig_need = d.l2 ? s.front_cap * (d.dig16 ? 2 : 1) + 16 : 0;
    if (s.front_cap < s.key_cap || s.key_cap < need_keys || s.key_cap < sh.key_cap || s.m1_n < m1n || s.m2_n < m2n ||
        s.sums_n < kc_scan_tmp_words(std::max(m1n, m2n)) || s.part_n < n_parts + 1 || !s.starts || !s.gcur ||
        (d.l2 && (!s.dig || s.dig_bytes < dig_need)))
        return KC_EINVAL;
    uint8_t *dig = d.l2 ? s.dig : nullptr;
    auto *gcur = reinterpret_cast<unsigned long long *>(s.gcur);
    // the pass filter (multi-pass counting) is compiled in only where it is used
    const bool pass = d.pass_bits != 0;
    auto f1 = pass ? &k_front_scatter_seq_staged<1, kFrontBlock1, kFrontChunk1, true>
                   : &k_front_scatter_seq_staged<1, kFrontBlock1, kFrontChunk1, false>;
    auto f2 = pass ? &k_front_scatter_seq_staged<2, kFrontBlock2, kFrontChunk2, true>
                   : &k_front_scatter_seq_staged<2, kFrontBlock2, kFrontChunk2, false>;

    // level 1, speculative: fixed-capacity regions, scatter without a count
    // pass; the exact front end below then runs only if a region overflowed
    const uint64_t spec_cap = from_seq ? spec_region_cap(t, sh, d, s.front_cap) : 0;
    const unsigned long long *gate = nullptr;
    if (spec_cap) {
        hipLaunchKernelGGL(k_spec_regions, dim3(1), dim3(256), 0, st, gcur, n_coarse, spec_cap);
        unsigned long long *flag = gcur + kGcurFlag;
        if (nw == 1) hipLaunchKernelGGL(f1, dim3(G), dim3(kFrontBlock1), 0, st, seq, t->cfg.k, d, gcur, s.keys_a, dig, tv.counters, nullptr, flag);
        else hipLaunchKernelGGL(f2, dim3(G), dim3(kFrontBlock2), 0, st, seq, t->cfg.k, d, gcur, s.keys_a, dig, tv.counters, nullptr, flag);
        KC_HIP(hipGetLastError());
        gate = flag;
    }

    // level 1, exact: count, scan, coarse starts, scatter
    if (from_seq) {
        auto fc = nw == 1 ? (pass ? &k_front_count_seq<1, true> : &k_front_count_seq<1, false>)
                          : (pass ? &k_front_count_seq<2, true> : &k_front_count_seq<2, false>);
        hipLaunchKernelGGL(fc, dim3(G), dim3(kParseBlock), 0, st, seq, t->cfg.k, d, s.m1, gate);
    } else {
        if (nw == 1) hipLaunchKernelGGL(k_front_count_keys<1>, dim3(G), dim3(kKeyBlock), 0, st, d_keys, n_keys, d, s.m1);
        else hipLaunchKernelGGL(k_front_count_keys<2>, dim3(G), dim3(kKeyBlock), 0, st, d_keys, n_keys, d, s.m1);
    }
    KC_HIP(hipGetLastError());
    if (t->fault_skew == 1 || t->fault_skew == 2)
        hipLaunchKernelGGL(k_skew_count, dim3(1), dim3(64), 0, st, s.m1, m1n, t->fault_skew == 1 ? -1 : 1);
    int rc;
    if ((rc = kc_scan_u32_on(st, s.m1, m1n, s.o1, s.sums, gate))) return rc;
    hipLaunchKernelGGL(k_coarse_starts, dim3(1), dim3(256), 0, st, s.o1, s.m1, G, n_coarse, s.starts, gate);
    if (from_seq) {
        hipLaunchKernelGGL(k_group_cursors, dim3(1), dim3(256), 0, st, s.o1, s.m1, G, n_coarse, gcur, gate, tv.counters);
        if (nw == 1) hipLaunchKernelGGL(f1, dim3(G), dim3(kFrontBlock1), 0, st, seq, t->cfg.k, d, gcur, s.keys_a, dig, tv.counters, gate, nullptr);
        else hipLaunchKernelGGL(f2, dim3(G), dim3(kFrontBlock2), 0, st, seq, t->cfg.k, d, gcur, s.keys_a, dig, tv.counters, gate, nullptr);
        hipLaunchKernelGGL(k_group_check, dim3(1), dim3(256), 0, st, gcur, n_coarse, tv.counters, gate);
    } else {
        if (nw == 1) hipLaunchKernelGGL(k_front_scatter_keys<1>, dim3(G), dim3(kKeyBlock), 0, st, d_keys, n_keys, d, s.o1, s.m1, s.keys_a, dig, tv.counters);
        else hipLaunchKernelGGL(k_front_scatter_keys<2>, dim3(G), dim3(kKeyBlock), 0, st, d_keys, n_keys, d, s.o1, s.m1, s.keys_a, dig, tv.counters);
        hipLaunchKernelGGL(k_group_from_starts, dim3(1), dim3(256), 0, st, s.starts, n_coarse, gcur);
    }
    KC_HIP(hipGetLastError());

    // level 2: S slices per coarse region
    const uint64_t *part_keys = s.keys_a;
    if (d.l2) {
        const dim3 g2(n_coarse * S);
        if (!d.dig16) hipLaunchKernelGGL(k_fine_count_dig<uint8_t>, g2, dim3(kKeyBlock), 0, st, s.dig, gcur, d, S, s.m2);
        else hipLaunchKernelGGL(k_fine_count_dig<uint16_t>, g2, dim3(kKeyBlock), 0, st, reinterpret_cast<const uint16_t *>(s.dig), gcur, d, S, s.m2);
        KC_HIP(hipGetLastError());
        if (t->fault_skew == 3 || t->fault_skew == 4)
            hipLaunchKernelGGL(k_skew_count, dim3(1), dim3(64), 0, st, s.m2, m2n, t->fault_skew == 3 ? -1 : 1);
        if ((rc = kc_scan_u32_on(st, s.m2, m2n, s.o2, s.sums))) return rc;
        unsigned long long *ctr = tv.counters;
        if (d.l2 <= 8) {
            if (nw == 1) hipLaunchKernelGGL((k_fine_scatter_staged<1, 256, 1024, 16>), g2, dim3(1024), 0, st, s.keys_a, gcur, d, S, s.o2, s.m2, s.keys_b, ctr);
            else hipLaunchKernelGGL((k_fine_scatter_staged<2, 256, 1024, 8>), g2, dim3(1024), 0, st, s.keys_a, gcur, d, S, s.o2, s.m2, s.keys_b, ctr);
        } else {
            if (nw == 1) hipLaunchKernelGGL((k_fine_scatter_staged<1, 1024, 1024, 16, false>), g2, dim3(1024), 0, st, s.keys_a, gcur, d, S, s.o2, s.m2, s.keys_b, ctr);
            else if (d.l2 == 9)  // 512 digits: their u16 copies fit next to the 8192 two-word keys
                hipLaunchKernelGGL((k_fine_scatter_staged<2, 512, 1024, 8, true>), g2, dim3(1024), 0, st, s.keys_a, gcur, d, S, s.o2, s.m2, s.keys_b, ctr);
            else hipLaunchKernelGGL((k_fine_scatter_staged<2, 1024, 1024, 8, false>), g2, dim3(1024), 0, st, s.keys_a, gcur, d, S, s.o2, s.m2, s.keys_b, ctr);
        }
        KC_HIP(hipGetLastError());
        part_keys = s.keys_b;
    }
    hipLaunchKernelGGL(k_part_offsets, dim3((uint32_t)std::min<uint64_t>((n_parts + 256) / 256, 4096)), dim3(256), 0,
                       st, s.o2, s.m2, s.starts, d, S, s.part_off);
    KC_HIP(hipGetLastError());

    // per-partition LDS counting
    const bool bloom = t->bloom != nullptr;
    const dim3 g((uint32_t)n_parts);
    uint64_t *first_out = part_keys == s.keys_b ? s.keys_a : s.keys_b;  // the idle key array
    using PartKernel = void (*)(const uint64_t *, const uint64_t *, TableView, uint64_t *);
    // [nw - 1][fresh][mode]: mode 0 = no Bloom split, 1..3 = the FIRST hand-over of the split
    static const PartKernel kernels[2][2][4] = {
        {{&k_count_parts<1, false, kFirstStaged, false>, &k_count_parts<1, true, kFirstStaged, false>,
          &k_count_parts<1, true, kFirstStep, false>, &k_count_parts<1, true, kFirstGroup, false>},
         {&k_count_parts<1, false, kFirstStaged, true>, &k_count_parts<1, true, kFirstStaged, true>,
          &k_count_parts<1, true, kFirstStep, true>, &k_count_parts<1, true, kFirstGroup, true>}},
        {{&k_count_parts<2, false, kFirstStaged, false>, &k_count_parts<2, true, kFirstStaged, false>,
          &k_count_parts<2, true, kFirstStep, false>, &k_count_parts<2, true, kFirstGroup, false>},
         {&k_count_parts<2, false, kFirstStaged, true>, &k_count_parts<2, true, kFirstStaged, true>,
          &k_count_parts<2, true, kFirstStep, true>, &k_count_parts<2, true, kFirstGroup, true>}}};
    const uint32_t mode = bloom ? std::min<uint32_t>(std::max<uint32_t>(sh.first_mode, 1), 3) : 0;
    hipLaunchKernelGGL(kernels[nw - 1][fresh ? 1 : 0][mode], g, dim3(kPartThreads), 0, st, part_keys, s.part_off, tv,
                       first_out);
    KC_HIP(hipGetLastError());
    return KC_OK;
}

// Counts the windows of d_seq (when d_keys is null) or the n_keys keys of
// d_keys into the table through the partitioned pipeline, on the table's
// stream, with the table's own scratch (grown to the largest call).  Inputs
// larger than the tune's piece_bytes (default 4 GB; keys) are counted in
// pieces, so the key arrays stay within 2 x 8N x piece_bytes bytes; byte
// pieces after the first start 64 bytes early and emit nothing there (the
// k-1 overlap rule of FastqReader.h:542-557).
constexpr uint64_t kPartPiece = 4ull << 30;
constexpr uint64_t kPartHalo = 64;

int kc_part_count(kc_table *t, const uint8_t *d_seq, uint64_t nbytes, uint64_t emit_from, const uint64_t *d_keys,
                  uint64_t n_keys, const TableView &tv)
{
    // a lazily cleared table (t->fresh) is written whole by the first piece's pass
    const bool from_seq = d_keys == nullptr;
    const uint64_t total = from_seq ? nbytes : n_keys;
    uint64_t pos = from_seq ? std::min(nbytes, emit_from) : 0;
    // pieces are whole 16-byte groups, so every piece starts 16-byte aligned
    const uint64_t piece = std::max<uint64_t>(1ull << 16, (t->tune.piece_bytes ? t->tune.piece_bytes : kPartPiece) & ~15ull);
    while (pos < total) {
        const uint64_t take = std::min(piece, total - pos);
        int rc;
        if (from_seq) {
            // piece [pos - halo, pos + take), windows ending before pos excluded
            const uint64_t halo = std::min(kPartHalo, pos);
            const kc_pipe_shape sh = kc_part_shape(t, t->tune, take);
            if ((rc = kc_part_reserve(t, t->ps, sh)) != KC_OK) return rc;
            SeqSrc src{d_seq + pos - halo, take + halo, halo, nullptr};
            rc = kc_part_enqueue(t, t->ps, sh, src, nullptr, 0, tv, t->ctx->stream, t->fresh);
        } else {
            const kc_pipe_shape sh = kc_part_shape(t, t->tune, take);
            if ((rc = kc_part_reserve(t, t->ps, sh)) != KC_OK) return rc;
            SeqSrc none{nullptr, 0, 0, nullptr};
            rc = kc_part_enqueue(t, t->ps, sh, none, d_keys + pos * t->nw, take, tv, t->ctx->stream, t->fresh);
        }
        if (rc != KC_OK) return rc;
        t->fresh = false;  // the first pass wrote every slot
        pos += take;
    }
    return KC_OK;
}
