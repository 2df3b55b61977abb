// Stable LSD radix sort of export records on device, the replacement of
// ExportRecordRadixSort (ExportRecordRadixSort.h:225-269, 8-bit digits) and
// radix_sort_export_records_12bit (ExportRecordRadixSort12.h:444-646).
//
// Both reference sorts are stable sorts on a contiguous bit range of the
// 64N-bit key (word 0 most significant).  A stable sort is unique, so any
// stable sort on the same range reproduces them record for record; this one
// uses 8-bit digits from the low end of the range:
//   upsweep    per-tile digit counts (LDS histogram) -> counts[digit][tile]
//   scan       exclusive scan of counts in digit-major order (global offsets)
//   downsweep  per-tile stable ranking (wave ballots, waves in order), the
//              tile is placed digit-sorted in LDS and written out in runs.
// Records are 32-bit-word arrays: {key words (2N u32), count u32, [pad]} at
// 12, 16, 20, 24, 36 or 40 bytes.  Integer work only, HBM-bound: each pass reads and
// writes every record once (plus the key words once in the upsweep).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "kc_internal.h"

namespace kcs {

constexpr int kThreads = 256;
// Digits of up to DB bits (2^DB bins); the host splits the key range into
// the fewest passes of at most kSortDigitBits bits.
constexpr int kSortDigitBits = 9;  // 5 passes over 42 bits; 11 (4 passes) measured slower: one block per CU

// Downsweep tiles: TB threads x ITEMS records (8192 records of up to 16
// bytes, 4096 of 20-24 bytes, 2048 of 36-40 bytes), staged in LDS (80-128
// KB) so that a digit's records of a tile leave as one run (16 records per
// digit at 9 bits, 4 for four-word keys).
template <int RW> struct Layout {
    static constexpr int NW = RW >= 9 ? 4 : RW >= 5 ? 2 : 1;       // key words (u64)
    static constexpr int TB = 1024;
    static constexpr int ITEMS = RW >= 9 ? 2 : RW >= 5 ? 4 : 8;    // records per thread per tile
    static constexpr int TILE = TB * ITEMS;
    static constexpr int UP = RW >= 9 ? 8 : 16;                    // upsweep records per thread per chunk
};

// Digit of bits [bit, bit+width) of the key (LSB-numbered over the whole
// 64*NW-bit number); the range may straddle two words.
template <int NW>
__device__ __forceinline__ uint32_t digit_of(const uint32_t *rec, uint32_t bit, uint32_t width)
{
    // u32 words 2w, 2w+1 of the record are key word w (little-endian halves);
    // the word pair is picked with selects so that register arrays are never
    // indexed dynamically.
    const uint64_t k0 = (uint64_t)rec[0] | ((uint64_t)rec[1] << 32);
    uint64_t v;
    if constexpr (NW == 1) {
        v = k0 >> bit;
    } else if constexpr (NW == 4) {
        // key word 3 - bit / 64 holds the digit's low bits, the next more
        // significant word any bits that straddle
        const uint64_t k1 = (uint64_t)rec[2] | ((uint64_t)rec[3] << 32);
        const uint64_t k2 = (uint64_t)rec[4] | ((uint64_t)rec[5] << 32);
        const uint64_t k3 = (uint64_t)rec[6] | ((uint64_t)rec[7] << 32);
        const uint32_t wi = bit >> 6, sh = bit & 63;
        const uint64_t lo = wi == 0 ? k3 : wi == 1 ? k2 : wi == 2 ? k1 : k0;
        const uint64_t hi = wi == 0 ? k2 : wi == 1 ? k1 : k0;
        v = lo >> sh;
        if (wi < 3 && sh + width > 64) v |= hi << (64 - sh);
    } else {
        const uint64_t k1 = (uint64_t)rec[2] | ((uint64_t)rec[3] << 32);
        const uint32_t sh = bit & 63;
        if (bit >= 64) {
            v = k0 >> sh;
        } else {
            v = k1 >> sh;
            if (sh + width > 64) v |= k0 << (64 - sh);
        }
    }
    return (uint32_t)(v & ((1u << width) - 1));
}

template <int RW>
__device__ __forceinline__ void load_rec(const uint32_t *src, uint64_t i, uint32_t *r)
{
#pragma unroll
    for (int j = 0; j < RW; ++j) r[j] = src[i * RW + j];
}

// Block b owns the contiguous tiles [b*T, (b+1)*T): block order is input
// order, so per-block digit counts scanned digit-major give every block the
// start of its records of each digit (and the sort stays stable).
__device__ __forceinline__ void block_tiles(uint64_t n_tiles, uint64_t &first, uint64_t &last)
{
    const uint64_t per = (n_tiles + gridDim.x - 1) / gridDim.x;
    first = blockIdx.x * per;
    last = first + per < n_tiles ? first + per : n_tiles;
}

// Digit counts of block b's records: the records of its downsweep tiles,
// read in chunks of kThreads x Layout::UP.

template <int RW, int DB>
__global__ __launch_bounds__(kThreads) void k_upsweep(const uint32_t *__restrict__ src, uint64_t n, uint32_t bit,
                                                      uint32_t width, uint64_t n_tiles, uint32_t *__restrict__ counts)
{
    using L = Layout<RW>;
    constexpr uint32_t kBins = 1u << DB;
    __shared__ uint32_t hist[kBins];
    for (uint32_t b = threadIdx.x; b < kBins; b += kThreads) hist[b] = 0;
    __syncthreads();
    uint64_t first, last;
    block_tiles(n_tiles, first, last);
    const uint64_t r0 = first * L::TILE, r1 = last * L::TILE < n ? last * L::TILE : n;
    for (uint64_t c0 = r0; c0 < r1; c0 += (uint64_t)kThreads * L::UP) {
        uint32_t r[L::UP][RW];
        // all loads first (index clamped into the range), then the counts
#pragma unroll
        for (int j = 0; j < L::UP; ++j) {
            const uint64_t i = c0 + (uint64_t)j * kThreads + threadIdx.x;
            load_rec<RW>(src, i < r1 ? i : r1 - 1, r[j]);
        }
#pragma unroll
        for (int j = 0; j < L::UP; ++j)
            if (c0 + (uint64_t)j * kThreads + threadIdx.x < r1)
                atomicAdd(&hist[digit_of<L::NW>(r[j], bit, width)], 1u);
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < kBins; b += kThreads) counts[(uint64_t)b * gridDim.x + blockIdx.x] = hist[b];
}

// Exclusive scan of a u32 array into u64 offsets: per-block sums, a scan of
// the block sums, then the per-block scan with its base.
constexpr int kScanItems = 16;
constexpr uint64_t kScanChunk = (uint64_t)kThreads * kScanItems;

__device__ __forceinline__ uint64_t block_excl_scan(uint64_t v, uint64_t &total)
{
    __shared__ uint64_t s[kThreads];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < kThreads; o <<= 1) {
        const uint64_t a = threadIdx.x >= (uint32_t)o ? s[threadIdx.x - o] : 0;
        __syncthreads();
        s[threadIdx.x] += a;
        __syncthreads();
    }
    total = s[kThreads - 1];
    const uint64_t r = s[threadIdx.x] - v;
    __syncthreads();
    return r;
}

// gate (the exact front end after a speculative scatter): the scan runs only
// when *gate is nonzero; a null gate always runs.
__global__ __launch_bounds__(kThreads) void k_scan_reduce(const uint32_t *in, uint64_t n, uint64_t *block_sums,
                                                          const unsigned long long *gate)
{
    if (gate && *gate == 0) return;
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
    uint64_t acc = 0;
    for (int j = 0; j < kScanItems; ++j) {
        const uint64_t i = base + (uint64_t)threadIdx.x * kScanItems + j;
        if (i < n) acc += in[i];
    }
    uint64_t total;
    block_excl_scan(acc, total);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kThreads) void k_scan_sums(uint64_t *sums, uint64_t nb, const unsigned long long *gate)
{
    if (gate && *gate == 0) return;
    // single block: sequential chunks of kThreads
    uint64_t carry = 0;
    for (uint64_t b = 0; b < nb; b += kThreads) {
        const uint64_t i = b + threadIdx.x;
        const uint64_t v = i < nb ? sums[i] : 0;
        uint64_t total;
        const uint64_t ex = block_excl_scan(v, total);
        if (i < nb) sums[i] = carry + ex;
        carry += total;
    }
}

__global__ __launch_bounds__(kThreads) void k_scan_apply(const uint32_t *in, uint64_t n, const uint64_t *block_base,
                                                         uint64_t *out, const unsigned long long *gate)
{
    if (gate && *gate == 0) return;
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
    uint64_t acc = 0;
    uint32_t v[kScanItems];
    for (int j = 0; j < kScanItems; ++j) {
        const uint64_t i = base + (uint64_t)threadIdx.x * kScanItems + j;
        v[j] = i < n ? in[i] : 0;
        acc += v[j];
    }
    uint64_t total;
    uint64_t run = block_excl_scan(acc, total) + block_base[blockIdx.x];
    for (int j = 0; j < kScanItems; ++j) {
        const uint64_t i = base + (uint64_t)threadIdx.x * kScanItems + j;
        if (i < n) out[i] = run;
        run += v[j];
    }
}

// ---- tile ranking and scatter (shared by both scatter kernels) ----------
// Inside a tile, wave w ranks its own contiguous share (records
// w*TILE/W .. (w+1)*TILE/W of W waves, 64 consecutive records per step) with
// wave-private digit counters: same-digit peers from one ballot per digit
// bit, rank = counter + peers before the lane, then the group's first lane
// bumps the counter.  One cross-wave prefix per digit and one scan over
// digits place every record; the tile is staged digit-sorted in LDS and
// written out in runs.  Order inside a digit = input order: the sort is
// stable.
template <int RW, int DB> struct TileLds {
    using L = Layout<RW>;
    static constexpr int kWaves = L::TB / 64;
    static constexpr uint32_t kBins = 1u << DB;
    uint32_t stage[L::TILE * RW];
    uint16_t wrun[kWaves][kBins];  // per wave: digit counts, then the wave's start in the digit
    uint32_t local_start[kBins];   // digit's start in the tile
    uint32_t wsum[kWaves];
    uint64_t gbase[kBins];         // digit's global start for this tile
};

template <int RW, int DB> struct TileRegs {
    using L = Layout<RW>;
    static constexpr uint32_t kPer = ((1u << DB) + L::TB - 1) / L::TB;  // digits owned by one thread
    uint32_t r[L::ITEMS][RW];
    uint32_t rank[L::ITEMS], dig[L::ITEMS];
    uint32_t tot[kPer];  // tile counts of the thread's digits [t*kPer, (t+1)*kPer)
};

// Loads and ranks the tile's tn records from t0; leaves local_start and the
// per-wave starts in S, the thread's digit totals in R.tot.  Ends with a
// barrier.
template <int RW, int DB>
__device__ __forceinline__ void tile_rank(TileLds<RW, DB> &S, TileRegs<RW, DB> &R, const uint32_t *__restrict__ src,
                                          uint64_t t0, uint32_t tn, uint32_t bit, uint32_t width)
{
    using L = Layout<RW>;
    constexpr int TB = L::TB;
    constexpr uint32_t kBins = 1u << DB;
    constexpr uint32_t kPer = TileRegs<RW, DB>::kPer;
    constexpr int kWaves = TB / 64;
    constexpr uint32_t kPerWave = L::TILE / kWaves;  // records per wave per tile
    static_assert(kPerWave <= 0xFFFF && L::TILE <= 0xFFFF, "16-bit wave counters");
    const uint32_t t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const uint64_t lane_lt = (1ULL << lane) - 1ULL;
    for (uint32_t d = t; d < kBins; d += TB)
        for (int w = 0; w < kWaves; ++w) S.wrun[w][d] = 0;
#pragma unroll
    for (int j = 0; j < L::ITEMS; ++j) {
        const uint32_t li = wave * kPerWave + (uint32_t)j * 64 + lane;
        const uint64_t i = t0 + (li < tn ? li : tn - 1);
        load_rec<RW>(src, i, R.r[j]);
    }
    __syncthreads();  // counters cleared
#pragma unroll
    for (int j = 0; j < L::ITEMS; ++j) {
        const uint32_t li = wave * kPerWave + (uint32_t)j * 64 + lane;
        const bool valid = li < tn;
        const uint32_t d = valid ? digit_of<L::NW>(R.r[j], bit, width) : 0;
        R.dig[j] = d;
        uint64_t peers = __ballot(valid);
        for (uint32_t b = 0; b < width; ++b) {
            const uint64_t m = __ballot((d >> b) & 1);
            peers &= ((d >> b) & 1) ? m : ~m;
        }
        const uint32_t before = __popcll(peers & lane_lt);
        R.rank[j] = S.wrun[wave][d] + before;  // read by every lane before the leader's update
        __builtin_amdgcn_wave_barrier();
        if (valid && before == 0) S.wrun[wave][d] = (uint16_t)(S.wrun[wave][d] + __popcll(peers));
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // digits [t*kPer, (t+1)*kPer): start of each wave inside the digit,
    // then the digit's start in the tile (block scan over the threads)
    uint32_t mine = 0;
#pragma unroll
    for (uint32_t q = 0; q < kPer; ++q) {
        const uint32_t d = t * kPer + q;
        uint32_t acc = 0;
        if (d < kBins) {
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                const uint32_t c = S.wrun[w][d];
                S.wrun[w][d] = (uint16_t)acc;
                acc += c;
            }
        }
        R.tot[q] = acc;
        mine += acc;
    }
    uint32_t x = mine;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) S.wsum[wave] = x;
    __syncthreads();
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) run += w < (int)wave ? S.wsum[w] : 0;
    run += x - mine;
#pragma unroll
    for (uint32_t q = 0; q < kPer; ++q) {
        const uint32_t d = t * kPer + q;
        if (d < kBins) S.local_start[d] = run;
        run += R.tot[q];
    }
    __syncthreads();
}

// Stages the ranked tile digit-sorted in LDS and writes it to dst at
// S.gbase[d] + (position - S.local_start[d]).  Ends with a barrier.
template <int RW, int DB>
__device__ __forceinline__ void tile_scatter(TileLds<RW, DB> &S, const TileRegs<RW, DB> &R,
                                             uint32_t *__restrict__ dst, uint32_t tn, uint32_t bit, uint32_t width)
{
    using L = Layout<RW>;
    constexpr int TB = L::TB;
    constexpr uint32_t kPerWave = L::TILE / (TB / 64);
    const uint32_t t = threadIdx.x, wave = t >> 6, lane = t & 63;
#pragma unroll
    for (int j = 0; j < L::ITEMS; ++j) {
        const uint32_t li = wave * kPerWave + (uint32_t)j * 64 + lane;
        if (li < tn) {
            const uint32_t p = S.local_start[R.dig[j]] + S.wrun[wave][R.dig[j]] + R.rank[j];
#pragma unroll
            for (int q = 0; q < RW; ++q) S.stage[p * RW + q] = R.r[j][q];
        }
    }
    __syncthreads();
    // consecutive positions of one digit land on consecutive global addresses
    for (uint32_t p = t; p < tn; p += TB) {
        const uint32_t *rec = S.stage + p * RW;
        const uint32_t d = digit_of<Layout<RW>::NW>(rec, bit, width);
        const uint64_t g = S.gbase[d] + (p - S.local_start[d]);
#pragma unroll
        for (int q = 0; q < RW; ++q) dst[g * RW + q] = rec[q];
    }
    __syncthreads();
}

// Scatter with per-block offsets from an upsweep: block b walks its
// contiguous tiles in order, carrying a running global offset per digit.
template <int RW, int DB>
__global__ __launch_bounds__(Layout<RW>::TB) void k_downsweep(const uint32_t *__restrict__ src,
                                                              uint32_t *__restrict__ dst, uint64_t n, uint32_t bit,
                                                              uint32_t width, uint64_t n_tiles,
                                                              const uint64_t *__restrict__ offsets)
{
    using L = Layout<RW>;
    constexpr uint32_t kBins = 1u << DB;
    constexpr uint32_t kPer = TileRegs<RW, DB>::kPer;
    __shared__ TileLds<RW, DB> S;
    TileRegs<RW, DB> R;
    const uint32_t t = threadIdx.x;
    for (uint32_t d = t; d < kBins; d += L::TB) S.gbase[d] = offsets[(uint64_t)d * gridDim.x + blockIdx.x];
    uint64_t first, last;
    block_tiles(n_tiles, first, last);
    for (uint64_t tile = first; tile < last; ++tile) {
        const uint64_t t0 = tile * L::TILE;
        const uint32_t tn = (uint32_t)(n - t0 < (uint64_t)L::TILE ? n - t0 : L::TILE);
        tile_rank<RW, DB>(S, R, src, t0, tn, bit, width);
        tile_scatter<RW, DB>(S, R, dst, tn, bit, width);
#pragma unroll
        for (uint32_t q = 0; q < kPer; ++q) {
            const uint32_t d = t * kPer + q;
            if (d < kBins) S.gbase[d] += R.tot[q];
        }
    }
}

// Abundance histogram of record counts (bins f-min for f in [min, max]).
template <int RW>
__global__ __launch_bounds__(kThreads) void k_count_hist(const uint32_t *__restrict__ recs, uint64_t n, uint32_t min_f,
                                                         uint32_t max_f, unsigned long long *bins)
{
    constexpr int NW = Layout<RW>::NW;
    __shared__ uint32_t lb[4096];
    for (int b = threadIdx.x; b < 4096; b += kThreads) lb[b] = 0;
    __syncthreads();
    for (uint64_t i = blockIdx.x * (uint64_t)kThreads + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kThreads) {
        const uint32_t c = recs[i * RW + 2 * NW];
        if (c < min_f || c > max_f) continue;
        const uint32_t b = c - min_f;
        if (b < 4096) atomicAdd(&lb[b], 1u);
        else atomicAdd(bins + b, 1ull);
    }
    __syncthreads();
    const uint64_t nb = (uint64_t)max_f - min_f + 1;
    for (uint32_t b = threadIdx.x; b < 4096 && b < nb; b += kThreads)
        if (lb[b]) atomicAdd(bins + b, (unsigned long long)lb[b]);
}

// Seeded export-record workload: key i = a 62-bit bijection of i (k=31
// layout: left-aligned in word 0, low 2 bits zero), count drawn from a
// Zipf-like law P(c >= x) = x^-(s-1) by inverse CDF, clipped to [1, 2^32-1].
__device__ __forceinline__ uint64_t mix62(uint64_t x, uint64_t seed)
{
    constexpr uint64_t M = (1ULL << 62) - 1;
    x = (x ^ seed) & M;
    x = (x * 0x9E3779B97F4A7C15ULL) & M;
    x ^= x >> 31;
    x = (x * 0xBF58476D1CE4E5B9ULL) & M;
    x ^= x >> 29;
    x = (x * 0x94D049BB133111EBULL) & M;
    x ^= x >> 32;
    return x;
}

template <int RW>
__global__ __launch_bounds__(kThreads) void k_synth_records(uint32_t *recs, uint64_t n, uint64_t seed, double inv_exp)
{
    for (uint64_t i = blockIdx.x * (uint64_t)kThreads + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kThreads) {
        const uint64_t key = mix62(i, seed) << 2;
        uint64_t h = (i + 1) * 0xD1B54A32D192ED03ULL ^ seed;
        h ^= h >> 33;
        h *= 0xff51afd7ed558ccdULL;
        h ^= h >> 33;
        const double u = ((h >> 11) + 1) * (1.0 / 9007199254740992.0);  // (0, 1]
        const double c = floor(pow(u, -inv_exp));
        const uint32_t cnt = c >= 4294967295.0 ? 0xFFFFFFFFu : (c < 1.0 ? 1u : (uint32_t)c);
        uint32_t *r = recs + i * RW;
        r[0] = (uint32_t)key;
        r[1] = (uint32_t)(key >> 32);
#pragma unroll
        for (int q = 2; q < RW; ++q) r[q] = 0;
        r[2 * Layout<RW>::NW] = cnt;
    }
}

}  // namespace kcs

using namespace kcs;

namespace {

struct SortScratch {
    uint32_t *counts = nullptr;
    uint64_t *offsets = nullptr;
    uint64_t *sums = nullptr;
    void release()
    {
        for (void *q : {(void *)counts, (void *)offsets, (void *)sums})
            if (q) (void)hipFree(q);
        *this = SortScratch();
    }
};

// One pass: per-block digit counts (upsweep), their scan, the scatter.
template <int RW, int DB>
int sort_pass(kc_ctx *ctx, SortScratch &s, const uint32_t *src, uint32_t *dst, uint64_t n, uint32_t bit,
              uint32_t width)
{
    using L = Layout<RW>;
    constexpr uint32_t kBins = 1u << DB;
    const uint64_t n_tiles = (n + L::TILE - 1) / L::TILE;
    // blocks own contiguous tile ranges: counts are [digit][block]
    const uint32_t grid = (uint32_t)std::min<uint64_t>(n_tiles, (uint64_t)ctx->cu_count * 8);
    const uint64_t m = (uint64_t)kBins * grid;
    const uint64_t nb = (m + kScanChunk - 1) / kScanChunk;
    if (!s.counts && (hipMalloc((void **)&s.counts, m * 4) != hipSuccess ||
                      hipMalloc((void **)&s.offsets, m * 8) != hipSuccess ||
                      hipMalloc((void **)&s.sums, nb * 8) != hipSuccess))
        return KC_ENOMEM;
    hipLaunchKernelGGL((k_upsweep<RW, DB>), dim3(grid), dim3(kThreads), 0, ctx->stream, src, n, bit, width, n_tiles, s.counts);
    hipLaunchKernelGGL(k_scan_reduce, dim3((uint32_t)nb), dim3(kThreads), 0, ctx->stream, s.counts, m, s.sums, nullptr);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kThreads), 0, ctx->stream, s.sums, nb, nullptr);
    hipLaunchKernelGGL(k_scan_apply, dim3((uint32_t)nb), dim3(kThreads), 0, ctx->stream, s.counts, m, s.sums, s.offsets, nullptr);
    hipLaunchKernelGGL((k_downsweep<RW, DB>), dim3(grid), dim3(L::TB), 0, ctx->stream, src, dst, n, bit, width, n_tiles, s.offsets);
    return hipGetLastError() == hipSuccess ? KC_OK : KC_EDEVICE;
}

// One LSD pass per digit of at most DB bits over [lo, hi), digits balanced
// (42 bits with DB = 9: 9, 9, 8, 8, 8).  A one-sweep form (decoupled
// look-back instead of the upsweep, one block per tile) measured slower on
// MI355X: 10.3-10.9 ms per pass against 9.6 for upsweep + scatter.
template <int RW, int DB>
int sort_passes(kc_ctx *ctx, uint32_t *a, uint32_t *b, uint64_t n, uint32_t lo, uint32_t hi, int *in_scratch)
{
    *in_scratch = 0;
    if (n <= 1 || hi <= lo) return KC_OK;
    const uint32_t range = hi - lo;
    const uint32_t npass = (range + DB - 1) / DB;
    SortScratch s;
    uint32_t *src = a, *dst = b;
    int rc = KC_OK;
    for (uint32_t p = 0, bit = lo; p < npass && rc == KC_OK; ++p) {
        // the first (range % npass) passes take one more bit
        const uint32_t width = range / npass + (p < range % npass ? 1 : 0);
        rc = sort_pass<RW, DB>(ctx, s, src, dst, n, bit, width);
        std::swap(src, dst);
        bit += width;
    }
    const hipError_t e = hipStreamSynchronize(ctx->stream);
    s.release();
    if (rc == KC_OK && e != hipSuccess) {
        kc_set_error("radix sort", e);
        rc = KC_EDEVICE;
    }
    if (rc == KC_OK) *in_scratch = npass & 1;
    return rc;
}

template <int RW>
int sort_impl(kc_ctx *ctx, uint32_t *a, uint32_t *b, uint64_t n, uint32_t lo, uint32_t hi, int *in_scratch)
{
    return sort_passes<RW, kSortDigitBits>(ctx, a, b, n, lo, hi, in_scratch);
}

// Key range of ExportRecordRadixSort (ExportRecordRadixSort.h:225-254):
// MSB-numbered positions from 2k-1 down in steps of at most 8 bits (never
// crossing a word), 2k-20 bits in total, each step reading 8 bits.
void range8(uint32_t nw, uint32_t k, uint32_t &lo, uint32_t &hi)
{
    lo = 64 * nw - 2 * k;
    hi = lo;
    uint32_t end_pos = 2 * k - 1, diff = 2 * k - 20;
    while (diff > 0) {
        uint32_t bits = end_pos + 1 - (end_pos / 64) * 64;
        bits = std::min(bits, 8u);
        bits = std::min(bits, diff);
        diff -= bits;
        const uint32_t word = end_pos / 64;
        const uint32_t lsb = 63 - end_pos % 64;
        hi = std::max(hi, 64 * (nw - 1 - word) + std::min(lsb + 8, 64u));
        end_pos -= bits;
    }
}

}  // namespace

extern "C" int kc_sort_records(kc_ctx *ctx, void *d_recs, void *d_scratch, uint64_t n, uint32_t nw, uint32_t k,
                               uint32_t cpb, uint32_t digit_bits, uint32_t record_bytes, int *result_in_scratch)
{
    if (!ctx || !result_in_scratch || (n && (!d_recs || !d_scratch)) || d_recs == d_scratch) return KC_EINVAL;
    if (nw != 1 && nw != 2 && nw != 4) return KC_EINVAL;
    if (record_bytes != 8 * nw + 4 && record_bytes != 8 * nw + 8) return KC_EINVAL;
    if ((reinterpret_cast<uintptr_t>(d_recs) | reinterpret_cast<uintptr_t>(d_scratch)) & 3) return KC_EINVAL;
    uint32_t lo, hi;
    if (digit_bits == 12) {
        // ExportRecordRadixSort12.h:481-503 argument checks
        if (k < cpb || k > 32 * nw) return KC_EINVAL;
        if (n > 0xFFFFFFFFULL) return KC_EINVAL;
        lo = 64 * nw - 2 * k;
        hi = 64 * nw - 2 * cpb;
    } else if (digit_bits == 8) {
        if (k < 10 || k > 32 * nw) return KC_EINVAL;
        range8(nw, k, lo, hi);
    } else {
        return KC_EINVAL;
    }
    KC_HIP(hipSetDevice(ctx->device));
    uint32_t *a = static_cast<uint32_t *>(d_recs), *b = static_cast<uint32_t *>(d_scratch);
    switch (record_bytes / 4) {
    case 3: return sort_impl<3>(ctx, a, b, n, lo, hi, result_in_scratch);
    case 4: return sort_impl<4>(ctx, a, b, n, lo, hi, result_in_scratch);
    case 5: return sort_impl<5>(ctx, a, b, n, lo, hi, result_in_scratch);
    case 6: return sort_impl<6>(ctx, a, b, n, lo, hi, result_in_scratch);
    case 9: return sort_impl<9>(ctx, a, b, n, lo, hi, result_in_scratch);
    case 10: return sort_impl<10>(ctx, a, b, n, lo, hi, result_in_scratch);
    default: return KC_EINVAL;
    }
}

extern "C" int kc_records_histogram(kc_ctx *ctx, const void *d_recs, uint64_t n, uint32_t nw, uint32_t record_bytes,
                                    uint32_t min_freq, uint32_t max_freq, int64_t *h_bins)
{
    if (!ctx || !h_bins || (n && !d_recs) || max_freq < min_freq || (nw != 1 && nw != 2 && nw != 4)) return KC_EINVAL;
    if (record_bytes != 8 * nw + 4 && record_bytes != 8 * nw + 8) return KC_EINVAL;
    const uint64_t nb = (uint64_t)max_freq - min_freq + 1;
    unsigned long long *gb = nullptr;
    KC_HIP(hipMalloc((void **)&gb, nb * 8));
    int rc = kc_launch_fill(ctx, reinterpret_cast<uint64_t *>(gb), nb, 0);
    if (rc == KC_OK && n) {
        const uint32_t g = (uint32_t)std::min<uint64_t>((n + kThreads - 1) / kThreads, (uint64_t)ctx->cu_count * 8);
        const uint32_t *r = static_cast<const uint32_t *>(d_recs);
        switch (record_bytes / 4) {
        case 3: hipLaunchKernelGGL(k_count_hist<3>, dim3(g), dim3(kThreads), 0, ctx->stream, r, n, min_freq, max_freq, gb); break;
        case 4: hipLaunchKernelGGL(k_count_hist<4>, dim3(g), dim3(kThreads), 0, ctx->stream, r, n, min_freq, max_freq, gb); break;
        case 5: hipLaunchKernelGGL(k_count_hist<5>, dim3(g), dim3(kThreads), 0, ctx->stream, r, n, min_freq, max_freq, gb); break;
        case 6: hipLaunchKernelGGL(k_count_hist<6>, dim3(g), dim3(kThreads), 0, ctx->stream, r, n, min_freq, max_freq, gb); break;
        case 9: hipLaunchKernelGGL(k_count_hist<9>, dim3(g), dim3(kThreads), 0, ctx->stream, r, n, min_freq, max_freq, gb); break;
        default: hipLaunchKernelGGL(k_count_hist<10>, dim3(g), dim3(kThreads), 0, ctx->stream, r, n, min_freq, max_freq, gb); break;
        }
        if (hipGetLastError() != hipSuccess) rc = KC_EDEVICE;
    }
    if (rc == KC_OK && (hipMemcpyAsync(h_bins, gb, nb * 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
                        hipStreamSynchronize(ctx->stream) != hipSuccess))
        rc = KC_EDEVICE;
    (void)hipFree(gb);
    return rc;
}

extern "C" int kc_synth_records(kc_ctx *ctx, void *d_recs, uint64_t n, uint32_t record_bytes, uint64_t seed,
                                double zipf_s)
{
    if (!ctx || (n && !d_recs) || zipf_s <= 1.0 || (record_bytes != 12 && record_bytes != 16)) return KC_EINVAL;
    if (!n) return KC_OK;
    const uint32_t g = (uint32_t)std::min<uint64_t>((n + kThreads - 1) / kThreads, (uint64_t)ctx->cu_count * 16);
    uint32_t *r = static_cast<uint32_t *>(d_recs);
    if (record_bytes == 12)
        hipLaunchKernelGGL(k_synth_records<3>, dim3(g), dim3(kThreads), 0, ctx->stream, r, n, seed, 1.0 / (zipf_s - 1.0));
    else
        hipLaunchKernelGGL(k_synth_records<4>, dim3(g), dim3(kThreads), 0, ctx->stream, r, n, seed, 1.0 / (zipf_s - 1.0));
    KC_HIP(hipGetLastError());
    return KC_OK;
}

// Exclusive scan of n u32 values into u64 offsets (out[n] is not written).
// sums must hold kc_scan_tmp_words(n) words.
uint64_t kc_scan_tmp_words(uint64_t n) { return (n + kScanChunk - 1) / kScanChunk + 1; }

int kc_scan_u32_on(hipStream_t st, const uint32_t *in, uint64_t n, uint64_t *out, uint64_t *sums,
                   const unsigned long long *gate)
{
    if (!n) return KC_OK;
    const uint64_t nb = (n + kScanChunk - 1) / kScanChunk;
    hipLaunchKernelGGL(k_scan_reduce, dim3((uint32_t)nb), dim3(kThreads), 0, st, in, n, sums, gate);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kThreads), 0, st, sums, nb, gate);
    hipLaunchKernelGGL(k_scan_apply, dim3((uint32_t)nb), dim3(kThreads), 0, st, in, n, sums, out, gate);
    KC_HIP(hipGetLastError());
    return KC_OK;
}

int kc_scan_u32(kc_ctx *ctx, const uint32_t *in, uint64_t n, uint64_t *out, uint64_t *sums)
{
    return kc_scan_u32_on(ctx->stream, in, n, out, sums);
}

// Sort of rw-word records on key bits [lo, hi) (internal entry for the dump).
int kc_sort_range(kc_ctx *ctx, uint32_t *a, uint32_t *b, uint64_t n, uint32_t rw, uint32_t lo, uint32_t hi,
                  int *in_scratch)
{
    switch (rw) {
    case 3: return sort_impl<3>(ctx, a, b, n, lo, hi, in_scratch);
    case 4: return sort_impl<4>(ctx, a, b, n, lo, hi, in_scratch);
    case 5: return sort_impl<5>(ctx, a, b, n, lo, hi, in_scratch);
    case 6: return sort_impl<6>(ctx, a, b, n, lo, hi, in_scratch);
    case 9: return sort_impl<9>(ctx, a, b, n, lo, hi, in_scratch);
    case 10: return sort_impl<10>(ctx, a, b, n, lo, hi, in_scratch);
    default: return KC_EINVAL;
    }
}
