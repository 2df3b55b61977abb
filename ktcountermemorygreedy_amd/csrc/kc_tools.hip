// The on-disk format from and to device memory, and the reference's tools
// over a temp_dir.
//
//   kc_export_low / kc_export_high   caller keys (and counts) packed into
//       low.bin / high.bin records: ExportWriter::byte_pack_to_buffer
//       (ExportWriter.h:296-313) and FinalDrainWriter::write_kmer_record /
//       write_map_record (FinalDrainWriter.h:71-146), the high records
//       filtered by the counter's rule (count + 1 < filter_min in u32, or
//       count > filter_max) and kept in input order
//   kc_table_load_dir   a temp_dir (infos.bin, high.bin, low.bin) into a table:
//       high.bin records become table counts and low.bin keys its FIRST list,
//       so the table's histogram, dump and lookups are the tools' precise
//       merges of the files (histogram_tool.cpp:734-816, dump_tool.cpp:
//       1202-1270); without low.bin they are the approximate modes, which
//       read high.bin alone (histogram_tool.cpp:840-880, dump_tool.cpp:
//       1294-1400)
//   kc_histogram_dir / kc_dump_dir   the two tools over a temp_dir
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <string>
#include <sys/stat.h>

#include "kc_dev_common.h"
#include "kc_guard.h"
#include "kc_pack_dev.h"

namespace kct {
using namespace kc;

constexpr int kB = kPackBlock;

// Exclusive scan of one value per thread of a kB-thread block.
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t &total)
{
    __shared__ uint32_t wsum[kB / 64];
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
    for (uint32_t w = 0; w < kB / 64; ++w) {
        base += w < wave ? wsum[w] : 0;
        total += wsum[w];
    }
    __syncthreads();
    return base + x - v;
}

__device__ __forceinline__ bool high_filtered(uint32_t c, uint32_t fmin, uint32_t fmax)
{
    return (uint32_t)(c + 1u) < fmin || c > fmax;
}

// low.bin records of keys[0, n): tile t packs its kExportTile keys (thread
// t takes kExportPer consecutive ones) in LDS and writes them as one range.
template <int NW>
__global__ __launch_bounds__(kB) void k_pack_low(const uint64_t *__restrict__ keys, uint64_t n, PackSpec ps,
                                                 uint8_t *__restrict__ out)
{
    __shared__ __align__(16) uint8_t stage[export_stage<NW>()];
    const uint64_t n_tiles = (n + kExportTile - 1) / kExportTile;
    for (uint64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const uint64_t t0 = tile * kExportTile;
        const uint32_t rn = (uint32_t)(n - t0 < kExportTile ? n - t0 : kExportTile);
        uint8_t *d = out + t0 * ps.kmer_bytes;
        uint8_t *sd = stage + stage_skew(d);
#pragma unroll
        for (int j = 0; j < kExportPer; ++j) {
            const uint32_t r = threadIdx.x * kExportPer + j;
            if (r < rn) pack_key<NW>(sd + r * ps.kmer_bytes, keys + (t0 + r) * NW, ps);
        }
        __syncthreads();
        stage_copy(d, sd, rn * ps.kmer_bytes);
        __syncthreads();
    }
}

// Records of each tile that pass the filter.
__global__ __launch_bounds__(kB) void k_high_keep(const uint32_t *__restrict__ counts, uint64_t n, uint32_t fmin,
                                                  uint32_t fmax, uint32_t *__restrict__ tile_n)
{
    const uint64_t n_tiles = (n + kExportTile - 1) / kExportTile;
    for (uint64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const uint64_t t0 = tile * kExportTile;
        uint32_t keep = 0;
#pragma unroll
        for (int j = 0; j < kExportPer; ++j) {
            const uint64_t i = t0 + threadIdx.x * kExportPer + j;
            keep += i < n && !high_filtered(counts[i], fmin, fmax);
        }
        uint32_t total;
        block_scan(keep, total);
        if (threadIdx.x == 0) tile_n[tile] = total;
    }
}

// high.bin records of the kept (key, count) pairs, in input order, at the
// tiles' scanned offsets.
template <int NW>
__global__ __launch_bounds__(kB) void k_pack_high(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ counts,
                                                  uint64_t n, PackSpec ps, uint32_t fmin, uint32_t fmax,
                                                  const uint64_t *__restrict__ tile_off, uint8_t *__restrict__ out)
{
    __shared__ __align__(16) uint8_t stage[export_stage<NW>()];
    const uint32_t rec = ps.kmer_bytes + ps.count_bytes;
    const uint64_t n_tiles = (n + kExportTile - 1) / kExportTile;
    for (uint64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const uint64_t t0 = tile * kExportTile;
        uint32_t c[kExportPer];
        bool keep[kExportPer];
        uint32_t mine = 0;
#pragma unroll
        for (int j = 0; j < kExportPer; ++j) {
            const uint64_t i = t0 + threadIdx.x * kExportPer + j;
            c[j] = i < n ? counts[i] : 0;
            keep[j] = i < n && !high_filtered(c[j], fmin, fmax);
            mine += keep[j];
        }
        uint32_t total;
        uint32_t at = block_scan(mine, total);
        uint8_t *dt = out + tile_off[tile] * rec;
        uint8_t *st = stage + stage_skew(dt);
#pragma unroll
        for (int j = 0; j < kExportPer; ++j) {
            if (!keep[j]) continue;
            uint8_t *d = st + at * rec;
            pack_key<NW>(d, keys + (t0 + threadIdx.x * kExportPer + j) * NW, ps);
            put_bytes(d + ps.kmer_bytes, c[j], ps.count_bytes, 0);
            ++at;
        }
        __syncthreads();
        stage_copy(dt, st, total * rec);
        __syncthreads();
    }
}

// Packed records (with a count_bytes count when `counted`) back to keys and
// counts.
template <int NW>
__global__ __launch_bounds__(kB) void k_unpack(const uint8_t *__restrict__ src, uint64_t n, PackSpec ps, bool counted,
                                               uint64_t *__restrict__ keys, uint32_t *__restrict__ counts)
{
    const uint32_t rec = ps.kmer_bytes + (counted ? ps.count_bytes : 0);
    for (uint64_t i = blockIdx.x * (uint64_t)kB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kB) {
        const uint8_t *r = src + i * rec;
        uint64_t key[NW];
        unpack_key<NW>(r, ps, key);
#pragma unroll
        for (int w = 0; w < NW; ++w) keys[i * NW + w] = key[w];
        if (counted) {
            uint32_t c = 0;
            for (uint32_t b = 0; b < ps.count_bytes; ++b) c |= (uint32_t)r[ps.kmer_bytes + b] << (8 * b);
            counts[i] = c;
        }
    }
}

// Keys [base, base + n) of a FIRST list loaded whole: key g goes to
// segment g % KC_LOW_SEGS at row g / KC_LOW_SEGS, widened from the caller's
// tv.rw words to the table's NW.
template <int NW>
__global__ __launch_bounds__(kB) void k_list_load(TableView tv, const uint64_t *__restrict__ keys, uint64_t n,
                                                  uint64_t base)
{
    for (uint64_t i = blockIdx.x * (uint64_t)kB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kB) {
        const uint64_t g = base + i;
        uint64_t *dst = tv.low + ((g % KC_LOW_SEGS) * tv.seg_cap + g / KC_LOW_SEGS) * NW;
        uint64_t key[NW];
        caller_key<NW>(keys, i, tv.rw, key);
#pragma unroll
        for (int w = 0; w < NW; ++w) dst[w] = key[w];
    }
}

// The segment cursors of a list of `total` keys loaded by k_list_load.
__global__ void k_list_cursors(unsigned long long *counters, uint64_t total)
{
    for (uint32_t s = threadIdx.x; s < KC_LOW_SEGS; s += blockDim.x)
        counters[KC_CURSOR_BASE + (uint64_t)s * KC_LINE_WORDS] = total / KC_LOW_SEGS + (s < total % KC_LOW_SEGS);
}

}  // namespace kct

using namespace kct;

kc::TableView kc_table_view(kc_table *t);

namespace {

uint32_t grid_of(uint64_t n, uint32_t cu)
{
    const uint64_t want = (n + kB - 1) / kB;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, (uint64_t)cu * 16));
}

uint64_t next_pow2(uint64_t v)
{
    uint64_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

template <typename T>
struct Dev {
    T *p = nullptr;
    ~Dev()
    {
        if (p) (void)hipFree(p);
    }
    bool alloc(uint64_t n) { return hipMalloc((void **)&p, std::max<uint64_t>(n, 1) * sizeof(T)) == hipSuccess; }
};

struct Pinned {
    uint8_t *p = nullptr;
    ~Pinned()
    {
        if (p) (void)hipHostFree(p);
    }
};

struct File {
    FILE *f = nullptr;
    ~File()
    {
        if (f) std::fclose(f);
    }
};

int file_size(const std::string &path, uint64_t *size)
{
    struct stat st {};
    if (::stat(path.c_str(), &st) != 0) return KC_EIO;
    *size = (uint64_t)st.st_size;
    return KC_OK;
}

}  // namespace

extern "C" int kc_export_low(kc_ctx *ctx, const uint64_t *d_keys, uint64_t n, uint32_t k, uint8_t *d_out)
{
    if (!ctx || k < 1 || k > 128 || (n && (!d_keys || !d_out))) return KC_EINVAL;
    if (!n) return KC_OK;
    KC_HIP(hipSetDevice(ctx->device));
    const PackSpec ps = kc_pack_spec(k, 0xFF);
    const uint32_t g = (uint32_t)std::min<uint64_t>((n + kExportTile - 1) / kExportTile, (uint64_t)ctx->cu_count * 8);
    if (k <= 32) hipLaunchKernelGGL(k_pack_low<1>, dim3(g), dim3(kB), 0, ctx->stream, d_keys, n, ps, d_out);
    else if (k <= 64) hipLaunchKernelGGL(k_pack_low<2>, dim3(g), dim3(kB), 0, ctx->stream, d_keys, n, ps, d_out);
    else hipLaunchKernelGGL(k_pack_low<4>, dim3(g), dim3(kB), 0, ctx->stream, d_keys, n, ps, d_out);
    KC_HIP(hipGetLastError());
    return KC_OK;
}

extern "C" int kc_export_high(kc_ctx *ctx, const uint64_t *d_keys, const uint32_t *d_counts, uint64_t n, uint32_t k,
                              uint32_t count_max, uint32_t filter_min, uint32_t filter_max, uint8_t *d_out,
                              uint64_t out_cap, uint64_t *out_bytes)
{
    if (!ctx || !out_bytes || k < 1 || k > 128 || count_max == 0 || (n && (!d_keys || !d_counts))) return KC_EINVAL;
    *out_bytes = 0;
    if (!n) return KC_OK;
    KC_HIP(hipSetDevice(ctx->device));
    const PackSpec ps = kc_pack_spec(k, count_max);
    const uint64_t rec = ps.kmer_bytes + ps.count_bytes;
    const uint64_t n_tiles = (n + kExportTile - 1) / kExportTile;
    Dev<uint32_t> tn;
    Dev<uint64_t> off, sums;
    if (!tn.alloc(n_tiles) || !off.alloc(n_tiles + 1) || !sums.alloc(kc_scan_tmp_words(n_tiles) + 1)) return KC_ENOMEM;
    const uint32_t g = (uint32_t)std::min<uint64_t>(n_tiles, (uint64_t)ctx->cu_count * 8);
    hipLaunchKernelGGL(k_high_keep, dim3(g), dim3(kB), 0, ctx->stream, d_counts, n, filter_min, filter_max, tn.p);
    KC_HIP(hipGetLastError());
    int rc = kc_scan_u32(ctx, tn.p, n_tiles, off.p, sums.p);
    if (rc != KC_OK) return rc;
    uint64_t last_off = 0;
    uint32_t last_n = 0;
    KC_HIP(hipMemcpyAsync(&last_off, off.p + n_tiles - 1, 8, hipMemcpyDeviceToHost, ctx->stream));
    KC_HIP(hipMemcpyAsync(&last_n, tn.p + n_tiles - 1, 4, hipMemcpyDeviceToHost, ctx->stream));
    KC_HIP(hipStreamSynchronize(ctx->stream));
    *out_bytes = (last_off + last_n) * rec;
    if (!d_out) return KC_OK;
    if (*out_bytes > out_cap) return KC_EINVAL;
    if (k <= 32)
        hipLaunchKernelGGL(k_pack_high<1>, dim3(g), dim3(kB), 0, ctx->stream, d_keys, d_counts, n, ps, filter_min,
                           filter_max, off.p, d_out);
    else if (k <= 64)
        hipLaunchKernelGGL(k_pack_high<2>, dim3(g), dim3(kB), 0, ctx->stream, d_keys, d_counts, n, ps, filter_min,
                           filter_max, off.p, d_out);
    else
        hipLaunchKernelGGL(k_pack_high<4>, dim3(g), dim3(kB), 0, ctx->stream, d_keys, d_counts, n, ps, filter_min,
                           filter_max, off.p, d_out);
    KC_HIP(hipGetLastError());
    KC_HIP(hipStreamSynchronize(ctx->stream));  // the scan scratch is freed on return
    return KC_OK;
}

// Streams one record file through page-locked memory to the device in
// chunks and hands each chunk's unpacked keys (and counts) to `use`.
template <typename Use>
static int stream_records(kc_ctx *ctx, const std::string &path, uint64_t n, uint32_t nw, const PackSpec &ps,
                          bool counted, Use &&use)
{
    if (!n) return KC_OK;
    const uint64_t rec = ps.kmer_bytes + (counted ? ps.count_bytes : 0);
    const uint64_t per = std::min<uint64_t>(n, std::max<uint64_t>(1, (64ull << 20) / rec));
    File f;
    if (!(f.f = std::fopen(path.c_str(), "rb"))) return KC_EIO;
    Pinned h;
    Dev<uint8_t> raw;
    Dev<uint64_t> keys;
    Dev<uint32_t> counts;
    if (hipHostMalloc((void **)&h.p, per * rec, hipHostMallocDefault) != hipSuccess) return KC_ENOMEM;
    if (!raw.alloc(per * rec) || !keys.alloc(per * nw) || !counts.alloc(per)) return KC_ENOMEM;
    for (uint64_t done = 0; done < n;) {
        const uint64_t m = std::min(per, n - done);
        if (std::fread(h.p, rec, m, f.f) != m) return KC_EIO;
        KC_HIP(hipMemcpyAsync(raw.p, h.p, m * rec, hipMemcpyHostToDevice, ctx->stream));
        if (nw == 1) hipLaunchKernelGGL(k_unpack<1>, dim3(grid_of(m, ctx->cu_count)), dim3(kB), 0, ctx->stream, raw.p, m, ps, counted, keys.p, counts.p);
        else if (nw == 2) hipLaunchKernelGGL(k_unpack<2>, dim3(grid_of(m, ctx->cu_count)), dim3(kB), 0, ctx->stream, raw.p, m, ps, counted, keys.p, counts.p);
        else hipLaunchKernelGGL(k_unpack<4>, dim3(grid_of(m, ctx->cu_count)), dim3(kB), 0, ctx->stream, raw.p, m, ps, counted, keys.p, counts.p);
        KC_HIP(hipGetLastError());
        const int rc = use(keys.p, counts.p, m, done);
        if (rc != KC_OK) return rc;
        KC_HIP(hipStreamSynchronize(ctx->stream));  // h.p and the device chunk are reused
        done += m;
    }
    return KC_OK;
}

static int load_dir(kc_ctx *ctx, const char *dir, int with_low, kc_table **out)
{
    std::string d(dir);
    if (!d.empty() && d.back() != '/') d += '/';
    uint32_t infos[2];
    {
        File f;
        if (!(f.f = std::fopen((d + "infos.bin").c_str(), "rb"))) return KC_EIO;
        if (std::fread(infos, 4, 2, f.f) != 2) return KC_EIO;
    }
    const uint32_t k = infos[0], count_max = infos[1];
    if (k < 1 || k > 128 || count_max == 0) return KC_EINVAL;
    const PackSpec ps = kc_pack_spec(k, count_max);
    const uint32_t nw = k <= 32 ? 1 : k <= 64 ? 2 : 4;  // the files' kmer<N> words
    const uint32_t tnw = k <= 32 ? 1 : k <= 63 ? 2 : 4;  // the table's key words
    uint64_t hsize = 0, lsize = 0;
    int rc = file_size(d + "high.bin", &hsize);
    if (rc != KC_OK) return rc;
    if (with_low && (rc = file_size(d + "low.bin", &lsize)) != KC_OK) return rc;
    if (hsize % (ps.kmer_bytes + ps.count_bytes) || lsize % ps.kmer_bytes) return KC_EINVAL;
    const uint64_t n_high = hsize / (ps.kmer_bytes + ps.count_bytes), n_low = lsize / ps.kmer_bytes;
    kc_table_config cfg{};
    cfg.k = k;
    cfg.count_max = count_max;
    cfg.filter_min = 0;
    cfg.filter_max = 0xFFFFFFFFu;
    cfg.capacity = std::max<uint64_t>(1024, next_pow2(n_high + n_high / 2 + 1));
    cfg.max_distinct = n_high + 1;
    cfg.bloom = with_low ? 1 : 0;  // the FIRST list holds low.bin
    cfg.low_capacity = n_low + 1;
    cfg.bloom_blocks = with_low ? cfg.capacity / kc_lds_part_slots(tnw, true) : 0;
    cfg.count_path = KC_PATH_DIRECT;
    kc_table *t = nullptr;
    if ((rc = kc_table_create(ctx, &cfg, &t)) != KC_OK) return rc;
    rc = stream_records(ctx, d + "high.bin", n_high, nw, ps, true, [&](uint64_t *keys, uint32_t *counts, uint64_t m, uint64_t) {
        return kc_table_increment(t, keys, m, counts);
    });
    if (rc == KC_OK && with_low) {
        const kc::TableView tv = kc_table_view(t);
        rc = stream_records(ctx, d + "low.bin", n_low, nw, ps, false, [&](uint64_t *keys, uint32_t *, uint64_t m, uint64_t base) {
            if (tnw == 1) hipLaunchKernelGGL(k_list_load<1>, dim3(grid_of(m, ctx->cu_count)), dim3(kB), 0, ctx->stream, tv, keys, m, base);
            else if (tnw == 2) hipLaunchKernelGGL(k_list_load<2>, dim3(grid_of(m, ctx->cu_count)), dim3(kB), 0, ctx->stream, tv, keys, m, base);
            else hipLaunchKernelGGL(k_list_load<4>, dim3(grid_of(m, ctx->cu_count)), dim3(kB), 0, ctx->stream, tv, keys, m, base);
            return hipGetLastError() == hipSuccess ? KC_OK : KC_EDEVICE;
        });
        if (rc == KC_OK) {
            hipLaunchKernelGGL(k_list_cursors, dim3(1), dim3(kB), 0, ctx->stream, tv.counters, n_low);
            if (hipGetLastError() != hipSuccess) rc = KC_EDEVICE;
        }
    }
    kc_stats st{};
    if (rc == KC_OK && (rc = kc_table_stats(t, &st)) == KC_OK && st.errors) rc = KC_EFULL;
    if (rc != KC_OK) {
        kc_table_destroy(t);
        return rc;
    }
    *out = t;
    return KC_OK;
}

extern "C" int kc_table_load_dir(kc_ctx *ctx, const char *dir, int with_low, kc_table **out)
{
    if (!ctx || !dir || !out) return KC_EINVAL;
    *out = nullptr;
    return kc_guard([&] { return load_dir(ctx, dir, with_low, out); });
}

extern "C" int kc_histogram_dir(kc_ctx *ctx, const char *dir, int precise, uint32_t min_freq, uint32_t max_freq,
                                int64_t *h_bins)
{
    if (!ctx || !dir || !h_bins || max_freq < min_freq) return KC_EINVAL;
    kc_table *t = nullptr;
    int rc = kc_table_load_dir(ctx, dir, precise ? 1 : 0, &t);
    if (rc != KC_OK) return rc;
    rc = kc_table_histogram(t, min_freq, max_freq, h_bins);
    kc_table_destroy(t);
    return rc;
}

extern "C" int kc_dump_dir(kc_ctx *ctx, const char *dir, int precise, uint32_t min_freq, uint32_t max_freq,
                           int sorted, const char *path)
{
    if (!ctx || !dir || !path || max_freq < min_freq) return KC_EINVAL;
    kc_table *t = nullptr;
    int rc = kc_table_load_dir(ctx, dir, precise ? 1 : 0, &t);
    if (rc != KC_OK) return rc;
    rc = kc_table_dump_file(t, min_freq, max_freq, sorted, path);
    kc_table_destroy(t);
    return rc;
}
