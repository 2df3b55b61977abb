// Device dump: dump_tool precise over the table's exported view
// (dump_tool.cpp:682-733 high insert when min<=c<=max, :824-909 low pass,
// :928-936 remaining entries, :345-354 + :1137-1198 lexicographic sort and
// "<bases>\t<count>\n" lines).
//
//   low pass   FIRST-list keys (or table singletons when no list is active):
//              found with an inserted high record -> min(c+1, count_max) and
//              the slot is flagged (the tool zeroes it); else count 1
//   high pass  unflagged high records in range
//   sort       records {string-order key, count} by the key's 2k bits
//              (keys are unique, so the order is total)
//   format     line lengths, exclusive scan, one thread writes one line
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "kc_dev_common.h"

namespace kcd {
using namespace kc;

constexpr int kThreads = 256;
constexpr uint64_t kFlag = 1ULL << 63;  // slot merged with a low record

struct DumpSpec {
    uint32_t min_freq, max_freq;
    bool singles_low;  // table singletons are low records (no active FIRST list)
};

__device__ __forceinline__ bool in_range(uint64_t v, const DumpSpec &d) { return v >= d.min_freq && v <= d.max_freq; }

// 2-bit-group reversal of a 64-bit word (base order reversal).
__device__ __forceinline__ uint64_t rev2(uint64_t x)
{
    x = ((x >> 2) & 0x3333333333333333ULL) | ((x & 0x3333333333333333ULL) << 2);
    x = ((x >> 4) & 0x0F0F0F0F0F0F0F0FULL) | ((x & 0x0F0F0F0F0F0F0F0FULL) << 4);
    return __builtin_bswap64(x);
}

// String-order key (first base most significant) of a stored key, as 64*NW
// bits with word 0 most significant, right-aligned in the 2k low bits.
template <int NW>
__device__ __forceinline__ void string_key(const uint64_t *key, uint32_t k, uint64_t *sk)
{
    // stored: position j at bit 64*NW - 2k + 2j of the (64*NW)-bit number
    if constexpr (NW == 1) {
        sk[0] = rev2(key[0] >> (64 - 2 * k)) >> (64 - 2 * k);
    } else if constexpr (NW == 4) {
        // the bits below position 0 are zero, so reversing the whole 256-bit
        // number by 2-bit groups leaves the string in its low 2k bits
#pragma unroll
        for (int w = 0; w < 4; ++w) sk[w] = rev2(key[3 - w]);
    } else {
        // 128-bit number hi:lo = key[0]:key[1]; shift right by 128-2k, reverse 2k bits
        const uint32_t sh = 128 - 2 * k;  // >= 2 for k <= 63
        uint64_t lo = key[1], hi = key[0];
        if (sh >= 64) {
            lo = hi >> (sh - 64);
            hi = 0;
        } else {
            lo = (lo >> sh) | (hi << (64 - sh));
            hi >>= sh;
        }
        // reverse the 128-bit value by 2-bit groups, then shift right by 128-2k
        uint64_t rhi = rev2(lo), rlo = rev2(hi);
        if (sh >= 64) {
            rlo = rhi >> (sh - 64);
            rhi = 0;
        } else {
            rlo = (rlo >> sh) | (rhi << (64 - sh));
            rhi >>= sh;
        }
        sk[0] = rhi;
        sk[1] = rlo;
    }
}

template <int NW>
__device__ __forceinline__ void put_record(uint32_t *rec, const uint64_t *key, uint32_t k, uint32_t count)
{
    uint64_t sk[NW];
    string_key<NW>(key, k, sk);
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        rec[2 * w] = (uint32_t)sk[w];
        rec[2 * w + 1] = (uint32_t)(sk[w] >> 32);
    }
    rec[2 * NW] = count;
}

// exported high count of a slot (FinalDrainWriter filter and clamp), or -1
template <int NW>
__device__ __forceinline__ int64_t slot_high(const TableView &tv, uint64_t c, bool singles_low)
{
    if (c == 0 || (singles_low && c == 1)) return -1;
    const uint32_t cc = (uint32_t)(c < tv.count_max ? c : tv.count_max);
    if (export_filtered(cc, tv)) return -1;
    return (int64_t)cc;
}

__device__ __forceinline__ uint64_t append(unsigned long long *cursor, bool want)
{
    const uint64_t m = __ballot(want);
    if (!m) return 0;
    const int leader = __ffsll((unsigned long long)m) - 1;
    unsigned long long b0 = 0;
    if ((int)__lane_id() == leader) b0 = atomicAdd(cursor, (unsigned long long)__popcll(m));
    b0 = __shfl(b0, leader, 64);
    return b0 + __popcll(m & ((1ULL << __lane_id()) - 1));
}

// low pass over the FIRST list
template <int NW>
__global__ __launch_bounds__(kThreads) void k_dump_low_list(TableView tv, DumpSpec ds, uint32_t *recs,
                                                            unsigned long long *cursor)
{
    constexpr int RW = 2 * NW + 1;
    const uint32_t seg = blockIdx.y;
    const uint64_t n = low_seg_fill(tv, seg);
    // the same iteration count for every lane of the block (append's ballot)
    const uint64_t iters = (n + (uint64_t)gridDim.x * kThreads - 1) / ((uint64_t)gridDim.x * kThreads);
    for (uint64_t it = 0; it < iters; ++it) {
        const uint64_t i = (it * gridDim.x + blockIdx.x) * (uint64_t)kThreads + threadIdx.x;
        bool out = false;
        uint32_t cnt = 0;
        uint64_t key[NW] = {};
        if (i < n) {
            const uint64_t *lk = low_seg_key(tv, seg, i, NW);
#pragma unroll
            for (int w = 0; w < NW; ++w) key[w] = lk[w];
            uint64_t *slot = table_locate<NW>(tv, key);
            const int64_t cc = slot ? slot_high<NW>(tv, slot[count_word<NW>()] & ~kFlag, false) : -1;
            if (cc >= 0 && in_range((uint64_t)cc, ds)) {
                uint64_t merged = (uint64_t)cc + 1;
                if (merged > tv.count_max) merged = tv.count_max;
                slot[count_word<NW>()] |= kFlag;
                out = in_range(merged, ds);
                cnt = (uint32_t)merged;
            } else {
                out = in_range(1, ds);
                cnt = 1;
            }
        }
        const uint64_t at = append(cursor, out);
        if (out) put_record<NW>(recs + at * RW, key, tv.k, cnt);
    }
}

// table pass: singles (when they are the low records) and unflagged highs;
// also clears the flags
template <int NW>
__global__ __launch_bounds__(kThreads) void k_dump_table(TableView tv, DumpSpec ds, uint32_t *recs,
                                                         unsigned long long *cursor)
{
    constexpr int SW = slot_words<NW>();
    constexpr int RW = 2 * NW + 1;
    const uint64_t cap = tv.mask + 1;
    const uint64_t iters = (cap + (uint64_t)gridDim.x * kThreads - 1) / ((uint64_t)gridDim.x * kThreads);
    for (uint64_t it = 0; it < iters; ++it) {
        const uint64_t i = (it * gridDim.x + blockIdx.x) * (uint64_t)kThreads + threadIdx.x;
        bool out = false;
        uint32_t cnt = 0;
        uint64_t key[NW] = {};
        if (i < cap) {
            uint64_t *s = tv.slots + i * SW;
            if (slot_used<NW>(s)) {
#pragma unroll
                for (int w = 0; w < NW; ++w) key[w] = s[w];
                const uint64_t raw = s[count_word<NW>()];
                const uint64_t c = raw & ~kFlag;
                if (raw & kFlag) {
                    s[count_word<NW>()] = c;  // merged with its low record
                } else if (ds.singles_low && c == 1) {
                    out = in_range(1, ds);
                    cnt = 1;
                } else {
                    const int64_t cc = slot_high<NW>(tv, c, ds.singles_low);
                    if (cc > 0 && in_range((uint64_t)cc, ds)) {
                        out = true;
                        cnt = (uint32_t)cc;
                    }
                }
            }
        }
        const uint64_t at = append(cursor, out);
        if (out) put_record<NW>(recs + at * RW, key, tv.k, cnt);
    }
}

__device__ __forceinline__ uint32_t ndigits(uint32_t v)
{
    uint32_t d = 1;
    while (v >= 10) {
        v /= 10;
        ++d;
    }
    return d;
}

template <int NW>
__global__ __launch_bounds__(kThreads) void k_line_lengths(const uint32_t *recs, uint64_t n, uint32_t k, uint32_t *len)
{
    constexpr int RW = 2 * NW + 1;
    for (uint64_t i = blockIdx.x * (uint64_t)kThreads + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kThreads)
        len[i] = k + 2 + ndigits(recs[i * RW + 2 * NW]);
}

template <int NW>
__global__ __launch_bounds__(kThreads) void k_write_lines(const uint32_t *recs, uint64_t n, uint32_t k,
                                                          const uint64_t *off, uint8_t *text)
{
    constexpr int RW = 2 * NW + 1;
    const char B[4] = {'A', 'C', 'G', 'T'};
    for (uint64_t i = blockIdx.x * (uint64_t)kThreads + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kThreads) {
        const uint32_t *r = recs + i * RW;
        uint8_t *o = text + off[i];
        // string-order key: base j (0 = first) at bits 2(k-1-j) of the 2k-bit value
        for (uint32_t j = 0; j < k; ++j) {
            const uint32_t bit = 2 * (k - 1 - j);
            uint32_t code;
            if constexpr (NW == 1) {
                const uint64_t v = (uint64_t)r[0] | ((uint64_t)r[1] << 32);
                code = (uint32_t)(v >> bit) & 3;
            } else if constexpr (NW == 4) {
                // word 3 - bit / 64 of the 256-bit key (word 0 most significant)
                const uint32_t q = 2 * (3 - bit / 64);
                const uint64_t v = (uint64_t)r[q] | ((uint64_t)r[q + 1] << 32);
                code = (uint32_t)(v >> (bit % 64)) & 3;
            } else {
                const uint64_t hi = (uint64_t)r[0] | ((uint64_t)r[1] << 32);
                const uint64_t lo = (uint64_t)r[2] | ((uint64_t)r[3] << 32);
                code = (uint32_t)((bit >= 64 ? hi >> (bit - 64) : lo >> bit) & 3);
            }
            o[j] = (uint8_t)B[code];
        }
        o[k] = '\t';
        uint32_t c = r[2 * NW];
        const uint32_t nd = ndigits(c);
        for (uint32_t d = 0; d < nd; ++d) {
            o[k + nd - d] = (uint8_t)('0' + c % 10);
            c /= 10;
        }
        o[k + 1 + nd] = '\n';
    }
}

}  // namespace kcd

using namespace kcd;

namespace {

struct DevMem {
    void *p = nullptr;
    ~DevMem()
    {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

kc::TableView kc_table_view(kc_table *t);
int kc_table_low_count(kc_table *t, uint64_t *n_low, bool *list_active);

// Builds the dump text in device memory; *text_out / *bytes_out own it.
static int dump_device(kc_table *t, uint32_t min_freq, uint32_t max_freq, int sorted, DevMem &text,
                       uint64_t &bytes)
{
    kc_ctx *c = t->ctx;
    hipStream_t st = c->stream;
    uint64_t n_low = 0;
    bool list = false;
    int rc = kc_table_low_count(t, &n_low, &list);
    if (rc != KC_OK) return rc;
    const TableView tv = kc_table_view(t);
    const uint32_t nw = t->nw, rw = 2 * nw + 1;
    const uint64_t max_recs = t->cap + n_low;
    DevMem recs, scratch, cursor;
    if (hipMalloc(&recs.p, max_recs * rw * 4 + 16) != hipSuccess || hipMalloc(&cursor.p, 8) != hipSuccess)
        return KC_ENOMEM;
    if ((rc = kc_launch_fill(c, static_cast<uint64_t *>(cursor.p), 1, 0)) != KC_OK) return rc;
    DumpSpec ds{min_freq, max_freq, !list};
    auto *cur = static_cast<unsigned long long *>(cursor.p);
    auto *r = static_cast<uint32_t *>(recs.p);
    const uint32_t g = (uint32_t)std::min<uint64_t>((std::max<uint64_t>(t->cap, n_low) + kThreads - 1) / kThreads,
                                                    (uint64_t)c->cu_count * 16);
    if (list && n_low) {
        // segments fill evenly (hashed appends): twice the mean covers the fullest
        const dim3 gl(kc_low_chunks(2 * n_low / KC_LOW_SEGS + 1, 4 * kThreads), KC_LOW_SEGS);
        if (nw == 1) hipLaunchKernelGGL(k_dump_low_list<1>, gl, dim3(kThreads), 0, st, tv, ds, r, cur);
        else if (nw == 2) hipLaunchKernelGGL(k_dump_low_list<2>, gl, dim3(kThreads), 0, st, tv, ds, r, cur);
        else hipLaunchKernelGGL(k_dump_low_list<4>, gl, dim3(kThreads), 0, st, tv, ds, r, cur);
    }
    if (nw == 1) hipLaunchKernelGGL(k_dump_table<1>, dim3(g), dim3(kThreads), 0, st, tv, ds, r, cur);
    else if (nw == 2) hipLaunchKernelGGL(k_dump_table<2>, dim3(g), dim3(kThreads), 0, st, tv, ds, r, cur);
    else hipLaunchKernelGGL(k_dump_table<4>, dim3(g), dim3(kThreads), 0, st, tv, ds, r, cur);
    KC_HIP(hipGetLastError());
    uint64_t n = 0;
    KC_HIP(hipMemcpyAsync(&n, cursor.p, 8, hipMemcpyDeviceToHost, st));
    KC_HIP(hipStreamSynchronize(st));
    if (sorted && n > 1) {
        if (hipMalloc(&scratch.p, n * rw * 4 + 16) != hipSuccess) return KC_ENOMEM;
        int in_s = 0;
        rc = kc_sort_range(c, r, static_cast<uint32_t *>(scratch.p), n, rw, 0, 2 * t->cfg.k, &in_s);
        if (rc != KC_OK) return rc;
        if (in_s) std::swap(recs.p, scratch.p);
        r = static_cast<uint32_t *>(recs.p);
    }
    // line lengths -> offsets -> text
    DevMem len, off, sums;
    if (hipMalloc(&len.p, n * 4 + 16) != hipSuccess || hipMalloc(&off.p, (n + 1) * 8) != hipSuccess ||
        hipMalloc(&sums.p, kc_scan_tmp_words(n) * 8 + 16) != hipSuccess)
        return KC_ENOMEM;
    const uint32_t g2 = (uint32_t)std::min<uint64_t>((n + kThreads - 1) / kThreads + 1, (uint64_t)c->cu_count * 16);
    if (n) {
        if (nw == 1) hipLaunchKernelGGL(k_line_lengths<1>, dim3(g2), dim3(kThreads), 0, st, r, n, t->cfg.k, static_cast<uint32_t *>(len.p));
        else if (nw == 2) hipLaunchKernelGGL(k_line_lengths<2>, dim3(g2), dim3(kThreads), 0, st, r, n, t->cfg.k, static_cast<uint32_t *>(len.p));
        else hipLaunchKernelGGL(k_line_lengths<4>, dim3(g2), dim3(kThreads), 0, st, r, n, t->cfg.k, static_cast<uint32_t *>(len.p));
        if ((rc = kc_scan_u32(c, static_cast<uint32_t *>(len.p), n, static_cast<uint64_t *>(off.p),
                              static_cast<uint64_t *>(sums.p))) != KC_OK)
            return rc;
        uint64_t last_off = 0;
        uint32_t last_len = 0;
        KC_HIP(hipMemcpyAsync(&last_off, static_cast<uint64_t *>(off.p) + (n - 1), 8, hipMemcpyDeviceToHost, st));
        KC_HIP(hipMemcpyAsync(&last_len, static_cast<uint32_t *>(len.p) + (n - 1), 4, hipMemcpyDeviceToHost, st));
        KC_HIP(hipStreamSynchronize(st));
        bytes = last_off + last_len;
    } else {
        bytes = 0;
    }
    if (hipMalloc(&text.p, bytes + 16) != hipSuccess) return KC_ENOMEM;
    if (n) {
        if (nw == 1) hipLaunchKernelGGL(k_write_lines<1>, dim3(g2), dim3(kThreads), 0, st, r, n, t->cfg.k, static_cast<uint64_t *>(off.p), static_cast<uint8_t *>(text.p));
        else if (nw == 2) hipLaunchKernelGGL(k_write_lines<2>, dim3(g2), dim3(kThreads), 0, st, r, n, t->cfg.k, static_cast<uint64_t *>(off.p), static_cast<uint8_t *>(text.p));
        else hipLaunchKernelGGL(k_write_lines<4>, dim3(g2), dim3(kThreads), 0, st, r, n, t->cfg.k, static_cast<uint64_t *>(off.p), static_cast<uint8_t *>(text.p));
        KC_HIP(hipGetLastError());
    }
    KC_HIP(hipStreamSynchronize(st));
    return KC_OK;
}

extern "C" int kc_table_dump_device(kc_table *t, uint32_t min_freq, uint32_t max_freq, int sorted, uint8_t *d_out,
                                    uint64_t out_cap, uint64_t *out_bytes)
{
    if (!t || !out_bytes || max_freq < min_freq) return KC_EINVAL;
    DevMem text;
    uint64_t bytes = 0;
    int rc = dump_device(t, min_freq, max_freq, sorted, text, bytes);
    if (rc != KC_OK) return rc;
    *out_bytes = bytes;
    if (!d_out) return KC_OK;
    if (bytes > out_cap) return KC_EINVAL;
    if (bytes) KC_HIP(hipMemcpyAsync(d_out, text.p, bytes, hipMemcpyDeviceToDevice, t->ctx->stream));
    KC_HIP(hipStreamSynchronize(t->ctx->stream));
    return KC_OK;
}

extern "C" int kc_table_dump_file(kc_table *t, uint32_t min_freq, uint32_t max_freq, int sorted, const char *path)
{
    if (!t || !path || max_freq < min_freq) return KC_EINVAL;
    DevMem text;
    uint64_t bytes = 0;
    int rc = dump_device(t, min_freq, max_freq, sorted, text, bytes);
    if (rc != KC_OK) return rc;
    constexpr uint64_t kChunk = 64ull << 20;
    std::vector<uint8_t> buf;
    try {
        buf.resize(std::min<uint64_t>(kChunk, std::max<uint64_t>(bytes, 1)));
    } catch (...) {
        return KC_ENOMEM;
    }
    FILE *f = std::fopen(path, "wb");
    if (!f) return KC_EIO;
    for (uint64_t o = 0; o < bytes && rc == KC_OK; o += kChunk) {
        const uint64_t m = std::min(kChunk, bytes - o);
        if (hipMemcpy(buf.data(), static_cast<uint8_t *>(text.p) + o, m, hipMemcpyDeviceToHost) != hipSuccess)
            rc = KC_EDEVICE;
        else if (std::fwrite(buf.data(), 1, m, f) != m)
            rc = KC_EIO;
    }
    if (std::fclose(f) != 0 && rc == KC_OK) rc = KC_EIO;
    return rc;
}
