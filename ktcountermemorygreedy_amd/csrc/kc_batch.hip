// Batched counting: many framed read chunks of uneven length packed into
// one buffer and counted by one launch of a captured HIP graph, and the
// autotuner that picks the launch shape.
//
// No counterpart in the reference, whose reader threads cut the input into
// fixed 32 KiB ring blocks with a k-1 base overlap (FastqReader.h:542-557,
// definition.h:53-54) and hand them to parser threads one by one.  Here the
// host packs whole chunks (a chunk is a "seq\n" byte stream: one request, one
// file part, one read group) into page-locked staging slots; each full slot
// is copied to HBM on a copy stream and counted by one graph launch on the
// table's stream.  The graph holds the persistent pipeline (partitioned) or
// the persistent fused kernel (direct); the batch's byte count lives in a
// device header word written right before the launch, so one captured graph
// serves every batch of its slot.  Slots rotate, so packing batch i+1, its
// copy and the counting of batch i overlap.
//
// Chunks that do not end in '\n' get one, so no window joins two chunks.  A
// chunk larger than a slot is sent in pieces of its own with a 128-byte halo
// and the halo excluded from emission (the reference's k-1 overlap rule).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

#include "kc_dev_common.h"

namespace {

constexpr uint64_t kHalo = 128;  // >= k-1 for k <= 128, multiple of 16
constexpr uint32_t kMaxSlots = 8;
constexpr size_t kMaxDevGraphs = 8;  // graphs cached by kc_batch_launch_device
constexpr uint64_t kPartMinBatch = 32ull << 20;

__global__ void k_set_hdr(uint64_t *hdr, uint64_t n, uint64_t emit_from)
{
    if (threadIdx.x == 0) {
        hdr[0] = n;
        hdr[1] = emit_from;
    }
}

struct Slot {
    uint8_t *pinned = nullptr;  // cap bytes
    uint8_t *dbuf = nullptr;    // cap + 16 bytes on the device
    uint64_t *hdr = nullptr;    // {n, emit_from} on the device
    hipGraphExec_t exec[2] = {nullptr, nullptr};  // [fresh]: see check_table
    hipEvent_t copied = nullptr, done = nullptr;
    uint64_t fill = 0;
    bool used = false;          // copied / done have been recorded
};

struct DevGraph {
    const uint8_t *ptr;
    uint64_t cap;
    uint64_t *hdr;
    hipGraphExec_t exec[2];  // [fresh]
    uint64_t last_use;
};

}  // namespace

struct kc_batch {
    kc_table *t = nullptr;
    kc_tune tune{};
    uint64_t cap = 0;
    uint32_t depth = 0;
    uint32_t path = 0;            // KC_PATH_DIRECT or KC_PATH_PARTITIONED
    uint32_t direct_grid = 0;
    kc_pipe_shape sh{};
    kc_part_scratch ps;
    hipStream_t copy_st = nullptr, cap_st = nullptr;
    Slot slot[kMaxSlots];
    uint32_t cur = 0;
    std::vector<DevGraph> dev;
    uint64_t tick = 0;
    kc_batch_info info{};
    uint32_t pass_key = 0;        // table pass (bits << 8 | pass) the graphs were captured with
};

namespace {

void destroy_graphs(kc_batch *b)
{
    for (uint32_t i = 0; i < b->depth; ++i)
        for (hipGraphExec_t &e : b->slot[i].exec)
            if (e) {
                (void)hipGraphExecDestroy(e);
                e = nullptr;
            }
    for (DevGraph &g : b->dev) {
        for (hipGraphExec_t e : g.exec)
            if (e) (void)hipGraphExecDestroy(e);
        if (g.hdr) (void)hipFree(g.hdr);
    }
    b->dev.clear();
}

// Scratch for up to key_cap keys; growing it invalidates captured graphs.
int ensure_scratch(kc_batch *b, uint64_t key_cap)
{
    if (b->path != KC_PATH_PARTITIONED) return KC_OK;
    if (b->ps.key_cap >= key_cap && b->ps.starts) return KC_OK;
    KC_HIP(hipStreamSynchronize(b->t->ctx->stream));
    destroy_graphs(b);
    b->sh = kc_part_shape(b->t, b->tune, key_cap);
    return kc_part_reserve(b->t, b->ps, b->sh);
}

// Captures the counting pass over seq[0, cap) with its size in *hdr; fresh:
// the partitioned pass that writes a lazily cleared table whole.
int capture(kc_batch *b, const uint8_t *seq, uint64_t cap, const uint64_t *hdr, bool fresh, hipGraphExec_t *out)
{
    const kc::SeqSrc src{seq, cap, 0, hdr};
    KC_HIP(hipStreamBeginCapture(b->cap_st, hipStreamCaptureModeThreadLocal));
    int rc;
    if (b->path == KC_PATH_PARTITIONED) {
        kc_pipe_shape sh = b->sh;
        sh.key_cap = cap;
        rc = kc_part_enqueue(b->t, b->ps, sh, src, nullptr, 0, kc_table_view(b->t), b->cap_st, fresh);
    } else {
        rc = kc_direct_enqueue(b->t, src, b->direct_grid, b->cap_st);
    }
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->cap_st, &g);
    if (rc != KC_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) {
        kc_set_error("hipStreamEndCapture", e);
        return KC_EDEVICE;
    }
    const hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        kc_set_error("hipGraphInstantiate", ei);
        return KC_EDEVICE;
    }
    ++b->info.graphs_captured;
    return KC_OK;
}

// The captured graphs carry the table's multi-pass filter as a kernel
// argument: after kc_table_set_pass they are dropped and captured again.
int sync_pass(kc_batch *b)
{
    const uint32_t key = (b->t->pass_bits << 8) | b->t->pass;
    if (key == b->pass_key) return KC_OK;
    KC_HIP(hipStreamSynchronize(b->t->ctx->stream));
    destroy_graphs(b);
    b->pass_key = key;
    return KC_OK;
}

// Before a launch: the table's counts change (see kc_table_touch), and the
// graphs must carry the table's current pass.  A lazily cleared table
// (kc_table::fresh) is counted by the fresh partitioned graph, which writes
// it whole; every other graph reads the table, so it is materialized first.
int check_table(kc_batch *b, bool &fresh)
{
    kc_table_touch(b->t);
    fresh = b->path == KC_PATH_PARTITIONED && b->t->fresh;
    const int rc = fresh ? KC_OK : kc_table_materialize(b->t);
    return rc != KC_OK ? rc : sync_pass(b);
}

// Launches slot i over its packed bytes (emit_from = halo of a piece).
int launch_slot(kc_batch *b, uint32_t i, uint64_t emit_from)
{
    Slot &s = b->slot[i];
    if (!s.fill) return KC_OK;
    bool fresh = false;
    int rc = check_table(b, fresh);
    if (rc != KC_OK) return rc;
    if ((rc = ensure_scratch(b, b->cap)) != KC_OK) return rc;
    if (!s.exec[fresh] && (rc = capture(b, s.dbuf, b->cap, s.hdr, fresh, &s.exec[fresh])) != KC_OK) return rc;
    hipStream_t ts = b->t->ctx->stream;
    // the device buffer is free once the previous launch of this slot is done
    if (s.used) KC_HIP(hipStreamWaitEvent(b->copy_st, s.done, 0));
    KC_HIP(hipMemcpyAsync(s.dbuf, s.pinned, s.fill, hipMemcpyHostToDevice, b->copy_st));
    KC_HIP(hipEventRecord(s.copied, b->copy_st));
    KC_HIP(hipStreamWaitEvent(ts, s.copied, 0));
    hipLaunchKernelGGL(k_set_hdr, dim3(1), dim3(64), 0, ts, s.hdr, s.fill, emit_from);
    KC_HIP(hipGetLastError());
    KC_HIP(hipGraphLaunch(s.exec[fresh], ts));
    if (fresh) b->t->fresh = false;  // this pass wrote every slot
    KC_HIP(hipEventRecord(s.done, ts));
    s.used = true;
    ++b->info.launches;
    b->info.bytes += s.fill;
    s.fill = 0;
    b->cur = (b->cur + 1) % b->depth;
    return KC_OK;
}

// Makes the current slot's page-locked buffer writable (its last copy
// done); slots are allocated on first use, so a batch that only launches
// over device-resident input holds no staging memory.
int slot_ready(kc_batch *b)
{
    Slot &s = b->slot[b->cur];
    if (!s.pinned) {
        Slot n;
        const bool ok = hipHostMalloc((void **)&n.pinned, b->cap, hipHostMallocDefault) == hipSuccess &&
                        hipMalloc((void **)&n.dbuf, b->cap + 16) == hipSuccess &&
                        hipMalloc((void **)&n.hdr, 16) == hipSuccess &&
                        hipEventCreateWithFlags(&n.copied, hipEventDisableTiming) == hipSuccess &&
                        hipEventCreateWithFlags(&n.done, hipEventDisableTiming) == hipSuccess;
        if (!ok) {
            if (n.pinned) (void)hipHostFree(n.pinned);
            if (n.dbuf) (void)hipFree(n.dbuf);
            if (n.hdr) (void)hipFree(n.hdr);
            if (n.copied) (void)hipEventDestroy(n.copied);
            if (n.done) (void)hipEventDestroy(n.done);
            return KC_ENOMEM;
        }
        s = n;
    }
    if (s.used && s.fill == 0) KC_HIP(hipEventSynchronize(s.copied));
    return KC_OK;
}

// Host copy into a staging slot, split over up to 8 threads for large spans
// (one thread does not reach the host's memory bandwidth).
void copy_par(uint8_t *dst, const uint8_t *src, uint64_t n)
{
    constexpr uint64_t kPerThread = 8ull << 20;
    const uint64_t hw = std::max(1u, std::thread::hardware_concurrency());
    const uint64_t T = std::min<uint64_t>(std::min<uint64_t>(8, hw), n / kPerThread);
    if (T <= 1) {
        std::memcpy(dst, src, n);
        return;
    }
    const uint64_t per = (n / T + 4095) & ~4095ull;
    std::vector<std::thread> th;
    uint64_t done = per;  // [0, per) is copied by this thread
    for (uint64_t lo = per; lo < n; lo += per) {
        const uint64_t len = std::min(per, n - lo);
        try {
            th.emplace_back([=] { std::memcpy(dst + lo, src + lo, len); });
            done = lo + len;
        } catch (...) {
            break;  // no thread: this one copies the rest
        }
    }
    std::memcpy(dst, src, std::min(per, n));
    if (done < n) std::memcpy(dst + done, src + done, n - done);
    for (std::thread &x : th) x.join();
}

void resolve_tune(const kc_table *t, const kc_tune *in, kc_tune &out, uint32_t &path)
{
    out = in ? *in : t->tune;
    if (!out.batch_bytes) out.batch_bytes = 256ull << 20;
    out.batch_bytes = std::max<uint64_t>(1ull << 20, (out.batch_bytes + 15) / 16 * 16);
    if (!out.queue_depth) out.queue_depth = 2;
    out.queue_depth = std::max<uint32_t>(2, std::min<uint32_t>(out.queue_depth, kMaxSlots));
    path = t->cfg.count_path != KC_PATH_AUTO ? t->cfg.count_path : out.path;  // a path fixed by the config wins
    if (path == KC_PATH_AUTO) path = out.batch_bytes >= kPartMinBatch ? KC_PATH_PARTITIONED : KC_PATH_DIRECT;
    if (path == KC_PATH_PARTITIONED && !kc_part_eligible(t)) path = KC_PATH_DIRECT;
}

}  // namespace

extern "C" int kc_table_set_tune(kc_table *t, const kc_tune *tune)
{
    if (!t || !tune) return KC_EINVAL;
    if (tune->path > KC_PATH_PARTITIONED || (tune->queue_depth && (tune->queue_depth < 2 || tune->queue_depth > kMaxSlots)))
        return KC_EINVAL;
    t->tune = *tune;
    return KC_OK;
}

extern "C" int kc_table_get_tune(kc_table *t, kc_tune *out)
{
    if (!t || !out) return KC_EINVAL;
    *out = t->tune;
    return KC_OK;
}

extern "C" int kc_batch_create(kc_table *t, const kc_tune *tune, kc_batch **out)
{
    if (!t || !out) return KC_EINVAL;
    *out = nullptr;
    if (tune && tune->path > KC_PATH_PARTITIONED) return KC_EINVAL;
    KC_HIP(hipSetDevice(t->ctx->device));
    kc_batch *b = new (std::nothrow) kc_batch();
    if (!b) return KC_ENOMEM;
    try {
        b->dev.reserve(kMaxDevGraphs);  // kc_batch_launch_device then never reallocates
    } catch (...) {
        delete b;
        return KC_ENOMEM;
    }
    b->t = t;
    resolve_tune(t, tune, b->tune, b->path);
    b->cap = b->tune.batch_bytes;
    b->depth = b->tune.queue_depth;
    b->direct_grid = t->ctx->cu_count * 8;
    b->info.path = b->path;
    b->info.capacity = b->cap;
    b->info.queue_depth = b->depth;
    const bool ok = hipStreamCreateWithFlags(&b->copy_st, hipStreamNonBlocking) == hipSuccess &&
                    hipStreamCreateWithFlags(&b->cap_st, hipStreamNonBlocking) == hipSuccess;
    if (!ok) {
        kc_batch_destroy(b);
        return KC_EDEVICE;
    }
    // staging slots and pipeline scratch are allocated on first use
    if (b->path == KC_PATH_PARTITIONED) {
        const kc_pipe_shape sh = kc_part_shape(t, b->tune, b->cap);
        b->info.front_blocks = sh.front_blocks;
        b->info.fine_slices = sh.fine_slices;
    } else {
        b->info.front_blocks = b->direct_grid;
    }
    *out = b;
    return KC_OK;
}

extern "C" void kc_batch_destroy(kc_batch *b)
{
    if (!b) return;
    (void)hipSetDevice(b->t->ctx->device);
    (void)hipStreamSynchronize(b->t->ctx->stream);
    if (b->copy_st) (void)hipStreamSynchronize(b->copy_st);
    destroy_graphs(b);
    for (uint32_t i = 0; i < kMaxSlots; ++i) {
        Slot &s = b->slot[i];
        if (s.pinned) (void)hipHostFree(s.pinned);
        if (s.dbuf) (void)hipFree(s.dbuf);
        if (s.hdr) (void)hipFree(s.hdr);
        if (s.copied) (void)hipEventDestroy(s.copied);
        if (s.done) (void)hipEventDestroy(s.done);
    }
    kc_part_scratch_free(b->ps);
    if (b->copy_st) (void)hipStreamDestroy(b->copy_st);
    if (b->cap_st) (void)hipStreamDestroy(b->cap_st);
    delete b;
}

extern "C" int kc_batch_add(kc_batch *b, const uint8_t *h_chunk, uint64_t len)
{
    if (!b || (!h_chunk && len)) return KC_EINVAL;
    if (!len) return KC_OK;
    bool fresh;
    int rc = check_table(b, fresh);
    if (rc != KC_OK) return rc;
    KC_HIP(hipSetDevice(b->t->ctx->device));
    ++b->info.chunks;
    const bool nl = h_chunk[len - 1] == '\n';
    const uint64_t need = len + (nl ? 0 : 1);
    if (need <= b->cap) {
        if (b->slot[b->cur].fill + need > b->cap && (rc = launch_slot(b, b->cur, 0)) != KC_OK) return rc;
        if ((rc = slot_ready(b)) != KC_OK) return rc;
        Slot &s = b->slot[b->cur];
        copy_par(s.pinned + s.fill, h_chunk, len);
        if (!nl) s.pinned[s.fill + len] = '\n';
        s.fill += need;
        return KC_OK;
    }
    // larger than a slot: pieces launched on their own, each after a halo
    if (b->slot[b->cur].fill && (rc = launch_slot(b, b->cur, 0)) != KC_OK) return rc;
    uint64_t pos = 0;
    while (pos < len) {
        const uint64_t halo = std::min(kHalo, pos);
        const uint64_t take = std::min(b->cap - halo, len - pos);
        if ((rc = slot_ready(b)) != KC_OK) return rc;
        Slot &s = b->slot[b->cur];
        copy_par(s.pinned, h_chunk + pos - halo, halo + take);
        s.fill = halo + take;
        ++b->info.pieces;
        if ((rc = launch_slot(b, b->cur, halo)) != KC_OK) return rc;
        pos += take;
    }
    return KC_OK;
}

extern "C" int kc_batch_add_chunks(kc_batch *b, const uint8_t *h_base, const uint64_t *ends, uint64_t n_chunks)
{
    if (!b || (n_chunks && !ends)) return KC_EINVAL;
    uint64_t prev = 0;
    for (uint64_t i = 0; i < n_chunks; ++i) {
        if (ends[i] < prev) return KC_EINVAL;
        prev = ends[i];
    }
    if (prev && !h_base) return KC_EINVAL;
    if (!prev) return KC_OK;
    bool fresh;
    int rc = check_table(b, fresh);
    if (rc != KC_OK) return rc;
    KC_HIP(hipSetDevice(b->t->ctx->device));
    uint64_t i = 0, pos = 0;  // chunk i is h_base[pos, ends[i])
    while (i < n_chunks) {
        const uint64_t len0 = ends[i] - pos;
        if (!len0) {
            ++i;
            continue;
        }
        const uint64_t need0 = len0 + (h_base[ends[i] - 1] == '\n' ? 0 : 1);
        if (need0 > b->cap) {  // pieces with halos, as kc_batch_add sends them
            if ((rc = kc_batch_add(b, h_base + pos, len0)) != KC_OK) return rc;
            pos = ends[i++];
            continue;
        }
        if (b->slot[b->cur].fill + need0 > b->cap && (rc = launch_slot(b, b->cur, 0)) != KC_OK) return rc;
        if ((rc = slot_ready(b)) != KC_OK) return rc;
        Slot &s = b->slot[b->cur];
        // the chunks that fit in this slot; a span of chunks that end in
        // '\n' is contiguous in h_base and goes in as one copy
        uint64_t span = pos;  // first source byte not copied yet (lands at s.pinned + s.fill)
        while (i < n_chunks) {
            const uint64_t len = ends[i] - pos;
            if (!len) {
                ++i;
                continue;
            }
            const bool nl = h_base[ends[i] - 1] == '\n';
            if (s.fill + (pos - span) + len + (nl ? 0 : 1) > b->cap) break;
            ++b->info.chunks;
            pos = ends[i++];
            if (!nl) {
                copy_par(s.pinned + s.fill, h_base + span, pos - span);
                s.fill += pos - span;
                s.pinned[s.fill++] = '\n';
                span = pos;
            }
        }
        copy_par(s.pinned + s.fill, h_base + span, pos - span);
        s.fill += pos - span;
    }
    return KC_OK;
}

extern "C" int kc_batch_flush(kc_batch *b)
{
    if (!b) return KC_EINVAL;
    KC_HIP(hipSetDevice(b->t->ctx->device));
    return launch_slot(b, b->cur, 0);
}

extern "C" int kc_batch_sync(kc_batch *b)
{
    if (!b) return KC_EINVAL;
    KC_HIP(hipStreamSynchronize(b->t->ctx->stream));
    return KC_OK;
}

// One graph launch over a batch that is already in HBM (packed framed
// bytes at d_seq, 16-byte aligned).  Graphs are kept per input pointer.
extern "C" int kc_batch_launch_device(kc_batch *b, const uint8_t *d_seq, uint64_t nbytes)
{
    if (!b || (!d_seq && nbytes)) return KC_EINVAL;
    if (reinterpret_cast<uintptr_t>(d_seq) % 16) return KC_EINVAL;
    if (!nbytes) return KC_OK;
    KC_HIP(hipSetDevice(b->t->ctx->device));
    bool fresh = false;
    int rc = check_table(b, fresh);
    if (rc != KC_OK) return rc;
    if ((rc = ensure_scratch(b, nbytes)) != KC_OK) return rc;
    DevGraph *g = nullptr;
    for (DevGraph &x : b->dev)
        if (x.ptr == d_seq && x.cap >= nbytes) g = &x;
    if (!g) {
        if (b->dev.size() >= kMaxDevGraphs) {  // drop the least recently used graph
            auto lru = std::min_element(b->dev.begin(), b->dev.end(),
                                        [](const DevGraph &a, const DevGraph &c) { return a.last_use < c.last_use; });
            KC_HIP(hipStreamSynchronize(b->t->ctx->stream));
            for (hipGraphExec_t e : lru->exec)
                if (e) (void)hipGraphExecDestroy(e);
            (void)hipFree(lru->hdr);
            b->dev.erase(lru);
        }
        DevGraph ng{d_seq, nbytes, nullptr, {nullptr, nullptr}, 0};
        if (hipMalloc((void **)&ng.hdr, 16) != hipSuccess) return KC_ENOMEM;
        b->dev.push_back(ng);
        g = &b->dev.back();
    }
    if (!g->exec[fresh] && (rc = capture(b, g->ptr, g->cap, g->hdr, fresh, &g->exec[fresh])) != KC_OK) return rc;
    g->last_use = ++b->tick;
    hipStream_t ts = b->t->ctx->stream;
    hipLaunchKernelGGL(k_set_hdr, dim3(1), dim3(64), 0, ts, g->hdr, nbytes, 0);
    KC_HIP(hipGetLastError());
    KC_HIP(hipGraphLaunch(g->exec[fresh], ts));
    if (fresh) b->t->fresh = false;  // this pass wrote every slot
    ++b->info.launches;
    b->info.bytes += nbytes;
    return KC_OK;
}

extern "C" int kc_batch_get_info(kc_batch *b, kc_batch_info *out)
{
    if (!b || !out) return KC_EINVAL;
    *out = b->info;
    return KC_OK;
}

// ---- autotuner -------------------------------------------------------------
// Times the counting call over a device-resident sample for each candidate
// launch shape (partitioned: persistent front blocks x fine slices; direct:
// one block per tile) and keeps the fastest.  Each trial starts from an
// empty table; the table is left empty.
extern "C" int kc_autotune(kc_table *t, const uint8_t *d_sample, uint64_t nbytes, uint32_t reps, kc_tune *best,
                           float *best_ms)
{
    if (!t || !d_sample || !nbytes || !best) return KC_EINVAL;
    if (reinterpret_cast<uintptr_t>(d_sample) % 16) return KC_EINVAL;
    KC_HIP(hipSetDevice(t->ctx->device));
    reps = std::max<uint32_t>(1, std::min<uint32_t>(reps, 8));
    const kc_tune saved = t->tune;
    hipEvent_t e0, e1;
    KC_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return KC_EDEVICE;
    }
    struct Cand {
        uint32_t path, blocks, slices;
    };
    Cand cands[17];
    uint32_t n_cands = 0;
    const uint32_t cu = t->ctx->cu_count;
    const uint32_t forced = t->cfg.count_path;  // a path fixed by the table config is kept
    if (kc_part_eligible(t) && forced != KC_PATH_DIRECT)
        for (uint32_t bpc : {2u, 4u, 8u, 16u})
            for (uint32_t sl : {2u, 4u, 8u, 16u}) cands[n_cands++] = {(uint32_t)KC_PATH_PARTITIONED, cu * bpc, sl};
    if (forced != KC_PATH_PARTITIONED || !kc_part_eligible(t)) cands[n_cands++] = {(uint32_t)KC_PATH_DIRECT, 0, 0};
    float best_t = 0;
    int rc = KC_OK;
    kc_tune win = saved;
    uint64_t win_redos = 0;  // speculative front ends the winning shape had to redo
    for (uint32_t ci = 0; ci < n_cands; ++ci) {
        const Cand &c = cands[ci];
        t->tune = saved;
        t->tune.front_blocks = c.blocks;
        t->tune.fine_slices = c.slices;
        t->tune.path = c.path;
        float tmin = 0;
        for (uint32_t r = 0; r <= reps && rc == KC_OK; ++r) {  // r = 0 warms up
            if ((rc = kc_table_clear(t)) != KC_OK) break;
            if (hipEventRecord(e0, t->ctx->stream) != hipSuccess) rc = KC_EDEVICE;
            if (rc == KC_OK) rc = kc_table_count_seq(t, d_sample, nbytes);
            if (rc == KC_OK && hipEventRecord(e1, t->ctx->stream) != hipSuccess) rc = KC_EDEVICE;
            float ms = 0;
            if (rc == KC_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                rc = KC_EDEVICE;
            if (rc == KC_OK && r > 0 && (tmin == 0 || ms < tmin)) tmin = ms;
        }
        if (rc != KC_OK) break;
        kc_stats st;
        if ((rc = kc_table_stats(t, &st)) != KC_OK) break;
        if (st.errors) {  // this shape overflowed: not a candidate
            continue;
        }
        uint64_t redos = 0;
        if ((rc = kc_table_front_redos(t, &redos)) != KC_OK) break;
        if (best_t == 0 || tmin < best_t) {
            best_t = tmin;
            win = t->tune;
            win_redos = redos;
        }
    }
    // input whose coarse regions overflow the speculative capacity (a few
    // keys repeated all over) pays the redo on every call: place it exactly
    if (rc == KC_OK && win_redos && win.front_mode == 0) win.front_mode = 1;
    // the other FIRST hand-overs of the partition pass on the winning shape
    const uint32_t first_default = kc_part_shape(t, win, 0).first_mode;
    for (uint32_t mode = 1; mode <= 3; ++mode) {
        if (rc != KC_OK || best_t <= 0 || win.path != KC_PATH_PARTITIONED || !t->bloom) break;
        if (mode == first_default) continue;
        t->tune = win;
        t->tune.first_mode = mode;
        float tmin = 0;
        for (uint32_t r = 0; r <= reps && rc == KC_OK; ++r) {
            if ((rc = kc_table_clear(t)) != KC_OK) break;
            if (hipEventRecord(e0, t->ctx->stream) != hipSuccess) rc = KC_EDEVICE;
            if (rc == KC_OK) rc = kc_table_count_seq(t, d_sample, nbytes);
            if (rc == KC_OK && hipEventRecord(e1, t->ctx->stream) != hipSuccess) rc = KC_EDEVICE;
            float ms = 0;
            if (rc == KC_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                rc = KC_EDEVICE;
            if (rc == KC_OK && r > 0 && (tmin == 0 || ms < tmin)) tmin = ms;
        }
        kc_stats st{};
        if (rc == KC_OK && (rc = kc_table_stats(t, &st)) == KC_OK && !st.errors && tmin > 0 && tmin < best_t) {
            best_t = tmin;
            win = t->tune;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    t->tune = rc == KC_OK ? win : saved;
    const int rc2 = kc_table_clear(t);
    if (rc != KC_OK) return rc;
    if (best_t == 0) return KC_EFULL;
    *best = t->tune;
    if (best_ms) *best_ms = best_t;
    return rc2;
}
