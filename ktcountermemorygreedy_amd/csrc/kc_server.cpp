// Serving loop: count jobs and point-query batches against device-resident
// tables, with continuous batching.  No counterpart in the reference (its
// only point lookup is FlatConcurrentHashMap::find_prepared in the tools,
// LowFrequencyQueryThreadPool.h:196-244).
//
// * Streams (all at default priority): count jobs run on a small pool of job
//   contexts (kJobStreams, one job per context at a time), lookups on one
//   lookup context of their own (kc_table_query_on), all created when the
//   server starts.  The process's streams share a few hardware queues (4 by
//   default) and work on one hardware queue runs in order, so a lookup on a
//   stream that shares its queue with a running job's stream waits for that
//   job's kernels; a lookup stream next to a bounded number of job streams
//   keeps lookups off the jobs' queues.  The job streams also carry a CU
//   mask that leaves kLookupCUs compute units to the lookup stream
//   (kc_ctx_create_leaving_cus), so a lookup launch never waits for a count
//   kernel's blocks to drain from every CU.  Requests on one table are
//   serialized by the table's mutex.
// * Worker threads take requests from one FIFO.  A count job is admitted
//   only when its device footprint fits the HBM measured free at that moment
//   (kc_device_mem_info) next to the running jobs, and the optional budget
//   next to the resident tables (kc_admit.h); a job that does not fit waits
//   while later requests that fit (queries, smaller jobs) go ahead.  When
//   nothing runs and a job still does not fit, least recently used tables are
//   dropped until it does; a job that cannot fit even then fails with
//   KC_ENOMEM instead of waiting forever.
// * Lookups are batched continuously: a worker that takes a query batch of a
//   table also takes every other query batch of that table waiting in the
//   queue (up to kMergeKeys keys) and answers them with one lookup launch
//   that reads the keys from and writes the counts to page-locked host
//   staging directly: no copy-engine transfer, which would queue behind the
//   count jobs' host-to-device streaming.
// * A count job with the Bloom split builds its table's FIRST-list index
//   before it completes, and every job allocates its table's page-locked
//   lookup staging (kStagingKeys keys), so lookups pay for neither; only a
//   merged lookup larger than the staging grows it.
// * Latency = completion time - submit time, per request.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/kc_api.h"
#include "../../include/kc_server.h"
#include "kc_admit.h"
#include "kc_guard.h"

namespace {

using Clock = std::chrono::steady_clock;

// Most keys one merged lookup launch answers.
constexpr uint64_t kMergeKeys = 4ull << 20;
// Count jobs running at once (one job context each).
constexpr uint32_t kJobStreams = 3;
// Keys of lookup staging a table gets when its job completes.
constexpr uint64_t kStagingKeys = 65536;

uint64_t next_pow2(uint64_t v);
// Compute units the job streams leave to lookups (CU masks on the job streams);
// KC_LOOKUP_CUS overrides it (0 = no mask).
constexpr uint32_t kLookupCUs = 16;

uint32_t lookup_cus()
{
    const char *e = std::getenv("KC_LOOKUP_CUS");
    if (!e || !*e) return kLookupCUs;
    char *end = nullptr;
    unsigned long v = std::strtoul(e, &end, 10);
    return (end && !*end && v <= 1024) ? (uint32_t)v : kLookupCUs;
}

struct Table {
    uint64_t id = 0;
    int nw = 1;
    kc_ctx *ctx = nullptr;   // the job context the table was counted on (the server's)
    kc_ctx *sctx = nullptr;  // the lookup context its staging belongs to (the server's)
    kc_table *t = nullptr;
    uint64_t bytes = 0;
    std::mutex mu;
    std::atomic<int> busy{0};
    Clock::time_point last_use;
    // lookup staging, grown on demand: page-locked host keys / counts that
    // the lookup kernel reads and writes in place (at their device addresses)
    void *hkeys = nullptr, *hcounts = nullptr, *dkeys = nullptr, *dcounts = nullptr;
    uint64_t qcap = 0;
    // Page-locked lookup staging for at least `keys` keys (65536 at least).
    int reserve_staging(uint64_t keys)
    {
        free_staging();
        const uint64_t cap = std::max<uint64_t>(next_pow2(keys), kStagingKeys);
        int rc = kc_host_alloc_pinned(sctx, cap * 8 * nw, &hkeys);
        if (rc == KC_OK) rc = kc_host_alloc_pinned(sctx, cap * 4, &hcounts);
        if (rc == KC_OK) rc = kc_host_device_ptr(sctx, hkeys, &dkeys);
        if (rc == KC_OK) rc = kc_host_device_ptr(sctx, hcounts, &dcounts);
        if (rc == KC_OK) qcap = cap;
        else free_staging();
        return rc;
    }
    void free_staging()
    {
        if (hkeys) kc_host_free_pinned(sctx, hkeys);
        if (hcounts) kc_host_free_pinned(sctx, hcounts);
        hkeys = hcounts = dkeys = dcounts = nullptr;
        qcap = 0;
    }
    ~Table()
    {
        if (sctx) free_staging();
        if (t) kc_table_destroy(t);  // after every lookup (kc_table_query_on orders them)
    }
};

struct Request {
    uint64_t id = 0;
    int kind = 0;  // 0 count, 1 query
    uint64_t table_id = 0;
    // count
    uint32_t k = 31, bloom = 0;
    const uint8_t *seq = nullptr;
    uint64_t nbytes = 0, max_distinct = 0, capacity = 0, resident = 0;
    kc_ticket ticket;
    int refused = KC_OK;  // set when the request is taken from the queue without running
    int job_slot = -1;    // the job context a running count holds
    // query
    const uint64_t *keys = nullptr;
    uint64_t n = 0;
    uint32_t *counts = nullptr;
    Clock::time_point submit, start, done;
    kc_request_result res{};
    bool finished = false;
};

uint64_t next_pow2(uint64_t v)
{
    uint64_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

// What a worker takes from the queue: one count job, or the query batches
// of one table.
struct Work {
    std::vector<Request *> reqs;
    std::shared_ptr<Table> tab;  // queries: the table (marked busy)
};

}  // namespace

struct kc_server {
    int device = 0;
    uint64_t total_hbm = 0;
    kc_admission adm;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> queue;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
    std::unordered_map<uint64_t, std::shared_ptr<Table>> tables;
    std::unordered_map<uint64_t, int> pending_counts;  // queued or running count jobs per table id
    kc_ctx *lookup_ctx = nullptr;
    std::vector<kc_ctx *> job_ctx;
    std::vector<char> job_busy;
    // tables dropped while jobs run are destroyed by a worker outside the
    // lock (kc_table_destroy waits for its job context's stream, which may be
    // running another job), or here once nothing runs
    std::vector<std::shared_ptr<Table>> dead;
    uint64_t running = 0;  // requests being executed
    uint64_t next_id = 1;
    bool stop = false;

    // Device bytes of a count job while it runs (footprint) and of its table
    // afterwards (resident): slots (16/32/64 B), Bloom filter, FIRST list and
    // its lookup index when the split is on; while counting also the device
    // staging and the partitioned pass's two key arrays.
    static uint64_t footprint(Request &r)
    {
        const uint64_t nw = r.k <= 32 ? 1 : r.k <= 63 ? 2 : 4;  // the table's key words
        const uint64_t want = std::max<uint64_t>(r.max_distinct, 1024);
        r.capacity = std::max<uint64_t>(1024, next_pow2(want + want / 2));
        r.resident = r.capacity * 16 * nw + (1 << 20);
        if (r.bloom) {
            const uint64_t low = want + want / 8 + 4096 * 1024;
            r.resident += 8 * r.capacity + 8 * nw * low + 8 * nw * next_pow2(2 * want);
        }
        const uint64_t chunk = std::min<uint64_t>(r.nbytes + 64, 256ull << 20);
        return r.resident + 2 * chunk + 20 * nw * chunk;
    }

    uint64_t free_hbm() const
    {
        uint64_t f = 0, t = 0;
        return kc_device_mem_info(device, &f, &t) == KC_OK ? f : 0;
    }

    // Takes the first request that can run now; mu held.
    bool pick(Work &w)
    {
        if (running == 0) dead.clear();  // every stream is idle: freeing them waits for nothing
        bool measured = false;
        uint64_t free_now = 0;
        for (auto it = queue.begin(); it != queue.end(); ++it) {
            Request *r = *it;
            if (r->kind == 1) {
                // a query waits for the count jobs of its table
                auto pc = pending_counts.find(r->table_id);
                if (pc != pending_counts.end() && pc->second > 0) continue;
                auto t = tables.find(r->table_id);
                if (t != tables.end() && t->second->busy.load()) continue;
                w.tab = t != tables.end() ? t->second : nullptr;
                // this batch and every later one of the same table, up to kMergeKeys keys
                const uint64_t id = r->table_id;
                uint64_t keys = 0;
                for (auto j = it; j != queue.end();) {
                    Request *q = *j;
                    if (q->kind == 1 && q->table_id == id && (w.reqs.empty() || keys + q->n <= kMergeKeys)) {
                        keys += q->n;
                        w.reqs.push_back(q);
                        j = queue.erase(j);
                    } else {
                        ++j;
                    }
                }
                return true;
            }
            if (tables.count(r->table_id)) {  // id in use: refused without a reservation
                r->refused = KC_EINVAL;
                w.reqs.push_back(r);
                queue.erase(it);
                return true;
            }
            int slot = -1;  // a job context to run on; none free: later queries go ahead
            for (size_t j = 0; j < job_busy.size() && slot < 0; ++j)
                if (!job_busy[j]) slot = (int)j;
            if (slot < 0) continue;
            if (!measured) {
                free_now = free_hbm();
                measured = true;
            }
            if (!adm.admit(r->ticket, free_now) && running == 0) {
                // nothing runs: drop idle tables, least recently used first
                while (!adm.admit(r->ticket, free_now) && evict_one()) free_now = free_hbm();
                if (!r->ticket.reserved) r->refused = KC_ENOMEM;  // cannot fit even with every table dropped
            }
            if (r->ticket.reserved || r->refused) {
                if (r->ticket.reserved) {
                    r->job_slot = slot;
                    job_busy[slot] = 1;
                }
                w.reqs.push_back(r);
                queue.erase(it);
                return true;
            }
        }
        return false;
    }

    // Drops the least recently used idle table; mu held.
    bool evict_one()
    {
        std::shared_ptr<Table> lru;
        for (auto &kv : tables)
            if (!kv.second->busy.load() && (!lru || kv.second->last_use < lru->last_use)) lru = kv.second;
        if (!lru) return false;
        adm.drop(lru->bytes);
        tables.erase(lru->id);
        return true;  // freed now (nothing runs) or by the last holder
    }

    void run_count(Request *r)
    {
        std::shared_ptr<Table> tab;
        int rc = KC_OK;
        try {
            tab = std::make_shared<Table>();
        } catch (...) {
            rc = KC_ENOMEM;
        }
        kc_stats st{};
        uint64_t table_bytes = 0;
        if (rc == KC_OK) {
            tab->id = r->table_id;
            tab->nw = r->k <= 32 ? 1 : r->k <= 64 ? 2 : 4;  // words of a lookup key (kmer<N>)
            tab->ctx = job_ctx[r->job_slot];
            tab->sctx = lookup_ctx;
        }
        if (rc == KC_OK) {
            kc_table_config cfg{};
            cfg.k = r->k;
            cfg.count_max = 0xFFFFFFFFu;
            cfg.filter_min = 0;
            cfg.filter_max = 0xFFFFFFFFu;
            cfg.capacity = r->capacity;
            cfg.max_distinct = std::max<uint64_t>(r->max_distinct, 1024);
            cfg.bloom = r->bloom;
            rc = kc_table_create(tab->ctx, &cfg, &tab->t);
        }
        if (rc == KC_OK) rc = kc_table_count_seq_host(tab->t, r->seq, r->nbytes);
        if (rc == KC_OK) rc = kc_ctx_sync(tab->ctx);
        if (rc == KC_OK) rc = kc_table_trim(tab->t);
        if (rc == KC_OK) rc = kc_table_stats(tab->t, &st);
        if (rc == KC_OK && st.errors) rc = KC_EFULL;
        if (rc == KC_OK && r->bloom) rc = kc_table_query(tab->t, nullptr, 0, nullptr);  // builds the FIRST-list index
        // the lookup staging too, so that a table's first lookup allocates nothing
        if (rc == KC_OK) rc = tab->reserve_staging(kStagingKeys);
        if (rc == KC_OK) table_bytes = r->resident;
        r->res.kmers = st.kmers;
        r->res.distinct = st.distinct;
        std::lock_guard<std::mutex> g(mu);
        if (rc == KC_OK && tables.count(r->table_id)) rc = KC_EINVAL;  // id taken meanwhile
        if (rc == KC_OK) {
            tab->bytes = table_bytes;
            tab->last_use = Clock::now();
            try {
                tables.emplace(tab->id, tab);
            } catch (...) {
                rc = KC_ENOMEM;
            }
        }
        adm.finish(r->ticket, rc == KC_OK ? table_bytes : 0);
        r->res.status = rc;
    }

    // One lookup launch over page-locked staging for every batch.
    void run_queries(const std::shared_ptr<Table> &tab, const std::vector<Request *> &batch)
    {
        if (!tab) {
            for (Request *r : batch) r->res.status = KC_EINVAL;
            return;
        }
        std::lock_guard<std::mutex> tg(tab->mu);
        const uint64_t nw = tab->nw;
        uint64_t total = 0;
        for (Request *r : batch) total += r->n;
        int rc = KC_OK;
        if (total > tab->qcap) rc = tab->reserve_staging(total);
        if (rc == KC_OK && total) {
            uint64_t at = 0;
            for (Request *r : batch) {
                if (r->n) std::memcpy(static_cast<uint64_t *>(tab->hkeys) + at * nw, r->keys, r->n * 8 * nw);
                at += r->n;
            }
            rc = kc_table_query_on(tab->t, tab->sctx, static_cast<const uint64_t *>(tab->dkeys), total,
                                   static_cast<uint32_t *>(tab->dcounts));
            if (rc == KC_OK) rc = kc_ctx_sync(tab->sctx);
            if (rc == KC_OK) {
                at = 0;
                for (Request *r : batch) {
                    if (r->n) std::memcpy(r->counts, static_cast<const uint32_t *>(tab->hcounts) + at, r->n * 4);
                    at += r->n;
                }
            }
        }
        for (Request *r : batch) {
            r->res.status = rc;
            r->res.batched = (int)batch.size();
        }
        tab->last_use = Clock::now();
    }

    void worker()
    {
        for (;;) {
            Work w;
            std::vector<std::shared_ptr<Table>> corpses;  // destroyed here, after the lock
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stop || pick(w); });
                corpses.swap(dead);
                if (w.reqs.empty()) return;
                ++running;
                if (w.tab) w.tab->busy.fetch_add(1);
            }
            corpses.clear();
            const Clock::time_point start = Clock::now();
            for (Request *r : w.reqs) r->start = start;
            Request *first = w.reqs[0];
            if (first->kind == 1) run_queries(w.tab, w.reqs);
            else if (first->refused) first->res.status = first->refused;  // holds no reservation
            else run_count(first);
            const Clock::time_point done = Clock::now();
            {
                std::lock_guard<std::mutex> g(mu);
                if (w.tab) w.tab->busy.fetch_sub(1);
                for (Request *r : w.reqs) {
                    r->done = done;
                    r->res.latency_ms = std::chrono::duration<double, std::milli>(r->done - r->submit).count();
                    r->res.queue_ms = std::chrono::duration<double, std::milli>(r->start - r->submit).count();
                    if (r->kind == 0 && --pending_counts[r->table_id] == 0) pending_counts.erase(r->table_id);
                    if (r->job_slot >= 0) job_busy[r->job_slot] = 0;
                    r->finished = true;
                }
                --running;
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

extern "C" int kc_server_create(int device, uint32_t workers, uint64_t hbm_budget_bytes, kc_server **out)
{
    return kc_guard([&]() -> int {
        if (!out || workers == 0 || workers > 64) return KC_EINVAL;
        uint64_t free_b = 0, total_b = 0;
        int rc = kc_device_mem_info(device, &free_b, &total_b);  // validates the device
        if (rc != KC_OK) return rc;
        auto *s = new (std::nothrow) kc_server();
        if (!s) return KC_ENOMEM;
        s->device = device;
        s->total_hbm = total_b;
        s->adm.budget = hbm_budget_bytes;
        s->adm.margin = std::min<uint64_t>(2ull << 30, total_b / 16);
        // the lookup context first, then the job contexts
        const uint32_t jobs = std::min(workers, kJobStreams);
        rc = kc_ctx_create(device, &s->lookup_ctx);
        const uint32_t left = lookup_cus();
        for (uint32_t i = 0; i < jobs && rc == KC_OK; ++i) {
            kc_ctx *c = nullptr;
            rc = kc_ctx_create_leaving_cus(device, left, &c);
            if (rc == KC_OK) {
                try {
                    s->job_ctx.push_back(c);
                    s->job_busy.push_back(0);
                } catch (...) {
                    kc_ctx_destroy(c);
                    rc = KC_ENOMEM;
                }
            }
        }
        if (rc != KC_OK) {
            kc_server_destroy(s);
            return rc;
        }
        try {
            s->workers.reserve(workers);
            for (uint32_t i = 0; i < workers; ++i) s->workers.emplace_back([s] { s->worker(); });
        } catch (...) {  // a thread that could not start: stop the ones that did
            kc_server_destroy(s);
            return KC_ENOMEM;
        }
        *out = s;
        return KC_OK;
    });
}

extern "C" void kc_server_destroy(kc_server *s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &w : s->workers) w.join();
    s->tables.clear();
    s->dead.clear();
    for (kc_ctx *c : s->job_ctx) kc_ctx_destroy(c);
    if (s->lookup_ctx) kc_ctx_destroy(s->lookup_ctx);
    delete s;
}

// Queues r (a count job also counts as pending on its table).  Throws on a
// failed allocation with the server left as it was.
static uint64_t enqueue(kc_server *s, std::unique_ptr<Request> r)
{
    std::lock_guard<std::mutex> g(s->mu);
    const uint64_t id = s->next_id++;
    const uint64_t table_id = r->table_id;
    const bool count = r->kind == 0;
    Request *raw = r.get();
    raw->id = id;
    raw->submit = Clock::now();
    s->reqs.emplace(id, std::move(r));
    try {
        s->queue.push_back(raw);
        try {
            if (count) ++s->pending_counts[table_id];
        } catch (...) {
            s->queue.pop_back();
            throw;
        }
    } catch (...) {
        s->reqs.erase(id);
        throw;
    }
    s->cv_work.notify_one();
    return id;
}

extern "C" int kc_server_submit_count(kc_server *s, uint64_t table_id, uint32_t k, const uint8_t *h_seq,
                                      uint64_t nbytes, uint64_t max_distinct, int bloom, uint64_t *req_id)
{
    return kc_guard([&]() -> int {
        if (!s || !req_id || (!h_seq && nbytes) || k < 1 || k > 128) return KC_EINVAL;
        auto r = std::make_unique<Request>();
        r->kind = 0;
        r->table_id = table_id;
        r->k = k;
        r->bloom = bloom ? 1 : 0;
        r->seq = h_seq;
        r->nbytes = nbytes;
        r->max_distinct = max_distinct ? max_distinct : std::max<uint64_t>(nbytes, 1024);
        r->ticket.footprint = kc_server::footprint(*r);
        if (!s->adm.could_fit(r->ticket.footprint, s->total_hbm)) return KC_ENOMEM;
        *req_id = enqueue(s, std::move(r));
        return KC_OK;
    });
}

extern "C" int kc_server_submit_query(kc_server *s, uint64_t table_id, const uint64_t *h_keys, uint64_t n,
                                      uint32_t *h_counts, uint64_t *req_id)
{
    return kc_guard([&]() -> int {
        if (!s || !req_id || (n && (!h_keys || !h_counts))) return KC_EINVAL;
        auto r = std::make_unique<Request>();
        r->kind = 1;
        r->table_id = table_id;
        r->keys = h_keys;
        r->n = n;
        r->counts = h_counts;
        *req_id = enqueue(s, std::move(r));
        return KC_OK;
    });
}

extern "C" int kc_server_wait(kc_server *s, uint64_t req_id, kc_request_result *out)
{
    if (!s || !out) return KC_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(req_id);
    if (it == s->reqs.end()) return KC_EINVAL;
    Request *r = it->second.get();
    s->cv_done.wait(lk, [&] { return r->finished; });
    *out = r->res;
    s->reqs.erase(it);
    return KC_OK;
}

extern "C" int kc_server_drop_table(kc_server *s, uint64_t table_id)
{
    if (!s) return KC_EINVAL;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->tables.find(table_id);
    if (it == s->tables.end()) return KC_EINVAL;
    if (it->second->busy.load()) return KC_EINVAL;
    s->adm.drop(it->second->bytes);
    if (s->running) s->dead.push_back(it->second);  // destroyed by a worker, outside the lock
    s->tables.erase(it);
    s->cv_work.notify_all();
    return KC_OK;
}

extern "C" int kc_server_table_histogram(kc_server *s, uint64_t table_id, uint32_t min_freq, uint32_t max_freq,
                                         int64_t *h_bins)
{
    if (!s) return KC_EINVAL;
    std::shared_ptr<Table> tab;
    {
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->tables.find(table_id);
        if (it == s->tables.end()) return KC_EINVAL;
        tab = it->second;
    }
    std::lock_guard<std::mutex> tg(tab->mu);
    return kc_table_histogram(tab->t, min_freq, max_freq, h_bins);
}

extern "C" int kc_server_resident(kc_server *s, uint64_t *tables, uint64_t *bytes)
{
    if (!s || !tables || !bytes) return KC_EINVAL;
    std::lock_guard<std::mutex> g(s->mu);
    *tables = s->tables.size();
    *bytes = s->adm.resident + s->adm.reserved;
    return KC_OK;
}
