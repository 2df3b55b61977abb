// Serving loop: count jobs and point-query batches against device-resident
// tables, with continuous batching.  No counterpart in the reference (its
// only point lookup is FlatConcurrentHashMap::find_prepared in the tools,
// LowFrequencyQueryThreadPool.h:196-244).
//
// * Every table owns a kc_ctx (its own HIP stream, default priority), so the
//   work of different tables overlaps on the device; requests on one table
//   are serialized by the table's mutex.
// * Worker threads take requests from one FIFO.  A count job is admitted
//   only when its device footprint fits the HBM budget (what is left after
//   resident tables and running jobs); a job that does not fit waits while
//   later requests that fit (queries, smaller jobs) go ahead.  When nothing
//   runs and a job still does not fit, least recently used tables are
//   dropped until it does.
// * Latency = completion time - submit time, per request.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/kc_api.h"
#include "../../include/kc_server.h"
#include "kc_guard.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Table {
    uint64_t id;
    int nw = 1;
    kc_ctx *ctx = nullptr;
    kc_table *t = nullptr;
    uint64_t bytes = 0;
    std::mutex mu;
    std::atomic<int> busy{0};
    Clock::time_point last_use;
    void *qkeys = nullptr, *qcounts = nullptr;  // query staging, grown on demand
    uint64_t qcap = 0;
    ~Table()
    {
        if (qkeys) kc_dev_free(ctx, qkeys);
        if (qcounts) kc_dev_free(ctx, qcounts);
        if (t) kc_table_destroy(t);
        if (ctx) kc_ctx_destroy(ctx);
    }
};

struct Request {
    uint64_t id;
    int kind;  // 0 count, 1 query
    uint64_t table_id;
    // count
    uint32_t k = 31, bloom = 0;
    const uint8_t *seq = nullptr;
    uint64_t max_distinct = 0, footprint = 0, resident = 0;
    uint64_t nbytes = 0;
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

}  // namespace

struct kc_server {
    int device;
    uint64_t budget;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> queue;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
    std::unordered_map<uint64_t, std::shared_ptr<Table>> tables;
    std::unordered_map<uint64_t, int> pending_counts;  // queued or running count jobs per table id
    uint64_t used = 0;      // bytes of resident tables + reservations
    uint64_t running = 0;   // requests being executed
    uint64_t next_id = 1;
    bool stop = false;

    // Device bytes of a count job while it runs (footprint) and of its
    // table afterwards (resident): slots (16/32 B at the sizing below),
    // Bloom filter and FIRST list when enabled; while counting also the
    // device staging and the partitioned pass's two key arrays.
    static uint64_t footprint(const Request &r, uint64_t &cap, uint64_t &resident)
    {
        const uint64_t nw = r.k <= 32 ? 1 : 2;
        const uint64_t want = std::max<uint64_t>(r.max_distinct, 1024);
        cap = std::max<uint64_t>(1024, next_pow2(want + want / 2));
        resident = cap * 16 * nw + (1 << 20);
        if (r.bloom) resident += 16 * (cap / 2) + 8 * nw * (want + want / 16);
        const uint64_t chunk = std::min<uint64_t>(r.nbytes + 64, 256ull << 20);
        return resident + 2 * chunk + 20 * nw * chunk;
    }

    // Picks the first request that can run now (queries of tables that are
    // not busy, count jobs whose footprint fits).  Called with mu held.
    Request *pick()
    {
        for (auto it = queue.begin(); it != queue.end(); ++it) {
            Request *r = *it;
            if (r->kind == 1) {
                // a query waits for the count jobs of its table submitted before it
                auto pc = pending_counts.find(r->table_id);
                if (pc != pending_counts.end() && pc->second > 0) continue;
                auto t = tables.find(r->table_id);
                if (t != tables.end() && t->second->busy.load()) continue;
                queue.erase(it);
                return r;
            }
            if (tables.count(r->table_id)) {  // id in use: fails without a reservation
                queue.erase(it);
                return r;
            }
            if (used + r->footprint <= budget) {
                used += r->footprint;
                queue.erase(it);
                return r;
            }
            if (running == 0 && evict_for(r->footprint)) {
                used += r->footprint;
                queue.erase(it);
                return r;
            }
        }
        return nullptr;
    }

    // LRU eviction of idle tables until need bytes fit.  mu held.
    bool evict_for(uint64_t need)
    {
        while (used + need > budget) {
            std::shared_ptr<Table> lru;
            for (auto &kv : tables)
                if (!kv.second->busy.load() && (!lru || kv.second->last_use < lru->last_use)) lru = kv.second;
            if (!lru) return false;
            used -= lru->bytes;
            tables.erase(lru->id);
        }
        return true;
    }

    void run_count(Request *r)
    {
        uint64_t cap = 0, resident = 0;
        footprint(*r, cap, resident);
        std::shared_ptr<Table> tab;
        try {
            tab = std::make_shared<Table>();
        } catch (...) {
            std::lock_guard<std::mutex> g(mu);
            r->res.status = KC_ENOMEM;
            used -= r->footprint;
            return;
        }
        tab->id = r->table_id;
        tab->nw = r->k <= 32 ? 1 : 2;
        int rc = kc_ctx_create(device, &tab->ctx);
        if (rc == KC_OK) {
            kc_table_config cfg{};
            cfg.k = r->k;
            cfg.count_max = 0xFFFFFFFFu;
            cfg.filter_min = 0;
            cfg.filter_max = 0xFFFFFFFFu;
            cfg.capacity = cap;
            cfg.max_distinct = std::max<uint64_t>(r->max_distinct, 1024);
            cfg.bloom = r->bloom;
            rc = kc_table_create(tab->ctx, &cfg, &tab->t);
        }
        if (rc == KC_OK) rc = kc_table_count_seq_host(tab->t, r->seq, r->nbytes);
        if (rc == KC_OK) rc = kc_ctx_sync(tab->ctx);
        if (rc == KC_OK) rc = kc_table_trim(tab->t);
        kc_stats st{};
        if (rc == KC_OK) rc = kc_table_stats(tab->t, &st);
        if (rc == KC_OK && st.errors) rc = KC_EFULL;
        r->res.kmers = st.kmers;
        r->res.distinct = st.distinct;
        r->res.status = rc;
        std::lock_guard<std::mutex> g(mu);
        if (rc == KC_OK && tables.count(tab->id)) rc = KC_EINVAL;  // id taken meanwhile
        if (rc == KC_OK) {
            tab->bytes = resident;
            tab->last_use = Clock::now();
            try {
                tables.emplace(tab->id, tab);
            } catch (...) {
                rc = KC_ENOMEM;
            }
        }
        r->res.status = rc;
        if (rc == KC_OK) used -= r->footprint - resident;
        else used -= r->footprint;
    }

    void run_query(Request *r)
    {
        std::shared_ptr<Table> tab;
        {
            std::lock_guard<std::mutex> g(mu);
            auto it = tables.find(r->table_id);
            if (it != tables.end()) tab = it->second;
        }
        if (!tab) {
            r->res.status = KC_EINVAL;
            return;
        }
        std::lock_guard<std::mutex> tg(tab->mu);
        const int nw = tab->nw;
        int rc = KC_OK;
        if (r->n > tab->qcap) {
            kc_dev_free(tab->ctx, tab->qkeys);
            kc_dev_free(tab->ctx, tab->qcounts);
            tab->qkeys = tab->qcounts = nullptr;
            tab->qcap = std::max<uint64_t>(r->n, 65536);
            rc = kc_dev_alloc(tab->ctx, tab->qcap * 8 * nw, &tab->qkeys);
            if (rc == KC_OK) rc = kc_dev_alloc(tab->ctx, tab->qcap * 4, &tab->qcounts);
            if (rc != KC_OK) tab->qcap = 0;
        }
        if (rc == KC_OK) rc = kc_copy_h2d(tab->ctx, tab->qkeys, r->keys, r->n * 8 * nw);
        if (rc == KC_OK)
            rc = kc_table_query(tab->t, static_cast<const uint64_t *>(tab->qkeys), r->n,
                                static_cast<uint32_t *>(tab->qcounts));
        if (rc == KC_OK) rc = kc_copy_d2h(tab->ctx, r->counts, tab->qcounts, r->n * 4);
        r->res.status = rc;
        tab->last_use = Clock::now();
    }

    void worker()
    {
        for (;;) {
            Request *r = nullptr;
            std::shared_ptr<Table> busy_tab;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stop || (r = pick()) != nullptr; });
                if (!r) return;
                ++running;
                if (r->kind == 1) {
                    auto it = tables.find(r->table_id);
                    if (it != tables.end()) {
                        busy_tab = it->second;
                        busy_tab->busy.fetch_add(1);
                    }
                }
            }
            r->start = Clock::now();
            if (r->kind == 0) {
                if (tables_has(r->table_id)) r->res.status = KC_EINVAL;  // picked without a reservation
                else run_count(r);
            } else {
                run_query(r);
            }
            r->done = Clock::now();
            r->res.latency_ms = std::chrono::duration<double, std::milli>(r->done - r->submit).count();
            r->res.queue_ms = std::chrono::duration<double, std::milli>(r->start - r->submit).count();
            {
                std::lock_guard<std::mutex> g(mu);
                if (busy_tab) busy_tab->busy.fetch_sub(1);
                if (r->kind == 0 && --pending_counts[r->table_id] == 0) pending_counts.erase(r->table_id);
                --running;
                r->finished = true;
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }

    bool tables_has(uint64_t id)
    {
        std::lock_guard<std::mutex> g(mu);
        return tables.count(id) != 0;
    }
};

extern "C" int kc_server_create(int device, uint32_t workers, uint64_t hbm_budget_bytes, kc_server **out)
{
    return kc_guard([&]() -> int {
        if (!out || workers == 0 || workers > 64) return KC_EINVAL;
        kc_ctx *probe = nullptr;
        int rc = kc_ctx_create(device, &probe);  // validates the device
        if (rc != KC_OK) return rc;
        kc_ctx_destroy(probe);
        auto *s = new (std::nothrow) kc_server();
        if (!s) return KC_ENOMEM;
        s->device = device;
        s->budget = hbm_budget_bytes;
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
        if (!s || !req_id || (!h_seq && nbytes) || k < 1 || k > 63) return KC_EINVAL;
        auto r = std::make_unique<Request>();
        r->kind = 0;
        r->table_id = table_id;
        r->k = k;
        r->bloom = bloom ? 1 : 0;
        r->seq = h_seq;
        r->nbytes = nbytes;
        r->max_distinct = max_distinct ? max_distinct : std::max<uint64_t>(nbytes, 1024);
        uint64_t cap;
        r->footprint = kc_server::footprint(*r, cap, r->resident);
        if (r->footprint > s->budget) return KC_ENOMEM;
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
    s->used -= it->second->bytes;
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
    *bytes = s->used;
    return KC_OK;
}
