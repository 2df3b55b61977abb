// Point-query timing harness built on the reference's FlatConcurrentHashMap
// (tool/FlatConcurrentHashMap.h) with the lookup loop of
// LowFrequencyQueryThreadPool.h:196-244: prepare_lookup, prefetch a fixed
// distance ahead (QUERY_PREFETCH_DISTANCE = 8, :25), find_prepared.
//
//   query_bench_ref <n_keys> <n_batches> <threads> <seed> [batch_keys]
//
// Inserts n_keys distinct k=31 keys (a 62-bit SplitMix64 stream shifted left
// by 2, the one-word layout of kmer.h) with counts 1.., seals, then answers
// n_batches query batches of log-uniform size in [1, 65536] (or of
// batch_keys keys each), half present keys and half absent ones, one batch
// per thread at a time.  Prints one JSON
// line: keys/s and the p50/p99 batch latency.  Exit 1 if any answer is wrong.
#include "FlatConcurrentHashMap.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

namespace {

uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

kmer<1> key_of(uint64_t stream, uint64_t i)
{
    uint64_t s = stream ^ (i * 0xD1B54A32D192ED03ULL);
    kmer<1> k;
    k.data[0] = (splitmix(s) >> 2) << 2;
    return k;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc != 5 && argc != 6) {
        std::fprintf(stderr, "usage: %s n_keys n_batches threads seed [batch_keys]\n", argv[0]);
        return 2;
    }
    const uint64_t fixed = argc == 6 ? std::strtoull(argv[5], nullptr, 10) : 0;
    const uint64_t n_keys = std::strtoull(argv[1], nullptr, 10);
    const uint64_t n_batches = std::strtoull(argv[2], nullptr, 10);
    const uint32_t threads = (uint32_t)std::strtoul(argv[3], nullptr, 10);
    const uint64_t seed = std::strtoull(argv[4], nullptr, 10);
    const uint64_t present_stream = seed * 0x9E3779B97F4A7C15ULL + 1, absent_stream = ~present_stream;
    constexpr uint64_t kPrefetch = 8;

    FlatConcurrentHashMap<1> map(n_keys, threads);
    {
        std::vector<std::thread> ws;
        for (uint32_t t = 0; t < threads; ++t)
            ws.emplace_back([&, t] {
                for (uint64_t i = n_keys * t / threads; i < n_keys * (t + 1) / threads; ++i)
                    if (!map.insert_unique(key_of(present_stream, i), (uint32_t)(i % 1000) + 1)) std::abort();
            });
        for (auto &w : ws) w.join();
    }
    map.seal();

    // batches: sizes log-uniform in [1, 65536]
    std::vector<uint64_t> size(n_batches), first(n_batches);
    uint64_t rs = seed ^ 0xABCDEF;
    uint64_t total = 0;
    for (uint64_t b = 0; b < n_batches; ++b) {
        const double u = (double)(splitmix(rs) >> 11) * 0x1.0p-53;
        size[b] = fixed ? fixed
                        : std::min<uint64_t>(65536, std::max<uint64_t>(1, (uint64_t)std::llround(std::exp(u * std::log(65536.0)))));
        first[b] = splitmix(rs) % n_keys;
        total += size[b];
    }

    std::vector<double> lat(n_batches);
    std::atomic<uint64_t> next{0}, wrong{0};
    const auto t0 = std::chrono::steady_clock::now();
    {
        std::vector<std::thread> ws;
        for (uint32_t t = 0; t < threads; ++t)
            ws.emplace_back([&] {
                std::vector<kmer<1>> keys;
                std::vector<uint32_t> want;
                for (;;) {
                    const uint64_t b = next.fetch_add(1);
                    if (b >= n_batches) break;
                    const uint64_t n = size[b], half = n / 2;
                    keys.resize(n);
                    want.resize(n);
                    for (uint64_t i = 0; i < n; ++i) {
                        if (i < half) {
                            const uint64_t j = (first[b] + i * 7919) % n_keys;
                            keys[i] = key_of(present_stream, j);
                            want[i] = (uint32_t)(j % 1000) + 1;
                        } else {
                            keys[i] = key_of(absent_stream, b * 65536 + i);
                            want[i] = 0;
                        }
                    }
                    const auto s = std::chrono::steady_clock::now();
                    using Lookup = FlatConcurrentHashMap<1>::PreparedLookup;
                    Lookup lk[kPrefetch];
                    uint64_t bad = 0;
                    for (uint64_t i = 0; i < std::min(n, kPrefetch); ++i) {
                        lk[i] = map.prepare_lookup(keys[i]);
                        map.prefetch(lk[i]);
                    }
                    for (uint64_t i = 0; i < n; ++i) {
                        uint32_t c = 0;
                        if (!map.find_prepared(keys[i], lk[i % kPrefetch], c)) c = 0;
                        bad += c != want[i];
                        if (i + kPrefetch < n) {
                            lk[i % kPrefetch] = map.prepare_lookup(keys[i + kPrefetch]);
                            map.prefetch(lk[i % kPrefetch]);
                        }
                    }
                    lat[b] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - s).count();
                    wrong += bad;
                }
            });
        for (auto &w : ws) w.join();
    }
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::vector<double> sorted = lat;
    std::sort(sorted.begin(), sorted.end());
    auto pct = [&](double p) { return sorted[std::min<uint64_t>(n_batches - 1, (uint64_t)(p / 100.0 * n_batches))]; };
    std::printf("{\"keys\": %llu, \"table_keys\": %llu, \"batches\": %llu, \"threads\": %u, \"keys_per_s\": %.6g, "
                "\"batch_p50_ms\": %.6g, \"batch_p99_ms\": %.6g, \"wrong\": %llu}\n",
                (unsigned long long)total, (unsigned long long)n_keys, (unsigned long long)n_batches, threads,
                total / wall, pct(50), pct(99), (unsigned long long)wrong.load());
    return wrong.load() ? 1 : 0;
}
