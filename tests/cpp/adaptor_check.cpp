// Exercises the C++ adaptors of include/kc/kc.hpp on the GPU: counting via
// increment, point queries, both export sorts against a std::stable_sort on
// the key bit range, batched counting of uneven read chunks against counting
// the whole stream, the 12-bit sort's argument checks.  Prints "ok".
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "kc/kc.hpp"

static uint64_t bits(const kc::kmer<1> &k, uint32_t lo, uint32_t hi)
{
    return (k.data[0] >> lo) & ((hi - lo == 64) ? ~0ULL : ((1ULL << (hi - lo)) - 1));
}

int main()
{
    kc::Context ctx(0);
    std::mt19937_64 rng(12345);

    // counting
    kc::ConcurrentCountingHashMap<1> map(ctx, 31, 1 << 16);
    std::vector<kc::kmer<1>> keys(1000);
    for (auto &k : keys) k.data[0] = rng() & ~3ULL;
    std::vector<kc::kmer<1>> stream;
    for (size_t i = 0; i < keys.size(); ++i)
        for (size_t r = 0; r <= i % 7; ++r) stream.push_back(keys[i]);
    map.increment(stream.data(), stream.size());
    map.increment(keys[0], 5);
    auto got = map.find(keys.data(), keys.size());
    for (size_t i = 0; i < keys.size(); ++i) {
        const uint32_t want = (uint32_t)(i % 7 + 1) + (i == 0 ? 5 : 0);
        if (got[i] != want) { std::printf("count mismatch at %zu: %u vs %u\n", i, got[i], want); return 1; }
    }

    // 12-bit sort: stable on bits [2, 44)
    std::vector<kc::ExportRecord<1>> recs(50000), scratch;
    for (size_t i = 0; i < recs.size(); ++i) {
        recs[i].key.data[0] = (rng() & ~3ULL);
        if (i % 5 == 0 && i) recs[i].key = recs[i - 1].key;
        recs[i].count = (uint32_t)i;
    }
    auto want = recs;
    std::stable_sort(want.begin(), want.end(), [](const auto &a, const auto &b) { return bits(a.key, 2, 44) < bits(b.key, 2, 44); });
    kc::radix_sort_export_records_12bit<1, 10>(ctx, recs, scratch, 31);
    for (size_t i = 0; i < recs.size(); ++i)
        if (recs[i].count != want[i].count) { std::printf("12-bit sort mismatch at %zu\n", i); return 1; }

    // 8-bit sort of k=31 reads bits [2, 50)
    auto r8 = want;
    std::shuffle(r8.begin(), r8.end(), rng);
    auto want8 = r8;
    std::stable_sort(want8.begin(), want8.end(), [](const auto &a, const auto &b) { return bits(a.key, 2, 50) < bits(b.key, 2, 50); });
    std::vector<kc::ExportRecord<1>> tmp(r8.size());
    auto *res = kc::export_record_radix_sort<1>(ctx, r8.data(), tmp.data(), r8.size(), 31);
    for (size_t i = 0; i < r8.size(); ++i)
        if (res[i].count != want8[i].count) { std::printf("8-bit sort mismatch at %zu\n", i); return 1; }

    // batched counting of uneven read chunks equals counting the whole stream
    {
        std::string genome(200000, 'A');
        for (auto &c : genome) c = "ACGT"[rng() & 3];
        std::string all;
        std::vector<std::string> chunks;
        for (int r = 0; r < 4000; ++r) {
            const size_t len = 100 + rng() % 900, pos = rng() % (genome.size() - len);
            std::string read = genome.substr(pos, len) + "\n";
            if (chunks.empty() || rng() % 4 == 0) chunks.emplace_back();
            chunks.back() += read;
            all += read;
        }
        kc::ConcurrentCountingHashMap<1> whole(ctx, 31, 1 << 22, 0xFFFFFFFFu, true);
        whole.count_reads(all.data(), all.size());
        kc::ConcurrentCountingHashMap<1> batched(ctx, 31, 1 << 22, 0xFFFFFFFFu, true);
        kc_tune tune{};
        tune.batch_bytes = 1 << 20;
        tune.queue_depth = 3;
        {
            kc::ReadBatch b(batched, &tune);
            for (const auto &c : chunks) b.add(c.data(), c.size());
            b.flush();
            b.sync();
            if (b.info().launches < 2) { std::printf("expected several batch launches\n"); return 1; }
        }
        if (whole.histogram(1, 1000) != batched.histogram(1, 1000) || whole.stats().kmers != batched.stats().kmers) {
            std::printf("batched counting differs\n");
            return 1;
        }
    }

    // a temp_dir written with the writer adaptors: the FIRST occurrence of
    // each key in low.bin, its other occurrences as the high.bin count;
    // loading it back merges to the true counts (histogram_tool precise)
    {
        const char *dir_env = std::getenv("KC_ADAPTOR_DIR");
        if (!dir_env) { std::printf("KC_ADAPTOR_DIR not set\n"); return 1; }
        const std::string dir = std::string(dir_env) + "/";
        kc::write_infos(dir, 31, 255);
        {
            kc::ExportWriter<1> low(ctx, 31, dir, 300);
            low.start();
            low.write(keys.data(), keys.size());
            low.join();
            if (low.written() != keys.size()) { std::printf("low records\n"); return 1; }
            kc::FinalDrainWriter<1> high(ctx, 31, 255, dir, 2, 5, 256);
            std::vector<kc::ExportRecord<1>> rs;
            for (size_t i = 0; i < keys.size(); ++i)
                if (i % 7) rs.push_back({keys[i], (uint32_t)(i % 7)});  // total count i % 7 + 1
            // the map path filters (count 1 < 2 - 1 is kept, count 6 > 5 is dropped)...
            high.write_map_record(rs.data(), (uint32_t)rs.size());
            // ...the k-mer path does not
            high.write_kmer_record(keys[6].data.data(), 6);
            high.close();
        }
        kc_table *t = nullptr;
        kc::check(kc_table_load_dir(ctx.get(), dir.c_str(), 1, &t), "kc_table_load_dir");
        std::vector<int64_t> bins(10);
        kc::check(kc_table_histogram(t, 1, 10, bins.data()), "kc_table_histogram");
        kc_table_destroy(t);
        // keys i with i % 7 in 1..5 merge to i % 7 + 1; i % 7 == 6 (dropped from
        // high.bin but for key 6) read 1; i % 7 == 0 read 1
        std::vector<int64_t> want_bins(10, 0);
        for (size_t i = 0; i < keys.size(); ++i) {
            const size_t r = i % 7;
            const uint32_t c = (r >= 1 && r <= 5) || i == 6 ? (uint32_t)r + 1 : 1;
            ++want_bins[c - 1];
        }
        if (bins != want_bins) { std::printf("writer round trip histogram differs\n"); return 1; }
    }

    bool threw = false;
    try { kc::radix_sort_export_records_12bit<1, 10>(ctx, recs, scratch, 9); } catch (const std::invalid_argument &) { threw = true; }
    if (!threw) { std::printf("missing invalid_argument\n"); return 1; }
    std::printf("ok\n");
    return 0;
}
