// Known-answer generator built from the reference's own headers.
// Compiled by oracle/Makefile against the sources under the reference tree;
// prints one line per key:
//   nw w0 [w1 ..] xxh3_64 xxh3_128_lo xxh3_128_hi rapidhash komihash
//   concurrent_map_hash hash_func bloom_block(cap=2^20) bloom_mask
// All values hex.  Keys come from a fixed SplitMix64 stream (seed argv[1]).
#include "BloomFilter.h"
#include "ConcurrentMap.h"
#include "ConcurrentMemoryPool.h"
#include "HashFunction.h"
#include "kmer.h"

#include <cstdio>
#include <cstdlib>

static uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

template <uint32_t N>
static void emit(ConcurrentBloomFilter<N> &bf, const kmer<N> &key)
{
    std::printf("%u", N);
    for (uint32_t i = 0; i < N; ++i) std::printf(" %016llx", (unsigned long long)key.data[i]);
    const uint64_t x64 = XXH3_64bits(&key, sizeof(key));
    const XXH128_hash_t x128 = XXH3_128bits(&key, sizeof(key));
    const uint64_t rh = rapidhash(&key, sizeof(key));
    const uint64_t kh = komihash(&key, sizeof(key), 0);
    const uint64_t cm = ConcurrentMap<N>::hash_func(key);
    const uint64_t hf = hash_func<N>(key);
    const auto probe = bf.prepare_insert(key);
    std::printf(" %016llx %016llx %016llx %016llx %016llx %016llx %016llx %016llx %016llx\n",
                (unsigned long long)x64, (unsigned long long)x128.low64,
                (unsigned long long)x128.high64, (unsigned long long)rh,
                (unsigned long long)kh, (unsigned long long)cm, (unsigned long long)hf,
                (unsigned long long)probe.block_idx, (unsigned long long)probe.insert_num);
}

template <uint32_t N>
static void run(ConcurrentMemoryPool &pool, uint64_t seed, int count)
{
    ConcurrentBloomFilter<N> bf(1ULL << 20, &pool);
    uint64_t s = seed;
    for (int i = 0; i < count; ++i) {
        kmer<N> key;
        for (uint32_t w = 0; w < N; ++w) key.data[w] = splitmix(s);
        // half the keys look like k-mers: zero padding below 2k bits
        if (i & 1) key.data[N - 1] &= ~3ULL;
        emit<N>(bf, key);
    }
    kmer<N> zero;
    zero.reset();
    emit<N>(bf, zero);
}

int main(int argc, char **argv)
{
    const uint64_t seed = argc > 1 ? std::strtoull(argv[1], nullptr, 0) : 7;
    const int count = argc > 2 ? std::atoi(argv[2]) : 64;
    ConcurrentMemoryPool pool(1ULL << 30);
    pool.init_arenas();
    run<1>(pool, seed, count);
    run<2>(pool, seed + 1, count);
    run<4>(pool, seed + 2, count);
    return 0;
}
