/*
 * kc_oracle.h — CPU restatement of KTCounterMemoryGreedy's canonical k-mer
 * counting path.  TEST INFRASTRUCTURE ONLY: this library is the checker for
 * the HIP engine.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product path never links or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Key layout (all functions): a k-mer is nw = words_for_k(k) little-endian
 * u64 words, word 0 most significant, string position j at bit
 * 64*nw - 2k + 2j of that 64*nw-bit number (GetKmer.h:71-77).
 *
 * Pinning: see the header comment of kc_oracle.c for which check pins which
 * function.
 */
#ifndef KC_ORACLE_H
#define KC_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

uint32_t kco_words_for_k(uint32_t k);

/* FASTQ line-2 framing: copies every sequence line followed by '\n'. */
uint64_t kco_fastq_frame(const uint8_t *fq, uint64_t len, uint8_t *out);

/* Rolling canonical extraction over a framed byte stream. */
uint64_t kco_extract(const uint8_t *bytes, uint64_t len, uint32_t k,
                     uint64_t *out, uint64_t out_cap);
/* The windows of kco_extract whose key is in hash slice `pass` of
 * 2^pass_bits (top pass_bits of kco_xxh3_64 of the key bytes). */
uint64_t kco_extract_slice(const uint8_t *bytes, uint64_t len, uint32_t k, uint32_t pass_bits,
                           uint32_t pass, uint64_t *out, uint64_t out_cap);
/* Closed-form encoder of one window (independent of the rolling code). */
int kco_encode_window(const uint8_t *s, uint32_t k, uint64_t *canon_out);

/* Exact counting: unique keys + counts, sorted by numeric key order. */
uint64_t kco_count(uint64_t *keys, uint64_t n, uint32_t nw,
                   uint64_t *uniq_out, uint64_t *counts_out);

/* Export packing. */
uint32_t kco_packed_kmer_bytes(uint32_t k);
uint32_t kco_count_bytes(uint32_t count_max);
void kco_pack_kmer(const uint64_t *key, uint32_t k, uint8_t *out);
void kco_unpack_kmer(const uint8_t *in, uint32_t k, uint64_t *key);

/* Merge of high/low into (key,count): dump_tool precise semantics. */
uint64_t kco_merge(const uint64_t *high_keys, const uint32_t *high_counts,
                   uint64_t n_high, const uint64_t *low_keys, uint64_t n_low,
                   uint32_t nw, uint32_t count_max, uint32_t min_freq,
                   uint32_t max_freq, uint64_t *keys_out, uint32_t *counts_out);
/* Histogram: histogram_tool precise semantics. */
void kco_histogram(const uint64_t *high_keys, const uint32_t *high_counts,
                   uint64_t n_high, const uint64_t *low_keys, uint64_t n_low,
                   uint32_t nw, uint32_t count_max, uint32_t min_freq,
                   uint32_t max_freq, int64_t *bins);
/* Lexicographic dump text. Returns bytes written (out may be NULL to size). */
uint64_t kco_dump_text(const uint64_t *keys, const uint32_t *counts, uint64_t n,
                       uint32_t k, char *out);
uint64_t kco_histogram_text(const int64_t *bins, uint32_t min_freq,
                            uint32_t max_freq, char *out);

/* Sorts over ExportRecord<N> arrays given as keys[n*nw] + counts[n]. */
void kco_sort_records8(uint64_t *keys, uint32_t *counts, uint64_t n,
                       uint32_t nw, uint32_t k);
int kco_sort_records12(uint64_t *keys, uint32_t *counts, uint64_t n,
                       uint32_t nw, uint32_t k, uint32_t common_prefix_bases);

/* Hashes of raw key bytes (the kmer<N> object bytes). */
uint64_t kco_xxh3_64(const void *p, uint64_t len);
void kco_xxh3_128(const void *p, uint64_t len, uint64_t *lo, uint64_t *hi);
uint64_t kco_rapidhash(const void *p, uint64_t len);
uint64_t kco_komihash(const void *p, uint64_t len, uint64_t seed);
uint64_t kco_mix_hash(uint64_t h1, uint64_t h2);
uint64_t kco_hash_func(const uint64_t *words, uint32_t nw);
uint64_t kco_concurrent_map_hash(const uint64_t *words, uint32_t nw);

/* Bloom probe (block, 3-bit mask) of BloomFilter.h:215-222, 287-297. */
void kco_bloom_probe(const uint64_t *key, uint32_t nw, uint64_t capacity,
                     uint64_t *block_out, uint64_t *mask_out);

#ifdef __cplusplus
}
#endif
#endif
