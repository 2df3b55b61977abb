/*
 * kc_oracle.c — CPU restatement of the KTCounterMemoryGreedy counting path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker for the HIP engine: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * The product library never links, calls or falls back to it.
 *
 * How each function is pinned (what would catch a plausible mistake):
 *  - kco_fastq_frame, kco_extract: the reference binaries' dump/histogram
 *    goldens (tests/golden, produced by oracle/_ref/kmc_ref + tools on seeded
 *    inputs) and a brute-force check against kco_encode_window, which builds
 *    each window from the string by the closed form, not by rolling.
 *  - kco_count, kco_merge, kco_histogram, kco_dump_text, kco_histogram_text:
 *    reference goldens (byte-identical dump --sort and histogram text) plus
 *    the merge-rule KAT of test_low_frequency_query_thread_pool.cpp:94-138.
 *    For k = 65..128 the reference's dump text is not a k-mer spelling (its
 *    sort key holds 128 bits) and at k = 96 its counter splits counts: there
 *    the extraction, counts and dump text are pinned by a count done on
 *    strings (tests/test_oracle.py) and the histograms by the reference's
 *    goldens at k = 65, 100, 127 and 128.
 *  - kco_pack_kmer / kco_unpack_kmer: the k=33 / k=64 vectors of
 *    test_export_reader.cpp:175-289 and round trips through the reference
 *    tools (their reader unpacks our files).
 *  - kco_sort_records8 / kco_sort_records12: oracle/_ref/sort_ref (the
 *    reference headers compiled into a harness) and a stable sort on the
 *    masked key computed independently in the tests.
 *  - hashes and Bloom probe: oracle/_ref/hash_kat_ref (the vendored headers
 *    and BloomFilter.h compiled into a harness) on 8/16/32-byte keys.
 *  - kco_extract_slice: kco_extract's windows filtered by the top bits of
 *    kco_xxh3_64 computed per key outside the oracle (tests/test_oracle.py);
 *    both are pinned above.
 */
#include "kc_oracle.h"

#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

uint32_t kco_words_for_k(uint32_t k)
{
    /* main.cpp:617-633 */
    if (k <= 32) return 1;
    if (k <= 64) return 2;
    return 4;
}

/* ------------------------------------------------------------------ */
/* FASTQ framing: FastqReader.h:478-487 (states), :657-687 (line 2 copy
 * followed by '\n').  The k-1 overlap between 32 KiB blocks
 * (FastqReader.h:542-557) makes every window of one sequence line appear
 * exactly once, which is what a single contiguous copy gives directly. */
uint64_t kco_fastq_frame(const uint8_t *fq, uint64_t len, uint8_t *out)
{
    uint64_t w = 0;
    uint32_t state = 0; /* 0 header, 1 sequence, 2 plus, 3 quality */
    uint64_t i = 0;
    while (i < len) {
        const uint8_t *nl = memchr(fq + i, '\n', len - i);
        uint64_t end = nl ? (uint64_t)(nl - fq) : len;
        if (state == 1) {
            memcpy(out + w, fq + i, end - i);
            w += end - i;
            if (nl) out[w++] = '\n';
        }
        if (!nl) break;
        i = end + 1;
        state = (state + 1) & 3;
    }
    return w;
}

/* ------------------------------------------------------------------ */
/* Rolling canonical extraction: GetKmer.h:19-49 (tables, masks), :58-82
 * (get_next_one), :181-207 (canonicalize = numeric min, word 0 first). */
static int code_of(uint8_t c)
{
    switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': case 'U': case 'u': return 3;
    default: return -1;
    }
}

static int cmp_words(const uint64_t *a, const uint64_t *b, uint32_t nw)
{
    for (uint32_t i = 0; i < nw; ++i) {
        if (a[i] < b[i]) return -1;
        if (a[i] > b[i]) return 1;
    }
    return 0;
}

/* Rolling extraction of every window; with pass_bits > 0 only the windows
 * whose key is in hash slice `pass` are kept (the engine's multi-pass
 * filter: the top pass_bits of XXH3_64 of the 8N key bytes). */
static uint64_t extract_impl(const uint8_t *bytes, uint64_t len, uint32_t k, uint32_t pass_bits,
                             uint32_t pass, uint64_t *out, uint64_t out_cap)
{
    const uint32_t nw = kco_words_for_k(k);
    const uint32_t last = (k - 1) / 32;
    const uint32_t rem = k % 32;
    const uint64_t back_mask = rem ? (((1ULL << (2 * rem)) - 1ULL) << (64 - 2 * rem)) : ~0ULL;
    const uint32_t rev_shift = rem ? 2 * (32 - rem) : 0;
    uint64_t fwd[4] = {0, 0, 0, 0}, rev[4] = {0, 0, 0, 0};
    uint64_t have = 0, n = 0;
    for (uint64_t p = 0; p < len; ++p) {
        int c = code_of(bytes[p]);
        if (c < 0) {
            memset(fwd, 0, sizeof fwd);
            memset(rev, 0, sizeof rev);
            have = 0;
            continue;
        }
        for (int i = (int)nw - 1; i > 0; --i) fwd[i] = (fwd[i] >> 2) | (fwd[i - 1] << 62);
        fwd[0] = (fwd[0] >> 2) | ((uint64_t)c << 62);
        fwd[last] &= back_mask;
        /* words after `last` are padding: zero (kmer<4> at k = 65..96 leaves
         * shifted-out bases there, GetKmer.h:71-73 masks word `last` only) */
        for (uint32_t i = last + 1; i < nw; ++i) fwd[i] = 0;
        for (uint32_t i = 0; i + 1 < nw; ++i) rev[i] = (rev[i] << 2) | (rev[i + 1] >> 62);
        rev[nw - 1] <<= 2;
        rev[last] |= (uint64_t)(c ^ 3) << rev_shift;
        if (++have >= k) {
            const uint64_t *m = cmp_words(fwd, rev, nw) < 0 ? fwd : rev;
            if (pass_bits && (uint32_t)(kco_xxh3_64(m, 8ULL * nw) >> (64 - pass_bits)) != pass) continue;
            if (n < out_cap) memcpy(out + n * nw, m, nw * sizeof(uint64_t));
            ++n;
        }
    }
    return n;
}

uint64_t kco_extract(const uint8_t *bytes, uint64_t len, uint32_t k,
                     uint64_t *out, uint64_t out_cap)
{
    return extract_impl(bytes, len, k, 0, 0, out, out_cap);
}

uint64_t kco_extract_slice(const uint8_t *bytes, uint64_t len, uint32_t k, uint32_t pass_bits,
                           uint32_t pass, uint64_t *out, uint64_t out_cap)
{
    return extract_impl(bytes, len, k, pass_bits, pass, out, out_cap);
}

/* Closed form for one window: position j sits at bit 64nw-2k+2j of the
 * 64nw-bit number (word 0 most significant); canonical = min(s, rc(s)). */
static void put_base(uint64_t *w, uint32_t nw, uint32_t bit, uint64_t code)
{
    /* bit counts from the least significant bit of the whole number */
    uint32_t word = nw - 1 - bit / 64;
    w[word] |= code << (bit % 64);
}

int kco_encode_window(const uint8_t *s, uint32_t k, uint64_t *canon_out)
{
    const uint32_t nw = kco_words_for_k(k);
    uint64_t f[4] = {0, 0, 0, 0}, r[4] = {0, 0, 0, 0};
    const uint32_t base0 = 64 * nw - 2 * k;
    for (uint32_t j = 0; j < k; ++j) {
        int c = code_of(s[j]);
        if (c < 0) return 0;
        put_base(f, nw, base0 + 2 * j, (uint64_t)c);
        put_base(r, nw, base0 + 2 * (k - 1 - j), (uint64_t)(3 - c));
    }
    memcpy(canon_out, cmp_words(f, r, nw) <= 0 ? f : r, nw * sizeof(uint64_t));
    return 1;
}

/* ------------------------------------------------------------------ */
/* LSD radix sort of nw-word keys (numeric order, word 0 most significant)
 * carrying an optional u64 payload.  Plain counting sort on 16-bit digits. */
static void radix_sort_words(uint64_t *keys, uint64_t *payload, uint64_t n, uint32_t nw)
{
    if (n < 2) return;
    uint64_t *tk = malloc(n * nw * sizeof(uint64_t));
    uint64_t *tp = payload ? malloc(n * sizeof(uint64_t)) : NULL;
    uint64_t *cnt = malloc(65536 * sizeof(uint64_t));
    for (int word = (int)nw - 1; word >= 0; --word) {
        for (uint32_t sh = 0; sh < 64; sh += 16) {
            /* skip digits that are constant across the array */
            uint64_t d0 = (keys[word] >> sh) & 0xFFFF;
            int constant = 1;
            for (uint64_t i = 1; i < n; ++i)
                if (((keys[i * nw + word] >> sh) & 0xFFFF) != d0) { constant = 0; break; }
            if (constant) continue;
            memset(cnt, 0, 65536 * sizeof(uint64_t));
            for (uint64_t i = 0; i < n; ++i) cnt[(keys[i * nw + word] >> sh) & 0xFFFF]++;
            uint64_t s = 0;
            for (uint32_t d = 0; d < 65536; ++d) { uint64_t c = cnt[d]; cnt[d] = s; s += c; }
            for (uint64_t i = 0; i < n; ++i) {
                uint64_t d = (keys[i * nw + word] >> sh) & 0xFFFF;
                uint64_t o = cnt[d]++;
                memcpy(tk + o * nw, keys + i * nw, nw * sizeof(uint64_t));
                if (payload) tp[o] = payload[i];
            }
            memcpy(keys, tk, n * nw * sizeof(uint64_t));
            if (payload) memcpy(payload, tp, n * sizeof(uint64_t));
        }
    }
    free(tk); free(tp); free(cnt);
}

/* Exact counts.  The reference's merged views equal min(true count,
 * count_max) per distinct canonical k-mer (ConcurrentMap.h:245-253,
 * NewKmerTree.h:1369,1385); clamping is left to the caller. */
uint64_t kco_count(uint64_t *keys, uint64_t n, uint32_t nw,
                   uint64_t *uniq_out, uint64_t *counts_out)
{
    radix_sort_words(keys, NULL, n, nw);
    uint64_t u = 0;
    for (uint64_t i = 0; i < n;) {
        uint64_t j = i + 1;
        while (j < n && cmp_words(keys + i * nw, keys + j * nw, nw) == 0) ++j;
        memcpy(uniq_out + u * nw, keys + i * nw, nw * sizeof(uint64_t));
        counts_out[u] = j - i;
        ++u;
        i = j;
    }
    return u;
}

/* ------------------------------------------------------------------ */
/* Packing: ExportWriter.h:296-313 and FinalDrainWriter.h:43-88.  Full words
 * as LE u64, then the top ceil(2(k%32)/8) bytes of the masked tail word. */
uint32_t kco_packed_kmer_bytes(uint32_t k)
{
    return (k / 32) * 8 + (2 * (k % 32) + 7) / 8;
}

uint32_t kco_count_bytes(uint32_t count_max)
{
    /* main.cpp:599-615 */
    if (count_max <= 0xFF) return 1;
    if (count_max <= 0xFFFF) return 2;
    if (count_max <= 0xFFFFFF) return 3;
    return 4;
}

void kco_pack_kmer(const uint64_t *key, uint32_t k, uint8_t *out)
{
    const uint32_t full = k / 32, tail_bits = 2 * (k % 32), tail_bytes = (tail_bits + 7) / 8;
    memcpy(out, key, full * 8);
    if (tail_bytes) {
        uint64_t t = key[full] & (~0ULL << (64 - tail_bits));
        memcpy(out + full * 8, (const uint8_t *)&t + (8 - tail_bytes), tail_bytes);
    }
}

void kco_unpack_kmer(const uint8_t *in, uint32_t k, uint64_t *key)
{
    /* ExportReader / dump_tool.cpp:864-876: the inverse of kco_pack_kmer */
    const uint32_t nw = kco_words_for_k(k);
    const uint32_t full = k / 32, tail_bytes = (2 * (k % 32) + 7) / 8;
    memset(key, 0, nw * 8);
    memcpy(key, in, full * 8);
    if (tail_bytes) {
        uint64_t t = 0;
        memcpy((uint8_t *)&t + (8 - tail_bytes), in + full * 8, tail_bytes);
        key[full] = t;
    }
}

/* ------------------------------------------------------------------ */
/* A sorted view of the high records for exact lookup (plays the role of
 * FlatConcurrentHashMap in the tools; lookups are exact either way). */
typedef struct {
    uint64_t *keys;
    uint64_t *idx;
    uint64_t n;
    uint32_t nw;
} sorted_view;

static void view_build(sorted_view *v, const uint64_t *keys, const uint8_t *take,
                       uint64_t n, uint32_t nw)
{
    uint64_t m = 0;
    for (uint64_t i = 0; i < n; ++i) m += take ? take[i] : 1;
    v->keys = malloc((m ? m : 1) * nw * 8);
    v->idx = malloc((m ? m : 1) * 8);
    v->n = m;
    v->nw = nw;
    uint64_t j = 0;
    for (uint64_t i = 0; i < n; ++i) {
        if (take && !take[i]) continue;
        memcpy(v->keys + j * nw, keys + i * nw, nw * 8);
        v->idx[j++] = i;
    }
    radix_sort_words(v->keys, v->idx, m, nw);
}

static int64_t view_find(const sorted_view *v, const uint64_t *key)
{
    uint64_t lo = 0, hi = v->n;
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo) / 2;
        int c = cmp_words(v->keys + mid * v->nw, key, v->nw);
        if (c == 0) return (int64_t)v->idx[mid];
        if (c < 0) lo = mid + 1; else hi = mid;
    }
    return -1;
}

static void view_free(sorted_view *v) { free(v->keys); free(v->idx); }

static int in_range(uint64_t v, uint32_t lo, uint32_t hi) { return v >= lo && v <= hi; }

/* dump_tool.cpp:682-733 (insert high only if min<=c<=max), :824-909
 * (low: found -> min(c+1,count_max) and zero the slot; missing -> 1),
 * :928-936 (remaining entries with count>0 in range). */
uint64_t kco_merge(const uint64_t *high_keys, const uint32_t *high_counts,
                   uint64_t n_high, const uint64_t *low_keys, uint64_t n_low,
                   uint32_t nw, uint32_t count_max, uint32_t min_freq,
                   uint32_t max_freq, uint64_t *keys_out, uint32_t *counts_out)
{
    uint8_t *take = malloc(n_high ? n_high : 1);
    uint32_t *cur = malloc((n_high ? n_high : 1) * 4);
    for (uint64_t i = 0; i < n_high; ++i) {
        take[i] = (uint8_t)in_range(high_counts[i], min_freq, max_freq);
        cur[i] = high_counts[i];
    }
    sorted_view v;
    view_build(&v, high_keys, take, n_high, nw);
    uint64_t m = 0;
    for (uint64_t i = 0; i < n_low; ++i) {
        int64_t f = view_find(&v, low_keys + i * nw);
        if (f >= 0) {
            uint64_t merged = (uint64_t)cur[f] + 1;
            if (merged > count_max) merged = count_max;
            if (in_range(merged, min_freq, max_freq)) {
                memcpy(keys_out + m * nw, low_keys + i * nw, nw * 8);
                counts_out[m++] = (uint32_t)merged;
            }
            cur[f] = 0;
        } else if (in_range(1, min_freq, max_freq)) {
            memcpy(keys_out + m * nw, low_keys + i * nw, nw * 8);
            counts_out[m++] = 1;
        }
    }
    for (uint64_t i = 0; i < n_high; ++i) {
        if (!take[i]) continue;
        if (cur[i] > 0 && in_range(cur[i], min_freq, max_freq)) {
            memcpy(keys_out + m * nw, high_keys + i * nw, nw * 8);
            counts_out[m++] = cur[i];
        }
    }
    view_free(&v);
    free(take);
    free(cur);
    return m;
}

/* histogram_tool.cpp:734-790 with HighFrequencyInsertThreadPool.h:150-177
 * and LowFrequencyQueryThreadPool.h:186-245. */
void kco_histogram(const uint64_t *high_keys, const uint32_t *high_counts,
                   uint64_t n_high, const uint64_t *low_keys, uint64_t n_low,
                   uint32_t nw, uint32_t count_max, uint32_t min_freq,
                   uint32_t max_freq, int64_t *bins)
{
    const uint64_t nb = (uint64_t)max_freq - min_freq + 1;
    memset(bins, 0, nb * sizeof(int64_t));
    uint8_t *take = malloc(n_high ? n_high : 1);
    for (uint64_t i = 0; i < n_high; ++i) {
        uint64_t c = high_counts[i];
        take[i] = 0;
        if (c + 1 < min_freq) continue;
        if (in_range(c, min_freq, max_freq)) bins[c - min_freq]++;
        take[i] = 1;
    }
    sorted_view v;
    view_build(&v, high_keys, take, n_high, nw);
    for (uint64_t i = 0; i < n_low; ++i) {
        int64_t f = view_find(&v, low_keys + i * nw);
        if (f >= 0) {
            uint64_t c = high_counts[f];
            if (in_range(c, min_freq, max_freq)) bins[c - min_freq]--;
            uint64_t merged = c + 1;
            if (merged > count_max) merged = count_max;
            if (in_range(merged, min_freq, max_freq)) bins[merged - min_freq]++;
        } else if (in_range(1, min_freq, max_freq)) {
            bins[1 - min_freq]++;
        }
    }
    view_free(&v);
    free(take);
}

/* ------------------------------------------------------------------ */
/* Text outputs.  dump --sort: lexicographic (A<C<G<T) order of the string
 * (dump_tool.cpp:345-354 make_sort_key, :373-454 sort), line = bases, '\t',
 * decimal count, '\n' (dump_tool.cpp:1172-1193).  The string-order key is
 * a 256-bit number (word 0 most significant) whose low 2k bits hold the
 * bases, first base most significant; for k > 64 that is wider than the
 * reference's 128-bit sort key (SortKeyType<4>, dump_tool.cpp:320-324),
 * which keeps only the last 64 bases of each line. */
static void string_order_key(const uint64_t *key, uint32_t k, uint64_t sk[4])
{
    const uint32_t nw = kco_words_for_k(k);
    const uint32_t base0 = 64 * nw - 2 * k;
    sk[0] = sk[1] = sk[2] = sk[3] = 0;
    for (uint32_t j = 0; j < k; ++j) {
        uint32_t bit = base0 + 2 * j;
        uint64_t c = (key[nw - 1 - bit / 64] >> (bit % 64)) & 3;
        uint32_t at = 2 * (k - 1 - j); /* the string key's bit for base j */
        sk[3 - at / 64] |= c << (at % 64);
    }
}

uint64_t kco_dump_text(const uint64_t *keys, const uint32_t *counts, uint64_t n,
                       uint32_t k, char *out)
{
    static const char B[4] = {'A', 'C', 'G', 'T'};
    uint64_t *sk = malloc((n ? n : 1) * 4 * 8);
    uint64_t *ord = malloc((n ? n : 1) * 8);
    for (uint64_t i = 0; i < n; ++i) {
        string_order_key(keys + i * kco_words_for_k(k), k, sk + 4 * i);
        ord[i] = i;
    }
    radix_sort_words(sk, ord, n, 4);
    uint64_t w = 0;
    char digits[16];
    for (uint64_t r = 0; r < n; ++r) {
        const uint64_t *s = sk + 4 * r;
        uint32_t c = counts[ord[r]];
        if (out) {
            for (uint32_t j = 0; j < k; ++j) {
                uint32_t at = 2 * (k - 1 - j);
                out[w + j] = B[(s[3 - at / 64] >> (at % 64)) & 3];
            }
        }
        w += k;
        if (out) out[w] = '\t';
        w++;
        int nd = 0;
        do { digits[nd++] = (char)('0' + c % 10); c /= 10; } while (c);
        for (int d = nd - 1; d >= 0; --d) { if (out) out[w] = digits[d]; w++; }
        if (out) out[w] = '\n';
        w++;
    }
    free(sk);
    free(ord);
    return w;
}

/* histogram_tool.cpp:462-479: "<f>\t<n>\n" for f = min..max, zeros included. */
uint64_t kco_histogram_text(const int64_t *bins, uint32_t min_freq,
                            uint32_t max_freq, char *out)
{
    uint64_t w = 0;
    char line[64];
    for (uint64_t f = min_freq; f <= max_freq; ++f) {
        int64_t v = bins[f - min_freq];
        int len = 0;
        char tmp[32];
        int t = 0;
        uint64_t x = f;
        do { tmp[t++] = (char)('0' + x % 10); x /= 10; } while (x);
        while (t) line[len++] = tmp[--t];
        line[len++] = '\t';
        uint64_t ax = v < 0 ? (uint64_t)(-(v + 1)) + 1 : (uint64_t)v;
        if (v < 0) line[len++] = '-';
        do { tmp[t++] = (char)('0' + ax % 10); ax /= 10; } while (ax);
        while (t) line[len++] = tmp[--t];
        line[len++] = '\n';
        if (out) memcpy(out + w, line, len);
        w += len;
    }
    return w;
}

/* ------------------------------------------------------------------ */
/* Stable counting-sort pass on a digit function. */
typedef uint32_t (*digit_fn)(const uint64_t *key, const void *ctx);

static void stable_pass(uint64_t *keys, uint32_t *counts, uint64_t n, uint32_t nw,
                        uint32_t buckets, digit_fn fn, const void *ctx,
                        uint64_t *tk, uint32_t *tc)
{
    uint64_t *cnt = calloc(buckets, sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) cnt[fn(keys + i * nw, ctx)]++;
    uint64_t s = 0;
    for (uint32_t d = 0; d < buckets; ++d) { uint64_t c = cnt[d]; cnt[d] = s; s += c; }
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t o = cnt[fn(keys + i * nw, ctx)]++;
        memcpy(tk + o * nw, keys + i * nw, nw * 8);
        tc[o] = counts[i];
    }
    memcpy(keys, tk, n * nw * 8);
    memcpy(counts, tc, n * 4);
    free(cnt);
}

typedef struct { uint32_t word, offset; } d8_ctx;
static uint32_t digit8(const uint64_t *key, const void *c)
{
    const d8_ctx *x = c;
    return (uint32_t)(key[x->word] >> x->offset) & 0xFF;
}

/* ExportRecordRadixSort.h:225-254: passes walk MSB-numbered bit positions
 * from 2k-1 down by at most 8 bits per pass, never crossing a word, until
 * 2k-20 bits are covered; each pass reads 8 bits at that position
 * (:28-34), so the last pass can also read up to 6 "ignored" bits. */
void kco_sort_records8(uint64_t *keys, uint32_t *counts, uint64_t n,
                       uint32_t nw, uint32_t k)
{
    const uint32_t total_bits = 2 * k;
    uint64_t *tk = malloc((n ? n : 1) * nw * 8);
    uint32_t *tc = malloc((n ? n : 1) * 4);
    uint32_t different = total_bits - 20;
    uint32_t end_pos = total_bits - 1;
    while (different > 0) {
        uint32_t bits = end_pos + 1 - (end_pos / 64) * 64;
        if (bits > 8) bits = 8;
        if (bits > different) bits = different;
        different -= bits;
        d8_ctx c = {end_pos / 64, 63 - (end_pos % 64)};
        stable_pass(keys, counts, n, nw, 256, digit8, &c, tk, tc);
        end_pos -= bits;
    }
    free(tk);
    free(tc);
}

typedef struct { uint32_t word, shift, width; } d12_ctx;
static uint32_t digit12(const uint64_t *key, const void *c)
{
    const d12_ctx *x = c;
    uint64_t v = key[x->word] >> x->shift;
    if (x->shift + x->width > 64) v |= key[x->word - 1] << (64 - x->shift);
    return (uint32_t)(v & ((1ULL << x->width) - 1));
}

/* ExportRecordRadixSort12.h:444-479 (pass list over LSB-numbered bits
 * [64N-2k, 64N-2*CPB) in 12-bit digits), :481-503 (argument checks). */
int kco_sort_records12(uint64_t *keys, uint32_t *counts, uint64_t n,
                       uint32_t nw, uint32_t k, uint32_t cpb)
{
    if (k < cpb || k > 32 * nw) return -1;           /* invalid_argument */
    if (n > 0xFFFFFFFFULL) return -2;                /* length_error */
    const uint32_t total = 64 * nw;
    const uint32_t low_pad = total - 2 * k;
    const uint32_t sort_end = total - 2 * cpb;
    const uint32_t sort_bits = sort_end - low_pad;
    const uint32_t passes = (sort_bits + 11) / 12;
    if (n <= 1 || passes == 0) return 0;
    uint64_t *tk = malloc(n * nw * 8);
    uint32_t *tc = malloc(n * 4);
    for (uint32_t p = 0; p < passes; ++p) {
        uint32_t off = low_pad + 12 * p;
        uint32_t width = sort_end - off < 12 ? sort_end - off : 12;
        d12_ctx c = {nw - 1 - off / 64, off % 64, width};
        stable_pass(keys, counts, n, nw, 1u << width, digit12, &c, tk, tc);
    }
    free(tk);
    free(tc);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Hashes.  XXH3 (include/xxhash.h 0.8.3: 64-bit :4696-4830, 128-bit
 * :6900-7070, default secret :4398), rapidhash V3 (include/rapidhash.h
 * :243-330, secret :130-138), komihash 5.29 (include/komihash.h :565-890),
 * HashFunction.h:9-63 (mix64/finalize/mix_hash/hash_func). */
static const uint8_t kSecret[192] = {
    0xb8, 0xfe, 0x6c, 0x39, 0x23, 0xa4, 0x4b, 0xbe, 0x7c, 0x01, 0x81, 0x2c, 0xf7, 0x21, 0xad, 0x1c,
    0xde, 0xd4, 0x6d, 0xe9, 0x83, 0x90, 0x97, 0xdb, 0x72, 0x40, 0xa4, 0xa4, 0xb7, 0xb3, 0x67, 0x1f,
    0xcb, 0x79, 0xe6, 0x4e, 0xcc, 0xc0, 0xe5, 0x78, 0x82, 0x5a, 0xd0, 0x7d, 0xcc, 0xff, 0x72, 0x21,
    0xb8, 0x08, 0x46, 0x74, 0xf7, 0x43, 0x24, 0x8e, 0xe0, 0x35, 0x90, 0xe6, 0x81, 0x3a, 0x26, 0x4c,
    0x3c, 0x28, 0x52, 0xbb, 0x91, 0xc3, 0x00, 0xcb, 0x88, 0xd0, 0x65, 0x8b, 0x1b, 0x53, 0x2e, 0xa3,
    0x71, 0x64, 0x48, 0x97, 0xa2, 0x0d, 0xf9, 0x4e, 0x38, 0x19, 0xef, 0x46, 0xa9, 0xde, 0xac, 0xd8,
    0xa8, 0xfa, 0x76, 0x3f, 0xe3, 0x9c, 0x34, 0x3f, 0xf9, 0xdc, 0xbb, 0xc7, 0xc7, 0x0b, 0x4f, 0x1d,
    0x8a, 0x51, 0xe0, 0x4b, 0xcd, 0xb4, 0x59, 0x31, 0xc8, 0x9f, 0x7e, 0xc9, 0xd9, 0x78, 0x73, 0x64,
    0xea, 0xc5, 0xac, 0x83, 0x34, 0xd3, 0xeb, 0xc3, 0xc5, 0x81, 0xa0, 0xff, 0xfa, 0x13, 0x63, 0xeb,
    0x17, 0x0d, 0xdd, 0x51, 0xb7, 0xf0, 0xda, 0x49, 0xd3, 0x16, 0x55, 0x26, 0x29, 0xd4, 0x68, 0x9e,
    0x2b, 0x16, 0xbe, 0x58, 0x7d, 0x47, 0xa1, 0xfc, 0x8f, 0xf8, 0xb8, 0xd1, 0x7a, 0xd0, 0x31, 0xce,
    0x45, 0xcb, 0x3a, 0x8f, 0x95, 0x16, 0x04, 0x28, 0xaf, 0xd7, 0xfb, 0xca, 0xbb, 0x4b, 0x40, 0x7e,
};

#define P64_1 0x9E3779B185EBCA87ULL
#define P64_2 0xC2B2AE3D27D4EB4FULL
#define P64_3 0x165667B19E3779F9ULL
#define P64_4 0x85EBCA77C2B2AE63ULL
#define P32_2 0x85EBCA77U
#define MX1 0x165667919E3779F9ULL
#define MX2 0x9FB21C651E98DF25ULL

static uint64_t rd64(const uint8_t *p) { uint64_t v; memcpy(&v, p, 8); return v; }
static uint32_t rd32(const uint8_t *p) { uint32_t v; memcpy(&v, p, 4); return v; }
static uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
static uint64_t bswap64(uint64_t x) { return __builtin_bswap64(x); }
static uint64_t fold64(uint64_t a, uint64_t b) { u128 p = (u128)a * b; return (uint64_t)p ^ (uint64_t)(p >> 64); }
static uint64_t xs(uint64_t v, int s) { return v ^ (v >> s); }
static uint64_t avalanche3(uint64_t h) { h = xs(h, 37); h *= MX1; return xs(h, 32); }

static uint64_t mix16(const uint8_t *in, const uint8_t *sec)
{
    return fold64(rd64(in) ^ rd64(sec), rd64(in + 8) ^ rd64(sec + 8));
}

uint64_t kco_xxh3_64(const void *vp, uint64_t len)
{
    const uint8_t *p = vp;
    if (len >= 4 && len <= 8) {
        uint32_t in1 = rd32(p), in2 = rd32(p + len - 4);
        uint64_t bitflip = rd64(kSecret + 8) ^ rd64(kSecret + 16);
        uint64_t h = ((uint64_t)in2 + ((uint64_t)in1 << 32)) ^ bitflip;
        h ^= rotl64(h, 49) ^ rotl64(h, 24);
        h *= MX2;
        h ^= (h >> 35) + len;
        h *= MX2;
        return xs(h, 28);
    }
    if (len >= 9 && len <= 16) {
        uint64_t b1 = rd64(kSecret + 24) ^ rd64(kSecret + 32);
        uint64_t b2 = rd64(kSecret + 40) ^ rd64(kSecret + 48);
        uint64_t lo = rd64(p) ^ b1, hi = rd64(p + len - 8) ^ b2;
        uint64_t acc = len + bswap64(lo) + hi + fold64(lo, hi);
        return avalanche3(acc);
    }
    if (len >= 17 && len <= 32) {
        uint64_t acc = len * P64_1;
        acc += mix16(p, kSecret);
        acc += mix16(p + len - 16, kSecret + 16);
        return avalanche3(acc);
    }
    return 0; /* other lengths are not on the path */
}

void kco_xxh3_128(const void *vp, uint64_t len, uint64_t *olo, uint64_t *ohi)
{
    const uint8_t *p = vp;
    if (len >= 4 && len <= 8) {
        uint32_t ilo = rd32(p), ihi = rd32(p + len - 4);
        uint64_t in64 = ilo + ((uint64_t)ihi << 32);
        uint64_t bitflip = rd64(kSecret + 16) ^ rd64(kSecret + 24);
        uint64_t keyed = in64 ^ bitflip;
        u128 m = (u128)keyed * (P64_1 + (len << 2));
        uint64_t lo = (uint64_t)m, hi = (uint64_t)(m >> 64);
        hi += lo << 1;
        lo ^= hi >> 3;
        lo = xs(lo, 35);
        lo *= MX2;
        lo = xs(lo, 28);
        hi = avalanche3(hi);
        *olo = lo; *ohi = hi;
        return;
    }
    if (len >= 9 && len <= 16) {
        uint64_t bl = rd64(kSecret + 32) ^ rd64(kSecret + 40);
        uint64_t bh = rd64(kSecret + 48) ^ rd64(kSecret + 56);
        uint64_t ilo = rd64(p), ihi = rd64(p + len - 8);
        u128 m = (u128)(ilo ^ ihi ^ bl) * P64_1;
        uint64_t mlo = (uint64_t)m, mhi = (uint64_t)(m >> 64);
        mlo += (uint64_t)(len - 1) << 54;
        ihi ^= bh;
        mhi += ihi + (uint64_t)(uint32_t)ihi * (uint64_t)(P32_2 - 1);
        mlo ^= bswap64(mhi);
        u128 h = (u128)mlo * P64_2;
        uint64_t hlo = (uint64_t)h, hhi = (uint64_t)(h >> 64);
        hhi += mhi * P64_2;
        *olo = avalanche3(hlo);
        *ohi = avalanche3(hhi);
        return;
    }
    if (len >= 17 && len <= 32) {
        uint64_t alo = len * P64_1, ahi = 0;
        const uint8_t *i1 = p, *i2 = p + len - 16;
        alo += mix16(i1, kSecret);
        alo ^= rd64(i2) + rd64(i2 + 8);
        ahi += mix16(i2, kSecret + 16);
        ahi ^= rd64(i1) + rd64(i1 + 8);
        uint64_t hlo = alo + ahi;
        uint64_t hhi = alo * P64_1 + ahi * P64_4 + len * P64_2;
        *olo = avalanche3(hlo);
        *ohi = 0 - avalanche3(hhi);
        return;
    }
    *olo = *ohi = 0;
}

static const uint64_t rsec[8] = {
    0x2d358dccaa6c78a5ULL, 0x8bb84b93962eacc9ULL, 0x4b33a62ed433d4a3ULL, 0x4d5a2da51de1aa47ULL,
    0xa0761d6478bd642fULL, 0xe7037ed1a0b428dbULL, 0x90ed1765281c388cULL, 0xaaaaaaaaaaaaaaaaULL};

static void mum(uint64_t *a, uint64_t *b) { u128 r = (u128)*a * *b; *a = (uint64_t)r; *b = (uint64_t)(r >> 64); }
static uint64_t rmix(uint64_t a, uint64_t b) { mum(&a, &b); return a ^ b; }

uint64_t kco_rapidhash(const void *vp, uint64_t len)
{
    const uint8_t *p = vp;
    uint64_t seed = 0;
    seed ^= rmix(seed ^ rsec[2], rsec[1]);
    uint64_t a, b, i = len;
    if (len <= 16) {
        if (len >= 8) { seed ^= len; a = rd64(p); b = rd64(p + len - 8); }
        else if (len >= 4) { seed ^= len; a = rd32(p); b = rd32(p + len - 4); }
        else return 0; /* not on the path */
    } else if (len <= 32) {
        seed = rmix(rd64(p) ^ rsec[2], rd64(p + 8) ^ seed);
        if (i > 32) return 0;
        a = rd64(p + i - 16) ^ i;
        b = rd64(p + i - 8);
    } else {
        return 0;
    }
    a ^= rsec[1];
    b ^= seed;
    mum(&a, &b);
    return rmix(a ^ rsec[7], b ^ rsec[1] ^ i);
}

/* komihash's multiply adds the high half into the second seed. */
static void kh_m128(uint64_t u, uint64_t v, uint64_t *rl, uint64_t *rha)
{
    u128 r = (u128)u * v;
    *rl = (uint64_t)r;
    *rha += (uint64_t)(r >> 64);
}

uint64_t kco_komihash(const void *vp, uint64_t len, uint64_t seed)
{
    const uint8_t *m = vp;
    uint64_t s1 = 0x243F6A8885A308D3ULL ^ (seed & 0x5555555555555555ULL);
    uint64_t s5 = 0x452821E638D01377ULL ^ (seed & 0xAAAAAAAAAAAAAAAAULL);
    uint64_t r1, r2;
    kh_m128(s1, s5, &s1, &s5); s1 ^= s5;
    if (len == 8) {
        r1 = s1 ^ rd64(m);
        r2 = s5 ^ 1; /* the 1<<24 marker shifted by (64^88) */
    } else if (len == 16) {
        kh_m128(rd64(m) ^ s1, rd64(m + 8) ^ s5, &s1, &s5); s1 ^= s5;
        r1 = 1 ^ s1; /* final partial word is empty: only the marker bit */
        r2 = s5;
    } else if (len == 32) {
        kh_m128(rd64(m) ^ s1, rd64(m + 8) ^ s5, &s1, &s5); s1 ^= s5;
        kh_m128(rd64(m + 16) ^ s1, rd64(m + 24) ^ s5, &s1, &s5); s1 ^= s5;
        r1 = 1 ^ s1;
        r2 = s5;
    } else {
        return 0; /* other lengths are not on the path */
    }
    kh_m128(r1, r2, &s1, &s5); s1 ^= s5;
    kh_m128(s1, s5, &s1, &s5); s1 ^= s5;
    return s1;
}

uint64_t kco_mix_hash(uint64_t h1, uint64_t h2)
{
    const uint64_t F = 0x9ddfea08eb382d69ULL;
    uint64_t a = (h1 ^ h2) * F;
    a ^= a >> 47;
    uint64_t b = (h2 ^ a) * F;
    b ^= b >> 47;
    b *= F;
    return b;
}

static uint64_t mix64(uint64_t x)
{
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

uint64_t kco_hash_func(const uint64_t *w, uint32_t nw)
{
    uint64_t h = 0x9e3779b97f4a7c15ULL * nw + 0x85ebca77c2b2ae63ULL;
    for (uint32_t i = 0; i < nw; ++i) {
        h ^= mix64(w[i]);
        h = rotl64(h, 27);
        h = h * 0x9e3779b97f4a7c15ULL + 0x165667b19e3779f9ULL;
    }
    h ^= h >> 33; h *= 0xff51afd7ed558ccdULL;
    h ^= h >> 33; h *= 0xc4ceb9fe1a85ec53ULL;
    h ^= h >> 33;
    return h;
}

/* ConcurrentMap.h:141-148 */
uint64_t kco_concurrent_map_hash(const uint64_t *w, uint32_t nw)
{
    return kco_mix_hash(kco_komihash(w, 8ULL * nw, 0), kco_rapidhash(w, 8ULL * nw));
}

/* ------------------------------------------------------------------ */
/* Bloom probe: BloomFilter.h:215-222 (block), :287-297 (3-bit mask). */
void kco_bloom_probe(const uint64_t *key, uint32_t nw, uint64_t capacity,
                     uint64_t *block_out, uint64_t *mask_out)
{
    uint64_t lo, hi;
    kco_xxh3_128(key, 8ULL * nw, &lo, &hi);
    uint64_t h2 = hi | 1ULL, mask = 0;
    for (uint64_t i = 0; i < 3; ++i) mask |= 1ULL << ((lo + i * h2) & 63);
    *block_out = lo & (capacity - 1);
    *mask_out = mask;
}
