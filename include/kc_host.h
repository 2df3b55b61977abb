/*
 * kc_host.h — host-side ingest and synthetic-input helpers of libkc.so.
 *
 * FASTQ framing keeps line 2 of every 4-line record followed by '\n', the
 * byte stream the reference's reader hands to its parsers
 * (FastqReader.h:478-487, 657-687); gzip input is detected by the ".gz"
 * suffix as in FastqReader.h:423,575 (any other file is read as it is, also
 * when it holds gzip bytes) and inflated with zlib on the host.
 *
 * The synthetic generator has no counterpart in the reference; it produces
 * the seeded workloads the engine is measured and checked on.
 */
#ifndef KC_HOST_H
#define KC_HOST_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Frames an in-memory FASTQ buffer; out must hold len bytes. Returns bytes. */
uint64_t kc_fastq_frame(const uint8_t *fq, uint64_t len, uint8_t *out);
/* Reads and frames one FASTQ file (plain or .gz).  *out is malloc'ed and
 * released with kc_host_free.  Returns KC_OK or KC_EIO. */
int kc_fastq_frame_file(const char *path, uint8_t **out, uint64_t *out_len);
void kc_host_free(void *p);

/* Streaming ingest: counts the reads of n_paths FASTQ files (plain or .gz)
 * into table t.  `threads` host threads take files in turn, inflate each with
 * zlib and frame it (FastqReader.h:478-487, 575-616), and hand whole reads in
 * chunks of about 4 MB to a kc_batch built with `tune` (NULL: the table's
 * tune): the chunks go straight into the batch's page-locked staging slots
 * and each full slot is counted by one graph launch while inflating goes on.
 * Several files inflate in parallel (the reference reads a directory or a
 * comma-separated file list, main.cpp:501-530), and so do the parts of one
 * BGZF file (blocked gzip, as htslib's bgzip writes it: gzip members whose
 * "BC" extra subfield gives their size): threads inflate runs of about 1 MB
 * of members, each run is framed from the record phase the newline counts of
 * the runs before it give, and the runs are handed over in file order.  A
 * single gzip member of 64 MB or more is cut into shares of its deflate
 * stream that threads decode from guessed block starts, checked and joined in
 * order (the CRC-32 and size checked against the trailer).  The reference's
 * gzread inflates either file as one stream on one thread; the counts are the
 * same.  Returns when every read is counted.  stats may be
 * NULL. */
typedef struct kc_ingest_stats {
    uint64_t files;
    uint64_t threads;
    uint64_t input_bytes;     /* decompressed FASTQ bytes */
    uint64_t framed_bytes;    /* "seq\n" bytes counted */
    uint64_t chunks;
    uint64_t launches;        /* graph launches of the batch */
    double inflate_seconds;   /* summed over threads */
    double wall_seconds;
} kc_ingest_stats;
int kc_count_fastq_files(kc_table *t, const char *const *paths, uint32_t n_paths, uint32_t threads,
                         const kc_tune *tune, kc_ingest_stats *stats);
/* The same reader without a table: frames the files into one malloc'ed
 * buffer (released with kc_host_free), each file's reads in order, the files'
 * chunks in the order the threads finish them, a '\n' after a file's last
 * read if it lacks one.  Host only; no reference counterpart beyond
 * FastqReader's framing. */
int kc_fastq_frame_files(const char *const *paths, uint32_t n_paths, uint32_t threads, uint8_t **out,
                         uint64_t *out_len, kc_ingest_stats *stats);

/* Out-of-HBM counting: the files are counted in n_passes passes, pass p
 * keeping the k-mers of hash slice p (kc_table_set_pass) in the cleared
 * table t, which then needs room for one slice's distinct k-mers.  After
 * each pass its export is appended to export_dir (pass 0 starts it), its
 * dump (dump_tool precise over [dump_min, dump_max], sorted when
 * dump_sorted) written to a part file, and its histogram over
 * [hist_min, hist_max] added to hist_bins; the parts are merged at the end
 * (sorted parts by key, kc_merge_sorted_dumps; otherwise concatenated).
 * The reference's fallback partitions its exported files by key prefix and
 * merges the partition outputs the same way (dump_tool.cpp:1043-1134,
 * Partition.h:135-346).  Any of export_dir / dump_path / hist_bins may be
 * NULL.  The table keeps the last pass; its pass setting is reset to one
 * pass.  stats sum over the passes (every pass reads the files again). */
typedef struct kc_pass_plan {
    uint32_t n_passes;       /* power of two, 1..256 */
    const char *export_dir;  /* infos.bin, high.bin, low.bin */
    const char *dump_path;
    uint32_t dump_min, dump_max;
    int dump_sorted;
    uint32_t hist_min, hist_max;
    int64_t *hist_bins;      /* hist_max - hist_min + 1 host bins */
} kc_pass_plan;
int kc_count_fastq_files_passes(kc_table *t, const char *const *paths, uint32_t n_paths, uint32_t threads,
                                const kc_tune *tune, const kc_pass_plan *plan, kc_ingest_stats *stats);
/* k-way merge of dump texts sorted by k-mer (keys unique across the inputs)
 * into out_path (merge_sorted_outputs, dump_tool.cpp:1043-1134).  Host only. */
int kc_merge_sorted_dumps(const char *const *paths, uint32_t n_paths, const char *out_path);

typedef struct kc_genome_spec {
    uint64_t seed;
    uint64_t length;            /* bases */
    uint32_t n_segments;        /* repeat families (0 = uniform genome) */
    uint32_t segment_len;
    double repeat_fraction;     /* share of the genome written as copies */
    double segment_divergence;  /* per-base substitution rate of a copy */
} kc_genome_spec;

typedef struct kc_read_spec {
    uint64_t seed;
    uint64_t n_reads;
    uint32_t read_len;          /* short-read length */
    uint32_t both_strands;      /* 1: reverse-complement half the reads */
    double sub_rate;            /* short-read substitution rate */
    double n_rate;              /* per-base N rate */
    double long_fraction;       /* share of reads (by count) that are long */
    double long_mu, long_sigma; /* lognormal parameters of long-read length */
    double long_sub_rate;
    uint32_t min_len;           /* lower clamp of long-read length */
    uint32_t threads;           /* worker threads (0 = hardware) */
} kc_read_spec;

/* Writes `length` ACGT bytes. */
int kc_synth_genome(const kc_genome_spec *spec, uint8_t *out);
/* Bytes kc_synth_reads will write (fastq = 1: FASTQ records; 0: framed
 * "seq\n" stream).  Negative on error. */
int64_t kc_synth_reads_size(const uint8_t *genome, uint64_t glen, const kc_read_spec *spec, int fastq);
int64_t kc_synth_reads(const uint8_t *genome, uint64_t glen, const kc_read_spec *spec, int fastq,
                       uint8_t *out, uint64_t cap);

#ifdef __cplusplus
}
#endif
#endif
