/*
 * kc_server.h — serving loop of libkc.so: count jobs and point-query batches
 * against device-resident tables, admitted by free HBM, executed
 * concurrently (one HIP stream per table).  No counterpart in the
 * reference, whose only point lookup is FlatConcurrentHashMap::find_prepared
 * inside its tools (LowFrequencyQueryThreadPool.h:196-244).
 *
 * Host buffers passed to submit calls must stay valid until kc_server_wait
 * returns for that request.  Table ids are chosen by the caller and must be
 * unique among resident tables.  A query waits for the count jobs of its
 * table that are queued or running; a query of a table that is not resident
 * (never counted, dropped, or evicted to admit a job) fails with KC_EINVAL.
 * Wait for every request before kc_server_destroy.
 */
#ifndef KC_SERVER_H
#define KC_SERVER_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kc_server kc_server;

typedef struct kc_request_result {
    int status;          /* KC_OK or a KC_E* code */
    int batched;         /* query batches answered by the same lookup launch (1: alone) */
    double latency_ms;   /* submit -> completion */
    double queue_ms;     /* submit -> start of execution */
    uint64_t kmers;      /* count jobs: windows counted */
    uint64_t distinct;   /* count jobs: keys in the table */
} kc_request_result;

int kc_server_create(int device, uint32_t workers, uint64_t hbm_budget_bytes, kc_server **out);
void kc_server_destroy(kc_server *s);
/* Count job: framed sequence bytes (host; page-locked memory is copied from
 * directly) into a new resident table sized for max_distinct keys (0: one
 * per input byte).  KC_ENOMEM at once if the job cannot fit the budget. */
int kc_server_submit_count(kc_server *s, uint64_t table_id, uint32_t k, const uint8_t *h_seq,
                           uint64_t nbytes, uint64_t max_distinct, int bloom, uint64_t *req_id);
/* Query batch: merged counts of n keys (n*nw host words) into h_counts. */
int kc_server_submit_query(kc_server *s, uint64_t table_id, const uint64_t *h_keys, uint64_t n,
                           uint32_t *h_counts, uint64_t *req_id);
int kc_server_wait(kc_server *s, uint64_t req_id, kc_request_result *out);
int kc_server_drop_table(kc_server *s, uint64_t table_id);
int kc_server_table_histogram(kc_server *s, uint64_t table_id, uint32_t min_freq,
                              uint32_t max_freq, int64_t *h_bins);
int kc_server_resident(kc_server *s, uint64_t *tables, uint64_t *bytes);

#ifdef __cplusplus
}
#endif
#endif
