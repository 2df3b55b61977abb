// kc.hpp — C++ adaptors with the reference's names over the C ABI
// (include/kc_api.h).  Header only; link with libkc.so.  No HIP headers are
// needed: device memory is handled through the ABI.
//
//   kc::kmer<N>, kc::ExportRecord<N>     layout of kmer.h:11-119 and
//                                        definition.h ExportRecord
//   kc::ConcurrentCountingHashMap<N>     ConcurrentCountingHashMap.h:211-312
//   kc::export_record_radix_sort<N>      ExportRecordRadixSort.h:256-269
//   kc::radix_sort_export_records_12bit  ExportRecordRadixSort12.h:639-646
//   kc::ExportWriter<N>                  ExportWriter.h:70-313 (low.bin)
//   kc::ReadBatch                        batched counting of uneven read
//                                        chunks (kc_batch; no reference
//                                        counterpart: FastqReader.h:542-557
//                                        feeds fixed 32 KiB blocks)
//
// Errors surface as kc::Error (an std::runtime_error carrying the status).
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../kc_api.h"

namespace kc {

class Error : public std::runtime_error {
public:
    Error(int status, const char *what)
        : std::runtime_error(std::string(what) + ": " + kc_strerror(status)), status(status) {}
    int status;
};

inline void check(int status, const char *what)
{
    if (status != KC_OK) throw Error(status, what);
}

template <uint32_t N>
struct kmer {
    std::array<uint64_t, N> data;
};

template <uint32_t N>
struct ExportRecord {
    kmer<N> key;
    uint32_t count;
};

static_assert(sizeof(ExportRecord<1>) == 16 && sizeof(ExportRecord<2>) == 24, "ExportRecord layout");

class Context {
public:
    explicit Context(int device = 0) { check(kc_ctx_create(device, &h_), "kc_ctx_create"); }
    ~Context() { kc_ctx_destroy(h_); }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    kc_ctx *get() const { return h_; }
    void sync() const { check(kc_ctx_sync(h_), "kc_ctx_sync"); }

private:
    kc_ctx *h_ = nullptr;
};

// HBM buffer owned by a context.
class DeviceBuffer {
public:
    DeviceBuffer(Context &c, uint64_t bytes) : c_(c), n_(bytes) { check(kc_dev_alloc(c.get(), bytes, &p_), "kc_dev_alloc"); }
    ~DeviceBuffer() { kc_dev_free(c_.get(), p_); }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    void *get() const { return p_; }
    uint64_t size() const { return n_; }
    void upload(const void *h, uint64_t bytes) { check(kc_copy_h2d(c_.get(), p_, h, bytes), "kc_copy_h2d"); }
    void download(void *h, uint64_t bytes) const { check(kc_copy_d2h(c_.get(), h, p_, bytes), "kc_copy_d2h"); }

private:
    Context &c_;
    void *p_ = nullptr;
    uint64_t n_;
};

// Device-resident counting table with the reference's increment interface.
// increment() of single keys works but costs a round trip; batch calls are
// the intended use.
template <uint32_t N>
class ConcurrentCountingHashMap {
    static_assert(N == 1 || N == 2, "k up to 63");

public:
    ConcurrentCountingHashMap(Context &c, uint32_t k, uint64_t capacity, uint32_t count_max = 0xFFFFFFFFu,
                              bool bloom = false)
        : c_(c), k_(k)
    {
        kc_table_config cfg{};
        cfg.k = k;
        cfg.count_max = count_max;
        cfg.filter_min = 0;
        cfg.filter_max = 0xFFFFFFFFu;
        cfg.capacity = capacity;
        cfg.max_distinct = capacity / 2;
        cfg.bloom = bloom ? 1 : 0;
        check(kc_table_create(c.get(), &cfg, &t_), "kc_table_create");
    }
    ~ConcurrentCountingHashMap() { kc_table_destroy(t_); }
    ConcurrentCountingHashMap(const ConcurrentCountingHashMap &) = delete;
    ConcurrentCountingHashMap &operator=(const ConcurrentCountingHashMap &) = delete;

    void increment(const kmer<N> &key, uint32_t increment_value = 1) { increment(&key, 1, &increment_value); }

    void increment(const kmer<N> *keys, uint64_t n, const uint32_t *inc = nullptr)
    {
        if (!n) return;
        DeviceBuffer dk(c_, n * sizeof(kmer<N>));
        dk.upload(keys, n * sizeof(kmer<N>));
        if (inc) {
            DeviceBuffer di(c_, n * 4);
            di.upload(inc, n * 4);
            check(kc_table_increment(t_, static_cast<const uint64_t *>(dk.get()), n,
                                     static_cast<const uint32_t *>(di.get())), "kc_table_increment");
            c_.sync();
        } else {
            check(kc_table_increment(t_, static_cast<const uint64_t *>(dk.get()), n, nullptr), "kc_table_increment");
            c_.sync();
        }
    }

    // Framed sequence bytes (sequence lines followed by '\n') from host memory.
    void count_reads(const char *framed, uint64_t bytes)
    {
        check(kc_table_count_seq_host(t_, reinterpret_cast<const uint8_t *>(framed), bytes), "kc_table_count_seq_host");
        c_.sync();
    }

    std::vector<uint32_t> find(const kmer<N> *keys, uint64_t n)
    {
        std::vector<uint32_t> out(n);
        if (!n) return out;
        DeviceBuffer dk(c_, n * sizeof(kmer<N>)), dc(c_, n * 4);
        dk.upload(keys, n * sizeof(kmer<N>));
        check(kc_table_query(t_, static_cast<const uint64_t *>(dk.get()), n, static_cast<uint32_t *>(dc.get())),
              "kc_table_query");
        dc.download(out.data(), n * 4);
        return out;
    }

    kc_stats stats() const
    {
        kc_stats s;
        check(kc_table_stats(t_, &s), "kc_table_stats");
        return s;
    }

    // infos.bin / high.bin / low.bin into dir (the reference's temp_dir contract)
    void export_dir(const std::string &dir) { check(kc_table_export_dir(t_, dir.c_str()), "kc_table_export_dir"); }
    // Appends this table's records to an existing export (one pass of a
    // multi-pass count; the partition files of Partition.h:135-346).
    void export_dir_append(const std::string &dir)
    {
        check(kc_table_export_dir_append(t_, dir.c_str()), "kc_table_export_dir_append");
    }
    // Later counting calls count only hash slice `pass` of `n_passes`.
    void set_pass(uint32_t pass, uint32_t n_passes) { check(kc_table_set_pass(t_, pass, n_passes), "kc_table_set_pass"); }

    std::vector<int64_t> histogram(uint32_t min_freq = 1, uint32_t max_freq = 10000)
    {
        std::vector<int64_t> bins(max_freq - min_freq + 1);
        check(kc_table_histogram(t_, min_freq, max_freq, bins.data()), "kc_table_histogram");
        return bins;
    }

    // Framed bytes already in HBM (asynchronous on the context's stream).
    void count_device(const DeviceBuffer &seq, uint64_t bytes)
    {
        check(kc_table_count_seq(t_, static_cast<const uint8_t *>(seq.get()), bytes), "kc_table_count_seq");
    }

    // Launch-shape search on a device-resident sample; leaves the table empty.
    kc_tune autotune(const DeviceBuffer &sample, uint64_t bytes, uint32_t reps = 1)
    {
        kc_tune t{};
        float ms = 0;
        check(kc_autotune(t_, static_cast<const uint8_t *>(sample.get()), bytes, reps, &t, &ms), "kc_autotune");
        return t;
    }

    void set_tune(const kc_tune &t) { check(kc_table_set_tune(t_, &t), "kc_table_set_tune"); }
    void clear() { check(kc_table_clear(t_), "kc_table_clear"); }

    kc_table *get() const { return t_; }

private:
    Context &c_;
    uint32_t k_;
    kc_table *t_ = nullptr;
};

// Many read chunks of uneven length counted into a table, one captured-graph
// launch per staging slot (kc_batch).  Destroy before the table.
class ReadBatch {
public:
    template <uint32_t N>
    explicit ReadBatch(ConcurrentCountingHashMap<N> &table, const kc_tune *tune = nullptr)
    {
        check(kc_batch_create(table.get(), tune, &b_), "kc_batch_create");
    }
    ~ReadBatch() { kc_batch_destroy(b_); }
    ReadBatch(const ReadBatch &) = delete;
    ReadBatch &operator=(const ReadBatch &) = delete;
    void add(const char *chunk, uint64_t bytes)
    {
        check(kc_batch_add(b_, reinterpret_cast<const uint8_t *>(chunk), bytes), "kc_batch_add");
    }
    // chunk i = base[ends[i-1], ends[i]) (ends[-1] = 0)
    void add_chunks(const char *base, const uint64_t *ends, uint64_t n)
    {
        check(kc_batch_add_chunks(b_, reinterpret_cast<const uint8_t *>(base), ends, n), "kc_batch_add_chunks");
    }
    void flush() { check(kc_batch_flush(b_), "kc_batch_flush"); }
    void sync() { check(kc_batch_sync(b_), "kc_batch_sync"); }
    kc_batch_info info() const
    {
        kc_batch_info i;
        check(kc_batch_get_info(b_, &i), "kc_batch_get_info");
        return i;
    }

private:
    kc_batch *b_ = nullptr;
};

namespace detail {
template <uint32_t N>
void device_sort(Context &c, ExportRecord<N> *recs, ExportRecord<N> *out, uint64_t len, uint32_t k, uint32_t cpb,
                 uint32_t bits)
{
    const uint64_t bytes = len * sizeof(ExportRecord<N>);
    DeviceBuffer a(c, bytes + 16), b(c, bytes + 16);
    a.upload(recs, bytes);
    int in_scratch = 0;
    check(kc_sort_records(c.get(), a.get(), b.get(), len, N, k, cpb, bits, sizeof(ExportRecord<N>), &in_scratch),
          "kc_sort_records");
    (in_scratch ? b : a).download(out, bytes);
}
}  // namespace detail

// ExportRecordRadixSort.h:256-269: sorts records (using temp_records as
// scratch) and returns the buffer that holds the result.  The device sort
// writes the result back into `records`, which is returned.
template <uint32_t N>
ExportRecord<N> *export_record_radix_sort(Context &c, ExportRecord<N> *records, ExportRecord<N> *temp_records,
                                          uint64_t length, uint32_t k_len)
{
    (void)temp_records;
    if (length > 1) detail::device_sort<N>(c, records, records, length, k_len, 10, 8);
    return records;
}

// ExportRecordRadixSort12.h:639-646: the result ends up in `records`;
// invalid arguments throw std::invalid_argument, too many records
// std::length_error, as the reference does.
template <uint32_t N, uint32_t CommonPrefixBases = 10>
void radix_sort_export_records_12bit(Context &c, std::vector<ExportRecord<N>> &records,
                                     std::vector<ExportRecord<N>> &scratch, uint32_t k_length)
{
    if (k_length < CommonPrefixBases || k_length > 32u * N)
        throw std::invalid_argument("k_length is outside the representable/common-prefix range");
    if (&records == &scratch) throw std::invalid_argument("records and scratch must be different vectors");
    if (records.size() > 0xFFFFFFFFull) throw std::length_error("12-bit radix sort supports at most UINT32_MAX records");
    scratch.resize(records.size());
    if (records.size() > 1)
        detail::device_sort<N>(c, records.data(), records.data(), records.size(), k_length, CommonPrefixBases, 12);
}

}  // namespace kc
