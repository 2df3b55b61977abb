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
//   kc::FinalDrainWriter<N>              FinalDrainWriter.h:43-146 (high.bin)
//   kc::write_infos                      main.cpp:590-597 (infos.bin)
//   kc::ReadBatch                        batched counting of uneven read
//                                        chunks (kc_batch; no reference
//                                        counterpart: FastqReader.h:542-557
//                                        feeds fixed 32 KiB blocks)
//
// Errors surface as kc::Error (an std::runtime_error carrying the status).
#pragma once

#include <array>
#include <cstdint>
#include <cstdio>
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

// infos.bin of a temp_dir: {u32 k, u32 count_max} (main.cpp:590-597).
inline void write_infos(const std::string &temp_dir, uint32_t k, uint32_t count_max)
{
    const uint32_t v[2] = {k, count_max};
    FILE *f = std::fopen((temp_dir + "infos.bin").c_str(), "wb");
    if (!f) throw Error(KC_EIO, "write_infos");
    const bool ok = std::fwrite(v, 4, 2, f) == 2;
    if (std::fclose(f) != 0 || !ok) throw Error(KC_EIO, "write_infos");
}

namespace detail {
// An output file of a temp_dir, created (truncated) when opened.
class OutFile {
public:
    explicit OutFile(const std::string &path) : f_(std::fopen(path.c_str(), "wb"))
    {
        if (!f_) throw Error(KC_EIO, path.c_str());
    }
    ~OutFile() { close(); }
    OutFile(const OutFile &) = delete;
    OutFile &operator=(const OutFile &) = delete;
    void write(const std::vector<uint8_t> &b)
    {
        if (!b.empty() && std::fwrite(b.data(), 1, b.size(), f_) != b.size()) throw Error(KC_EIO, "fwrite");
    }
    void close()
    {
        if (f_ && std::fclose(f_) != 0) {
            f_ = nullptr;
            throw Error(KC_EIO, "fclose");
        }
        f_ = nullptr;
    }

private:
    FILE *f_;
};
}  // namespace detail

// ExportWriter.h:70-313: writes temp_dir + "low.bin".  The constructor
// creates the file (truncated), as the reference's does; write() takes the
// k-mers of one export block, which are packed on the device
// (kc_export_low, the byte_pack_to_buffer layout) in batches of up to
// `batch` keys and appended; join() writes what is buffered and closes the
// file.  start() is kept for the reference's call sequence (no writer thread
// is needed).  Errors throw kc::Error where the reference exits.
template <uint32_t N>
class ExportWriter {
    static_assert(N == 1 || N == 2, "k up to 63");

public:
    ExportWriter(Context &c, uint32_t k, const std::string &temp_dir, uint64_t batch = 1u << 20)
        : c_(c), k_(k), batch_(batch ? batch : 1), file_(temp_dir + "low.bin")
    {
        if (k < 1 || k > 32 * N) throw Error(KC_EINVAL, "ExportWriter k");
    }
    void start() {}
    void write(const kmer<N> *kmers, uint64_t n)
    {
        pending_.insert(pending_.end(), kmers, kmers + n);
        if (pending_.size() >= batch_) flush();
    }
    void join()
    {
        flush();
        file_.close();
    }
    uint64_t written() const { return written_; }

private:
    void flush()
    {
        const uint64_t n = pending_.size();
        if (!n) return;
        const uint64_t kb = (k_ / 32) * 8 + (2 * (k_ % 32) + 7) / 8;
        DeviceBuffer dk(c_, n * sizeof(kmer<N>)), out(c_, n * kb + 16);
        dk.upload(pending_.data(), n * sizeof(kmer<N>));
        check(kc_export_low(c_.get(), static_cast<const uint64_t *>(dk.get()), n, k_,
                            static_cast<uint8_t *>(out.get())), "kc_export_low");
        std::vector<uint8_t> bytes(n * kb);
        out.download(bytes.data(), bytes.size());
        file_.write(bytes);
        written_ += n;
        pending_.clear();
    }

    Context &c_;
    uint32_t k_;
    uint64_t batch_;
    detail::OutFile file_;
    std::vector<kmer<N>> pending_;
    uint64_t written_ = 0;
};

// FinalDrainWriter.h:43-146 (with the high.bin file of
// FinalDrainWriterThread.h): write_kmer_record() packs a (k-mer, count)
// record as given -- its caller filters, as NewKmerTree.h:1303-1311 does --
// and write_map_record() leaves out records whose count fails the counter's
// filter (count + 1 < filter_min in u32, or count > filter_max;
// FinalDrainWriter.h:95-104).  Counts are written in count_bytes(count_max)
// bytes (main.cpp:599-615).  The reference takes count_max, filter_min and
// filter_max from its globals (definition.h:93-96); here they are arguments.
// Records are packed on the device (kc_export_high) in batches; close()
// writes what is buffered and closes temp_dir + "high.bin".
template <uint32_t N>
class FinalDrainWriter {
    static_assert(N == 1 || N == 2, "k up to 63");

public:
    FinalDrainWriter(Context &c, uint32_t k, uint32_t count_max, const std::string &temp_dir,
                     uint32_t filter_min = 2, uint32_t filter_max = 0xFFFFFFFFu, uint64_t batch = 1u << 20)
        : c_(c), k_(k), count_max_(count_max), fmin_(filter_min), fmax_(filter_max), batch_(batch ? batch : 1),
          file_(temp_dir + "high.bin")
    {
        if (k < 1 || k > 32 * N || count_max == 0) throw Error(KC_EINVAL, "FinalDrainWriter");
    }
    ~FinalDrainWriter()
    {
        try {
            close();
        } catch (...) {
        }
    }
    void write_kmer_record(const uint64_t *kmer_data, uint32_t count)
    {
        kmer<N> key;
        for (uint32_t w = 0; w < N; ++w) key.data[w] = kmer_data[w];
        plain_.push_back({key, count});
        if (plain_.size() >= batch_) flush(plain_, 0, 0xFFFFFFFFu);
    }
    void write_map_record(const ExportRecord<N> *records, uint32_t count)
    {
        filtered_.insert(filtered_.end(), records, records + count);
        if (filtered_.size() >= batch_) flush(filtered_, fmin_, fmax_);
    }
    void close()
    {
        if (closed_) return;
        flush(plain_, 0, 0xFFFFFFFFu);
        flush(filtered_, fmin_, fmax_);
        closed_ = true;
        file_.close();
    }
    uint64_t written() const { return written_; }

private:
    void flush(std::vector<ExportRecord<N>> &recs, uint32_t fmin, uint32_t fmax)
    {
        const uint64_t n = recs.size();
        if (!n) return;
        std::vector<kmer<N>> keys(n);
        std::vector<uint32_t> counts(n);
        for (uint64_t i = 0; i < n; ++i) {
            keys[i] = recs[i].key;
            counts[i] = recs[i].count;
        }
        DeviceBuffer dk(c_, n * sizeof(kmer<N>)), dc(c_, n * 4);
        dk.upload(keys.data(), n * sizeof(kmer<N>));
        dc.upload(counts.data(), n * 4);
        uint64_t bytes = 0;
        const auto *kp = static_cast<const uint64_t *>(dk.get());
        const auto *cp = static_cast<const uint32_t *>(dc.get());
        check(kc_export_high(c_.get(), kp, cp, n, k_, count_max_, fmin, fmax, nullptr, 0, &bytes), "kc_export_high");
        DeviceBuffer out(c_, bytes + 16);
        check(kc_export_high(c_.get(), kp, cp, n, k_, count_max_, fmin, fmax, static_cast<uint8_t *>(out.get()),
                             bytes + 16, &bytes),
              "kc_export_high");
        std::vector<uint8_t> host(bytes);
        out.download(host.data(), bytes);
        file_.write(host);
        const uint64_t rec = (k_ / 32) * 8 + (2 * (k_ % 32) + 7) / 8 +
                             (count_max_ <= 0xFF ? 1 : count_max_ <= 0xFFFF ? 2 : count_max_ <= 0xFFFFFF ? 3 : 4);
        written_ += bytes / rec;
        recs.clear();
    }

    Context &c_;
    uint32_t k_, count_max_, fmin_, fmax_;
    uint64_t batch_;
    detail::OutFile file_;
    std::vector<ExportRecord<N>> plain_, filtered_;
    uint64_t written_ = 0;
    bool closed_ = false;
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
