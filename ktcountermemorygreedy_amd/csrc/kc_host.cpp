// Host ingest (FASTQ framing, zlib inflate, streaming multi-file ingest into
// batches) and the seeded synthetic workload generator.
#include <fcntl.h>
#include <sys/mman.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <string_view>
#include <thread>
#include <vector>

#include "../../include/kc_api.h"
#include "../../include/kc_host.h"
#include "kc_guard.h"
#include "kc_inflate.h"

namespace {

struct SplitMix {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    double u01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    uint64_t below(uint64_t m) { return (uint64_t)(((unsigned __int128)next() * m) >> 64); }
};

const char kBase[4] = {'A', 'C', 'G', 'T'};

uint64_t threshold(double p)
{
    if (p <= 0) return 0;
    if (p >= 1) return ~0ULL;
    return (uint64_t)std::ldexp(p, 64);
}

// Framing state machine over one buffer (FastqReader.h:478-487).
struct Framer {
    uint32_t state = 0;  // 0 header, 1 sequence, 2 plus, 3 quality
    uint64_t feed(const uint8_t *p, uint64_t n, uint8_t *out)
    {
        uint64_t w = 0, i = 0;
        while (i < n) {
            const uint8_t *nl = static_cast<const uint8_t *>(std::memchr(p + i, '\n', n - i));
            const uint64_t end = nl ? (uint64_t)(nl - p) : n;
            if (state == 1) {
                std::memcpy(out + w, p + i, end - i);
                w += end - i;
                if (nl) out[w++] = '\n';
            }
            if (!nl) break;
            i = end + 1;
            state = (state + 1) & 3;
        }
        return w;
    }
};

bool has_gz_suffix(const char *path)
{
    const size_t n = std::strlen(path);
    return n >= 3 && std::strcmp(path + n - 3, ".gz") == 0;
}

// An input file as the reference's reader opens it (FastqReader.h:423,575):
// a name ending in ".gz" is inflated with zlib (gzread passes plain bytes
// through as well), any other file is read as it is -- gzip bytes in a file
// without the suffix are framed as text, as there.
struct InFile {
    gzFile g = nullptr;
    FILE *f = nullptr;
    InFile() = default;
    InFile(const InFile &) = delete;
    InFile &operator=(const InFile &) = delete;
    bool open(const char *path)
    {
        if (has_gz_suffix(path)) {
            g = gzopen(path, "rb");
            if (g) gzbuffer(g, 1u << 20);
            return g != nullptr;
        }
        f = std::fopen(path, "rb");
        return f != nullptr;
    }
    // bytes read (0 at the end), or -1 on an error
    int read(uint8_t *p, unsigned n)
    {
        if (g) return gzread(g, p, n);
        const size_t r = std::fread(p, 1, n, f);
        return r == 0 && std::ferror(f) ? -1 : (int)r;
    }
    ~InFile()
    {
        if (g) gzclose(g);
        if (f) std::fclose(f);
    }
};

}  // namespace

extern "C" uint64_t kc_fastq_frame(const uint8_t *fq, uint64_t len, uint8_t *out)
{
    Framer f;
    return f.feed(fq, len, out);
}

extern "C" void kc_host_free(void *p) { std::free(p); }

extern "C" int kc_fastq_frame_file(const char *path, uint8_t **out, uint64_t *out_len)
{
    return kc_guard([&]() -> int {
        if (!path || !out || !out_len) return KC_EINVAL;
        std::vector<uint8_t> in(1u << 22);
        InFile g;
        if (!g.open(path)) return KC_EIO;
        uint64_t cap = 1ull << 24, used = 0;
        uint8_t *buf = static_cast<uint8_t *>(std::malloc(cap));
        if (!buf) return KC_ENOMEM;
        Framer fr;
        for (;;) {
            const int r = g.read(in.data(), (unsigned)in.size());
            if (r < 0) {
                std::free(buf);
                return KC_EIO;
            }
            if (r == 0) break;
            if (used + (uint64_t)r > cap) {
                while (used + (uint64_t)r > cap) cap *= 2;
                uint8_t *nb = static_cast<uint8_t *>(std::realloc(buf, cap));
                if (!nb) {
                    std::free(buf);
                    return KC_ENOMEM;
                }
                buf = nb;
            }
            used += fr.feed(in.data(), (uint64_t)r, buf + used);
        }
        *out = buf;
        *out_len = used;
        return KC_OK;
    });
}

extern "C" int kc_synth_genome(const kc_genome_spec *sp, uint8_t *out)
{
    return kc_guard([&]() -> int {
        if (!sp || !out || !sp->length) return KC_EINVAL;
        SplitMix rng{sp->seed};
        for (uint64_t i = 0; i < sp->length; ++i) out[i] = kBase[rng.next() & 3];
        if (!sp->n_segments || sp->repeat_fraction <= 0) return KC_OK;
        if (sp->segment_len == 0 || sp->segment_len > sp->length) return KC_EINVAL;
        // repeat families, then diverged copies until repeat_fraction is reached
        std::vector<uint8_t> seg((uint64_t)sp->n_segments * sp->segment_len);
        for (auto &b : seg) b = (uint8_t)(rng.next() & 3);
        const uint64_t target = (uint64_t)(sp->repeat_fraction * (double)sp->length);
        const uint64_t div = threshold(sp->segment_divergence);
        for (uint64_t copied = 0; copied < target; copied += sp->segment_len) {
            const uint64_t s = rng.below(sp->n_segments);
            const uint64_t pos = rng.below(sp->length - sp->segment_len + 1);
            for (uint32_t j = 0; j < sp->segment_len; ++j) {
                uint8_t b = seg[s * sp->segment_len + j];
                const uint64_t r = rng.next();
                if (r < div) b = (uint8_t)((b + 1 + (r & 0xFFFF) % 3) & 3);
                out[pos + j] = kBase[b];
            }
        }
        return KC_OK;
    });
}

namespace {

struct ReadPlan {
    uint64_t len;
    uint64_t start;
    bool rc;
    bool is_long;
    SplitMix rng;  // state after the plan draws, continues with bases
};

ReadPlan plan_read(const kc_read_spec *sp, uint64_t glen, uint64_t i)
{
    ReadPlan p;
    p.rng.s = sp->seed ^ ((i + 1) * 0xD1B54A32D192ED03ULL);
    p.rng.next();
    p.is_long = sp->long_fraction > 0 && p.rng.u01() < sp->long_fraction;
    if (p.is_long) {
        double u1 = p.rng.u01(), u2 = p.rng.u01();
        if (u1 < 1e-300) u1 = 1e-300;
        const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
        double l = std::exp(sp->long_mu + sp->long_sigma * z);
        l = std::max<double>(l, (double)std::max<uint32_t>(sp->min_len, 1));
        p.len = std::min<uint64_t>((uint64_t)std::llround(l), glen);
    } else {
        p.len = std::min<uint64_t>(sp->read_len, glen);
    }
    p.start = p.rng.below(glen - p.len + 1);
    p.rc = sp->both_strands ? (p.rng.next() & 1) : false;
    return p;
}

uint64_t digits(uint64_t v)
{
    uint64_t d = 1;
    while (v >= 10) {
        v /= 10;
        ++d;
    }
    return d;
}

uint64_t record_bytes(const ReadPlan &p, uint64_t i, bool fastq)
{
    if (!fastq) return p.len + 1;
    return 2 + digits(i) + 1 + p.len + 1 + 2 + p.len + 1;  // "@r<i>\n" seq "\n+\n" qual "\n"
}

void write_read(const uint8_t *genome, const kc_read_spec *sp, ReadPlan p, uint64_t i, bool fastq, uint8_t *o)
{
    const uint64_t sub = threshold(p.is_long ? sp->long_sub_rate : sp->sub_rate);
    const uint64_t nth = sub + threshold(sp->n_rate);
    if (fastq) {
        *o++ = '@';
        *o++ = 'r';
        char tmp[24];
        int t = 0;
        uint64_t v = i;
        do {
            tmp[t++] = (char)('0' + v % 10);
            v /= 10;
        } while (v);
        while (t) *o++ = (uint8_t)tmp[--t];
        *o++ = '\n';
    }
    for (uint64_t j = 0; j < p.len; ++j) {
        uint8_t c = p.rc ? genome[p.start + p.len - 1 - j] : genome[p.start + j];
        uint32_t b = c == 'A' ? 0 : c == 'C' ? 1 : c == 'G' ? 2 : 3;
        if (p.rc) b ^= 3;
        const uint64_t r = p.rng.next();
        uint8_t out;
        if (r < sub) out = (uint8_t)kBase[(b + 1 + (r & 0xFFFF) % 3) & 3];
        else if (r < nth) out = 'N';
        else out = (uint8_t)kBase[b];
        o[j] = out;
    }
    o += p.len;
    *o++ = '\n';
    if (fastq) {
        *o++ = '+';
        *o++ = '\n';
        uint64_t q = 0;
        for (uint64_t j = 0; j < p.len; ++j) {
            if ((j & 7) == 0) q = p.rng.next();
            *o++ = (uint8_t)('5' + ((q >> (8 * (j & 7))) & 0xFF) % 21);  // Phred33 20..40
        }
        *o++ = '\n';
    }
}

// Splits [0, n) into one contiguous range per thread; `grain` items at least
// per thread.
template <typename F>
void parallel_for(uint64_t n, uint32_t threads, F fn, uint64_t grain = 4096)
{
    uint32_t th = threads ? threads : std::max(1u, std::thread::hardware_concurrency());
    th = (uint32_t)std::min<uint64_t>(th, std::max<uint64_t>(1, n / grain));
    if (th <= 1) {
        fn(0, n);
        return;
    }
    std::vector<std::thread> ws;
    uint32_t started = 0;
    try {
        ws.reserve(th);
        for (; started < th; ++started) {
            const uint32_t t = started;
            ws.emplace_back([=] { fn(n * t / th, n * (t + 1) / th); });
        }
    } catch (...) {
        // a thread that could not start: this one takes the ranges left
    }
    for (uint32_t t = started; t < th; ++t) fn(n * t / th, n * (t + 1) / th);
    for (auto &w : ws) w.join();
}

}  // namespace

extern "C" int64_t kc_synth_reads_size(const uint8_t *genome, uint64_t glen, const kc_read_spec *sp, int fastq)
{
    return kc_guard([&]() -> int64_t {
        if (!genome || !sp || !glen) return KC_EINVAL;
        uint64_t total = 0;
        parallel_for(sp->n_reads, sp->threads, [&](uint64_t a, uint64_t b) {
            uint64_t s = 0;
            for (uint64_t i = a; i < b; ++i) s += record_bytes(plan_read(sp, glen, i), i, fastq != 0);
            __atomic_fetch_add(&total, s, __ATOMIC_RELAXED);
        });
        return (int64_t)total;
    });
}

extern "C" int64_t kc_synth_reads(const uint8_t *genome, uint64_t glen, const kc_read_spec *sp, int fastq,
                                  uint8_t *out, uint64_t cap)
{
    return kc_guard([&]() -> int64_t {
        if (!genome || !sp || !out || !glen) return KC_EINVAL;
        const uint64_t n = sp->n_reads;
        // offsets: per-block sums, then a second pass writes
        const uint64_t kBlk = 1 << 14;
        const uint64_t nblk = (n + kBlk - 1) / kBlk;
        std::vector<uint64_t> boff(nblk + 1, 0);
        parallel_for(nblk, sp->threads, [&](uint64_t a, uint64_t b) {
            for (uint64_t bi = a; bi < b; ++bi) {
                uint64_t s = 0;
                for (uint64_t i = bi * kBlk; i < std::min(n, (bi + 1) * kBlk); ++i)
                    s += record_bytes(plan_read(sp, glen, i), i, fastq != 0);
                boff[bi + 1] = s;
            }
        }, 1);
        for (uint64_t bi = 0; bi < nblk; ++bi) boff[bi + 1] += boff[bi];
        if (boff[nblk] > cap) return KC_EINVAL;
        parallel_for(nblk, sp->threads, [&](uint64_t a, uint64_t b) {
            for (uint64_t bi = a; bi < b; ++bi) {
                uint64_t o = boff[bi];
                for (uint64_t i = bi * kBlk; i < std::min(n, (bi + 1) * kBlk); ++i) {
                    const ReadPlan p = plan_read(sp, glen, i);
                    write_read(genome, sp, p, i, fastq != 0, out + o);
                    o += record_bytes(p, i, fastq != 0);
                }
            }
        }, 1);
        return (int64_t)boff[nblk];
    });
}

// ---- streaming ingest: FASTQ(.gz) files -> framed chunks -> a sink --------
// Files are cut into segments that host threads take in turn.  A plain or
// single-member gzip file is one segment: its thread reads it (a ".gz" name
// inflated with zlib, as FastqReader.h:423,575-616 does) and frames it (line
// 2 of every record + '\n') as it goes.  A BGZF file
// (htslib's blocked gzip: a chain of gzip members whose "BC" extra subfield
// gives each member's size, so members are found without inflating) is cut
// into runs of members that threads inflate in parallel; the record phase at
// the start of a run (its line number mod 4) comes from the newline counts
// of the runs before it, handed on in file order, so each run is framed
// independently and the framed runs are handed to the sink in file order,
// whole lines only (a line cut by a run boundary is carried to the next).
// The sink gets whole reads in chunks (a file's last chunk may lack its
// final '\n'); chunks of different files may interleave, each file's reads
// stay in order.
namespace {

constexpr uint64_t kIngestChunk = 4ull << 20;
constexpr unsigned kGzRead = 1u << 20;
constexpr uint64_t kBgzfRun = 1ull << 20;  // compressed bytes per BGZF segment

using Sink = std::function<int(const uint8_t *, uint64_t)>;

struct Member {
    uint64_t off;   // file offset of the gzip member
    uint32_t size;  // bytes of the member
};

uint16_t le16(const uint8_t *p) { return (uint16_t)(p[0] | (p[1] << 8)); }
uint32_t le32(const uint8_t *p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24); }

// Header of a BGZF member at p (n bytes available): the member's size and
// where its deflate data starts, or false if it is no BGZF member (FLG must
// be FEXTRA alone and the extra field must hold a "BC" subfield, SAM/BAM
// format specification section 4.1).
bool bgzf_header(const uint8_t *p, uint64_t n, uint32_t &size, uint32_t &data_at)
{
    if (n < 18 || p[0] != 0x1f || p[1] != 0x8b || p[2] != 8 || p[3] != 4) return false;
    const uint32_t xlen = le16(p + 10);
    if (n < 12ull + xlen) return false;
    for (uint32_t i = 0; i + 4 <= xlen;) {
        const uint8_t *sf = p + 12 + i;
        const uint32_t slen = le16(sf + 2);
        if (sf[0] == 'B' && sf[1] == 'C' && slen == 2 && i + 6 <= xlen) {
            size = (uint32_t)le16(sf + 4) + 1;
            data_at = 12 + xlen;
            return size >= data_at + 8;
        }
        i += 4 + slen;
    }
    return false;
}

// The members of a BGZF file, or false if any part of it is not BGZF.
bool bgzf_members(FILE *f, uint64_t file_size, std::vector<Member> &out)
{
    uint8_t h[12 + 65535];
    uint64_t off = 0;
    while (off < file_size) {
        const uint64_t want = std::min<uint64_t>(sizeof(h), file_size - off);
        if (fseeko(f, (off_t)off, SEEK_SET) != 0 || std::fread(h, 1, 12, f) != 12) return false;
        const uint64_t hdr = std::min<uint64_t>(want, 12ull + le16(h + 10));
        if (hdr > 12 && std::fread(h + 12, 1, hdr - 12, f) != hdr - 12) return false;
        uint32_t size = 0, data_at = 0;
        if (!bgzf_header(h, hdr, size, data_at) || off + size > file_size) return false;
        out.push_back({off, size});
        off += size;
    }
    return !out.empty();
}

enum SegmentKind : uint8_t { kStreamed = 0, kBgzfMembers = 1, kGzipChunk = 2 };

struct Segment {
    uint32_t file;
    uint32_t index;   // among the file's segments
    bool last;        // the file's last segment
    uint8_t kind;     // kBgzfMembers: members [m0, m1); kGzipChunk: chunk `index`; kStreamed: the whole file
    uint32_t m0, m1;
};

// Hand-over state of a file cut into segments: the next segment whose phase
// and whose framed bytes are due, the file's newlines so far, the partial
// line carried.
struct FileOrder {
    uint32_t next_phase = 0, next_hand = 0;
    uint64_t lines = 0;
    std::vector<uint8_t> carry;
};

// One gzip member inflated in chunks on several threads (kc_inflate.h).
// The deflate stream is cut into equal shares; chunk 0 decodes from the
// stream's start, chunk i > 0 from the first dynamic block it finds in its
// share (one that decodes to its end as text with the 32 KB before it
// unknown).  Every chunk decodes block by block until a block starts in the
// next share (or the member ends).  In chunk order, each chunk's start is
// checked against where the output before it ends: equal, its markers are
// answered by the last 32 KB of that output; different (the guess was no
// block of the stream, or nothing was found), the chunk decodes its share
// again from there with that window.  So the output is the stream's, byte
// for byte, and the member's CRC-32 and size are checked as gzread checks
// them; members after the first go through zlib in order.
constexpr uint64_t kGzipChunkMin = 64ull << 20;  // compressed bytes from which a member is cut

// Test-only override (KC_TEST_GZIP_CHUNK=<bytes>): files from that size are
// cut into shares of that size, so small test files take this path.
uint64_t test_gzip_chunk()
{
    const char *e = std::getenv("KC_TEST_GZIP_CHUNK");
    return e ? std::strtoull(e, nullptr, 10) : 0;
}

struct GzipFile {
    const uint8_t *map = nullptr;
    uint64_t size = 0;
    uint64_t d0 = 0;     // first byte of the deflate stream
    uint64_t chunk = 0;  // compressed bytes per share
    uint32_t n_chunks = 0;
    // in chunk order (Ingest::omu): the next chunk due, where the output so
    // far ends (bit), whether the member ended, the output's last <= 32 KB,
    // the CRC-32 and size of the output so far
    uint32_t next_window = 0;
    uint64_t stop_bit = 0;
    bool ended = false;
    std::vector<uint8_t> window;
    uint32_t crc = 0;
    uint64_t isize = 0;
    GzipFile() = default;
    GzipFile(const GzipFile &) = delete;
    GzipFile &operator=(const GzipFile &) = delete;
    ~GzipFile()
    {
        if (map) munmap(const_cast<uint8_t *>(map), size);
    }
};

// The first byte of the deflate stream after a gzip header at p, 0 if p
// holds no gzip header.
uint64_t gzip_header_end(const uint8_t *p, uint64_t n)
{
    if (n < 18 || p[0] != 0x1f || p[1] != 0x8b || p[2] != 8 || (p[3] & 0xE0)) return 0;
    const uint8_t flg = p[3];
    uint64_t o = 10;
    if (flg & 4) {
        if (o + 2 > n) return 0;
        o += 2 + (uint64_t)(p[o] | (p[o + 1] << 8));
    }
    for (uint8_t bit : {(uint8_t)8, (uint8_t)16})  // file name, comment: zero-terminated
        if (flg & bit) {
            while (o < n && p[o]) ++o;
            ++o;
        }
    if (flg & 2) o += 2;
    return o < n ? o : 0;
}

// Frames p[0, n) in place (the framed bytes never pass the bytes read).
uint64_t frame_in_place(uint8_t *p, uint64_t n, uint32_t state)
{
    uint64_t w = 0, i = 0;
    while (i < n) {
        const uint8_t *nl = static_cast<const uint8_t *>(std::memchr(p + i, '\n', n - i));
        const uint64_t end = nl ? (uint64_t)(nl - p) : n;
        if (state == 1) {
            std::memmove(p + w, p + i, end - i);
            w += end - i;
            if (nl) p[w++] = '\n';
        }
        if (!nl) break;
        i = end + 1;
        state = (state + 1) & 3;
    }
    return w;
}

struct Ingest {
    const char *const *paths;
    uint32_t n_paths;
    Sink sink;
    std::vector<std::vector<Member>> members;  // per file (empty: streamed)
    std::vector<std::unique_ptr<GzipFile>> gzip;  // per file cut into gzip shares
    std::vector<Segment> segs;
    std::vector<FileOrder> order;
    std::mutex omu;  // order, the results and the error
    std::condition_variable ocv;
    std::atomic<uint64_t> next{0};
    std::atomic<int64_t> spare_threads{0};
    int first_error = KC_OK;
    uint64_t in_bytes = 0, framed = 0, chunks = 0;
    double inflate_s = 0;

    void fail(int e)
    {
        std::lock_guard<std::mutex> lk(omu);
        if (first_error == KC_OK) first_error = e;
        ocv.notify_all();
    }
    bool failed()
    {
        std::lock_guard<std::mutex> lk(omu);
        return first_error != KC_OK;
    }

    // A large gzip file: mapped, its deflate stream cut into shares of 4 to
    // 32 MB (about four per thread).  False when it has no gzip header.
    bool plan_gzip_chunks(uint32_t fi, uint64_t size, uint32_t threads, uint64_t tchunk)
    {
        const int fd = ::open(paths[fi], O_RDONLY);
        if (fd < 0) return false;
        void *m = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
        ::close(fd);
        if (m == MAP_FAILED) return false;
        auto g = std::make_unique<GzipFile>();
        g->map = static_cast<const uint8_t *>(m);
        g->size = size;
        g->d0 = gzip_header_end(g->map, size);
        if (!g->d0) return false;
        const uint64_t body = size - g->d0;
        g->chunk = tchunk ? tchunk : std::max<uint64_t>(4ull << 20, std::min<uint64_t>(32ull << 20, body / (4ull * threads)));
        g->n_chunks = (uint32_t)std::max<uint64_t>(1, (body + g->chunk - 1) / g->chunk);
        g->stop_bit = g->d0 * 8;
        for (uint32_t i = 0; i < g->n_chunks; ++i) segs.push_back({fi, i, i + 1 == g->n_chunks, kGzipChunk, 0, 0});
        gzip[fi] = std::move(g);
        return true;
    }

    // Plans the segments: member runs of BGZF files, shares of large gzip
    // files, one per other file.
    int plan(uint32_t threads)
    {
        members.resize(n_paths);
        order.resize(n_paths);
        gzip.resize(n_paths);
        for (uint32_t fi = 0; fi < n_paths; ++fi) {
            FILE *f = std::fopen(paths[fi], "rb");
            if (!f) return KC_EIO;
            uint64_t size = 0;
            if (fseeko(f, 0, SEEK_END) == 0) size = (uint64_t)ftello(f);
            std::vector<Member> m;
            // only worth the walk when there are threads to share the file
            const bool gz = threads > 1 && has_gz_suffix(paths[fi]);
            const bool bgzf = gz && size > kBgzfRun && bgzf_members(f, size, m);
            std::fclose(f);
            const uint64_t tchunk = test_gzip_chunk();
            if (!bgzf && gz && size >= (tchunk ? tchunk : kGzipChunkMin) && plan_gzip_chunks(fi, size, threads, tchunk))
                continue;
            if (!bgzf) {
                segs.push_back({fi, 0, true, kStreamed, 0, 0});
                continue;
            }
            uint32_t idx = 0;
            for (uint32_t a = 0; a < m.size();) {
                uint64_t bytes = 0;
                uint32_t b = a;
                while (b < m.size() && (b == a || bytes + m[b].size <= kBgzfRun)) bytes += m[b++].size;
                segs.push_back({fi, idx++, b == m.size(), kBgzfMembers, a, b});
                a = b;
            }
            members[fi] = std::move(m);
        }
        return KC_OK;
    }

    // Whole lines of p[0, n) (after the file's carried partial line) to the
    // sink; the rest is carried.  The file's last piece also hands over the
    // carry.  omu is held by the caller (file order).
    int hand(FileOrder &fo, const uint8_t *p, uint64_t n, bool last, uint64_t &n_chunks)
    {
        int rc = KC_OK;
        const uint8_t *lnl = n ? static_cast<const uint8_t *>(memrchr(p, '\n', n)) : nullptr;
        if (!lnl) {
            fo.carry.insert(fo.carry.end(), p, p + n);
        } else {
            uint64_t from = 0;
            const uint64_t cut = (uint64_t)(lnl - p) + 1;
            if (!fo.carry.empty()) {
                const uint64_t first = (uint64_t)(static_cast<const uint8_t *>(std::memchr(p, '\n', n)) - p) + 1;
                fo.carry.insert(fo.carry.end(), p, p + first);
                rc = sink(fo.carry.data(), fo.carry.size());
                ++n_chunks;
                fo.carry.clear();
                from = first;
            }
            if (rc == KC_OK && cut > from) {
                rc = sink(p + from, cut - from);
                ++n_chunks;
            }
            fo.carry.assign(p + cut, p + n);
        }
        if (rc == KC_OK && last && !fo.carry.empty()) {
            rc = sink(fo.carry.data(), fo.carry.size());  // the file's last read (newline added if missing)
            ++n_chunks;
            fo.carry.clear();
        }
        return rc;
    }

    // One BGZF member run: inflate (CRC and size checked, as gzread does),
    // count its lines, take its phase in file order, frame, hand over in
    // file order.
    int run_bgzf(const Segment &sg, std::vector<uint8_t> &cbuf, std::vector<uint8_t> &out, z_stream &z)
    {
        const std::vector<Member> &m = members[sg.file];
        const uint64_t c0 = m[sg.m0].off, c1 = m[sg.m1 - 1].off + m[sg.m1 - 1].size;
        cbuf.resize(c1 - c0);
        FILE *f = std::fopen(paths[sg.file], "rb");
        if (!f) return KC_EIO;
        const bool rd = fseeko(f, (off_t)c0, SEEK_SET) == 0 && std::fread(cbuf.data(), 1, cbuf.size(), f) == cbuf.size();
        std::fclose(f);
        if (!rd) return KC_EIO;
        uint64_t total = 0;
        for (uint32_t i = sg.m0; i < sg.m1; ++i) total += le32(cbuf.data() + (m[i].off - c0) + m[i].size - 4);
        if (total > (4ull << 30)) return KC_EIO;  // not what ~1 MB of BGZF members inflate to
        out.resize(total);
        const auto a = std::chrono::steady_clock::now();
        uint64_t o = 0;
        for (uint32_t i = sg.m0; i < sg.m1; ++i) {
            const uint8_t *p = cbuf.data() + (m[i].off - c0);
            uint32_t size = 0, data_at = 0;
            if (!bgzf_header(p, m[i].size, size, data_at)) return KC_EIO;
            const uint32_t isize = le32(p + size - 4), crc = le32(p + size - 8);
            uint8_t dummy = 0;
            if (inflateReset(&z) != Z_OK) return KC_EIO;
            z.next_in = const_cast<Bytef *>(p + data_at);
            z.avail_in = size - data_at - 8;
            z.next_out = isize ? out.data() + o : &dummy;
            z.avail_out = isize;
            if (inflate(&z, Z_FINISH) != Z_STREAM_END || z.avail_out != 0 || z.total_out != isize) return KC_EIO;
            if (crc32(0L, isize ? out.data() + o : &dummy, isize) != crc) return KC_EIO;
            o += isize;
        }
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
        const uint64_t lines = (uint64_t)std::count(out.begin(), out.end(), (uint8_t)'\n');
        FileOrder &fo = order[sg.file];
        uint32_t state;
        {
            std::unique_lock<std::mutex> lk(omu);
            ocv.wait(lk, [&] { return fo.next_phase == sg.index || first_error != KC_OK; });
            if (first_error != KC_OK) return first_error;
            state = (uint32_t)(fo.lines & 3);
            fo.lines += lines;
            ++fo.next_phase;
            ocv.notify_all();
        }
        const uint64_t w = frame_in_place(out.data(), out.size(), state);
        std::unique_lock<std::mutex> lk(omu);
        ocv.wait(lk, [&] { return fo.next_hand == sg.index || first_error != KC_OK; });
        if (first_error != KC_OK) return first_error;
        uint64_t n_chunks = 0;
        const int rc = hand(fo, out.data(), w, sg.last, n_chunks);
        ++fo.next_hand;
        in_bytes += total;
        framed += w;
        chunks += n_chunks;
        inflate_s += secs;
        ocv.notify_all();
        return rc;
    }

    // Frames one inflated block into buf (at used), handing whole reads to the
    // sink once a chunk has gathered.
    int frame_block(Framer &fr, const uint8_t *p, uint64_t r, std::vector<uint8_t> &buf, uint64_t &used,
                    uint64_t &my_framed, uint64_t &my_chunks)
    {
        used += fr.feed(p, r, buf.data() + used);
        if (used >= kIngestChunk) {
            // hand over whole reads; a partial sequence line waits for the next block
            const uint8_t *nl = static_cast<const uint8_t *>(memrchr(buf.data(), '\n', used));
            const uint64_t cut = nl ? (uint64_t)(nl - buf.data()) + 1 : 0;
            if (cut) {
                const int err = sink(buf.data(), cut);
                if (err != KC_OK) return err;
            }
            std::memmove(buf.data(), buf.data() + cut, used - cut);
            my_framed += cut;
            my_chunks += cut ? 1 : 0;
            used -= cut;
        }
        // room for the next block (more than a chunk when one sequence line is longer)
        if (buf.size() < used + kGzRead + (used > kIngestChunk ? kIngestChunk : 0))
            buf.resize(used + kGzRead + (used > kIngestChunk ? kIngestChunk : 0));
        return KC_OK;
    }

    // A whole file streamed (InFile).  With a spare thread (helper),
    // this thread only inflates, into a ring of kRing blocks, and the helper
    // frames the blocks in order and hands the reads to the sink, so a single
    // gzip stream inflates back to back; otherwise one thread does both.
    static constexpr unsigned kRing = 4;
    int run_stream(const Segment &sg, std::vector<uint8_t> &in, std::vector<uint8_t> &buf)
    {
        InFile g;
        if (!g.open(paths[sg.file])) return KC_EIO;
        Framer fr;
        uint64_t used = 0, my_in = 0, my_framed = 0, my_chunks = 0;
        double my_inflate = 0;
        int err = KC_OK;
        bool helper = false;
        if (spare_threads.fetch_sub(1) > 0) helper = true;
        else spare_threads.fetch_add(1);
        if (helper) {
            struct Ring {
                std::mutex mu;
                std::condition_variable cv;
                std::vector<std::vector<uint8_t>> blk;
                int len[kRing] = {};
                uint64_t head = 0, tail = 0;  // blocks inflated / framed
                bool done = false;
                int err = KC_OK;
            } ring;
            std::thread th;
            try {
                ring.blk.assign(kRing, std::vector<uint8_t>(kGzRead));
                th = std::thread([&]() {
                    int e = KC_OK;
                    try {
                        for (;;) {
                            std::unique_lock<std::mutex> lk(ring.mu);
                            ring.cv.wait(lk, [&] { return ring.tail < ring.head || ring.done || ring.err; });
                            if (ring.err || ring.tail == ring.head) break;  // failed, or inflated and framed
                            const unsigned b = (unsigned)(ring.tail % kRing);
                            const int r = ring.len[b];
                            lk.unlock();
                            e = frame_block(fr, ring.blk[b].data(), (uint64_t)r, buf, used, my_framed, my_chunks);
                            lk.lock();
                            ++ring.tail;
                            if (e != KC_OK) ring.err = e;
                            ring.cv.notify_all();
                            if (e != KC_OK) break;
                        }
                    } catch (...) {
                        std::lock_guard<std::mutex> lk(ring.mu);
                        ring.err = KC_ENOMEM;
                        ring.cv.notify_all();
                    }
                });
            } catch (...) {
                helper = false;
            }
            if (helper) {
                for (;;) {
                    unsigned b;
                    {
                        std::unique_lock<std::mutex> lk(ring.mu);
                        ring.cv.wait(lk, [&] { return ring.head - ring.tail < kRing || ring.err; });
                        if (ring.err) break;
                        b = (unsigned)(ring.head % kRing);
                    }
                    const auto a = std::chrono::steady_clock::now();
                    const int r = g.read(ring.blk[b].data(), kGzRead);
                    my_inflate += std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
                    std::lock_guard<std::mutex> lk(ring.mu);
                    if (r < 0) ring.err = KC_EIO;
                    else if (r == 0) ring.done = true;
                    else {
                        my_in += (uint64_t)r;
                        ring.len[b] = r;
                        ++ring.head;
                    }
                    ring.cv.notify_all();
                    if (r <= 0) break;
                }
                th.join();
                err = ring.err;
            }
            spare_threads.fetch_add(1);
        }
        if (!helper) {
            for (;;) {
                const auto a = std::chrono::steady_clock::now();
                const int r = g.read(in.data(), kGzRead);
                my_inflate += std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
                if (r < 0) {
                    err = KC_EIO;
                    break;
                }
                if (r == 0) break;
                my_in += (uint64_t)r;
                if ((err = frame_block(fr, in.data(), (uint64_t)r, buf, used, my_framed, my_chunks)) != KC_OK) break;
            }
        }
        if (err == KC_OK && used) {
            err = sink(buf.data(), used);  // the file's last read (newline added if missing)
            my_framed += used;
            ++my_chunks;
        }
        std::lock_guard<std::mutex> lk(omu);
        in_bytes += my_in;
        framed += my_framed;
        chunks += my_chunks;
        inflate_s += my_inflate;
        return err;
    }

    // Members after the first of a gzip file cut into shares: inflated in
    // order with zlib (gzread reads concatenated members too, and stops at
    // bytes that start no gzip header), framed on from the file's line phase
    // and handed over.  omu held (file order).
    int rest_members(const GzipFile &g, uint64_t from, FileOrder &fo, uint64_t &n_in, uint64_t &n_framed,
                     uint64_t &n_chunks)
    {
        z_stream z{};
        if (inflateInit2(&z, 16 + MAX_WBITS) != Z_OK) return KC_ENOMEM;
        std::vector<uint8_t> out(kGzRead), framed(kGzRead);
        Framer fr;
        fr.state = (uint32_t)(fo.lines & 3);
        int rc = KC_OK;
        uint64_t at = from;
        while (rc == KC_OK && at + 18 <= g.size && g.map[at] == 0x1f && g.map[at + 1] == 0x8b) {
            inflateReset(&z);
            z.next_in = const_cast<Bytef *>(g.map + at);
            z.avail_in = (uInt)std::min<uint64_t>(g.size - at, 1u << 30);
            int zr = Z_OK;
            while (zr == Z_OK) {
                z.next_out = out.data();
                z.avail_out = (uInt)out.size();
                zr = inflate(&z, Z_NO_FLUSH);
                if (zr != Z_OK && zr != Z_STREAM_END) {
                    rc = KC_EIO;
                    break;
                }
                const uint64_t got = out.size() - z.avail_out;
                n_in += got;
                const uint64_t w = fr.feed(out.data(), got, framed.data());
                n_framed += w;
                rc = hand(fo, framed.data(), w, false, n_chunks);
                if (rc != KC_OK) break;
                if (zr == Z_OK && z.avail_in == 0) {
                    if (at + (uint64_t)(z.next_in - (g.map + at)) >= g.size) {
                        rc = KC_EIO;  // a member cut short
                        break;
                    }
                    z.avail_in = (uInt)std::min<uint64_t>(g.size - (uint64_t)(z.next_in - g.map), 1u << 30);
                }
            }
            at = (uint64_t)(z.next_in - g.map);
        }
        inflateEnd(&z);
        return rc;
    }

    // Share i of a gzip member (GzipFile): speculative decode, then in chunk
    // order the start check (and a redo from the known output if it fails),
    // the resolve of the symbols to bytes, the record phase, framing, and
    // the hand-over in file order.
    int run_gzip_chunk(const Segment &sg, kcz::Out16 &o, std::vector<uint8_t> &prev)
    {
        GzipFile &g = *gzip[sg.file];
        FileOrder &fo = order[sg.file];
        const uint32_t i = sg.index;
        const uint64_t total_bits = g.size * 8;
        const uint64_t sbit = (g.d0 + (uint64_t)i * g.chunk) * 8;
        const uint64_t nbit = i + 1 < g.n_chunks ? (g.d0 + (uint64_t)(i + 1) * g.chunk) * 8 : UINT64_MAX;
        const auto a = std::chrono::steady_clock::now();
        kcz::Decoder dec;
        dec.in.base = g.map;
        dec.in.nbytes = g.size;
        o.n = 0;
        uint64_t start = sbit, stop = 0;
        bool ok = true, fin = false;
        if (i > 0) {
            const uint64_t lim = std::min(nbit, total_bits);
            start = kcz::find_block_start(g.map, g.size, sbit, lim);
            ok = start < lim;
            dec.unknown_window = true;
            // a guess gone wrong must not decode without bound: FASTQ inflates
            // a few times its size; more than 64x is redone from the known output
            dec.max_out = 64 * g.chunk + (1u << 20);
        }
        // blocks from `from` until one starts in the next share or the member ends
        auto decode_from = [&](uint64_t from) {
            dec.in.pos = from;
            for (;;) {
                if (dec.in.pos >= nbit) {
                    stop = dec.in.pos;
                    return true;
                }
                const int r = dec.block(o);
                if (r == kcz::kBlockFinal) {
                    fin = true;
                    stop = dec.in.pos;
                    return true;
                }
                if (r != kcz::kBlockOk) return false;
            }
        };
        if (ok) ok = decode_from(start);
        double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
        uint64_t skip = 0;  // leading symbols of o that are the window (a redo)
        bool dead = false;
        {
            std::unique_lock<std::mutex> lk(omu);
            ocv.wait(lk, [&] { return g.next_window == i || first_error != KC_OK; });
            if (first_error != KC_OK) return first_error;
            if (g.ended) {
                dead = true;  // the member ended in an earlier share (more members follow it)
            } else {
                if (!ok || start != g.stop_bit) {
                    // no block of the stream starts there: this share again,
                    // from where the output ends, with its window known
                    const auto b = std::chrono::steady_clock::now();
                    o.n = 0;
                    o.reserve_more(g.window.size());
                    for (uint8_t c : g.window) o.v[o.n++] = c;
                    skip = g.window.size();
                    dec.unknown_window = false;
                    dec.text_only = false;
                    dec.max_out = ~0ull;
                    fin = false;
                    if (!decode_from(g.stop_bit)) {
                        first_error = KC_EIO;
                        ocv.notify_all();
                        return KC_EIO;
                    }
                    secs += std::chrono::duration<double>(std::chrono::steady_clock::now() - b).count();
                }
                // the markers answer to the output before this share; the next
                // share's to the last 32 KB of the output with this one
                prev = g.window;
                const uint64_t n = o.n - skip;
                std::vector<uint8_t> nw;
                bool rok;
                if (n >= kcz::kWindow) {
                    nw.resize(kcz::kWindow);
                    rok = kcz::resolve(o.v.data() + o.n - kcz::kWindow, kcz::kWindow, prev.data(), (uint32_t)prev.size(),
                                       nw.data());
                } else {
                    const uint64_t keep = std::min<uint64_t>(prev.size(), kcz::kWindow - n);
                    nw.assign(prev.end() - keep, prev.end());
                    nw.resize(keep + n);
                    rok = kcz::resolve(o.v.data() + skip, n, prev.data(), (uint32_t)prev.size(), nw.data() + keep);
                }
                if (!rok) {
                    first_error = KC_EIO;
                    ocv.notify_all();
                    return KC_EIO;
                }
                g.window.swap(nw);
                g.stop_bit = stop;
                if (fin) g.ended = true;
            }
            ++g.next_window;
            ocv.notify_all();
        }
        // the share's bytes, in place over its symbols
        uint8_t *bytes = reinterpret_cast<uint8_t *>(o.v.data());
        const uint64_t n = dead ? 0 : o.n - skip;
        if (!dead && !kcz::resolve(o.v.data() + skip, n, prev.data(), (uint32_t)prev.size(), bytes)) return KC_EIO;
        uint32_t crc_n = 0;
        for (uint64_t k = 0; k < n; k += 1u << 30)
            crc_n = (uint32_t)crc32(crc_n, bytes + k, (uInt)std::min<uint64_t>(n - k, 1u << 30));
        const uint64_t lines = (uint64_t)std::count(bytes, bytes + n, (uint8_t)'\n');
        uint32_t state;
        {
            std::unique_lock<std::mutex> lk(omu);
            ocv.wait(lk, [&] { return fo.next_phase == i || first_error != KC_OK; });
            if (first_error != KC_OK) return first_error;
            state = (uint32_t)(fo.lines & 3);
            fo.lines += lines;
            ++fo.next_phase;
            ocv.notify_all();
        }
        const uint64_t w = frame_in_place(bytes, n, state);
        std::unique_lock<std::mutex> lk(omu);
        ocv.wait(lk, [&] { return fo.next_hand == i || first_error != KC_OK; });
        if (first_error != KC_OK) return first_error;
        uint64_t n_chunks = 0, n_in = n, n_framed = w;
        int rc = KC_OK;
        if (!dead) {
            g.crc = (uint32_t)crc32_combine(g.crc, crc_n, (z_off_t)n);
            g.isize += n;
            rc = hand(fo, bytes, w, false, n_chunks);
            if (rc == KC_OK && fin) {
                // the member's trailer: CRC-32 and size (mod 2^32) of its bytes
                const uint64_t t = (stop + 7) / 8;
                if (t + 8 > g.size || le32(g.map + t) != g.crc || le32(g.map + t + 4) != (uint32_t)g.isize) rc = KC_EIO;
                if (rc == KC_OK && t + 8 < g.size) rc = rest_members(g, t + 8, fo, n_in, n_framed, n_chunks);
                if (rc == KC_OK) rc = hand(fo, nullptr, 0, true, n_chunks);  // the file's last read
            }
        }
        ++fo.next_hand;
        in_bytes += n_in;
        framed += n_framed;
        chunks += n_chunks;
        inflate_s += secs;
        ocv.notify_all();
        return rc;
    }

    void work()
    {
        std::vector<uint8_t> in, buf, cbuf, out, prev;
        kcz::Out16 sym;
        z_stream z{};
        bool z_ok = false;
        for (;;) {
            const uint64_t si = next.fetch_add(1);
            if (si >= segs.size() || failed()) break;
            const Segment &sg = segs[si];
            int rc;
            if (sg.kind == kGzipChunk) {
                rc = run_gzip_chunk(sg, sym, prev);
            } else if (sg.kind == kBgzfMembers) {
                if (!z_ok) {
                    if (inflateInit2(&z, -15) != Z_OK) {
                        fail(KC_ENOMEM);
                        break;
                    }
                    z_ok = true;
                }
                rc = run_bgzf(sg, cbuf, out, z);
            } else {
                if (in.empty()) {
                    in.resize(kGzRead);
                    buf.resize(kIngestChunk + kGzRead);
                }
                rc = run_stream(sg, in, buf);
            }
            if (rc != KC_OK) {
                fail(rc);
                break;
            }
        }
        if (z_ok) inflateEnd(&z);
    }

    // Runs the segments on up to `threads` threads; returns the first error.
    int run(uint32_t threads)
    {
        int rc = plan(threads);
        if (rc != KC_OK) return rc;
        // an allocation that fails in a worker ends the ingest with KC_ENOMEM
        auto worker = [&]() {
            try {
                work();
            } catch (...) {
                fail(KC_ENOMEM);
            }
        };
        const uint32_t nt = std::max<uint32_t>(
            1, std::min<uint64_t>(threads ? threads : 1, std::max<uint64_t>(segs.size(), 1)));
        // threads beyond one per segment frame streamed files beside their inflating thread
        spare_threads = (int64_t)(threads ? threads : 1) - (int64_t)nt;
        std::vector<std::thread> pool;
        try {
            pool.reserve(nt);
            for (uint32_t i = 0; i < nt; ++i) pool.emplace_back(worker);
        } catch (...) {
            // fewer threads: the ones running take every segment from the cursor
        }
        if (pool.empty()) worker();
        for (auto &th : pool) th.join();
        used_threads = pool.empty() ? 1 : (uint32_t)pool.size();
        return first_error;
    }
    uint32_t used_threads = 1;
};

}  // namespace

extern "C" int kc_count_fastq_files(kc_table *t, const char *const *paths, uint32_t n_paths, uint32_t threads,
                                    const kc_tune *tune, kc_ingest_stats *stats)
{
    return kc_guard([&]() -> int {
        if (!t || (n_paths && !paths)) return KC_EINVAL;
        for (uint32_t i = 0; i < n_paths; ++i)
            if (!paths[i]) return KC_EINVAL;
        kc_batch *b = nullptr;
        int rc = kc_batch_create(t, tune, &b);
        if (rc != KC_OK) return rc;
        const auto t0 = std::chrono::steady_clock::now();
        std::mutex bmu;  // guards the batch
        Ingest ing;
        ing.paths = paths;
        ing.n_paths = n_paths;
        ing.sink = [&](const uint8_t *p, uint64_t n) {
            std::lock_guard<std::mutex> lk(bmu);
            return kc_batch_add(b, p, n);
        };
        int first_error = ing.run(threads);
        if (first_error == KC_OK) first_error = kc_batch_flush(b);
        const int rs = kc_batch_sync(b);
        if (first_error == KC_OK) first_error = rs;
        kc_batch_info bi{};
        kc_batch_get_info(b, &bi);
        kc_batch_destroy(b);
        if (stats) {
            stats->files = n_paths;
            stats->threads = ing.used_threads;
            stats->input_bytes = ing.in_bytes;
            stats->framed_bytes = ing.framed;
            stats->chunks = ing.chunks;
            stats->launches = bi.launches;
            stats->inflate_seconds = ing.inflate_s;
            stats->wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
        return first_error;
    });
}

extern "C" int kc_fastq_frame_files(const char *const *paths, uint32_t n_paths, uint32_t threads, uint8_t **out,
                                    uint64_t *out_len, kc_ingest_stats *stats)
{
    return kc_guard([&]() -> int {
        if (!out || !out_len || (n_paths && !paths)) return KC_EINVAL;
        for (uint32_t i = 0; i < n_paths; ++i)
            if (!paths[i]) return KC_EINVAL;
        const auto t0 = std::chrono::steady_clock::now();
        std::mutex bmu;  // guards the output
        std::vector<uint8_t> acc;
        Ingest ing;
        ing.paths = paths;
        ing.n_paths = n_paths;
        ing.sink = [&](const uint8_t *p, uint64_t n) {
            std::lock_guard<std::mutex> lk(bmu);
            acc.insert(acc.end(), p, p + n);
            if (n && p[n - 1] != '\n') acc.push_back('\n');  // as a batch completes a chunk's last line
            return KC_OK;
        };
        const int rc = ing.run(threads);
        if (rc != KC_OK) return rc;
        uint8_t *buf = static_cast<uint8_t *>(std::malloc(std::max<size_t>(acc.size(), 1)));
        if (!buf) return KC_ENOMEM;
        if (!acc.empty()) std::memcpy(buf, acc.data(), acc.size());
        *out = buf;
        *out_len = acc.size();
        if (stats) {
            stats->files = n_paths;
            stats->threads = ing.used_threads;
            stats->input_bytes = ing.in_bytes;
            stats->framed_bytes = ing.framed;
            stats->chunks = ing.chunks;
            stats->launches = 0;
            stats->inflate_seconds = ing.inflate_s;
            stats->wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
        return KC_OK;
    });
}

// ---- out-of-HBM counting: several passes over the files ------------------
// Merge of sorted dump texts ("<bases>\t<count>\n", one key per line, keys
// unique across the inputs) by the bases field, the order dump_tool --sort
// writes (dump_tool.cpp:345-354); the reference merges its partition outputs
// the same way (merge_sorted_outputs, dump_tool.cpp:1043-1134).
namespace {

struct DumpReader {
    FILE *f = nullptr;
    std::string line;  // current line without its '\n'
    std::vector<char> buf;
    char *raw = nullptr;
    size_t raw_cap = 0;
    DumpReader() = default;
    DumpReader(const DumpReader &) = delete;
    DumpReader &operator=(const DumpReader &) = delete;
    ~DumpReader()
    {
        if (f) std::fclose(f);
        std::free(raw);
    }
    bool next()
    {
        const ssize_t r = getline(&raw, &raw_cap, f);
        if (r <= 0) {
            line.clear();
            return false;
        }
        line.assign(raw, (size_t)r - (raw[r - 1] == '\n' ? 1 : 0));
        return true;
    }
    std::string_view key() const
    {
        const size_t tab = line.find('\t');
        return std::string_view(line).substr(0, tab == std::string::npos ? line.size() : tab);
    }
};

}  // namespace

extern "C" int kc_merge_sorted_dumps(const char *const *paths, uint32_t n_paths, const char *out_path)
{
    return kc_guard([&]() -> int {
        if (!out_path || (n_paths && !paths)) return KC_EINVAL;
        for (uint32_t i = 0; i < n_paths; ++i)
            if (!paths[i]) return KC_EINVAL;
        std::vector<DumpReader> in(n_paths);
        int rc = KC_OK;
        for (uint32_t i = 0; i < n_paths && rc == KC_OK; ++i) {
            in[i].f = std::fopen(paths[i], "rb");
            if (!in[i].f) rc = KC_EIO;
            else {
                in[i].buf.resize(1 << 20);
                std::setvbuf(in[i].f, in[i].buf.data(), _IOFBF, in[i].buf.size());
            }
        }
        FILE *out = rc == KC_OK ? std::fopen(out_path, "wb") : nullptr;
        if (rc == KC_OK && !out) rc = KC_EIO;
        if (rc == KC_OK) {
            std::vector<char> obuf;  // outlives the stream that buffers into it
            try {
                obuf.resize(1 << 20);
                std::setvbuf(out, obuf.data(), _IOFBF, obuf.size());
                // min-heap of the readers by their current key (the heap holds
                // the readers that still have a line)
                std::vector<uint32_t> heap;
                auto greater = [&](uint32_t a, uint32_t b) { return in[a].key() > in[b].key(); };
                for (uint32_t i = 0; i < n_paths; ++i)
                    if (in[i].next()) heap.push_back(i);
                std::make_heap(heap.begin(), heap.end(), greater);
                while (!heap.empty()) {
                    std::pop_heap(heap.begin(), heap.end(), greater);
                    const uint32_t i = heap.back();
                    const std::string &l = in[i].line;
                    if (std::fwrite(l.data(), 1, l.size(), out) != l.size() || std::fputc('\n', out) == EOF) {
                        rc = KC_EIO;
                        break;
                    }
                    if (in[i].next()) std::push_heap(heap.begin(), heap.end(), greater);
                    else heap.pop_back();
                }
            } catch (...) {
                rc = KC_ENOMEM;
            }
            if (std::fclose(out) != 0 && rc == KC_OK) rc = KC_EIO;
        }
        return rc;  // the readers close their files
    });
}

static int append_file(const std::string &from, FILE *to)
{
    std::vector<char> b(1 << 20);
    FILE *f = std::fopen(from.c_str(), "rb");
    if (!f) return KC_EIO;
    size_t r;
    int rc = KC_OK;
    while ((r = std::fread(b.data(), 1, b.size(), f)) > 0)
        if (std::fwrite(b.data(), 1, r, to) != r) {
            rc = KC_EIO;
            break;
        }
    std::fclose(f);
    return rc;
}

extern "C" int kc_count_fastq_files_passes(kc_table *t, const char *const *paths, uint32_t n_paths, uint32_t threads,
                                           const kc_tune *tune, const kc_pass_plan *plan, kc_ingest_stats *stats)
{
    return kc_guard([&]() -> int {
        if (!t || !plan) return KC_EINVAL;
        const uint32_t P = plan->n_passes;
        if (P == 0 || P > 256 || (P & (P - 1))) return KC_EINVAL;
        if (plan->hist_bins && plan->hist_max < plan->hist_min) return KC_EINVAL;
        if (plan->dump_path && plan->dump_max < plan->dump_min) return KC_EINVAL;
        const uint64_t nb = plan->hist_bins ? (uint64_t)plan->hist_max - plan->hist_min + 1 : 0;
        std::vector<int64_t> acc(nb, 0), bins(nb, 0);
        std::vector<std::string> parts;
        parts.reserve(P);
        kc_ingest_stats sum{};
        int rc = KC_OK;
        try {
            for (uint32_t p = 0; p < P && rc == KC_OK; ++p) {
                if ((rc = kc_table_clear(t)) != KC_OK || (rc = kc_table_set_pass(t, p, P)) != KC_OK) break;
                kc_ingest_stats st{};
                if ((rc = kc_count_fastq_files(t, paths, n_paths, threads, tune, &st)) != KC_OK) break;
                sum.files = st.files;
                sum.threads = st.threads;
                sum.input_bytes += st.input_bytes;
                sum.framed_bytes += st.framed_bytes;
                sum.chunks += st.chunks;
                sum.launches += st.launches;
                sum.inflate_seconds += st.inflate_seconds;
                sum.wall_seconds += st.wall_seconds;
                if (plan->export_dir)
                    rc = p == 0 ? kc_table_export_dir(t, plan->export_dir)
                                : kc_table_export_dir_append(t, plan->export_dir);
                if (rc == KC_OK && plan->dump_path) {
                    parts.push_back(P == 1 ? std::string(plan->dump_path)
                                           : std::string(plan->dump_path) + ".pass" + std::to_string(p));
                    rc = kc_table_dump_file(t, plan->dump_min, plan->dump_max, plan->dump_sorted, parts.back().c_str());
                }
                if (rc == KC_OK && nb) {
                    rc = kc_table_histogram(t, plan->hist_min, plan->hist_max, bins.data());
                    for (uint64_t i = 0; i < nb && rc == KC_OK; ++i) acc[i] += bins[i];
                }
            }
            if (rc == KC_OK && plan->dump_path && P > 1) {
                // slices are hash-disjoint: sorted parts merge, unsorted parts concatenate
                if (plan->dump_sorted) {
                    std::vector<const char *> pp;
                    for (auto &s : parts) pp.push_back(s.c_str());
                    rc = kc_merge_sorted_dumps(pp.data(), (uint32_t)pp.size(), plan->dump_path);
                } else {
                    FILE *out = std::fopen(plan->dump_path, "wb");
                    if (!out) rc = KC_EIO;
                    for (size_t i = 0; i < parts.size() && rc == KC_OK; ++i) rc = append_file(parts[i], out);
                    if (out && std::fclose(out) != 0 && rc == KC_OK) rc = KC_EIO;
                }
            }
        } catch (...) {
            rc = KC_ENOMEM;
        }
        kc_table_set_pass(t, 0, 1);
        if (P > 1)
            for (auto &s : parts) std::remove(s.c_str());
        if (rc == KC_OK && nb) std::memcpy(plan->hist_bins, acc.data(), nb * sizeof(int64_t));
        if (stats) *stats = sum;
        return rc;
    });
}

// ---- table sizing from a device sample -----------------------------------
namespace {

// Slots and 64-bit Bloom blocks of one table partition of the LDS counting
// pass (Bloom split on).
uint64_t kc_lds_slots(uint32_t nw) { return nw == 1 ? 4096 : 2048; }
uint64_t kc_lds_bloom(uint32_t) { return 4096; }

uint64_t pow2_at_least(uint64_t v)
{
    uint64_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

// False positive rate of the table's Bloom probe (bloom_in_part: three
// distinct bits of a 64-bit block) at n distinct keys per block on average:
// a fresh key finds its three bits set in a block holding a Poisson(n)
// number of keys, each of which set three of the 64 bits.
double bloom_fp(double n)
{
    double fp = 0.0, p = std::exp(-n);  // p = P(j keys in the block)
    for (int j = 0; j < 256; ++j) {
        if (j) p *= n / j;
        const double f = 1.0 - std::pow(61.0 / 64.0, j);
        fp += p * f * f * f;
    }
    return std::min(1.0, fp);
}

// The same averaged over the insertions: the i-th of n keys per block meets a
// block that holds (i/n) of the final keys, so the false positives a count
// makes are the mean of bloom_fp over [0, n] (Simpson's rule).  Against the
// table keys measured on MI355X (k = 21/31/63, 1.3-6.8 keys per block) this
// predicts within 5%; the final-state rate above is 3x too high.
double bloom_fp_avg(double n)
{
    if (n <= 0) return 0.0;
    constexpr int m = 64;
    const double h = n / m;
    double s = bloom_fp(0.0) + bloom_fp(n);
    for (int i = 1; i < m; ++i) s += (i % 2 ? 4.0 : 2.0) * bloom_fp(i * h);
    return s * h / 3.0 / n;
}

}  // namespace

extern "C" int kc_plan_sizes(uint64_t distinct_total, uint64_t repeated_total, uint64_t hbm_budget,
                             kc_table_config *cfg, kc_plan *plan)
{
    if (!cfg || !plan) return KC_EINVAL;
    if (cfg->k < 1 || cfg->k > 128) return KC_EINVAL;
    const uint32_t nw = cfg->k <= 32 ? 1 : cfg->k <= 63 ? 2 : 4;  // the table's key words
    plan->distinct = distinct_total;
    plan->repeated = std::min(repeated_total, distinct_total);
    // one pass, then halve per pass until the device bytes fit the budget
    const uint64_t slot_bytes = 8ull * (nw == 1 ? 2 : nw == 2 ? 4 : 8);
    const uint64_t lds_bloom = kc_lds_bloom(nw);  // Bloom blocks of one partition in the LDS pass
    const uint64_t max_bloom = lds_bloom << 18;  // the partition pass takes 2^18 Bloom parts
    uint32_t P = 1;
    for (;;) {
        const uint64_t distinct = plan->distinct / P + 1024, repeated = plan->repeated / P + 1024;
        uint64_t cap, bloom = 0, low = 0;
        if (cfg->bloom) {
            // 3.7-7.4 distinct keys per 64-bit block, up to what the
            // partition pass takes (then more keys per block, at most 8);
            // its false positives send singletons to the table as well
            bloom = std::min(max_bloom, std::max<uint64_t>(pow2_at_least((uint64_t)((double)distinct / 7.4)), 1 << 12));
            const double fp = bloom_fp_avg((double)distinct / (double)bloom);
            const uint64_t table_keys = repeated + (uint64_t)(fp * (double)(distinct - std::min(distinct, repeated)));
            cap = pow2_at_least(std::max<uint64_t>(table_keys + table_keys * 3 / 5, 1 << 16));
            // and at least the LDS-sized partitions the filter's parts need
            // (the table is split into more, smaller partitions otherwise)
            bloom = std::max(bloom, cap / kc_lds_slots(nw) * lds_bloom);
            // partitions of >= 512 slots: the load of a small partition
            // strays further from the table's mean (Poisson), and a full
            // partition cannot borrow from its neighbours
            cap = std::max(cap, bloom * 512 / lds_bloom);
            low = distinct + distinct / 20 + (1 << 16);
        } else {
            cap = pow2_at_least(std::max<uint64_t>(distinct + distinct * 3 / 5, 1 << 16));
        }
        const uint64_t bytes = cap * slot_bytes + 8 * bloom + 8ull * nw * low;
        const bool fits = (!hbm_budget || bytes <= hbm_budget) && (!cfg->bloom || distinct <= 8 * bloom);
        if (fits || P == 256) {
            cfg->capacity = cap;
            cfg->bloom_blocks = bloom;
            cfg->low_capacity = low;
            cfg->max_distinct = distinct;
            plan->device_bytes = bytes;
            plan->passes = P;
            return fits ? KC_OK : KC_ENOMEM;
        }
        P <<= 1;
    }
}

extern "C" int kc_plan_table(kc_ctx *ctx, const uint8_t *d_seq, uint64_t nbytes, uint32_t sample_slices,
                             uint64_t hbm_budget, kc_table_config *cfg, kc_plan *plan)
{
    if (!ctx || !cfg || !plan || (nbytes && !d_seq)) return KC_EINVAL;
    if (cfg->k < 1 || cfg->k > 128) return KC_EINVAL;
    uint32_t S = sample_slices ? sample_slices : 16;
    if (S > 256 || (S & (S - 1))) return KC_EINVAL;
    // the probe holds every distinct key of its slice: at most nbytes / S
    while (S < 256 && nbytes / S > (1ull << 26)) S <<= 1;
    kc_table_config pc{};
    pc.k = cfg->k;
    pc.count_max = 0xFFFFFFFFu;
    pc.filter_max = 0xFFFFFFFFu;
    pc.capacity = pow2_at_least(std::max<uint64_t>(nbytes / S + nbytes / S / 2, 1 << 16));
    pc.count_path = KC_PATH_AUTO;
    kc_table *probe = nullptr;
    int rc = kc_table_create(ctx, &pc, &probe);
    if (rc != KC_OK) return rc;
    kc_stats st{};
    int64_t singles = 0;
    if ((rc = kc_table_set_pass(probe, 0, S)) == KC_OK && (rc = kc_table_count_seq(probe, d_seq, nbytes)) == KC_OK &&
        (rc = kc_table_stats(probe, &st)) == KC_OK && (rc = kc_table_histogram(probe, 1, 1, &singles)) == KC_OK &&
        st.errors)
        rc = KC_EFULL;
    kc_table_destroy(probe);
    if (rc != KC_OK) return rc;
    rc = kc_plan_sizes(st.distinct * S, (st.distinct - std::min<uint64_t>(st.distinct, (uint64_t)singles)) * S,
                       hbm_budget, cfg, plan);
    plan->sample_slices = S;
    plan->kmers = st.kmers * S;
    return rc;
}
