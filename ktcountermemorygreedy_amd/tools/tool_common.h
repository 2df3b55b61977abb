// Shared by the two command-line tools: argv parsing and the temp_dir
// checks of the reference's histogram_tool / dump_tool
// (histogram_tool.cpp:130-175, 239-330; dump_tool.cpp:456-575, 639-678),
// with the same messages and exit status (-1) on bad input.  Unparsable
// numbers are reported the same way instead of ending in an uncaught
// exception.
#pragma once
#include <cerrno>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <limits>
#include <string>
#include <sys/stat.h>

#include "../../include/kc_api.h"

namespace kctool {

[[noreturn]] inline void die(const std::string &msg)
{
    std::cerr << msg << std::endl;
    std::exit(-1);
}

inline std::string with_trailing_slash(std::string path)
{
    if (path.empty()) return "./";
    const char tail = path.back();
    if (tail != '/' && tail != '\\') path.push_back('/');
    return path;
}

inline uint32_t parse_u32(const char *text, const char *name)
{
    char *end = nullptr;
    errno = 0;
    const unsigned long long v = std::strtoull(text, &end, 10);
    if (!*text || *end || errno || v > std::numeric_limits<uint32_t>::max() || text[0] == '-')
        die(std::string("invalid ") + name + ": " + text);
    return (uint32_t)v;
}

// max_memory_gb: a positive number of GiB.  The device tools size their
// device memory from the input files, so the value is checked, not used.
inline void parse_memory_gib(const char *text)
{
    char *end = nullptr;
    errno = 0;
    const long double v = std::strtold(text, &end);
    if (!*text || *end || errno || !(v > 0.0L)) die(std::string("invalid max_memory_gb: ") + text);
    if (v * 1073741824.0L > (long double)std::numeric_limits<uint64_t>::max()) die("max_memory_gb is too large");
}

struct Infos {
    uint32_t k_len = 0, count_max = 0, count_bytes = 4;
};

inline Infos read_infos(const std::string &dir)
{
    const std::string path = dir + "infos.bin";
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) die("failed to open " + path + ": " + std::strerror(errno));
    uint32_t buf[2];
    const bool ok = std::fread(buf, 4, 2, f) == 2;
    std::fclose(f);
    if (!ok) die("failed to read " + path);
    Infos i;
    i.k_len = buf[0];
    i.count_max = buf[1];
    i.count_bytes = i.count_max <= 0xFF ? 1 : i.count_max <= 0xFFFF ? 2 : i.count_max <= 0xFFFFFF ? 3 : 4;
    if (i.k_len == 0 || i.k_len > 128) die("invalid k_len from infos.bin: " + std::to_string(i.k_len));
    return i;
}

inline uint64_t packed_kmer_bytes(uint32_t k) { return (k / 32) * 8 + (2 * (k % 32) + 7) / 8; }

inline uint64_t size_of(const std::string &path)
{
    struct stat st {};
    if (::stat(path.c_str(), &st) != 0) die("failed to open " + path + ": " + std::strerror(errno));
    return (uint64_t)st.st_size;
}

// high.bin must hold whole records and at least one (the reference tools
// refuse an empty one); low.bin, when read, whole packed k-mers.
inline void check_files(const std::string &dir, const Infos &in, bool with_low)
{
    const uint64_t rec = packed_kmer_bytes(in.k_len) + in.count_bytes;
    const uint64_t hs = size_of(dir + "high.bin");
    if (hs == 0) die("empty high.bin file");
    if (hs % rec)
        die("invalid high.bin file size: " + std::to_string(hs) + " is not divisible by compact record size " +
            std::to_string(rec));
    if (with_low) {
        const uint64_t ls = size_of(dir + "low.bin"), kb = packed_kmer_bytes(in.k_len);
        if (ls % kb)
            die("invalid low.bin size: " + std::to_string(ls) + " is not divisible by packed k-mer size " +
                std::to_string(kb));
    }
}

inline kc_ctx *device_context()
{
    kc_ctx *ctx = nullptr;
    const int rc = kc_ctx_create(0, &ctx);
    if (rc != KC_OK) die(std::string("no device: ") + kc_strerror(rc));
    return ctx;
}

}  // namespace kctool
