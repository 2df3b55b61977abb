// histogram_tool on the MI355X: the reference's command line
// (histogram_tool.cpp:176-224)
//
//   histogram_tool <precise/approximate> <tmp_dir> <max_threads> <max_memory_gb> <output_file>
//                  [min_freq=1] [max_freq=10000]
//
// over a temp_dir written by the reference's counter or by this engine, with
// the same output bytes ("<f>\t<n>\n" for f = min..max, histogram_tool.cpp:
// 462-479).  The merge runs on the device (kc_histogram_dir): precise merges
// high.bin and low.bin, approximate counts high.bin alone.  max_threads and
// max_memory_gb are checked as the reference checks them; the device sizes
// its memory from the files.
#include <cstdio>
#include <vector>

#include "tool_common.h"

int main(int argc, char *argv[])
{
    using namespace kctool;
    if (argc < 6 || argc > 8)
        die("Usage: histogram_tool <precise/approximate> <tmp_dir> <max_threads> <max_memory_gb> <output_file> "
            "[min_freq=1] [max_freq=10000]");
    const std::string dir = with_trailing_slash(argv[2]);
    const uint32_t threads = parse_u32(argv[3], "max_threads");
    parse_memory_gib(argv[4]);
    const std::string out = argv[5];
    bool precise;
    if (std::strcmp(argv[1], "precise") == 0) precise = true;
    else if (std::strcmp(argv[1], "approximate") == 0) precise = false;
    else die("invalid mode: must be 'precise' or 'approximate'");
    const uint32_t min_freq = argc >= 7 ? parse_u32(argv[6], "min_freq") : 1;
    const uint32_t max_freq = argc >= 8 ? parse_u32(argv[7], "max_freq") : 10000;
    if (threads == 0) die("invalid max_threads: must be greater than 0");
    if (max_freq < min_freq) die("invalid max_freq: must be greater than or equal to min_freq");
    const Infos in = read_infos(dir);
    check_files(dir, in, precise);

    std::vector<int64_t> bins((size_t)max_freq - min_freq + 1);
    kc_ctx *ctx = device_context();
    const int rc = kc_histogram_dir(ctx, dir.c_str(), precise ? 1 : 0, min_freq, max_freq, bins.data());
    kc_ctx_destroy(ctx);
    if (rc != KC_OK) die(std::string("histogram failed: ") + kc_strerror(rc));

    FILE *f = std::fopen(out.c_str(), "wb");
    if (!f) {
        std::cerr << "failed to open output file: " << out << std::endl;
        return 1;
    }
    std::vector<char> buf;
    buf.reserve(1 << 20);
    char line[64];
    for (size_t i = 0; i < bins.size(); ++i) {
        const int n = std::snprintf(line, sizeof line, "%llu\t%lld\n", (unsigned long long)min_freq + i,
                                    (long long)bins[i]);
        buf.insert(buf.end(), line, line + n);
        if (buf.size() > (1 << 20) - 64) {
            std::fwrite(buf.data(), 1, buf.size(), f);
            buf.clear();
        }
    }
    const bool ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    if (std::fclose(f) != 0 || !ok) die("failed to write " + out);
    return 0;
}
