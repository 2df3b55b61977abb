// dump_tool on the MI355X: the reference's command line
// (dump_tool.cpp:493-537)
//
//   dump_tool <precise/approximate> <tmp_dir> <max_threads> <max_memory_gb> <output_file>
//             [min_freq] [max_freq] [--sort]
//
// over a temp_dir written by the reference's counter or by this engine, with
// the same output bytes ("<bases>\t<count>\n" lines; --sort gives the
// lexicographic order of dump_tool.cpp:345-354, 1137-1198).  The merge, the
// sort and the text run on the device (kc_dump_dir): precise merges high.bin
// and low.bin (dump_tool.cpp:1202-1270), approximate writes the high.bin
// records in range (dump_tool.cpp:1294-1400).  Without --sort the line order
// is the device's (the reference's unsorted order depends on its threads).
#include "tool_common.h"

int main(int argc, char *argv[])
{
    using namespace kctool;
    if (argc < 6 || argc > 9)
        die("Usage: dump_tool <precise/approximate> <tmp_dir> <max_threads> <max_memory_gb> <output_file> "
            "[min_freq] [max_freq] [--sort]");
    const std::string dir = with_trailing_slash(argv[2]);
    const uint32_t threads = parse_u32(argv[3], "max_threads");
    parse_memory_gib(argv[4]);
    const std::string out = argv[5];
    bool precise;
    if (std::strcmp(argv[1], "precise") == 0) precise = true;
    else if (std::strcmp(argv[1], "approximate") == 0) precise = false;
    else die("mode must be precise or approximate");
    uint32_t min_freq = 1, max_freq = 0xFFFFFFFFu;
    bool sorted = false;
    int next = 6;
    if (argc > next && argv[next][0] != '-') min_freq = parse_u32(argv[next++], "min_freq");
    if (argc > next && argv[next][0] != '-') max_freq = parse_u32(argv[next++], "max_freq");
    if (argc > next && std::strcmp(argv[next], "--sort") == 0) sorted = true;
    if (threads == 0) die("invalid max_threads");
    if (max_freq < min_freq) die("max_freq < min_freq");
    const Infos in = read_infos(dir);
    check_files(dir, in, precise);

    kc_ctx *ctx = device_context();
    const int rc = kc_dump_dir(ctx, dir.c_str(), precise ? 1 : 0, min_freq, max_freq, sorted ? 1 : 0, out.c_str());
    kc_ctx_destroy(ctx);
    if (rc != KC_OK) die(std::string("dump failed: ") + kc_strerror(rc));
    return 0;
}
