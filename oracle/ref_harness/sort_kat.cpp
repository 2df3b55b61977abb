// Sort known-answer harness built from the reference's own sort headers.
// Reads binary records from stdin: u32 nw, u32 k, u32 cpb, u32 digit_bits,
// u64 n, then n records of (nw u64 words + u32 count).  Sorts them with
// export_record_radix_sort (digit_bits 8) or
// tree::sort::radix_sort_export_records_12bit<N, 10> (digit_bits 12) and
// writes the sorted records back to stdout in the same layout.
#include "ExportRecordRadixSort.h"
#include "ExportRecordRadixSort12.h"

#include <cstdio>
#include <cstdlib>
#include <vector>

template <uint32_t N>
static int run(uint32_t k, uint32_t cpb, uint32_t bits, uint64_t n)
{
    std::vector<ExportRecord<N>> recs(n), scratch(n);
    for (uint64_t i = 0; i < n; ++i) {
        if (std::fread(recs[i].key.data.data(), 8, N, stdin) != N) return 2;
        if (std::fread(&recs[i].count, 4, 1, stdin) != 1) return 2;
    }
    const ExportRecord<N> *out = nullptr;
    if (bits == 8) {
        out = export_record_radix_sort<N>(recs.data(), scratch.data(), n, k);
    } else {
        if (cpb != 10) return 3;
        try {
            tree::sort::radix_sort_export_records_12bit<N, 10>(recs, scratch, k);
        } catch (const std::exception &) {
            return 4;
        }
        out = recs.data();
    }
    for (uint64_t i = 0; i < n; ++i) {
        std::fwrite(out[i].key.data.data(), 8, N, stdout);
        std::fwrite(&out[i].count, 4, 1, stdout);
    }
    return 0;
}

int main()
{
    uint32_t hdr[4];
    uint64_t n;
    if (std::fread(hdr, 4, 4, stdin) != 4 || std::fread(&n, 8, 1, stdin) != 1) return 2;
    switch (hdr[0]) {
    case 1: return run<1>(hdr[1], hdr[2], hdr[3], n);
    case 2: return run<2>(hdr[1], hdr[2], hdr[3], n);
    default: return 3;
    }
}
