// ita_caller <flags> <nfa> <requests> <out> [scan]
//
// A user program of the reference's C++ surface (include/ITA_regex.h): builds
// ita_scratch (one kernel bit) or ad_scratch (AD_MATCHER set in <flags>),
// allocScratch, then scans every request of <requests> into a
// vector<int>[n] whose entries arrive pre-filled -- with BatchedScan, or with
// Scan one string at a time when the last argument is "scan" -- and writes
// each vector after the call, one line per request, to <out>.  The results
// are the union of the pre-filled ids and the accepted rules, sorted and
// de-duplicated (src/iNFA_kernel.cu:298-306).
//
// <requests> (little endian): u32 n, then per request u32 len, len bytes,
// u32 k, k int32 pre-filled ids.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ITA_regex.h"

namespace {
bool read_u32(FILE* f, uint32_t& v) { return std::fread(&v, 4, 1, f) == 1; }

template <class S>
void run(S& scratch, std::vector<std::string>& text, std::vector<vector<int>>& acc, bool single) {
    allocScratch(scratch);
    if (single) {
        // Scan takes a NUL-terminated string (length strlen, Interface.cu:155)
        for (size_t i = 0; i < text.size(); ++i) Scan(scratch, &text[i][0], &acc[i]);
    } else {
        std::vector<char*> ptr(text.size());
        std::vector<int> len(text.size());
        for (size_t i = 0; i < text.size(); ++i) {
            ptr[i] = &text[i][0];
            len[i] = int(text[i].size());
        }
        BatchedScan(scratch, ptr.data(), len.data(), int(text.size()), acc.data());
    }
    freeScratch(scratch);
}
}  // namespace

int main(int argc, char** argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: ita_caller <flags> <nfa> <requests> <out> [scan]\n");
        return 2;
    }
    const ITA_FLAGS flags = ITA_FLAGS(std::strtoul(argv[1], nullptr, 0));
    const bool single = argc > 5 && std::strcmp(argv[5], "scan") == 0;
    FILE* f = std::fopen(argv[3], "rb");
    if (!f) return 3;
    uint32_t n = 0;
    if (!read_u32(f, n)) return 3;
    std::vector<std::string> text(n);
    std::vector<vector<int>> acc(n);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t len = 0, k = 0;
        if (!read_u32(f, len)) return 3;
        text[i].resize(len);
        if (len && std::fread(&text[i][0], 1, len, f) != len) return 3;
        if (!read_u32(f, k)) return 3;
        acc[i].resize(k);
        if (k && std::fread(acc[i].data(), 4, k, f) != k) return 3;
    }
    std::fclose(f);
    std::vector<char> path(argv[2], argv[2] + std::strlen(argv[2]) + 1);
    if (flags & AD_MATCHER) {
        ad_scratch scratch(flags, path.data());
        run(scratch, text, acc, single);
    } else {
        ita_scratch scratch(flags, path.data());
        run(scratch, text, acc, single);
    }
    FILE* o = std::fopen(argv[4], "w");
    if (!o) return 4;
    for (uint32_t i = 0; i < n; ++i) {
        for (size_t j = 0; j < acc[i].size(); ++j) std::fprintf(o, j ? " %d" : "%d", acc[i][j]);
        std::fprintf(o, "\n");
    }
    std::fclose(o);
    return 0;
}
