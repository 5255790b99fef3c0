// sequence <nfa> <string>: scans one string with the plain iNFA kernel and
// prints profile and results (the reference's tools/sequence.cpp:11-17).
#include <vector>

#include "ITA_regex.h"

int main(int argc, char** argv) {
    if (argc < 3) return 1;
    ita_scratch scratch(INFA_KERNEL | PROFILER_MODE | SHOW_RESULTS, argv[1]);
    allocScratch(scratch);
    vector<int> acc[1];
    Scan(scratch, argv[2], acc);
    freeScratch(scratch);
    return 0;
}
