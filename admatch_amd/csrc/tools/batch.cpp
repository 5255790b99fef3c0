// batch -f <nfa> -i <input> -n <count> [-k iNFA|TKO|AS|AD|HEUR|LEARNED] [-p] [-s]
// The reference's tools/batch.cpp:32-113: the input file's lines are joined
// without separators and the same string is scanned <count> times.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "ITA_regex.h"

static std::string join_lines(const char* path) {
    std::ifstream fin(path);
    std::string line, all;
    while (std::getline(fin, line)) all += line;  // batch.cpp:14-30
    return all;
}

int main(int argc, char** argv) {
    std::string nfa, input, kernel = "iNFA";
    int count = 0;
    ITA_FLAGS flag = 0;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> const char* { return i + 1 < argc ? argv[++i] : ""; };
        if (a == "-f" || a == "--nfa") nfa = next();
        else if (a == "-i" || a == "--input") input = next();
        else if (a == "-n" || a == "--num") count = std::atoi(next());
        else if (a == "-k") kernel = next();
        else if (a == "-p") flag |= PROFILER_MODE;
        else if (a == "-s") flag |= SHOW_RESULTS;
        else if (a == "-h" || a == "--help") {
            std::printf("batch -f nfa -i input -n count [-k iNFA|TKO|AS|AD|HEUR|LEARNED] [-p] [-s]\n");
            return 0;
        } else {
            std::printf("Undefined Arguments\n");
            return EXIT_FAILURE;
        }
    }
    if (count <= 0) {
        std::fprintf(stderr, "Error: invalid string_count value %d\n", count);
        return EXIT_FAILURE;
    }
    std::string str = join_lines(input.c_str());
    std::vector<char*> texts(static_cast<size_t>(count), &str[0]);
    std::vector<int> lens(static_cast<size_t>(count), int(str.size()));
    std::vector<vector<int>> acc(static_cast<size_t>(count));
    std::vector<char> path(nfa.begin(), nfa.end());
    path.push_back(0);
    if (kernel == "AD" || kernel == "HEUR" || kernel == "LEARNED") {
        if (kernel == "HEUR") flag |= HEURISTIC_SELECTOR;
        if (kernel == "LEARNED") flag |= LEARNED_SELECTOR;
        ad_scratch scratch(flag, path.data());
        allocScratch(scratch);
        BatchedScan(scratch, texts.data(), lens.data(), count, acc.data());
        freeScratch(scratch);
    } else {
        if (kernel == "iNFA") flag |= INFA_KERNEL;
        if (kernel == "TKO") flag |= TKO_KERNEL;
        if (kernel == "AS") flag |= AS_KERNEL;
        ita_scratch scratch(flag, path.data());
        allocScratch(scratch);
        BatchedScan(scratch, texts.data(), lens.data(), count, acc.data());
        freeScratch(scratch);
    }
    return EXIT_SUCCESS;
}
