// The reference's C++ surface (include/ITA_regex.h) over the C ABI.
#include "../../include/ITA_regex.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

[[noreturn]] void die(const char* what) {
    std::fprintf(stderr, "Error: %s: %s\n", what, ita_last_error());
    std::exit(-1);
}

ita_handle* make(ITA_FLAGS flags, char* nfa) {
    ita_handle* h = nullptr;
    const int rc = ita_create(flags, nfa, &h);
    if (rc == ITA_ERR_ARG) {
        std::fprintf(stderr, "Error: invalid flags 0x%x: %s\n", unsigned(flags), ita_last_error());
        std::abort();  // the reference asserts (Interface.cu:17)
    }
    if (rc != ITA_OK) {
        std::fprintf(stderr, "Error: load NFA file %s\n", nfa);  // Interface.cu:28-31
        std::exit(-1);
    }
    return h;
}

void batched(ita_handle* h, char** text, int* text_len, int n, vector<int>* out) {
    if (n <= 0) return;
    std::vector<int32_t> off(size_t(n) + 1);
    std::vector<int32_t> ids(size_t(n) * 4 + 16);
    int rc = ita_batched_scan(h, reinterpret_cast<const uint8_t* const*>(text), text_len, n, off.data(), ids.data(),
                              int64_t(ids.size()));
    if (rc == ITA_ERR_CAPACITY) {
        // the scan's results are kept by the handle: fetch them, no rescan
        ids.resize(size_t(off[n]));
        if (ita_last_results(h, ids.data(), int64_t(ids.size())) != int64_t(off[n])) die("BatchedScan");
        rc = ITA_OK;
    }
    if (rc != ITA_OK) die("BatchedScan");
    // append, then sort + unique the whole vector (src/iNFA_kernel.cu:298-306)
    for (int i = 0; i < n; ++i) {
        vector<int>& v = out[i];
        v.insert(v.end(), ids.begin() + off[i], ids.begin() + off[i + 1]);
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
}

void single(ita_handle* h, char* text, vector<int>* out) {
    int len = int(std::strlen(text));  // Interface.cu:155
    batched(h, &text, &len, 1, out);
}

}  // namespace

ita_scratch::ita_scratch(ITA_FLAGS flags, char* nfa) : flag(flags), handle(nullptr) {
    if (flags & AD_MATCHER) {
        std::fprintf(stderr, "Error: AD_MATCHER needs ad_scratch\n");
        std::abort();
    }
    handle = make(flags, nfa);
}
ita_scratch::~ita_scratch() { ita_destroy(handle); }

ad_scratch::ad_scratch(ITA_FLAGS flags, char* nfa) : flag(flags), handle(nullptr) {
    handle = make(ITA_FLAGS(flags | AD_MATCHER), nfa);
}
ad_scratch::~ad_scratch() { ita_destroy(handle); }

void allocScratch(struct ita_scratch& s) {
    if (ita_alloc(s.handle) != ITA_OK) die("allocScratch");
}
void allocScratch(struct ad_scratch& s) {
    if (ita_alloc(s.handle) != ITA_OK) die("allocScratch");
}
void freeScratch(struct ita_scratch& s) { ita_free(s.handle); }
void freeScratch(struct ad_scratch& s) { ita_free(s.handle); }

void Scan(struct ita_scratch& s, char* text, vector<int>* out) { single(s.handle, text, out); }
void Scan(struct ad_scratch& s, char* text, vector<int>* out) { single(s.handle, text, out); }

void BatchedScan(struct ita_scratch& s, char** text, int* text_len, int n, vector<int>* out) {
    batched(s.handle, text, text_len, n, out);
}
void BatchedScan(struct ad_scratch& s, char** text, int* text_len, int n, vector<int>* out) {
    batched(s.handle, text, text_len, n, out);
}
