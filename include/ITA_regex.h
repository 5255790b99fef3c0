// ITA_regex.h -- the reference's C++ API (include/ITA_regex.h:1-70 of
// QimingZheng/AdMatch), same names, signatures and flag values, implemented
// over the C ABI in admatch.h (admatch_amd/csrc/ita_regex.cpp).
//
// Differences a caller can observe:
//   * the scratch structs hold an opaque handle instead of the reference's
//     TransitionGraph pointer and raw device pointers (Interface.cu:30-46);
//   * a rule set may have more than 65,535 states;
//   * ad_scratch accepts two extra flag bits, LEARNED_SELECTOR and
//     HEURISTIC_SELECTOR, to pick the engine's step selector.
// Error behaviour follows the reference: a bad flag combination aborts (the
// reference asserts, Interface.cu:17), a load failure prints
// "Error: load NFA file <path>" and exits with -1 (Interface.cu:28-31);
// device errors, which the reference does not check, print and exit too.
#ifndef ITA_REGEX_H
#define ITA_REGEX_H

#include <vector>

#include "admatch.h"

using namespace std;

// flags: contain basic configurations
typedef unsigned short int ITA_FLAGS;

// Kernel Selection
#define INFA_KERNEL ITA_INFA_KERNEL
#define AS_KERNEL ITA_AS_KERNEL
#define TKO_KERNEL ITA_TKO_KERNEL

// Matcher Selection (default not selected)
#define AD_MATCHER ITA_AD_MATCHER
#define LEARNED_SELECTOR ITA_LEARNED_SELECTOR
#define HEURISTIC_SELECTOR ITA_HEURISTIC_SELECTOR

// Profiler mode on/off
#define PROFILER_MODE ITA_PROFILER_MODE

// Show results on/off
#define SHOW_RESULTS ITA_SHOW_RESULTS

struct ita_scratch {
    ita_scratch(ITA_FLAGS flags, char *nfa);
    ~ita_scratch();
    ITA_FLAGS flag;
    ita_handle *handle;
};

struct ad_scratch {
    ad_scratch(ITA_FLAGS flags, char *nfa);
    ~ad_scratch();
    ITA_FLAGS flag;
    ita_handle *handle;
};

// alloc gpu-mem and record them in scratch
void allocScratch(struct ita_scratch &scratch);
void allocScratch(struct ad_scratch &scratch);

// free gpu-mem
void freeScratch(struct ita_scratch &scratch);
void freeScratch(struct ad_scratch &scratch);

// single input string scan (length = strlen(text)); accepted rule ids are
// appended to accepted_rules[0], which is then sorted and de-duplicated
void Scan(struct ita_scratch &scratch, char *text, vector<int> *accepted_rules);
void Scan(struct ad_scratch &scratch, char *text, vector<int> *accepted_rules);

// batched scan of str_count strings; accepted_rules[i] receives string i's
// rule ids (appended, then sorted and de-duplicated)
void BatchedScan(struct ita_scratch &scratch, char **text, int *text_len, int str_count,
                 vector<int> *accepted_rules);
void BatchedScan(struct ad_scratch &scratch, char **text, int *text_len, int str_count,
                 vector<int> *accepted_rules);

#endif
