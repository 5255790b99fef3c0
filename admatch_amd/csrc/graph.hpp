// Host-side compiled rule set: the .nfa parser and the per-method tables.
//
// Restates the reference's TransitionGraph (src/transition_graph.h:55-104,
// src/transition_graph.cpp:82-411) for one process-wide graph that feeds all
// three step methods, with these MI355X-side choices:
//   * 32-bit state ids (the reference's unsigned short caps a rule set at
//     65,535 states, src/common.h:6);
//   * state bit vectors in u64 words, bit i in word i/64, LSB first (the
//     reference uses u32 words with the same bit order, state_vector.cpp:52-62);
//   * the active-state table is sparse: one row of destination ids per
//     (state, byte) pair instead of the dense [257][S][W] bit table
//     (transition_graph.cpp:119-126), which would need 257*S^2/8 bytes.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <algorithm>
#include <iterator>
#include <vector>

namespace admatch {

constexpr int kSymbols = 257;      // 256 byte values + word boundary (src/common.h:13)
constexpr int kWordBoundary = 256;  // src/common.h:15
constexpr int kMaxK = 1024;        // src/common.h:21
constexpr size_t kMaxPairEntries = size_t(1) << 24;
constexpr int32_t kMaxStates = 1 << 24;  // ids fit the 24-bit payload of the AS table
// AS table count of a row of two destinations one of which is the source
// state itself: the other one is the entry's payload, so the row needs no
// load from as_multi (a persistent state that advances on the byte).
constexpr uint32_t kAsSelfRow = 254;

struct Edge {
    uint32_t src, dst;
};

struct Graph {
    int32_t S = 0;  // states
    int32_t W = 0;  // u64 words per state vector
    std::vector<uint64_t> init, persis, accept;

    // iNFA list: per-symbol (src,dst) sorted unique, CSR over 257 symbols
    // (merge_transitions, transition_graph.cpp:369-411).
    std::array<int32_t, kSymbols + 1> sym_off{};
    std::vector<Edge> list;

    // accept map: state -> rule ids, first accepting line per state wins
    // (transition_graph.cpp:282-283).  CSR over states.
    std::vector<int32_t> acc_off, acc_rules;

    // TKO partition (transition_graph.cpp:298-357).  koff: 258 prefix sums
    // of the top-k group count per symbol; shift[K]: offset dst-src of each
    // group; lim[K*W]: source mask of each group; loff/left: the leftover
    // transitions per symbol (word-boundary transitions are all leftover).
    int32_t tko_threshold = 0;
    std::vector<int32_t> tko_koff, tko_shift, tko_loff;
    std::vector<uint64_t> tko_lim;
    std::vector<Edge> tko_left;

    // Sparse active-state rows: destinations of state s on byte c are
    // as_dst[as_off[s*256+c] .. as_off[s*256+c+1]), ascending.  Persistent
    // 0|255 self-loops are not stored (every method applies the persistent
    // mask instead).
    std::vector<uint32_t> as_off, as_dst;

    // Device form of the AS rows, one 32-bit entry per (state, byte), state
    // major (as_tab[s*256 + c], so the 256 entries of an active state share
    // one kilobyte of cache lines): count in the top 8 bits (0 = empty, 255 =
    // long row whose count is as_multi[idx], kAsSelfRow = the source state
    // plus the payload), then either the single destination (count 1 and
    // kAsSelfRow) or the row's index into as_multi.  Rows of
    // persistent states include the state itself, so the AS step needs no
    // separate persistent mask.  Always-active destinations are left out:
    // the AS step holds those states implicitly.
    std::vector<uint32_t> as_tab, as_multi;
    bool as_ok = false;  // false when the rows do not fit the 24-bit index
    // Always-active states (initial and persistent: active at every step).
    // Their combined rows per byte, ascending, with bit 31 marking a
    // destination that is itself always active: alw_dst[alw_off[c] ..
    // alw_off[c+1]).
    std::vector<int32_t> alw_off;
    std::vector<uint32_t> alw_dst;
    std::vector<uint32_t> alw_states;    // the always-active states
    std::vector<uint64_t> init_explicit; // initial states that are not always active
    // Two-byte rows: the states the always-active rows reach on byte c1
    // (the first positions, D1(c1)) are active at the next step whatever
    // else happened, so their rows on the next byte c2 are merged ahead of
    // time: pair_dst[pair_off[c1*256+c2] .. pair_off[c1*256+c2+1]).  Empty
    // when the merged rows would exceed kMaxPairEntries (or the limit in
    // ADM_MAX_PAIR_ENTRIES); the engine then lists the first positions
    // before each AS step.
    std::vector<uint32_t> pair_off, pair_dst;
    // Word-boundary rows, built when the boundary list is chain-free (one
    // parallel pass equals the in-place pass): the boundary successors of
    // each state, wbs_dst[wbs_off[s] .. wbs_off[s+1]), and of each implicit
    // set, wbi_dst[wbi_off[r] .. wbi_off[r+1]) with r = the byte of an
    // always-active row (the always-active states and the first positions
    // reached on it) or r = 256 for the always-active states alone.  Empty
    // otherwise.
    std::vector<uint32_t> wbs_off, wbs_dst, wbi_dst;
    std::vector<int32_t> wbi_off;

    // parse statistics
    int64_t n_persis_lines = 0, n_bad_lines = 0, n_raw_transitions = 0;

    int32_t wb_count() const { return sym_off[kWordBoundary + 1] - sym_off[kWordBoundary]; }
    // true when no word-boundary destination is a word-boundary source: the
    // in-place pass in (src, dst) order then equals one parallel pass
    bool wb_chain_free() const {
        std::vector<uint32_t> src, dst;
        for (int32_t t = sym_off[kWordBoundary]; t < sym_off[kWordBoundary + 1]; ++t) {
            src.push_back(list[size_t(t)].src);
            dst.push_back(list[size_t(t)].dst);
        }
        std::sort(src.begin(), src.end());
        std::sort(dst.begin(), dst.end());
        std::vector<uint32_t> both;
        std::set_intersection(src.begin(), src.end(), dst.begin(), dst.end(), std::back_inserter(both));
        return both.empty();
    }
    uint64_t content_hash() const;
};

// Parse a .nfa file (format of transition_graph.cpp:82-294).  Returns 0 on
// success, otherwise a nonzero code and a message in err.
int load_nfa(const char* path, Graph& g, std::string& err);

// TKO top-k partition; threshold < 0 selects the reference's S/16.
void build_tko(Graph& g, int32_t threshold = -1);

// Sparse active-state rows from the iNFA lists, and their device form.
void build_as(Graph& g);

// Word-boundary rows (after build_as); left empty unless chain-free.
void build_wb_rows(Graph& g);

}  // namespace admatch
