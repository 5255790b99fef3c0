// .nfa parser and per-method table construction (see graph.hpp).
#include "graph.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <unordered_map>

namespace admatch {
namespace {

bool blank(char c) { return c == ' ' || c == '\t'; }

struct Span {
    const char* b;
    const char* e;
    bool empty() const { return b >= e; }
    Span trim() const {
        const char* x = b;
        const char* y = e;
        while (x < y && blank(*x)) ++x;
        while (y > x && blank(y[-1])) --y;
        return {x, y};
    }
    bool has(const char* needle) const {
        size_t n = std::strlen(needle);
        for (const char* p = b; p + n <= e; ++p)
            if (std::memcmp(p, needle, n) == 0) return true;
        return false;
    }
};

// Whole-token unsigned decimal (optional '+'), as boost::lexical_cast on a
// state id; negatives are rejected.
bool to_u32(Span s, uint32_t& out) {
    const char* p = s.b;
    if (p < s.e && *p == '+') ++p;
    if (p >= s.e) return false;
    uint64_t v = 0;
    for (; p < s.e; ++p) {
        if (*p < '0' || *p > '9') return false;
        v = v * 10 + uint64_t(*p - '0');
        if (v > 0xffffffffull) return false;
    }
    out = uint32_t(v);
    return true;
}

bool to_i64(Span s, int64_t& out) {
    const char* p = s.b;
    bool neg = false;
    if (p < s.e && (*p == '+' || *p == '-')) neg = *p++ == '-';
    if (p >= s.e) return false;
    int64_t v = 0;
    for (; p < s.e; ++p) {
        if (*p < '0' || *p > '9') return false;
        v = v * 10 + (*p - '0');
        if (v > (int64_t(1) << 40)) return false;
    }
    out = neg ? -v : v;
    return true;
}

template <class F>
void for_each_token(Span s, F&& f) {
    const char* p = s.b;
    while (p < s.e) {
        while (p < s.e && blank(*p)) ++p;
        if (p >= s.e) break;
        const char* q = p;
        while (q < s.e && !blank(*q)) ++q;
        if (!f(Span{p, q})) return;
        p = q;
    }
}

void set_bit(std::vector<uint64_t>& v, int32_t S, uint32_t i) {
    if (i >= uint32_t(S)) return;  // StateVector::set_bit ignores (state_vector.cpp:53)
    v[i >> 6] |= uint64_t(1) << (i & 63);
}

bool edge_less(const Edge& a, const Edge& b) { return a.src != b.src ? a.src < b.src : a.dst < b.dst; }
bool edge_eq(const Edge& a, const Edge& b) { return a.src == b.src && a.dst == b.dst; }

}  // namespace

uint64_t Graph::content_hash() const {
    // FNV-1a over the tables that decide scan results and kernel timing.
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void* p, size_t n) {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 1099511628211ull;
    };
    mix(&S, sizeof S);
    mix(sym_off.data(), sizeof(int32_t) * sym_off.size());
    mix(list.data(), sizeof(Edge) * list.size());
    mix(persis.data(), 8 * persis.size());
    mix(init.data(), 8 * init.size());
    return h;
}

int load_nfa(const char* path, Graph& g, std::string& err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = std::string("cannot open ") + path;
        return 1;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    g = Graph();
    std::vector<std::vector<Edge>> per(kSymbols);
    std::vector<std::vector<int32_t>> acc_rules_of;
    std::vector<uint8_t> has_acc;
    bool have_count = false;
    int64_t lineno = 0;
    size_t pos = 0;
    auto fail = [&](const char* what) {
        err = std::string(what) + " at line " + std::to_string(lineno);
        return 2;
    };
    while (pos < text.size()) {
        size_t nl = text.find('\n', pos);
        if (nl == std::string::npos) nl = text.size();
        Span line{text.data() + pos, text.data() + nl};
        pos = nl + 1;
        ++lineno;
        if (!have_count) {
            // first non-comment line carries the state count (:91-103)
            if (!line.empty() && *line.b == '#') continue;
            const char* p = line.b;
            while (p < line.e && (blank(*p) || *p == '\r')) ++p;
            const char* q = p;
            while (q < line.e && *q >= '0' && *q <= '9') ++q;
            int64_t s = 0;
            if (!to_i64(Span{p, q}, s) || s <= 0 || s > kMaxStates) return fail("bad state count");
            g.S = int32_t(s);
            g.W = (g.S + 63) / 64;
            g.init.assign(g.W, 0);
            g.persis.assign(g.W, 0);
            g.accept.assign(g.W, 0);
            acc_rules_of.assign(g.S, {});
            has_acc.assign(g.S, 0);
            have_count = true;
            continue;
        }
        if (line.empty() || *line.b == '#') continue;
        if (line.e[-1] == '\r') --line.e;
        if (line.empty()) continue;
        // "<lhs> : <rhs>" with exactly one ':' (:142-148)
        const char* colon = nullptr;
        int ncolon = 0;
        for (const char* p = line.b; p < line.e; ++p)
            if (*p == ':') colon = p, ++ncolon;
        if (ncolon != 1) {
            ++g.n_bad_lines;
            continue;
        }
        Span lhs = Span{line.b, colon};
        while (lhs.e > lhs.b && blank(lhs.e[-1])) --lhs.e;
        Span rhs = Span{colon + 1, line.e};
        while (rhs.b < rhs.e && blank(*rhs.b)) ++rhs.b;
        const char* arrow = nullptr;
        int narrow = 0;
        for (const char* p = lhs.b; p + 1 < lhs.e; ++p)
            if (p[0] == '-' && p[1] == '>') arrow = p, ++narrow;
        if (narrow == 1) {
            Span a{lhs.b, arrow}, b{arrow + 2, lhs.e};
            while (a.e > a.b && blank(a.e[-1])) --a.e;
            while (b.b < b.e && blank(*b.b)) ++b.b;
            uint32_t src, dst;
            if (!to_u32(a, src) || !to_u32(b, dst)) return fail("bad state id");
            if (src >= uint32_t(g.S) || dst >= uint32_t(g.S)) return fail("state id out of range");
            const char* bad = nullptr;
            for_each_token(rhs, [&](Span atom) {
                const char* bar = static_cast<const char*>(std::memchr(atom.b, '|', size_t(atom.e - atom.b)));
                int64_t lo, hi;
                if (!bar) {
                    if (!to_i64(atom, lo)) return (bad = "bad symbol"), false;
                    hi = lo;
                } else {
                    const char* bar2 = static_cast<const char*>(std::memchr(bar + 1, '|', size_t(atom.e - bar - 1)));
                    if (!to_i64(Span{atom.b, bar}, lo) || !to_i64(Span{bar + 1, bar2 ? bar2 : atom.e}, hi))
                        return (bad = "bad symbol range"), false;
                }
                if (src == dst && lo == 0 && hi == 255) {  // persistent state (:184-189)
                    set_bit(g.persis, g.S, src);
                    ++g.n_persis_lines;
                    return true;
                }
                if (lo <= hi && (lo < 0 || hi > kWordBoundary)) return (bad = "symbol out of range"), false;
                for (int64_t c = lo; c <= hi; ++c) per[size_t(c)].push_back({src, dst});
                g.n_raw_transitions += hi >= lo ? hi - lo + 1 : 0;
                return true;
            });
            if (bad) return fail(bad);
        } else if (rhs.has("accepting")) {  // (:261-283)
            uint32_t st;
            if (narrow != 0 || !to_u32(lhs, st)) return fail("bad accept state");
            set_bit(g.accept, g.S, st);
            std::vector<int32_t> rules;
            int tok = 0;
            const char* bad = nullptr;
            for_each_token(rhs, [&](Span t) {
                if (tok++ == 0) return true;
                int64_t r;
                if (!to_i64(t, r)) return (bad = "bad rule id"), false;
                rules.push_back(int32_t(r));
                return true;
            });
            if (bad) return fail(bad);
            if (tok <= 1) continue;  // "No accept rules" (:271-274)
            if (st < uint32_t(g.S) && !has_acc[st]) {
                has_acc[st] = 1;
                acc_rules_of[st] = std::move(rules);
            }
        } else if (rhs.has("initial")) {
            uint32_t st;
            if (narrow != 0 || !to_u32(lhs, st)) return fail("bad initial state");
            set_bit(g.init, g.S, st);
        } else {
            ++g.n_bad_lines;  // "Cannot parse" (:291-293)
        }
    }
    if (!have_count) {
        err = "no state count line";
        return 3;
    }
    // merge_transitions (:369-411)
    int64_t total = 0;
    for (int c = 0; c < kSymbols; ++c) {
        auto& v = per[c];
        std::sort(v.begin(), v.end(), edge_less);
        v.erase(std::unique(v.begin(), v.end(), edge_eq), v.end());
        total += int64_t(v.size());
    }
    if (total >= (int64_t(1) << 31)) {
        err = "too many transitions";
        return 4;
    }
    g.list.reserve(size_t(total));
    for (int c = 0; c < kSymbols; ++c) {
        g.sym_off[c] = int32_t(g.list.size());
        g.list.insert(g.list.end(), per[c].begin(), per[c].end());
    }
    g.sym_off[kSymbols] = int32_t(g.list.size());
    g.acc_off.assign(size_t(g.S) + 1, 0);
    for (int32_t s = 0; s < g.S; ++s) g.acc_off[s + 1] = g.acc_off[s] + int32_t(acc_rules_of[s].size());
    g.acc_rules.reserve(size_t(g.acc_off[g.S]));
    for (int32_t s = 0; s < g.S; ++s) g.acc_rules.insert(g.acc_rules.end(), acc_rules_of[s].begin(), acc_rules_of[s].end());
    return 0;
}

void build_tko(Graph& g, int32_t threshold) {
    // Top-k offset partition of transition_graph.cpp:197-245,298-357: per
    // byte symbol, transitions (plus a (s,s) self-loop per persistent state)
    // grouped by d = dst - src; groups larger than the threshold, largest
    // first (ties: smaller offset first), at most kMaxK of them, become
    // shift groups; the rest is leftover.
    const int32_t thr = threshold >= 0 ? threshold : g.S / 16;
    g.tko_threshold = thr;
    g.tko_koff.assign(kSymbols + 1, 0);
    g.tko_loff.assign(kSymbols + 1, 0);
    g.tko_shift.clear();
    g.tko_lim.clear();
    g.tko_left.clear();
    std::vector<uint32_t> pstates;
    for (int32_t w = 0; w < g.W; ++w)
        for (uint64_t v = g.persis[w]; v; v &= v - 1) pstates.push_back(uint32_t(w * 64 + __builtin_ctzll(v)));
    struct Item {
        int64_t d;
        Edge e;
    };
    std::vector<Item> items;
    for (int c = 0; c < kSymbols; ++c) {
        g.tko_koff[c] = int32_t(g.tko_shift.size());
        g.tko_loff[c] = int32_t(g.tko_left.size());
        const Edge* b = g.list.data() + g.sym_off[c];
        const Edge* e = g.list.data() + g.sym_off[c + 1];
        if (c == kWordBoundary) {
            g.tko_left.insert(g.tko_left.end(), b, e);
            continue;
        }
        std::vector<Edge> all(b, e);
        for (uint32_t s : pstates) all.push_back({s, s});
        std::sort(all.begin(), all.end(), edge_less);
        all.erase(std::unique(all.begin(), all.end(), edge_eq), all.end());
        items.clear();
        for (const Edge& x : all) items.push_back({int64_t(x.dst) - int64_t(x.src), x});
        std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& b) { return a.d < b.d; });
        struct Grp {
            int64_t d;
            size_t b, n;
        };
        std::vector<Grp> grps;
        for (size_t i = 0; i < items.size();) {
            size_t j = i;
            while (j < items.size() && items[j].d == items[i].d) ++j;
            grps.push_back({items[i].d, i, j - i});
            i = j;
        }
        std::stable_sort(grps.begin(), grps.end(), [](const Grp& a, const Grp& b) { return a.n > b.n; });
        std::vector<Edge> rest;
        int taken = 0;
        for (const Grp& gr : grps) {
            if (gr.n > size_t(thr) && taken < kMaxK) {
                ++taken;
                g.tko_shift.push_back(int32_t(gr.d));
                size_t base = g.tko_lim.size();
                g.tko_lim.resize(base + size_t(g.W), 0);
                for (size_t k = gr.b; k < gr.b + gr.n; ++k) {
                    uint32_t s = items[k].e.src;
                    g.tko_lim[base + (s >> 6)] |= uint64_t(1) << (s & 63);
                }
            } else {
                for (size_t k = gr.b; k < gr.b + gr.n; ++k) rest.push_back(items[k].e);
            }
        }
        std::sort(rest.begin(), rest.end(), edge_less);
        g.tko_left.insert(g.tko_left.end(), rest.begin(), rest.end());
    }
    g.tko_koff[kSymbols] = int32_t(g.tko_shift.size());
    g.tko_loff[kSymbols] = int32_t(g.tko_left.size());
}

// Entries the two-byte rows may hold: kMaxPairEntries, or less when the
// environment variable ADM_MAX_PAIR_ENTRIES says so (tests set 0 to run the
// engine without the rows).
static size_t pair_limit() {
    const char* e = std::getenv("ADM_MAX_PAIR_ENTRIES");
    if (!e || !*e) return kMaxPairEntries;
    return std::min<size_t>(kMaxPairEntries, size_t(std::strtoull(e, nullptr, 10)));
}

void build_as(Graph& g) {
    const size_t rows = size_t(g.S) * 256;
    g.as_off.assign(rows + 1, 0);
    for (int c = 0; c < 256; ++c)
        for (int32_t t = g.sym_off[c]; t < g.sym_off[c + 1]; ++t) g.as_off[size_t(g.list[t].src) * 256 + c + 1]++;
    for (size_t r = 0; r < rows; ++r) g.as_off[r + 1] += g.as_off[r];
    g.as_dst.assign(g.as_off[rows], 0);
    std::vector<uint32_t> fill(g.as_off.begin(), g.as_off.end() - 1);
    // each symbol's list is sorted by (src,dst), so rows come out ascending
    for (int c = 0; c < 256; ++c)
        for (int32_t t = g.sym_off[c]; t < g.sym_off[c + 1]; ++t) {
            const Edge& e = g.list[t];
            g.as_dst[fill[size_t(e.src) * 256 + c]++] = e.dst;
        }

    // device form: rows plus the persistent self-loop
    auto persistent = [&](uint32_t s) { return (g.persis[s >> 6] >> (s & 63)) & 1ull; };
    auto always = [&](uint32_t s) { return persistent(s) && ((g.init[s >> 6] >> (s & 63)) & 1ull); };
    g.as_tab.assign(size_t(g.S) * 256, 0);
    g.as_multi.clear();
    g.as_ok = true;
    std::vector<uint32_t> row;
    for (int32_t s = 0; s < g.S && g.as_ok; ++s) {
        const bool p = persistent(uint32_t(s));
        for (int c = 0; c < 256; ++c) {
            const uint32_t* b = g.as_dst.data() + g.as_off[size_t(s) * 256 + c];
            const uint32_t* e = g.as_dst.data() + g.as_off[size_t(s) * 256 + c + 1];
            row.assign(b, e);
            if (p && !std::binary_search(row.begin(), row.end(), uint32_t(s)))
                row.insert(std::lower_bound(row.begin(), row.end(), uint32_t(s)), uint32_t(s));
            // always-active destinations are held implicitly by the AS step
            row.erase(std::remove_if(row.begin(), row.end(), always), row.end());
            uint32_t& ent = g.as_tab[size_t(s) * 256 + c];
            if (row.empty()) continue;
            if (row.size() == 1) {
                ent = (1u << 24) | row[0];
                continue;
            }
            if (row.size() == 2 && (row[0] == uint32_t(s) || row[1] == uint32_t(s))) {
                ent = (kAsSelfRow << 24) | (row[0] == uint32_t(s) ? row[1] : row[0]);
                continue;
            }
            const size_t idx = g.as_multi.size();
            if (idx + row.size() + 1 >= (size_t(1) << 24)) {
                g.as_ok = false;
                break;
            }
            if (row.size() >= kAsSelfRow) {
                g.as_multi.push_back(uint32_t(row.size()));
                ent = (255u << 24) | uint32_t(idx);
            } else {
                ent = (uint32_t(row.size()) << 24) | uint32_t(idx);
            }
            g.as_multi.insert(g.as_multi.end(), row.begin(), row.end());
        }
    }
    if (!g.as_ok) {
        g.as_tab.clear();
        g.as_multi.clear();
    }
    // rows of the always-active states, merged per byte
    g.alw_off.assign(257, 0);
    g.alw_dst.clear();
    std::vector<uint32_t> alw;
    g.init_explicit.assign(size_t(g.W), 0);
    for (int32_t w = 0; w < g.W; ++w) {
        g.init_explicit[w] = g.init[w] & ~g.persis[w];
        for (uint64_t v = g.persis[w] & g.init[w]; v; v &= v - 1) alw.push_back(uint32_t(w * 64 + __builtin_ctzll(v)));
    }
    g.alw_states = alw;
    for (int c = 0; c < 256; ++c) {
        g.alw_off[c] = int32_t(g.alw_dst.size());
        row.clear();
        for (uint32_t s : alw) {
            row.push_back(s);
            const uint32_t* b = g.as_dst.data() + g.as_off[size_t(s) * 256 + c];
            const uint32_t* e = g.as_dst.data() + g.as_off[size_t(s) * 256 + c + 1];
            row.insert(row.end(), b, e);
        }
        std::sort(row.begin(), row.end());
        row.erase(std::unique(row.begin(), row.end()), row.end());
        // bit 31: always active; bit 30: a first position (reached from an
        // always-active state, not itself always active)
        for (uint32_t d : row) g.alw_dst.push_back(always(d) ? (d | 0x80000000u) : (d | 0x40000000u));
    }
    g.alw_off[256] = int32_t(g.alw_dst.size());

    // two-byte rows of the first positions
    g.pair_off.assign(size_t(256) * 256 + 1, 0);
    g.pair_dst.clear();
    bool pair_ok = true;
    const size_t max_pair = pair_limit();
    std::vector<uint32_t> d1;
    for (int c1 = 0; c1 < 256 && pair_ok; ++c1) {
        d1.clear();
        for (int32_t k = g.alw_off[c1]; k < g.alw_off[c1 + 1]; ++k)
            if (!(g.alw_dst[k] & 0x80000000u)) d1.push_back(g.alw_dst[k] & 0x00ffffffu);
        for (int c2 = 0; c2 < 256; ++c2) {
            g.pair_off[size_t(c1) * 256 + c2] = uint32_t(g.pair_dst.size());
            row.clear();
            for (uint32_t s : d1) {
                const uint32_t* b = g.as_dst.data() + g.as_off[size_t(s) * 256 + c2];
                const uint32_t* e = g.as_dst.data() + g.as_off[size_t(s) * 256 + c2 + 1];
                row.insert(row.end(), b, e);
                if (persistent(s)) row.push_back(s);
            }
            std::sort(row.begin(), row.end());
            row.erase(std::unique(row.begin(), row.end()), row.end());
            if (g.pair_dst.size() + row.size() > max_pair) {
                pair_ok = false;
                break;
            }
            for (uint32_t d : row)
                if (!always(d)) g.pair_dst.push_back(d);
        }
    }
    if (!pair_ok) {
        g.pair_off.clear();
        g.pair_dst.clear();
    } else {
        g.pair_off[size_t(256) * 256] = uint32_t(g.pair_dst.size());
    }
}

void build_wb_rows(Graph& g) {
    g.wbs_off.clear();
    g.wbs_dst.clear();
    g.wbi_off.clear();
    g.wbi_dst.clear();
    if (g.wb_count() == 0 || !g.wb_chain_free()) return;
    // CSR by source (the boundary list is sorted by (src, dst))
    const int32_t b0 = g.sym_off[kWordBoundary], b1 = g.sym_off[kWordBoundary + 1];
    g.wbs_off.assign(size_t(g.S) + 1, 0);
    for (int32_t t = b0; t < b1; ++t) g.wbs_off[size_t(g.list[size_t(t)].src) + 1] += 1;
    for (int32_t s = 0; s < g.S; ++s) g.wbs_off[size_t(s) + 1] += g.wbs_off[size_t(s)];
    g.wbs_dst.resize(size_t(b1 - b0));
    for (int32_t t = b0; t < b1; ++t) g.wbs_dst[size_t(t - b0)] = g.list[size_t(t)].dst;
    // rows of the implicit sets
    std::vector<uint32_t> row;
    g.wbi_off.assign(258, 0);
    for (int r = 0; r <= 256; ++r) {
        g.wbi_off[size_t(r)] = int32_t(g.wbi_dst.size());
        row.clear();
        auto add = [&](uint32_t st) {
            for (uint32_t k = g.wbs_off[st]; k < g.wbs_off[size_t(st) + 1]; ++k) row.push_back(g.wbs_dst[k]);
        };
        if (r < 256) {
            for (int32_t k = g.alw_off[size_t(r)]; k < g.alw_off[size_t(r) + 1]; ++k) add(g.alw_dst[size_t(k)] & 0x00ffffffu);
        } else {
            for (uint32_t st : g.alw_states) add(st);
        }
        std::sort(row.begin(), row.end());
        row.erase(std::unique(row.begin(), row.end()), row.end());
        g.wbi_dst.insert(g.wbi_dst.end(), row.begin(), row.end());
    }
    g.wbi_off[257] = int32_t(g.wbi_dst.size());
}

}  // namespace admatch
