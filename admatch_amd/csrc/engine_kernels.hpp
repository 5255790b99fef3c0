// Device code of the adaptive scan engine for gfx950 (wave64, 160 KB LDS per
// CU), shared by the translation units that instantiate its kernels.
//
//  * engine_kernel: a persistent grid whose waves pull strings longest-first
//    from a work queue; one wave owns one string, so a step needs no
//    workgroup barrier.  Each step runs one of three methods, chosen per
//    step by the selector:
//      iNFA  scatter over the byte's (src,dst) list      (src/iNFA_kernel.cu:89-132)
//      TKO   top-k offset groups as word shifts + leftover scatter
//                                                        (src/TKO_kernel.cu:89-188)
//      AS    destination rows of the active states       (src/AS_kernel.cu:74-125)
//    The active set is kept both as a bitmap and, while it has at most CAP
//    states, as a list, so the AS and iNFA methods never touch all W words.
//  * engine_h_kernel: the same step loop with the set held as an active list
//    plus an LDS hash table, the bitmaps in device memory.
//  * learned selector: the MLP is evaluated once per (byte, active-count
//    bucket, list validity) into a 2-bit policy table (policy_kernel in
//    kernels.hip); a step reads its method with one scalar load.
#pragma once

#include "device.hpp"
#include "graph.hpp"

namespace admatch {

// ------------------------------------------------------------------ common
// List entries carry bit 31 for always-active states (their rows come from
// the per-byte always-active list instead of the AS table).
constexpr uint32_t kTag = 0x80000000u;
constexpr uint32_t kIdMask = 0x00ffffffu;

__device__ __forceinline__ bool set_new(uint64_t* bits, uint32_t d) {
    const unsigned long long m = 1ull << (d & 63);
    const unsigned long long old = atomicOr(reinterpret_cast<unsigned long long*>(&bits[d >> 6]), m);
    return !(old & m);
}

__device__ __forceinline__ void set_bit_atomic(uint64_t* bits, uint32_t d) {
    atomicOr(reinterpret_cast<unsigned long long*>(&bits[d >> 6]), 1ull << (d & 63));
}

// Destination sets of a step.  BitSet: a bitmap of all S states.  HashSet: an
// open-addressing table of H slots in LDS (linear probing, kEmpty marks a free
// slot); it holds at most CAP + 64 keys, so steps that use it stop early
// (kBounded) once the list passes CAP and are redone on the bitmaps.
constexpr uint32_t kEmpty = 0xffffffffu;

struct BitSet {
    static constexpr bool kBounded = false;
    uint64_t* bits;
    __device__ __forceinline__ bool insert(uint32_t d) const { return set_new(bits, d); }
};

// Each table is followed by 64 sink slots, one per lane, that permanently
// hold kSink: a lane with nothing to insert compare-and-swaps its sink slot,
// so inserts need no lane masking.
constexpr uint32_t kSink = 0xfffffffeu;

template <int H>
struct HashSet {
    static_assert((H & (H - 1)) == 0, "table size must be a power of two");
    static constexpr bool kBounded = true;
    static constexpr int kSlots = H + 64;  // u32 per table including the sink slots
    uint32_t* tab;
    __device__ __forceinline__ static uint32_t slot(uint32_t d) {
        return (d * 2654435761u) >> (32 - __builtin_ctz(uint32_t(H)));
    }
    // true if d was not in the set.  One compare-and-swap on the home slot;
    // the probe loop runs only after a collision with another key.
    __device__ __forceinline__ bool insert(uint32_t d) const {
        uint32_t i = slot(d);
        uint32_t old = atomicCAS(&tab[i], kEmpty, d);
        while (old != kEmpty && old != d) {
            i = (i + 1) & uint32_t(H - 1);
            old = atomicCAS(&tab[i], kEmpty, d);
        }
        return old == kEmpty;
    }
    // insert for every lane, d only where pred: true if d was new
    __device__ __forceinline__ bool insert_if(bool pred, uint32_t d) const {
        const uint32_t key = pred ? d : kSink;
        // the home slot for every lane, selected after (no lane masking)
        const uint32_t h = slot(d);
        uint32_t i = pred ? h : uint32_t(H + lane_id());
        uint32_t old = atomicCAS(&tab[i], kEmpty, key);
        while (old != kEmpty && old != key) {
            i = (i + 1) & uint32_t(H - 1);
            old = atomicCAS(&tab[i], kEmpty, key);
        }
        return old == kEmpty;
    }
    // the sink slots (once per table)
    __device__ __forceinline__ void init_sink() const { tab[H + lane_id()] = kSink; }
    __device__ __forceinline__ bool contains(uint32_t d) const {
        uint32_t i = slot(d);
        for (;;) {
            const uint32_t v = tab[i];
            if (v == d) return true;
            if (v == kEmpty) return false;
            i = (i + 1) & uint32_t(H - 1);
        }
    }
    // every slot free again (whole wave)
    __device__ __forceinline__ void clear() const {
        static_assert(H % 256 == 0, "the wave clears 256 slots per store");
        uint4* t4 = reinterpret_cast<uint4*>(tab);
#pragma unroll
        for (int r = 0; r < H / 256; ++r) t4[r * 64 + lane_id()] = make_uint4(kEmpty, kEmpty, kEmpty, kEmpty);
    }
};

// Byte i of a string, read through a 256-byte window held one u32 per lane
// (strings shorter than 2^31 bytes).
struct ByteWindow {
    const uint8_t* p;
    int len;
    int base;    // first byte index of the window
    uint32_t w;  // lane l holds bytes base+4l .. base+4l+3

    __device__ void load(int b) {
        base = b;
        const int i = b + 4 * lane_id();
        uint32_t v = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (i + k < len) v |= uint32_t(p[i + k]) << (8 * k);
        w = v;
    }
    __device__ int at(int i) {  // i uniform, base <= i
        if (uint32_t(i - base) >= 256u) load(i);
        const int off = i - base;
        const uint32_t v = uint32_t(__builtin_amdgcn_readlane(int(w), off >> 2));
        return int((v >> (8 * (off & 3))) & 0xffu);
    }
};

// Appends v for every lane with pred; returns the new (uniform) count,
// which may exceed cap (the list is then incomplete).
template <int CAP>
__device__ __forceinline__ int list_append(uint32_t* list, int n, bool pred, uint32_t v) {
    const uint64_t m = __ballot(pred);
    if (pred) {
        const int idx = n + int(mbcnt(m));
        if (unsigned(idx) < unsigned(CAP)) list[idx] = v;
    }
    return n + __popcll(m);
}

// list_append for lists followed by 64 sink entries: every lane stores (the
// ones without a new entry into their sink entry), so no lane masking.
template <int CAP>
__device__ __forceinline__ int list_append_all(uint32_t* list, int n, bool pred, uint32_t v) {
    const uint64_t m = __ballot(pred);
    const uint32_t idx = uint32_t(n) + mbcnt(m);
    list[(pred && idx < uint32_t(CAP)) ? idx : uint32_t(CAP + lane_id())] = v;
    return n + __popcll(m);
}

// Ascending ids of the set bits of bits[0..W) into list; returns the count.
template <int CAP>
__device__ int rebuild_list(const uint64_t* bits, int W, uint32_t* list) {
    int n = 0;
    for (int j0 = 0; j0 < W; j0 += 64) {
        const int j = j0 + lane_id();
        uint64_t v = j < W ? bits[j] : 0ull;
        const int cnt = __popcll(v);
        const int incl = wave_scan(cnt);
        const int tot = uniform(__shfl(incl, 63, 64));
        if (n + tot <= CAP) {
            int pos = n + incl - cnt;
            while (v) {
                const int b = __ffsll(static_cast<unsigned long long>(v)) - 1;
                v &= v - 1;
                list[pos++] = uint32_t(j * 64 + b);
            }
        }
        n += tot;
    }
    return n;
}

// Writes the accepted states (cur & accept, ascending) of string sid, or
// every state of cur when a.emit_all.
static inline __device__ void emit_accepts(const DevGraph& g, const ScanArgs& a, int sid, const uint64_t* cur) {
    const uint64_t keep = a.emit_all ? ~0ull : 0ull;
    int total = 0;
    for (int j0 = 0; j0 < g.W; j0 += 64) {
        const int j = j0 + lane_id();
        const uint64_t v = j < g.W ? (cur[j] & (g.accept[j] | keep)) : 0ull;
        total += wave_sum(__popcll(v));
    }
    total = uniform(total);
    int64_t start = 0;
    if (total > 0) {
        unsigned long long s0 = 0;
        if (lane_id() == 0) s0 = atomicAdd(a.pool_top, (unsigned long long)total);
        start = int64_t(__shfl(static_cast<long long>(s0), 0, 64));
        if (start + total > a.pool_cap) {
            if (lane_id() == 0) atomicOr(a.overflow, 1);
            total = -total;  // marks a truncated result
        } else {
            int n = 0;
            for (int j0 = 0; j0 < g.W; j0 += 64) {
                const int j = j0 + lane_id();
                uint64_t v = j < g.W ? (cur[j] & (g.accept[j] | keep)) : 0ull;
                const int cnt = __popcll(v);
                const int incl = wave_scan(cnt);
                int64_t pos = start + n + incl - cnt;
                while (v) {
                    const int b = __ffsll(static_cast<unsigned long long>(v)) - 1;
                    v &= v - 1;
                    a.res_pool[pos++] = uint32_t(j * 64 + b);
                }
                n += uniform(__shfl(incl, 63, 64));
            }
        }
    }
    if (lane_id() == 0) {
        a.res_count[sid] = total;
        a.res_start[sid] = start;
    }
}

// ------------------------------------------------------------ engine steps
// Each helper is executed by one whole wave; counts are wave-uniform.

// fut |= cur & persistent, via the active list (fut empty on entry).
template <int CAP>
__device__ int persist_sparse(const DevGraph& g, const uint32_t* cl, int n_cur, uint64_t* fut, uint32_t* fl) {
    int n = 0;
    for (int k0 = 0; k0 < n_cur; k0 += 64) {
        const int k = k0 + lane_id();
        const bool ok = k < n_cur;
        const uint32_t s = ok ? (cl[k] & kIdMask) : 0u;
        const bool p = ok && test_bit(g.persis, s);
        if (p) set_bit_atomic(fut, s);
        n = list_append<CAP>(fl, n, p, s);
    }
    return n;
}

// Scatter step over (src,dst) edges: fut |= {dst : src in cur}.
// U edges per lane, and then the U source words they test, are loaded
// before any is used (memory-level parallelism: a round trip to L2 costs
// hundreds of cycles).  Loads use a clamped index and are unconditional: a
// guarded load compiles to a branch that waits for it before the next load
// is issued; lanes past the end drop their result instead.
template <int CAP, bool TRACK, int U = 4>
__device__ int scatter(const uint2* edges, int cnt, const uint64_t* cur, uint64_t* fut, uint32_t* fl, int n) {
    for (int t0 = 0; t0 < cnt; t0 += 64 * U) {
        uint2 e[U];
#pragma unroll
        for (int u = 0; u < U; ++u) e[u] = edges[min(t0 + u * 64 + lane_id(), cnt - 1)];
        uint64_t w[U];
#pragma unroll
        for (int u = 0; u < U; ++u) w[u] = cur[e[u].x >> 6];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool hit = t0 + u * 64 + lane_id() < cnt && ((w[u] >> (e[u].x & 63)) & 1ull);
            if (TRACK) {
                const bool nw = hit && set_new(fut, e[u].y);
                n = list_append<CAP>(fl, n, nw, e[u].y);
            } else if (hit) {
                set_bit_atomic(fut, e[u].y);
            }
        }
    }
    return n;
}

// scatter with the first round's U edges per lane already loaded
template <int U>
__device__ __forceinline__ void scatter_pre(const uint2 (&e0)[U], const uint2* edges, int cnt, const uint64_t* cur,
                                            uint64_t* fut) {
    for (int t0 = 0; t0 < cnt; t0 += 64 * U) {
        uint2 e[U];
#pragma unroll
        for (int u = 0; u < U; ++u) e[u] = t0 == 0 ? e0[u] : edges[min(t0 + u * 64 + lane_id(), cnt - 1)];
        uint64_t w[U];
#pragma unroll
        for (int u = 0; u < U; ++u) w[u] = cur[e[u].x >> 6];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool hit = t0 + u * 64 + lane_id() < cnt && ((w[u] >> (e[u].x & 63)) & 1ull);
            if (hit) set_bit_atomic(fut, e[u].y);
        }
    }
}

// A byte's TKO record as one wave holds it (one int2 per lane): lanes
// 0..kTkoRecGroups-1 the (shift, row) of its first groups, lane
// kTkoRecGroups its groups [kb, ke), the next lane its leftover edges [lb, le).
struct TkoRec {
    int2 v;
    __device__ __forceinline__ int kb() const { return __builtin_amdgcn_readlane(v.x, kTkoRecGroups); }
    __device__ __forceinline__ int ke() const { return __builtin_amdgcn_readlane(v.y, kTkoRecGroups); }
    __device__ __forceinline__ int lb() const { return __builtin_amdgcn_readlane(v.x, kTkoRecGroups + 1); }
    __device__ __forceinline__ int le() const { return __builtin_amdgcn_readlane(v.y, kTkoRecGroups + 1); }
};

// Makes the compiler wait for a loaded value at this point (an empty asm
// statement that reads it).
__device__ __forceinline__ void wait_ready(int2 v) { asm volatile("" ::"v"(v.x), "v"(v.y)); }

__device__ __forceinline__ TkoRec tko_rec(const DevGraph& g, int c) {
    return TkoRec{g.tko_rec[(c << 6) + lane_id()]};
}

// x of lane (l - 1) mod 64 in lane l: a DPP row rotate over the whole wave
// (wave_ror:1), a VALU move that needs no LDS round trip.
__device__ __forceinline__ uint64_t wave_ror1(uint64_t x) {
    const int lo = __builtin_amdgcn_update_dpp(0, int(uint32_t(x)), 0x13C, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, int(uint32_t(x >> 32)), 0x13C, 0xf, 0xf, false);
    return uint64_t(uint32_t(hi)) << 32 | uint32_t(lo);
}

// TKO shift groups of one byte (record tk): fut[j] = OR over groups i of
// shift(lim_i & cur, d_i)[j], fut zero on entry.  Lane l owns word j = j0 +
// 64u + l of U consecutive 64-word chunks and accumulates every group in
// registers, so each word is stored once and no atomics are needed; the U
// chunks' loads are issued together.  Shifts are split as d = 64q + r with
// 0 <= r < 64 (floor division): word j takes source word j-q shifted left by
// r and, when r != 0, source word j-q-1 shifted right by 64-r (the r == 0
// case never shifts by 64).  Word j-q-1 is the lane below's word j-q, moved
// up one lane by a wave rotate; lane 0 takes it from lane 63 of the previous
// chunk, or loads it for the first chunk.  The rows are zero-padded by
// kTkoPad words on each side, so for |q| within the padding the row words
// are read unchecked and the cur word at a clamped index (the zero row word
// masks it); the funnel shifts are two 32-bit alignbits per word.
template <int U>
__device__ __forceinline__ void tko_groups(const DevGraph& g, const TkoRec& tk, const uint64_t* cur, uint64_t* fut) {
    const int W = g.W;
    const int kb = tk.kb(), ke = tk.ke();
    if (kb == ke) return;
    const int l = lane_id();
    const bool l0 = l == 0;
    auto lo = [](uint64_t v) { return uint32_t(v); };
    auto hi = [](uint64_t v) { return uint32_t(v >> 32); };
    for (int j0 = 0; j0 < W; j0 += 64 * U) {
        uint64_t acc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) acc[u] = 0;
        for (int i = 0; i < ke - kb; ++i) {
            // the record's groups come from its lanes, later ones from memory
            int d, row;
            if (i < kTkoRecGroups) {
                d = __builtin_amdgcn_readlane(tk.v.x, i);
                row = __builtin_amdgcn_readlane(tk.v.y, i);
            } else {
                const int2 t = g.tko_grp[kb + i];
                d = uniform(t.x);
                row = uniform(t.y);
            }
            const int q = d >> 6;  // arithmetic shift: floor(d / 64)
            const int r = d & 63;
            const uint64_t* lim = g.tko_lim + (int64_t(row) * g.tko_stride + kTkoPad);
            uint64_t a[U];
            uint64_t prev;  // lane 0's word below the first chunk (the same address in every lane)
            if (q >= -kTkoPad && q < kTkoPad) {
                const uint64_t* lrow = lim + (j0 - q);
#pragma unroll
                for (int u = 0; u < U; ++u) a[u] = lrow[64 * u + l] & cur[min(max(j0 + 64 * u + l - q, 0), W - 1)];
                prev = lrow[-1] & cur[min(max(j0 - q - 1, 0), W - 1)];
            } else {
                // far shifts: clamped reads, words outside [0, W) dropped
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int s0 = j0 + 64 * u + l - q;
                    const int sc = min(max(s0, 0), W - 1);
                    const uint64_t v = lim[sc] & cur[sc];
                    a[u] = s0 == sc ? v : 0ull;
                }
                const int s1 = j0 - q - 1;
                const int sc1 = min(max(s1, 0), W - 1);
                const uint64_t v1 = lim[sc1] & cur[sc1];
                prev = s1 == sc1 ? v1 : 0ull;
            }
            if (r == 0) {
#pragma unroll
                for (int u = 0; u < U; ++u) acc[u] |= a[u];
            } else if (r <= 32) {
                // (a << r) | (b >> (64 - r)) needs only the high half of the word below b
                const uint32_t s = uint32_t(32 - r);
                uint32_t bprev = hi(prev);
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t ah = hi(a[u]), al = lo(a[u]);
                    const uint32_t rot = uint32_t(__builtin_amdgcn_update_dpp(0, int(ah), 0x13C, 0xf, 0xf, false));
                    const uint32_t bh = l0 ? bprev : rot;
                    acc[u] |= uint64_t(__builtin_amdgcn_alignbit(ah, al, s)) << 32 | __builtin_amdgcn_alignbit(al, bh, s);
                    bprev = rot;  // lane 0 of the next chunk: lane 63 of this one
                }
            } else {
                const uint32_t t = uint32_t(64 - r);
                uint64_t bprev = prev;
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint64_t rot = wave_ror1(a[u]);
                    const uint64_t b = l0 ? bprev : rot;
                    const uint32_t bh = hi(b);
                    acc[u] |= uint64_t(__builtin_amdgcn_alignbit(lo(a[u]), bh, t)) << 32 |
                              __builtin_amdgcn_alignbit(bh, lo(b), t);
                    bprev = rot;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = j0 + 64 * u + l;
            if (j < W && acc[u]) fut[j] = acc[u];
        }
    }
}

// Flattened processing of multi-destination rows: lane l owns a row of
// len_l destinations starting at as_multi[start_l]; each iteration every
// lane fetches one destination (coalesced within a row).
template <int CAP, typename SET>
__device__ int multi_rows(const uint32_t* as_multi, uint32_t start, int len, const SET& fut, uint32_t* fl, int n) {
    const int incl = wave_scan(len);
    const int excl = incl - len;
    const int tot = uniform(__shfl(incl, 63, 64));
    for (int t0 = 0; t0 < tot; t0 += 64) {
        const int t = t0 + lane_id();
        int lo = 0, hi = 63;  // owner lane: smallest r with incl[r] > t
#pragma unroll
        for (int it = 0; it < 6; ++it) {
            const int mid = (lo + hi) >> 1;
            const int v = __shfl(incl, mid, 64);
            if (v > t) hi = mid;
            else lo = mid + 1;
        }
        const uint32_t so = uint32_t(__shfl(int(start), lo, 64));
        const int exo = __shfl(excl, lo, 64);
        const bool ok = t < tot;
        const uint32_t d = ok ? as_multi[so + uint32_t(t - exo)] : 0u;
        const bool nw = ok && fut.insert(d);
        n = list_append<CAP>(fl, n, nw, d);
        if (SET::kBounded && n > CAP) return n;
    }
    return n;
}

// The tables an AS step reads, loaded once per string by the caller (kept in
// scalar registers across the step loop rather than reloaded per step).
struct AsRows {
    const uint32_t* tab;       // DevGraph::as_tab
    const uint32_t* multi;     // DevGraph::as_multi
    const uint32_t* pair_dst;  // DevGraph::pair_dst
};

// Prefetched slice of a precomputed row list: the first 64 entries, one per
// lane, plus where the list lives.
struct RowSlice {
    uint32_t v;  // entry lane_id() (valid when lane_id() < n)
    int b, n;    // start and length in the backing array
};

// AS step (src/AS_kernel.cu:74-125 with sparse rows): fut = union of the
// rows of the active states on byte c.  Rows of persistent states contain
// the state itself.  The always-active states and the first positions they
// reached on the previous byte are held implicitly (the caller records the
// byte): their rows on c are the precomputed two-byte rows (pair, merged
// here when use_pair) and the always-active rows of c, which become the
// next step's implicit set.  Always-active states (bit 31) are therefore
// neither gathered nor stored.  clr (or null): the bitmap of the listed
// states, whose words are cleared as the list is read.
template <int CAP, typename SET>
__device__ int as_step(const AsRows& g, int c, const RowSlice& pair, bool use_pair, const uint32_t* cl, int n_cur,
                       const SET& fut, uint32_t* fl, uint64_t* clr) {
    // state-major table: entry (s, c) at s*256 + c
    const uint32_t* tab = g.tab + c;
    // the first chunk's table gathers are in flight while the two-byte rows
    // are merged
    const uint32_t sv0 = lane_id() < n_cur ? cl[lane_id()] : kTag;
    const uint32_t e0 = (sv0 & kTag) ? 0u : tab[(sv0 & kIdMask) << 8];
    if (clr != nullptr) {
        if (lane_id() < n_cur) clr[(sv0 & kIdMask) >> 6] = 0ull;
        for (int k = 64 + lane_id(); k < n_cur; k += 64) clr[(cl[k] & kIdMask) >> 6] = 0ull;
    }
    // common case: one chunk of states, one of two-byte row entries, rows of
    // at most two destinations (a persistent state's row is itself plus the
    // state it advances to, held in the entry): every lane's candidates in
    // one pass
    const uint32_t cnt0 = e0 >> 24;
    const bool self0 = cnt0 == kAsSelfRow;
    if (n_cur <= 64 && (!use_pair || pair.n <= 64) && !__ballot(cnt0 >= 3 && !self0)) {
        const bool okp = use_pair && lane_id() < pair.n;
        uint32_t cnt = cnt0;
        uint32_t x0 = e0 & kIdMask, x1 = 0;
        if (__ballot(cnt == 2)) {  // a row in the store
            if (cnt == 2) {
                const uint32_t o = x0;
                x0 = g.multi[o];
                x1 = g.multi[o + 1];
            }
        }
        if (self0) {
            x1 = sv0 & kIdMask;
            cnt = 2;
        }
        const bool two = __ballot(cnt == 2) != 0;
        if constexpr (SET::kBounded) {
            const bool n1 = fut.insert_if(okp, pair.v);
            const bool n2 = fut.insert_if(cnt != 0, x0);
            int n = list_append_all<CAP>(fl, 0, n1, pair.v);
            n = list_append_all<CAP>(fl, n, n2, x0);
            if (two) {
                const bool n3 = fut.insert_if(cnt == 2, x1);
                n = list_append_all<CAP>(fl, n, n3, x1);
            }
            return n;
        } else {
            const bool n1 = okp && fut.insert(pair.v);
            const bool n2 = cnt != 0 && fut.insert(x0);
            int n = list_append<CAP>(fl, 0, n1, pair.v);
            n = list_append<CAP>(fl, n, n2, x0);
            if (two) {
                const bool n3 = cnt == 2 && fut.insert(x1);
                n = list_append<CAP>(fl, n, n3, x1);
            }
            return n;
        }
    }
    int n = 0;
    if (use_pair) {
        for (int t0 = 0; t0 < pair.n; t0 += 64) {
            const int t = t0 + lane_id();
            const bool ok = t < pair.n;
            const uint32_t v = t0 == 0 ? pair.v : (ok ? g.pair_dst[pair.b + t] : 0u);
            const bool nw = ok && fut.insert(v);
            n = list_append<CAP>(fl, n, nw, v);
            if (SET::kBounded && n > CAP) return n;
        }
    }
    for (int k0 = 0; k0 < n_cur; k0 += 64) {
        uint32_t e = e0;
        if (k0 > 0) {
            const int k = k0 + lane_id();
            const uint32_t sv = k < n_cur ? cl[k] : kTag;
            e = (sv & kTag) ? 0u : tab[(sv & kIdMask) << 8];
        }
        const uint32_t cnt = e >> 24;
        const uint32_t d = e & kIdMask;
        const bool self = cnt == kAsSelfRow;
        const bool nw = (cnt == 1 || self) && fut.insert(d);
        n = list_append<CAP>(fl, n, nw, d);
        if (SET::kBounded && n > CAP) return n;
        if (__ballot(self)) {
            // the source state itself, the entry's second destination
            const uint32_t own = (k0 == 0 ? sv0 : (self ? cl[k0 + lane_id()] : 0u)) & kIdMask;
            const bool ns = self && fut.insert(own);
            n = list_append<CAP>(fl, n, ns, own);
            if (SET::kBounded && n > CAP) return n;
        }
        // short rows: each lane walks its own row, two or four loads in flight
        const int len = (cnt >= 2 && cnt < kAsSelfRow) ? int(cnt) : 0;
        if (__ballot(len > 0) && !__ballot(len > 2)) {
            // rows of two in the store (neither entry the state itself)
            const uint32_t y0 = len ? g.multi[d] : 0u;
            const uint32_t y1 = len ? g.multi[d + 1] : 0u;
            const bool n0 = len && fut.insert(y0);
            n = list_append<CAP>(fl, n, n0, y0);
            if (SET::kBounded && n > CAP) return n;
            const bool n1 = len && fut.insert(y1);
            n = list_append<CAP>(fl, n, n1, y1);
            if (SET::kBounded && n > CAP) return n;
        } else if (__ballot(len > 0)) {
            uint32_t x[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) x[j] = j < len ? g.multi[d + j] : 0u;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const bool nj = j < len && fut.insert(x[j]);
                n = list_append<CAP>(fl, n, nj, x[j]);
                if (SET::kBounded && n > CAP) return n;
            }
            for (int j = 4; __ballot(j < len); ++j) {
                const uint32_t y = j < len ? g.multi[d + j] : 0u;
                const bool nj = j < len && fut.insert(y);
                n = list_append<CAP>(fl, n, nj, y);
                if (SET::kBounded && n > CAP) return n;
            }
        }
        // long rows (kAsSelfRow or more destinations, count in the store):
        // flattened over the wave
        const bool lng = cnt == 255;
        if (__ballot(lng)) n = multi_rows<CAP>(g.multi, d + 1, lng ? int(g.multi[d]) : 0, fut, fl, n);
        if (SET::kBounded && n > CAP) return n;
    }
    return n;
}

// ------------------------------------------------------------ selector
struct StepCtx {
    int c, c_prev, n_act, sparse, T, K, Lft, last_m;
    int pos, len;
    // out-degree of the step's listed states on c: the share whose row is
    // empty (-1 for a bitmap-only set) and the rows' entries; cycles of the
    // step before
    float deg_empty = -1.f, deg_entries = 0.f, last_cyc = 0.f;
};

// Out-degree bins of a step for the selector features: over the listed
// states (always-active ones excluded: their rows are the implicit set),
// the share with an empty row on byte c and the rows' total entries.  A
// bitmap-only set gets -1 and the expected entries |T_c| * n / S.
static inline __device__ void out_degree(const DevGraph& g, int c, const uint32_t* cl, int n_act, bool sparse,
                                         StepCtx& ctx) {
    if (!sparse || g.as_tab == nullptr) {
        ctx.deg_empty = -1.f;
        ctx.deg_entries = float(g.sym_off[c + 1] - g.sym_off[c]) * float(n_act) / float(g.S);
        return;
    }
    int empty = 0, listed = 0, entries = 0;
    for (int k0 = 0; k0 < n_act; k0 += 64) {
        const int k = k0 + lane_id();
        const uint32_t sv = k < n_act ? cl[k] : kTag;
        const bool ok = !(sv & kTag);
        const uint32_t e = ok ? g.as_tab[((sv & kIdMask) << 8) + uint32_t(c)] : 0u;
        uint32_t len = e >> 24;
        if (len == 255u) len = g.as_multi[e & kIdMask];
        if (len == kAsSelfRow) len = 2;
        listed += ok ? 1 : 0;
        empty += (ok && len == 0) ? 1 : 0;
        entries += ok ? int(len) : 0;
    }
    listed = wave_sum(listed);
    ctx.deg_empty = listed > 0 ? float(wave_sum(empty)) / float(listed) : 0.f;
    ctx.deg_entries = float(wave_sum(entries));
}

__device__ __forceinline__ int heuristic_choice(const DevGraph& g, const StepCtx& s) {
    // AdMatch rule (src/AdMatch_kernel.cu:87) extended with an iNFA cost term
    if (s.sparse && s.K > s.n_act) return M_AS;
    const int64_t tko_cost = int64_t(s.K) * g.W + s.Lft;
    return int64_t(s.T) <= tko_cost ? M_INFA : M_TKO;
}

// Selector feature i (0..15) of a step; the wave holds feature l in lane l.
__device__ __forceinline__ float feature(const DevGraph& g, const StepCtx& s, int i) {
    switch (i) {
        case 0: return __log2f(1.f + float(s.n_act));
        case 1: return 100.f * float(s.n_act) / float(g.S);
        case 2: return __log2f(1.f + float(s.T));
        case 3: return float(s.K);
        case 4: return __log2f(1.f + float(s.Lft));
        case 5: return s.deg_empty;
        case 6: return word_char(uint32_t(s.c)) ? 1.f : 0.f;
        case 7: return float(s.c >> 5);
        case 8: return float(s.pos) / float(s.len > 0 ? s.len : 1);
        case 9: return __log2f(1.f + s.deg_entries);
        case 10: return float(s.sparse);
        case 11: return s.last_m == M_INFA ? 1.f : 0.f;
        case 12: return s.last_m == M_TKO ? 1.f : 0.f;
        case 13: return s.last_m == M_AS ? 1.f : 0.f;
        case 14: return __log2f(1.f + float(s.len - s.pos));
        case 15: return __log2f(1.f + s.last_cyc);
        default: return 0.f;
    }
}

// 16 -> 64 -> 64 -> 3 MLP evaluated by one wave: lane j is hidden unit j;
// fl holds feature l in lane l.  W1 is stored input-major [16][64], W2 as
// [64 in][64 out], so every weight load is one coalesced 256-byte row.
// Returns argmin of the three predicted log step costs among the allowed
// methods.
static inline __device__ int mlp_choice(const float* w, float fl, bool allow_as) {
    const int j = lane_id();
    float h = w[kMlpB1 + j];
#pragma unroll 4
    for (int i = 0; i < kFeatures; ++i)
        h += w[kMlpW1 + i * kHidden + j] * __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fl), i));
    h = fmaxf(h, 0.f);
    float h2 = w[kMlpB2 + j];
#pragma unroll 8
    for (int k = 0; k < kHidden; ++k) {
        const float hk = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(h), k));
        h2 += w[kMlpW2 + k * kHidden + j] * hk;
    }
    h2 = fmaxf(h2, 0.f);
    float best = 3.0e38f;
    int arg = M_INFA;
#pragma unroll
    for (int m = 0; m < kOutputs; ++m) {
        const float o = wave_sum_f(w[kMlpW3 + m * kHidden + j] * h2) + w[kMlpB3 + m];
        const float ou = __int_as_float(uniform(__float_as_int(o)));
        if ((m != M_AS || allow_as) && ou < best) {
            best = ou;
            arg = m;
        }
    }
    return arg;
}

// Learned policy table: entry (bucket, sparse, c) is the method the network
// picks for a step on byte c with an explicit active count n in bucket
// min(15, floor(log2(n + 1))) and list validity `sparse`; 2 bits per entry,
// 16 entries per u32, [kPolicyBuckets][2][16] words.
constexpr int kPolicyBuckets = 16;
constexpr int kPolicyWords = kPolicyBuckets * 2 * 16;
static_assert(kPolicyWords == kPolicyTableWords, "policy table size");
// Step context at which the network is evaluated for a table entry: the
// bucket's middle count plus the mean always-active row, a step in the
// middle of a kPolicyLen-byte string, the method a step of that list
// validity usually follows, half of the listed rows empty and one row entry
// per active state, and a step before of kPolicyStepCycles cycles.
constexpr int kPolicyLen = 1024;
constexpr int kPolicyStepCycles = 1024;  // the step before, at a synthetic context

__device__ __forceinline__ int policy_bucket(int n_act) { return min(kPolicyBuckets - 1, 31 - __clz(n_act + 1)); }

__device__ __forceinline__ StepCtx policy_ctx(const DevGraph& g, int c, int bucket, bool sparse) {
    StepCtx s;
    s.c = c;
    s.c_prev = -1;
    s.n_act = (bucket == 0 ? 0 : (3 << (bucket - 1)) - 1) + g.alw_total / 256;
    s.sparse = sparse;
    const int4 st = g.sym_stats[c];
    s.T = st.x;
    s.K = st.y;
    s.Lft = st.z;
    s.last_m = sparse ? M_AS : M_INFA;
    s.pos = kPolicyLen / 2;
    s.len = kPolicyLen;
    s.deg_empty = sparse ? 0.5f : -1.f;
    s.deg_entries = float(s.n_act);
    s.last_cyc = float(kPolicyStepCycles);
    return s;
}

// One wave per table entry; table zeroed beforehand.
// The policy table's method for a step on a bitmap-only set of n_act states.
__device__ __forceinline__ int dense_policy(const ScanArgs& a, int n_act, int c) {
    if (a.policy_dense >= 0) return a.policy_dense;
    using CU32 = const __attribute__((address_space(4))) uint32_t;
    const CU32* pt = (CU32*)(a.policy);
    const uint32_t wv = pt[uint32_t(policy_bucket(n_act) * 2 * 16 + (c >> 4))];
    return int((wv >> ((c & 15) * 2)) & 3u);
}

// Method for the step on ctx.c.  total_active() counts the active states
// including the implicitly held ones (selector features).  cached: the last
// network decision of the every-k-steps mode.
template <int MODE, typename TOTAL>
__device__ __forceinline__ int choose_method(const DevGraph& g, const ScanArgs& a, int sid, StepCtx& ctx, int n_act,
                                             bool sparse, const TOTAL& total_active, int& cached, bool& log_step) {
    const int c = ctx.c;
    const int b = ctx.pos;
    // per-byte statistics, read only by the selectors that use them
    auto load_stats = [&]() {
        const int4 st = g.sym_stats[c];
        ctx.T = st.x;
        ctx.K = st.y;
        ctx.Lft = st.z;
    };
    int m;
    switch (MODE) {
        case MODE_INFA: m = M_INFA; break;
        case MODE_TKO: m = M_TKO; break;
        case MODE_AS: m = sparse ? M_AS : M_INFA; break;
        case MODE_ADMATCH:
            load_stats();
            ctx.n_act = total_active();
            m = (sparse && ctx.K > ctx.n_act) ? M_AS : M_TKO;
            break;
        case MODE_LEARNED: {
            // policy table (set whenever this mode is launched): one scalar
            // load, or none when the table holds one method for every entry
            // of this list validity
            m = sparse ? a.policy_sparse : a.policy_dense;
            if (m < 0) {
                using CU32 = const __attribute__((address_space(4))) uint32_t;
                const CU32* pt = (CU32*)(a.policy);
                const uint32_t wv = pt[uint32_t((policy_bucket(n_act) * 2 + int(sparse)) * 16 + (c >> 4))];
                m = int((wv >> ((c & 15) * 2)) & 3u);
            }
            break;
        }
        case MODE_LEARNED_AS:
            // the table picks AS whenever the list is valid; a step on a
            // bitmap-only set reads it in the dense path (dense_policy)
            m = sparse ? M_AS : M_INFA;
            break;
        case MODE_LEARNED_EVERY:
            // the network itself, re-evaluated every mlp_every steps
            load_stats();
            ctx.n_act = total_active();
            if (cached < 0 || (b % a.mlp_every) == 0) cached = mlp_choice(a.mlp, feature(g, ctx, lane_id()), true);
            m = cached;
            if (m == M_AS && !sparse) m = heuristic_choice(g, ctx);
            break;
        case MODE_LOG:
            // steps that are not timed run what a learned engine mostly runs
            // (AS on a listed set, the heuristic's pick on a dense one), so
            // the timed steps see the memory traffic of a real scan
            load_stats();
            ctx.n_act = total_active();
            m = sparse ? M_AS : heuristic_choice(g, ctx);
            log_step = ((uint64_t(sid) * 2654435761ull + uint64_t(b)) % uint64_t(a.log_stride)) == 0;
            break;
        default:
            load_stats();
            ctx.n_act = total_active();
            m = heuristic_choice(g, ctx);
            break;
    }
    return m;
}

// ------------------------------------------------------------ engine

template <int CAP>
struct Lane {
    // u64 words per wave: two bitmaps, two lists
    static constexpr size_t words(int W) { return size_t(2) * W + CAP; }
};

// fut = cur & persistent over all W words, U words per lane in flight.
template <int U>
__device__ __forceinline__ void persist_dense(const uint64_t* cur, const uint64_t* per, uint64_t* fut, int W) {
    for (int j0 = lane_id(); j0 < W; j0 += 64 * U) {
        uint64_t v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = min(j0 + 64 * u, W - 1);  // clamped, unconditional (see scatter)
            v[u] = cur[j] & per[j];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = j0 + 64 * u;
            if (j < W) fut[j] = v[u];
        }
    }
}

// Runs one step with method m from (cur, cl, n_cur, sparse) into fut/fl.
// Returns the fut count (list valid iff *fut_sparse).  NARROW: the kernel is
// built for state vectors of at most 64 * kTkoPassChunks words and more waves
// per SIMD, so the TKO step keeps fewer loads in flight.
template <int CAP, bool NARROW>
__device__ int run_method(const DevGraph& g, int m, int c, const TkoRec& tk, const uint64_t* cur, const uint32_t* cl,
                          int n_cur, bool sparse, uint64_t* fut, uint32_t* fl, bool* fut_sparse) {
    constexpr int UE = NARROW ? 8 : 16;  // leftover edges per lane loaded with the record
    int n = 0;
    const int W = g.W;
    if (m == M_TKO) {
        // the leftover edges' first round is loaded before the shift rows, so
        // its round trip to L2 overlaps theirs (the guard keeps an empty
        // leftover list from reading past the table's end)
        const int lb = tk.lb();
        const uint2* le = g.tko_left + lb;
        const int lc = tk.le() - lb;
        uint2 e0[UE];
        if (lc > 0) {
#pragma unroll
            for (int u = 0; u < UE; ++u) e0[u] = le[min(u * 64 + lane_id(), lc - 1)];
        } else {
#pragma unroll
            for (int u = 0; u < UE; ++u) e0[u] = make_uint2(0u, 0u);
        }
        // one pass when the vector fits in kTkoOnePass chunks per lane (every
        // load of a row in flight at once); two sizes keep the kernel's code
        // small
        const int nch = (W + 63) >> 6;
        if (!NARROW && nch > kTkoPassChunks && nch <= kTkoOnePass)
            tko_groups<kTkoOnePass>(g, tk, cur, fut);
        else
            tko_groups<kTkoPassChunks>(g, tk, cur, fut);
        wave_sync();
        scatter_pre<UE>(e0, le, lc, cur, fut);
        *fut_sparse = false;
        return -1;
    }
    if (sparse) {
        n = persist_sparse<CAP>(g, cl, n_cur, fut, fl);
    } else {
        persist_dense<4>(cur, g.persis, fut, W);
    }
    wave_sync();
    if (sparse) {
        n = scatter<CAP, true, 8>(g.list + g.sym_off[c], g.sym_off[c + 1] - g.sym_off[c], cur, fut, fl, n);
    } else {
        scatter<CAP, false, 8>(g.list + g.sym_off[c], g.sym_off[c + 1] - g.sym_off[c], cur, fut, fl, 0);
        n = -1;
    }
    *fut_sparse = sparse && n <= CAP;
    return n;
}

// Set bits of bits[0..W) (whole wave), four words per lane in flight.
static inline __device__ int count_bits(const uint64_t* bits, int W) {
    int n = 0;
    for (int j0 = lane_id(); j0 < W; j0 += 256) {
        uint64_t v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = bits[min(j0 + 64 * u, W - 1)];
#pragma unroll
        for (int u = 0; u < 4; ++u) n += j0 + 64 * u < W ? __popcll(v[u]) : 0;
    }
    return uniform(wave_sum(n));
}

template <int CAP>
__device__ __forceinline__ void clear_bits(uint64_t* bits, const uint32_t* list, int n, bool sparse, int W) {
    if (sparse) {
        for (int k = lane_id(); k < n; k += 64) bits[(list[k] & kIdMask) >> 6] = 0ull;
    } else {
        for (int j = lane_id(); j < W; j += 64) bits[j] = 0ull;
    }
}

template <int CAP, int MODE, bool NARROW, bool LEAN>
__device__ void scan_string(const DevGraph& g, const ScanArgs& a, int sid, const uint8_t* x, int L, uint64_t* bitsA,
                            uint64_t* bitsB, uint32_t* listA, uint32_t* listB, int* msteps) {
    const int W = g.W;
    msteps[3] += L;
    uint64_t* cur = bitsA;
    uint64_t* fut = bitsB;
    uint32_t* cl = listA;
    uint32_t* fl = listB;
    // streaming: the string continues a stream whose state and last byte
    // class were saved by the previous chunk (LEAN: no streams, no boundary
    // transitions, two-byte rows present)
    const int slot = (!LEAN && a.stream_ids != nullptr) ? a.stream_ids[sid] : -1;
    const uint32_t sflags = slot >= 0 ? uint32_t(a.stream_flags[slot]) : 0u;
    const uint64_t* saved = slot >= 0 ? a.stream_state + size_t(slot) * W : nullptr;
    // invariant: both bitmaps are zero here.  The always-active states
    // (initial and persistent) start implicit; the other initial states are
    // listed.  A continued stream starts from its saved set, all listed.
    const uint64_t* start = (sflags & 1u) ? saved : g.init_explicit;
    for (int j = lane_id(); j < W; j += 64) cur[j] = start[j];
    wave_sync();
    int n_act = uniform(rebuild_list<CAP>(cur, W, cl));
    bool sparse = n_act <= CAP;
    ByteWindow win{x, L, 0, 0u};
    if (L > 0) win.load(0);
    const bool has_wb = !LEAN && g.wb_count > 0;
    const bool have_pair = LEAN || g.pair_off != nullptr;
    // States active but held implicitly (in neither the bitmap nor the list):
    // imp >= 0: the always-active row of byte imp (the always-active states
    // and the first positions they reached on that byte); imp == -2: the
    // always-active states alone; imp == -1: none.  Set by AS steps;
    // everything that needs the full set (other methods, the word-boundary
    // pass, acceptance) first calls materialize().
    int imp = (sflags & 1u) ? -1 : -2;
    auto alw_range = [&](int cc, int& b0, int& n0) {
        b0 = g.sym_stats[cc].w;
        n0 = (cc < 255 ? g.sym_stats[cc + 1].w : g.alw_total) - b0;
    };
    auto materialize = [&]() {
        if (imp == -1) return;
        int ab = 0, an = g.n_alw_states;
        const uint32_t* src = g.alw_states;
        if (imp >= 0) {
            alw_range(imp, ab, an);
            src = g.alw_dst;
        }
        for (int t0 = 0; t0 < an; t0 += 64) {
            const int t = t0 + lane_id();
            const bool ok = t < an;
            const uint32_t v = ok ? (imp >= 0 ? src[ab + t] : (src[t] | kTag)) : 0u;
            const bool nw = ok && set_new(cur, v & kIdMask);
            // always-active states keep bit 31: their rows are the always-active rows
            n_act = list_append<CAP>(cl, n_act, nw, v & (kIdMask | kTag));
        }
        n_act = uniform(n_act);
        if (n_act > CAP) sparse = false;
        imp = -1;
        wave_sync();
    };

    auto wb_pass = [&]() {
        if (g.wbs_off != nullptr && sparse) {
            // chain-free list with boundary rows: the boundary successors of
            // the listed states and of the implicit set (precomputed per
            // implicit row), without materialising the implicit set
            const int n0 = n_act;
            if (imp != -1) {
                const int r = imp >= 0 ? imp : 256;
                const int rb = g.wbi_off[r], rn = g.wbi_off[r + 1] - rb;
                for (int t0 = 0; t0 < rn; t0 += 64) {
                    const int t = t0 + lane_id();
                    const bool ok = t < rn;
                    const uint32_t v = ok ? g.wbi_dst[rb + t] : 0u;
                    const bool nw = ok && set_new(cur, v);
                    n_act = list_append<CAP>(cl, n_act, nw, v);
                }
            }
            for (int k0 = 0; k0 < n0; k0 += 64) {
                const int k = k0 + lane_id();
                const uint32_t st = k < n0 ? (cl[k] & kIdMask) : 0u;
                const uint32_t b = k < n0 ? g.wbs_off[st] : 0u;
                const uint32_t e = k < n0 ? g.wbs_off[st + 1] : 0u;
                for (uint32_t j = 0; __ballot(b + j < e); ++j) {
                    const bool ok = b + j < e;
                    const uint32_t v = ok ? g.wbs_dst[b + j] : 0u;
                    const bool nw = ok && set_new(cur, v);
                    n_act = list_append<CAP>(cl, n_act, nw, v);
                }
            }
            n_act = uniform(n_act);
            if (n_act > CAP) sparse = false;
            wave_sync();
            return;
        }
        materialize();
        if (g.wb_parallel) {
            // no destination is a source: one parallel pass, the wave's lanes
            // over the edges
            for (int t0 = 0; t0 < g.wb_count; t0 += 64) {
                const int t = t0 + lane_id();
                const bool ok = t < g.wb_count;
                const uint2 e = ok ? g.list[g.wb_begin + t] : make_uint2(0u, 0u);
                const bool nw = ok && test_bit(cur, e.x) && set_new(cur, e.y);
                n_act = list_append<CAP>(cl, n_act, nw, e.y);
            }
            n_act = uniform(n_act);
            if (n_act > CAP) sparse = false;
            wave_sync();
            return;
        }
        // in place, list order, one lane (thread_count == 1 semantics of
        // src/iNFA_kernel.cu:145-164)
        int n = n_act;
        if (lane_id() == 0) {
            for (int t = 0; t < g.wb_count; ++t) {
                const uint2 e = g.list[g.wb_begin + t];
                if (test_bit(cur, e.x)) {
                    const uint64_t m = 1ull << (e.y & 63);
                    if (!(cur[e.y >> 6] & m)) {
                        cur[e.y >> 6] |= m;
                        if (sparse && n < CAP) cl[n] = e.y;
                        ++n;
                    }
                }
            }
        }
        n_act = uniform(__builtin_amdgcn_readlane(n, 0));
        if (n_act > CAP) sparse = false;
        wave_sync();
    };
    int c_next = L > 0 ? win.at(0) : 0;
    // a boundary before the first byte: the stream's previous byte (or the
    // start, a non-word position) against it
    if (has_wb && L > 0 && word_char(uint32_t(c_next)) != ((sflags & 2u) != 0)) wb_pass();

    // the tables of the AS steps, loaded once per string
    const uint32_t* const pair_off = g.pair_off;
    const AsRows rows{g.as_tab, g.as_multi, g.pair_dst};
    // two-byte rows of (byte, next byte), first 64 entries loaded a step ahead
    auto load_pair = [&](int c1, int c2) -> RowSlice {
        RowSlice r;
        // read-only offsets: scalar loads
        using CU32 = const __attribute__((address_space(4))) uint32_t;
        const CU32* o = (CU32*)(pair_off) + (c1 * 256 + c2);
        r.b = int(o[0]);
        r.n = int(o[1]) - r.b;
        r.v = lane_id() < r.n ? rows.pair_dst[r.b + lane_id()] : 0u;
        return r;
    };
    RowSlice pair{0u, 0, 0};
    const bool can_as = g.as_tab != nullptr;
    // After a step that leaves a dense set, the next byte's TKO record is
    // loaded (tk_next, valid when tk_ready) for a TKO step to use; other TKO
    // steps load their own.
    constexpr bool kTko = MODE != MODE_INFA && MODE != MODE_AS;
    TkoRec tk_next{make_int2(0, 0)};
    bool tk_ready = false;

    StepCtx ctx;
    ctx.len = L;
    ctx.last_m = M_INFA;
    ctx.c_prev = -1;
    int cached = -1;
    uint64_t t_step = 0;  // clock at the last step's start (feature-evaluating modes)
    for (int b = 0; b < L; ++b) {
        // the next byte is read one step ahead (two-byte rows, word boundaries)
        const int c = c_next;
        TkoRec tk = tk_next;
        // the record is waited for here, before this step issues loads of
        // its own (loads complete in order: a wait for it later would also
        // wait for those)
        if (kTko) wait_ready(tk.v);
        const bool tk_ok = tk_ready;
        RowSlice pair_next{0u, 0, 0};
        if (b + 1 < L) {
            c_next = win.at(b + 1);
            if (have_pair) pair_next = load_pair(c, c_next);
        }
        auto tk_of = [&]() {
            if (!tk_ok) {
                tk = tko_rec(g, c);
                wait_ready(tk.v);
            }
            return tk;
        };
        // active states including the implicit ones (selector features)
        auto total_active = [&]() {
            int n_tot = n_act;
            if (imp >= 0) {
                int ab, an;
                alw_range(imp, ab, an);
                n_tot += an;
            } else if (imp == -2) {
                n_tot += g.n_alw_states;
            }
            return n_tot;
        };
        ctx.c = c;
        ctx.sparse = sparse;
        ctx.pos = b;
        if constexpr (MODE == MODE_LOG || MODE == MODE_LEARNED_EVERY) {
            // the modes that evaluate features: the step before's cycles, and
            // the listed states' out-degree on this byte when the network or
            // the log will read it
            const uint64_t now = clock64();
            ctx.last_cyc = b > 0 ? float(now - t_step) : 0.f;
            t_step = now;
            const bool need = MODE == MODE_LOG
                                  ? ((uint64_t(sid) * 2654435761ull + uint64_t(b)) % uint64_t(a.log_stride)) == 0
                                  : (cached < 0 || (b % a.mlp_every) == 0);
            if (need) out_degree(g, c, cl, n_act, sparse, ctx);
        }
        bool log_step = false;
        int m = choose_method<MODE>(g, a, sid, ctx, n_act, sparse, total_active, cached, log_step);
        // without two-byte rows the first positions of the implicit set are
        // not reachable from the table: list them before an AS step
        if (m == M_AS && !have_pair && imp >= 0) materialize();
        if (m == M_AS && !(sparse && can_as)) m = M_INFA;
        bool fsp;
        int n_fut;
        int next_imp = -1;
        if (!log_step) {
            if (m == M_AS) {
                n_fut = as_step<CAP>(rows, c, pair, have_pair && imp >= 0, cl, n_act, BitSet{fut}, fl, cur);
                fsp = n_fut <= CAP;
                next_imp = c;
            } else {
                if (MODE == MODE_LEARNED_AS && !sparse) m = dense_policy(a, n_act, c);
                materialize();
                n_fut = run_method<CAP, NARROW>(g, m, c, tk_of(), cur, cl, n_act, sparse, fut, fl, &fsp);
            }
        } else {
            // time every method on the same state: AS on the implicit form,
            // then TKO and iNFA on the materialised set; keep the iNFA result
            float tm[3] = {-1.f, -1.f, -1.f};
            if (sparse && can_as) {
                __builtin_amdgcn_s_waitcnt(0);
                const uint64_t t0 = clock64();
                n_fut = as_step<CAP>(rows, c, pair, have_pair && imp >= 0, cl, n_act, BitSet{fut}, fl, nullptr);
                fsp = n_fut <= CAP;
                wave_sync();
                __builtin_amdgcn_s_waitcnt(0);
                tm[M_AS] = float(clock64() - t0);
                clear_bits<CAP>(fut, fl, n_fut, fsp, W);
                wave_sync();
            }
            // listing the implicitly held states is what leaving the AS
            // form costs: charged to both other methods
            __builtin_amdgcn_s_waitcnt(0);
            const uint64_t tm0 = clock64();
            materialize();
            wave_sync();
            __builtin_amdgcn_s_waitcnt(0);
            const float t_mat = float(clock64() - tm0);
            for (int mm = M_TKO; mm >= M_INFA; --mm) {
                __builtin_amdgcn_s_waitcnt(0);
                const uint64_t t0 = clock64();
                n_fut = run_method<CAP, NARROW>(g, mm, c, tko_rec(g, c), cur, cl, n_act, sparse, fut, fl, &fsp);
                wave_sync();
                if (!fsp) {  // a bitmap-only result pays for its list rebuild
                    n_fut = uniform(rebuild_list<CAP>(fut, W, fl));
                    fsp = n_fut <= CAP;
                }
                wave_sync();
                __builtin_amdgcn_s_waitcnt(0);
                tm[mm] = float(clock64() - t0) + t_mat;
                if (mm == M_TKO) {
                    clear_bits<CAP>(fut, fl, n_fut, fsp, W);
                    wave_sync();
                }
            }
            const float fl = feature(g, ctx, lane_id());
            unsigned long long slot = 0;
            if (lane_id() == 0) slot = atomicAdd(a.log_top, 1ull);
            slot = __shfl(static_cast<long long>(slot), 0, 64);
            if (int64_t(slot) < a.log_cap) {
                float* rec = a.log_buf + int64_t(slot) * kLogRecord;
                const int l = lane_id();
                if (l < kFeatures) rec[l] = fl;
                if (l < 3) rec[kFeatures + l] = tm[l];
                if (l == 0) rec[kFeatures + 3] = float(m);
                if (l == 1) rec[kFeatures + 4] = float(c);
                if (l == 2) rec[kFeatures + 5] = float(n_act);
            }
            m = M_INFA;
        }
        tk_ready = kTko && !fsp && b + 1 < L;
        if (tk_ready) tk_next = tko_rec(g, c_next);
        wave_sync();
        // an AS step cleared the listed states' words as it read them
        if (log_step || m != M_AS) clear_bits<CAP>(cur, cl, n_act, sparse, W);
        // swap
        uint64_t* tb = cur;
        cur = fut;
        fut = tb;
        uint32_t* tl = cl;
        cl = fl;
        fl = tl;
        wave_sync();
        imp = next_imp;
        ctx.c_prev = c;
        if (fsp) {
            n_act = n_fut;
            sparse = true;
        } else {
            // bitmap only: fold the implicit set in, then count (and list
            // the states if they fit)
            n_act = 0;
            sparse = false;
            materialize();
            // count first: a set that stays dense is not listed
            n_act = count_bits(cur, W);
            if (n_act <= CAP) {
                n_act = uniform(rebuild_list<CAP>(cur, W, cl));
                sparse = true;
            }
        }
        // constant indices keep the counters in registers (AS steps: the
        // string length minus these)
        msteps[0] += m == M_INFA ? 1 : 0;
        msteps[1] += m == M_TKO ? 1 : 0;
        ctx.last_m = m;
        if (has_wb) {
            const uint32_t xb = uint32_t(c);
            // a stream defers its last byte's boundary to the next chunk
            const bool boundary = (b < L - 1) ? (word_char(xb) != word_char(uint32_t(c_next)))
                                              : (slot < 0 && word_char(xb));
            if (boundary) wb_pass();
        }
        pair = pair_next;
    }
    materialize();
    wave_sync();
    if (slot >= 0) {
        // save the set before the end-of-string boundary, then report as if
        // the stream ended here
        uint64_t* out = a.stream_out_state + size_t(sid) * W;
        for (int j = lane_id(); j < W; j += 64) out[j] = cur[j];
        const bool last_word = L > 0 ? word_char(uint32_t(c_next)) : ((sflags & 2u) != 0);
        if (lane_id() == 0) a.stream_out_flags[sid] = uint8_t(1u | (last_word ? 2u : 0u));
        if (has_wb && last_word) wb_pass();
        wave_sync();
    }
    emit_accepts(g, a, sid, cur);
    wave_sync();
    clear_bits<CAP>(cur, cl, n_act, sparse, W);
    wave_sync();
}

// ------------------------------------------------------------ serving
// Device-side admission (ScanArgs::serve): each wave takes the next ticket
// and waits until the host has published that many requests.  Only the wave
// holding the lowest unpublished ticket reads the host's counter (over PCIe)
// and mirrors what it saw into device memory; the others poll the mirror.
// Every wait is bounded: tickets past the trace's end exit at once, a stop
// flag (the host's abort, or a wait that outlived serve_wait_ticks) ends
// every wait, so every wave exits.

// Lane 0 polls; true when request t may start.
__device__ bool serve_wait(const ScanArgs& a, int t) {
    if (t >= a.n) return false;
    int go = 0;
    if (lane_id() == 0) {
        const uint64_t t0 = wall_clock64();
        // Polls are relaxed loads and the acquire is one fence once the
        // ticket is admitted: an acquire load at agent or system scope
        // invalidates caches, and hundreds of waiting waves polling with one
        // would keep evicting the table rows of the waves that are scanning.
        for (;;) {
            const int m = __hip_atomic_load(&a.serve_mirror[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (m > t) {
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                go = 1;
                break;
            }
            if (__hip_atomic_load(&a.serve_mirror[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
            if (m == t) {
                // the lowest unpublished ticket: read the host's counter
                const int p = __hip_atomic_load(a.serve_published, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                if (p > t) {
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
                    atomicMax(&a.serve_mirror[0], p);
                    go = 1;
                    break;
                }
                if (__hip_atomic_load(a.serve_abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0) {
                    atomicOr(&a.serve_mirror[1], 1);
                    break;
                }
            }
            if (wall_clock64() - t0 > uint64_t(a.serve_wait_ticks)) {
                atomicOr(&a.serve_mirror[1], 2);
                break;
            }
            // the head polls often; waves further back sleep longer (all
            // waiting waves read one mirror word, so their reads are spread)
            if (m == t) {
                __builtin_amdgcn_s_sleep(2);
            } else if (t - m < 256) {
                __builtin_amdgcn_s_sleep(32);
            } else {
                __builtin_amdgcn_s_sleep(127);
                __builtin_amdgcn_s_sleep(127);
                __builtin_amdgcn_s_sleep(127);
                __builtin_amdgcn_s_sleep(127);
            }
        }
    }
    return __builtin_amdgcn_readlane(go, 0) != 0;
}

// Copies request bytes [beg, beg+L) of host memory into the wave's staging
// slot with 16-byte loads from the aligned-down start (the host buffer is
// 16-byte aligned and padded past its end); returns where the string starts
// in the slot.
__device__ const uint8_t* serve_fetch(const ScanArgs& a, int64_t beg, int L, uint8_t* slot) {
    const int64_t b0 = beg & ~int64_t(15);
    const int nvec = int((beg + L - b0 + 15) >> 4);
    const uint4* src = reinterpret_cast<const uint4*>(a.text + b0);
    uint4* dst = reinterpret_cast<uint4*>(slot);
    for (int k0 = 0; k0 < nvec; k0 += 256) {
        uint4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = src[min(k0 + 64 * u + lane_id(), nvec - 1)];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (k0 + 64 * u + lane_id() < nvec) dst[k0 + 64 * u + lane_id()] = v[u];
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    return slot + (beg - b0);
}

// The request's results (written to host memory by the scan) before its
// done flag.
__device__ void serve_finish(const ScanArgs& a, int sid) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
    if (lane_id() == 0) __hip_atomic_store(&a.serve_done[sid], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// ------------------------------------------------------------ hashed storage
// The same step loop with the set held as the active list plus an LDS hash
// table of 2*CAP slots per list (instead of two S-bit bitmaps), so a wave's
// LDS no longer grows with the rule set and many more waves fit on a CU.
// Two S-bit bitmaps per wave live in HBM (ScanArgs::dense) and are touched
// only by dense steps: a set of more than CAP states, a step whose method is
// iNFA or TKO, an AS step that overflowed the list (redone as iNFA), and
// acceptance.  Invariants between steps: the fut table is empty and the fut
// bitmap is zero; a sparse set has its states in cl and hc (cur bitmap zero),
// a dense set in the cur bitmap (hc empty).

// Orders this wave's global-memory accesses across lanes (the HBM bitmaps).
__device__ __forceinline__ void dense_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}


// iNFA or TKO step on whole bitmaps (fut zero on entry).
static inline __device__ void dense_method(const DevGraph& g, int m, int c, const uint64_t* cur, uint64_t* fut) {
    if (m == M_TKO) {
        const TkoRec tk = tko_rec(g, c);
        tko_groups<1>(g, tk, cur, fut);
        dense_sync();
        scatter<64, false>(g.tko_left + tk.lb(), tk.le() - tk.lb(), cur, fut, nullptr, 0);
    } else {
        for (int j = lane_id(); j < g.W; j += 64) fut[j] = cur[j] & g.persis[j];
        dense_sync();
        scatter<64, false>(g.list + g.sym_off[c], g.sym_off[c + 1] - g.sym_off[c], cur, fut, nullptr, 0);
    }
    dense_sync();
}

template <int CAP, int MODE>
__device__ void scan_string_h(const DevGraph& g, const ScanArgs& a, int sid, const uint8_t* x, int L, uint64_t* gA,
                              uint64_t* gB, uint32_t* hA, uint32_t* hB, uint32_t* listA, uint32_t* listB,
                              int* msteps) {
    using HS = HashSet<2 * CAP>;
    // the selector mode, and whether the boundary code and the path for
    // missing two-byte rows are compiled in
    constexpr int M = MODE & ~kModeLean;
    constexpr bool kLean = (MODE & kModeLean) != 0;
    const int W = g.W;
    msteps[3] += L;
    uint64_t* gcur = gA;
    uint64_t* gfut = gB;
    HS hc{hA}, hf{hB};
    uint32_t* cl = listA;
    uint32_t* fl = listB;
    // the always-active states start implicit, the other initial states listed
    int n_act = uniform(rebuild_list<CAP>(g.init_explicit, W, cl));
    bool sparse = n_act <= CAP;
    if (sparse) {
        for (int k = lane_id(); k < n_act; k += 64) hc.insert(cl[k]);
    } else {
        for (int j = lane_id(); j < W; j += 64) gcur[j] = g.init_explicit[j];
        dense_sync();
    }
    wave_sync();
    ByteWindow win{x, L, 0, 0u};
    if (L > 0) win.load(0);
    const bool has_wb = !kLean && g.wb_count > 0;
    const bool have_pair = kLean || g.pair_off != nullptr;
    int imp = -2;  // implicitly held states, as in scan_string
    auto alw_range = [&](int cc, int& b0, int& n0) {
        b0 = g.sym_stats[cc].w;
        n0 = (cc < 255 ? g.sym_stats[cc + 1].w : g.alw_total) - b0;
    };
    // sparse -> dense: the listed states into the cur bitmap
    auto to_dense = [&]() {
        if (!sparse) return;
        for (int k = lane_id(); k < n_act; k += 64) set_bit_atomic(gcur, cl[k] & kIdMask);
        hc.clear();
        sparse = false;
        dense_sync();
        wave_sync();
    };
    auto materialize = [&]() {
        if (imp == -1) return;
        int ab = 0, an = g.n_alw_states;
        const uint32_t* src = g.alw_states;
        if (imp >= 0) {
            alw_range(imp, ab, an);
            src = g.alw_dst;
        }
        if (sparse && n_act + an <= CAP) {
            for (int t0 = 0; t0 < an; t0 += 64) {
                const int t = t0 + lane_id();
                const bool ok = t < an;
                const uint32_t v = ok ? (imp >= 0 ? src[ab + t] : (src[t] | kTag)) : 0u;
                const bool nw = ok && hc.insert(v & kIdMask);
                n_act = list_append<CAP>(cl, n_act, nw, v & (kIdMask | kTag));
            }
            n_act = uniform(n_act);
            wave_sync();
        } else {
            to_dense();
            for (int t = lane_id(); t < an; t += 64) set_bit_atomic(gcur, src[ab + t] & kIdMask);
            dense_sync();
            n_act = count_bits(gcur, W);
        }
        imp = -1;
    };
    auto wb_pass = [&]() {
        materialize();
        if (g.wb_parallel) {
            // no destination is a source: one parallel pass.  The edges whose
            // source is active bound the states it adds; the list form is kept
            // when they fit.
            int hits = CAP + 1;
            if (sparse) {
                hits = 0;
                for (int t0 = 0; t0 < g.wb_count; t0 += 64) {
                    const int t = t0 + lane_id();
                    const bool ok = t < g.wb_count;
                    const uint2 e = ok ? g.list[g.wb_begin + t] : make_uint2(0u, 0u);
                    hits += __popcll(__ballot(ok && hc.contains(e.x)));
                }
            }
            if (n_act + hits <= CAP) {
                for (int t0 = 0; t0 < g.wb_count; t0 += 64) {
                    const int t = t0 + lane_id();
                    const bool ok = t < g.wb_count;
                    const uint2 e = ok ? g.list[g.wb_begin + t] : make_uint2(0u, 0u);
                    const bool nw = ok && hc.contains(e.x) && hc.insert(e.y);
                    n_act = list_append<CAP>(cl, n_act, nw, e.y);
                }
                n_act = uniform(n_act);
                wave_sync();
            } else {
                to_dense();
                for (int t0 = 0; t0 < g.wb_count; t0 += 64) {
                    const int t = t0 + lane_id();
                    if (t < g.wb_count) {
                        const uint2 e = g.list[g.wb_begin + t];
                        if (test_bit(gcur, e.x)) set_bit_atomic(gcur, e.y);
                    }
                }
                dense_sync();
                n_act = count_bits(gcur, W);
            }
            return;
        }
        // in place, list order, one lane (thread_count == 1 semantics of
        // src/iNFA_kernel.cu:145-164)
        if (sparse && n_act + g.wb_count <= CAP) {
            int n = n_act;
            if (lane_id() == 0) {
                for (int t = 0; t < g.wb_count; ++t) {
                    const uint2 e = g.list[g.wb_begin + t];
                    if (hc.contains(e.x) && hc.insert(e.y)) cl[n++] = e.y;
                }
            }
            n_act = uniform(__builtin_amdgcn_readlane(n, 0));
            wave_sync();
        } else {
            to_dense();
            if (lane_id() == 0) {
                for (int t = 0; t < g.wb_count; ++t) {
                    const uint2 e = g.list[g.wb_begin + t];
                    if (test_bit(gcur, e.x)) gcur[e.y >> 6] |= 1ull << (e.y & 63);
                }
            }
            dense_sync();
            n_act = count_bits(gcur, W);
        }
    };
    int c_next = L > 0 ? win.at(0) : 0;
    if (has_wb && L > 0 && word_char(uint32_t(c_next))) wb_pass();

    auto load_pair = [&](int c1, int c2) -> RowSlice {
        RowSlice r;
        // read-only offsets: scalar loads
        using CU32 = const __attribute__((address_space(4))) uint32_t;
        // a 32-bit byte offset lets the load take it as its scalar offset
        const uint32_t off = ((uint32_t(c1) << 8) | uint32_t(c2)) << 2;
        const CU32* o = (const CU32*)((const __attribute__((address_space(4))) char*)(g.pair_off) + off);
        r.b = int(o[0]);
        r.n = int(o[1]) - r.b;
        r.v = lane_id() < r.n ? g.pair_dst[r.b + lane_id()] : 0u;
        return r;
    };
    RowSlice pair{0u, 0, 0};
    const bool can_as = g.as_tab != nullptr;

    StepCtx ctx;
    ctx.len = L;
    ctx.last_m = M_INFA;
    ctx.c_prev = -1;
    int cached = -1;
    for (int b = 0; b < L; ++b) {
        // the next byte is read one step ahead (two-byte rows, word boundaries)
        // (at the last byte the window reads 0 past the end and the row is unused)
        const int c = c_next;
        RowSlice pair_next{0u, 0, 0};
        c_next = win.at(b + 1);
        if (have_pair) pair_next = load_pair(c, c_next);
        auto total_active = [&]() {
            int n_tot = n_act;
            if (imp >= 0) {
                int ab, an;
                alw_range(imp, ab, an);
                n_tot += an;
            } else if (imp == -2) {
                n_tot += g.n_alw_states;
            }
            return n_tot;
        };
        ctx.c = c;
        ctx.sparse = sparse;
        ctx.pos = b;
        bool log_step = false;
        int m = choose_method<M>(g, a, sid, ctx, n_act, sparse, total_active, cached, log_step);
        if (m == M_AS && !(sparse && can_as)) m = M_INFA;
        int n_fut = 0;
        bool fsp = false;
        if (m == M_AS) {
            n_fut = as_step<CAP>(AsRows{g.as_tab, g.as_multi, g.pair_dst}, c, pair, have_pair && imp >= 0, cl, n_act,
                                 hf, fl, nullptr);
            wave_sync();
            if (n_fut <= CAP) {
                fsp = true;
            } else {
                // the list overflowed: drop the partial result, redo as iNFA
                hf.clear();
                wave_sync();
                m = M_INFA;
            }
        }
        if (fsp) {
            hc.clear();  // retire cur (its bitmap is zero)
            HS th = hc;
            hc = hf;
            hf = th;
            uint32_t* tl = cl;
            cl = fl;
            fl = tl;
            n_act = n_fut;
            imp = c;
            wave_sync();
            // without two-byte rows the next step could not reach the first
            // positions of the implicit set from the table: list them now
            if (!have_pair) materialize();
        } else {
            if (M == MODE_LEARNED_AS && !sparse) m = dense_policy(a, n_act, c);
            // bitmap steps counted here, off the AS path (AS steps: the string
            // length minus these)
            msteps[0] += m == M_TKO ? 0 : 1;
            msteps[1] += m == M_TKO ? 1 : 0;
            materialize();
            to_dense();
            dense_method(g, m, c, gcur, gfut);
            for (int j = lane_id(); j < W; j += 64) gcur[j] = 0ull;  // retire cur
            uint64_t* tb = gcur;
            gcur = gfut;
            gfut = tb;
            dense_sync();
            // count first: a set that stays dense is not listed
            n_act = count_bits(gcur, W);
            if (n_act <= CAP) n_act = uniform(rebuild_list<CAP>(gcur, W, cl));
            if (n_act <= CAP) {
                // back to the list form
                for (int k = lane_id(); k < n_act; k += 64) {
                    hc.insert(cl[k]);
                    gcur[cl[k] >> 6] = 0ull;
                }
                sparse = true;
                dense_sync();
            }
            wave_sync();
            imp = -1;
        }
        ctx.c_prev = c;
        ctx.last_m = m;
        if (has_wb) {
            const uint32_t xb = uint32_t(c);
            const bool boundary = (b < L - 1) ? (word_char(xb) != word_char(uint32_t(c_next))) : word_char(xb);
            if (boundary) wb_pass();
        }
        pair = pair_next;
    }
    // accepted states of a valid list: the listed ones and the implicitly
    // held ones (not in the table), compacted into the free list; up to 64
    // are emitted by rank without touching the bitmaps
    int cnt = -1;
    if (sparse && !a.emit_all) {
        cnt = 0;
        for (int k0 = 0; k0 < n_act; k0 += 64) {
            const int k = k0 + lane_id();
            const uint32_t id = k < n_act ? (cl[k] & kIdMask) : 0u;
            cnt = list_append<CAP>(fl, cnt, k < n_act && test_bit(g.accept, id), id);
        }
        if (imp != -1) {
            int ab = 0, an = g.n_alw_states;
            const uint32_t* src = g.alw_states;
            if (imp >= 0) {
                alw_range(imp, ab, an);
                src = g.alw_dst;
            }
            for (int t0 = 0; t0 < an; t0 += 64) {
                const int t = t0 + lane_id();
                const uint32_t id = t < an ? (src[ab + t] & kIdMask) : 0u;
                cnt = list_append<CAP>(fl, cnt, t < an && test_bit(g.accept, id) && !hc.contains(id), id);
            }
        }
        cnt = uniform(cnt);
        wave_sync();
    }
    if (cnt == 0) {
        if (lane_id() == 0) {
            a.res_count[sid] = 0;
            a.res_start[sid] = 0;
        }
        hc.clear();
    } else if (cnt > 0 && cnt <= 64) {
        // ascending by rank: lane l's id is larger than `rank` others (the
        // ids are distinct)
        const uint32_t v = lane_id() < cnt ? fl[lane_id()] : 0xffffffffu;
        int rank = 0;
        for (int j = 0; j < cnt; ++j) rank += uint32_t(__builtin_amdgcn_readlane(int(v), j)) < v ? 1 : 0;
        unsigned long long s0 = 0;
        if (lane_id() == 0) s0 = atomicAdd(a.pool_top, (unsigned long long)cnt);
        const int64_t start = int64_t(__shfl(static_cast<long long>(s0), 0, 64));
        int total = cnt;
        if (start + cnt > a.pool_cap) {
            if (lane_id() == 0) atomicOr(a.overflow, 1);
            total = -cnt;  // marks a truncated result
        } else if (lane_id() < cnt) {
            a.res_pool[start + rank] = v;
        }
        if (lane_id() == 0) {
            a.res_count[sid] = total;
            a.res_start[sid] = start;
        }
        hc.clear();
    } else {
        // many accepted states, or a bitmap-only set: the whole set, then the
        // accepted states through a bitmap in ascending order
        materialize();
        if (sparse) {
            for (int k = lane_id(); k < n_act; k += 64) {
                const uint32_t id = cl[k] & kIdMask;
                if (a.emit_all || test_bit(g.accept, id)) set_bit_atomic(gfut, id);
            }
            dense_sync();
            emit_accepts(g, a, sid, gfut);
            dense_sync();
            for (int k = lane_id(); k < n_act; k += 64) gfut[(cl[k] & kIdMask) >> 6] = 0ull;
            dense_sync();
            hc.clear();
        } else {
            emit_accepts(g, a, sid, gcur);
            dense_sync();
            for (int j = lane_id(); j < W; j += 64) gcur[j] = 0ull;
            dense_sync();
        }
    }
    wave_sync();
}

// The LDS-storage engine kernel's (DevGraph, ScanArgs) arguments, read
// through the kernarg segment pointer made opaque once per string.  Reading
// the by-value parameters directly makes every field the kernel uses a load at
// kernel entry, live in scalar registers for the whole kernel, and the step
// loop then spills dozens of them to vector lanes and reloads them every step;
// through the opaque pointer a field is loaded where it is used, at most once
// per string.  The explicit arguments start at offset 0 of the segment, each
// at its own alignment (the HSA kernel argument layout).  The hashed kernel
// reads its parameters directly: it has fewer spills to begin with and ran
// 3% slower through the pointer on the sparse large set.
struct KernelArgs {
    const DevGraph* g;
    const ScanArgs* a;
};
__device__ __forceinline__ KernelArgs kernel_args() {
    using KG = const __attribute__((address_space(4))) DevGraph;
    using KA = const __attribute__((address_space(4))) ScanArgs;
    using KC = const __attribute__((address_space(4))) char;
    constexpr size_t kArgsOff = (sizeof(DevGraph) + alignof(ScanArgs) - 1) / alignof(ScanArgs) * alignof(ScanArgs);
    KG* g = (KG*)__builtin_amdgcn_kernarg_segment_ptr();
    KA* a = (KA*)((KC*)__builtin_amdgcn_kernarg_segment_ptr() + kArgsOff);
    asm volatile("" : "+s"(g), "+s"(a));
    return KernelArgs{(const DevGraph*)g, (const ScanArgs*)a};
}

template <int CAP>
struct LaneH {
    // u64 words per wave: two lists of CAP + 64 u32 and two tables of
    // 2*CAP + 64 u32 (each followed by its 64 sink entries)
    static constexpr size_t words() { return size_t(3) * CAP + 128; }
};

// Build knob: ADM_HASHED_WAVES > 0 asks the compiler to fit that many waves
// per SIMD in registers (it may spill to scratch to do so).
#ifndef ADM_HASHED_WAVES
#define ADM_HASHED_WAVES 0
#endif
#if ADM_HASHED_WAVES > 0
#define ADM_HASHED_ATTR __attribute__((amdgpu_waves_per_eu(ADM_HASHED_WAVES)))
#else
#define ADM_HASHED_ATTR
#endif

// The per-byte statistics stay in device memory here (they are read off the
// AS fast path only), keeping the workgroup's LDS to the waves' own.
template <int WPB, int CAP, int MODE>
__global__ __launch_bounds__(64 * WPB) ADM_HASHED_ATTR void engine_h_kernel(DevGraph g, ScanArgs a) {
    extern __shared__ uint64_t lds[];
    const int wave = threadIdx.x >> 6;
    uint64_t* base = lds + LaneH<CAP>::words() * wave;
    uint32_t* listA = reinterpret_cast<uint32_t*>(base);
    uint32_t* listB = listA + CAP + 64;
    uint32_t* hA = listB + CAP + 64;
    uint32_t* hB = hA + HashSet<2 * CAP>::kSlots;
    uint64_t* gA = a.dense + (size_t(blockIdx.x) * WPB + wave) * 2 * size_t(g.W);
    uint64_t* gB = gA + g.W;
    HashSet<2 * CAP>{hA}.clear();
    HashSet<2 * CAP>{hB}.clear();
    HashSet<2 * CAP>{hA}.init_sink();
    HashSet<2 * CAP>{hB}.init_sink();
    wave_sync();
    int msteps[4] = {0, 0, 0, 0};  // iNFA, TKO, AS (filled below), bytes scanned
    uint8_t* stage = a.serve_stage + (size_t(blockIdx.x) * WPB + wave) * size_t(a.serve_slot);
    for (;;) {
        int idx = 0;
        if (lane_id() == 0) idx = atomicAdd(a.queue, 1);
        idx = uniform(__builtin_amdgcn_readlane(idx, 0));
        if (a.serve) {
            if (!serve_wait(a, idx)) break;
        } else if (idx >= a.n) {
            break;
        }
        const int sid = a.order ? a.order[idx] : idx;
        const int64_t beg = a.offsets[sid];
        const int L = int(a.offsets[sid + 1] - beg);
        const uint8_t* x = a.serve ? serve_fetch(a, beg, L, stage) : a.text + beg;
        scan_string_h<CAP, MODE>(g, a, sid, x, L, gA, gB, hA, hB, listA, listB, msteps);
        if (a.serve) serve_finish(a, sid);
    }
    msteps[2] = msteps[3] - msteps[0] - msteps[1];
    if (a.method_steps != nullptr && lane_id() == 0)
        for (int m = 0; m < 3; ++m)
            if (msteps[m]) atomicAdd(&a.method_steps[m], (unsigned long long)msteps[m]);
}

// Two waves per SIMD: large rule sets fill the LDS with eight waves per CU,
// and the register budget that leaves pays for the TKO step's one-pass rows
// (kTkoOnePass words per lane in flight) and 16-edge scatter rounds.
// (MODE & kModeNarrow: state vectors of at most 64 * kTkoPassChunks words;
// the smaller TKO step lets four waves per SIMD fit in the registers)
template <int WPB, int CAP, int MODE>
__global__ __launch_bounds__(64 * WPB) __attribute__((amdgpu_waves_per_eu((MODE & kModeNarrow) ? 4 : 1, (MODE & kModeNarrow) ? 4 : 2))) void
engine_kernel(DevGraph g, ScanArgs a) {
    constexpr int M = MODE & ~(kModeNarrow | kModeLean);
    constexpr bool kNarrow = (MODE & kModeNarrow) != 0;
    constexpr bool kLean = (MODE & kModeLean) != 0;
    // the per-byte statistics stay in device memory (scalar loads): a 4 KB
    // copy per workgroup in LDS would cost large rule sets a workgroup per CU
    extern __shared__ uint64_t lds[];
    const int wave = threadIdx.x >> 6;
    const size_t per = Lane<CAP>::words(g.W);
    uint64_t* base = lds + per * wave;
    uint64_t* bitsA = base;
    uint64_t* bitsB = base + g.W;
    uint32_t* listA = reinterpret_cast<uint32_t*>(base + 2 * g.W);
    uint32_t* listB = listA + CAP;
    for (int j = lane_id(); j < 2 * g.W; j += 64) base[j] = 0ull;
    wave_sync();
    int msteps[4] = {0, 0, 0, 0};  // iNFA, TKO, AS (filled below), bytes scanned
    for (;;) {
        const KernelArgs k = kernel_args();
        const ScanArgs& ka = *k.a;
        int idx = 0;
        if (lane_id() == 0) idx = atomicAdd(ka.queue, 1);
        idx = uniform(__builtin_amdgcn_readlane(idx, 0));
        if (ka.serve) {
            if (!serve_wait(ka, idx)) break;
        } else if (idx >= ka.n) {
            break;
        }
        const int sid = ka.order ? ka.order[idx] : idx;
        const int64_t beg = ka.offsets[sid];
        const int L = int(ka.offsets[sid + 1] - beg);
        uint8_t* stage = ka.serve_stage + (size_t(blockIdx.x) * WPB + wave) * size_t(ka.serve_slot);
        const uint8_t* x = ka.serve ? serve_fetch(ka, beg, L, stage) : ka.text + beg;
        scan_string<CAP, M, kNarrow, kLean>(*k.g, *k.a, sid, x, L, bitsA, bitsB, listA, listB, msteps);
        if (ka.serve) serve_finish(ka, sid);
    }
    msteps[2] = msteps[3] - msteps[0] - msteps[1];
    const ScanArgs& ka = *kernel_args().a;
    if (ka.method_steps != nullptr && lane_id() == 0)
        for (int m = 0; m < 3; ++m)
            if (msteps[m]) atomicAdd(&ka.method_steps[m], (unsigned long long)msteps[m]);
}

}  // namespace admatch
