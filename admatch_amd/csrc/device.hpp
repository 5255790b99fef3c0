// Device-side types and wave64 helpers shared by the scan kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace admatch {

// Read-only compiled rule set as the kernels see it (all pointers device).
struct DevGraph {
    int32_t S, W;           // states, u64 words per state vector
    const uint64_t* init;   // [W]
    const uint64_t* persis; // [W]
    const uint64_t* accept; // [W]
    const int32_t* sym_off; // [258] iNFA CSR offsets
    const uint2* list;      // (src,dst) per symbol, sorted
    const int32_t* tko_koff;  // [258]
    const int2* tko_grp;      // [K] group i: shift d and its row in tko_lim
    const uint64_t* tko_lim;  // distinct source-mask rows: word j of row r at r * tko_stride + kTkoPad + j
    int32_t tko_stride;       // zero outside [0, W), readable for j in [-kTkoPad, tko_stride - kTkoPad)
    const int32_t* tko_loff;  // [258]
    const uint2* tko_left;    // leftover (src,dst)
    // [256][64] per byte, one int2 per lane: lanes 0..59 the (shift, row) of
    // the byte's first 60 groups, lane 60 its groups [x, y), lane 61 its
    // leftover edges [x, y)
    const int2* tko_rec;
    const uint32_t* as_tab;   // [256*S] AS rows, byte major (null: AS unavailable)
    const uint32_t* as_multi; // destinations of multi-entry rows
    const uint32_t* alw_dst;  // always-active rows per byte
    int32_t alw_total;        // alw_dst entries
    const uint32_t* pair_off; // [65537] two-byte rows of first positions (null: none)
    const uint32_t* pair_dst;
    const uint32_t* alw_states;  // always-active states (initial and persistent)
    int32_t n_alw_states;
    const uint64_t* init_explicit;  // [W] initial states that are not always active
    const int4* sym_stats;    // [256]: |T_c|, k_c, |leftover_c|, start in alw_dst
    int32_t wb_begin, wb_count;  // word-boundary slice of list
    int32_t wb_parallel;         // no boundary destination is a boundary source
    // boundary successors per state and per implicit set (null unless
    // wb_parallel; read by the LDS-storage engine): wbs_dst[wbs_off[s] ..),
    // wbi_dst[wbi_off[r] ..), r = byte of the implicit row or 256
    const uint32_t* wbs_off;  // [S+1]
    const uint32_t* wbs_dst;
    const int32_t* wbi_off;   // [258]
    const uint32_t* wbi_dst;
};

// Per-launch arguments of the batched engine.
struct ScanArgs {
    const uint8_t* text;      // concatenated strings
    const int64_t* offsets;   // [n+1]
    const int32_t* order;     // [n] processing order (longest first) or null
    int32_t n;
    int32_t mode;             // step selector, see SelectMode
    int32_t* queue;           // [1] work counter (zeroed before launch)
    int32_t* res_count;       // [n] accepted-state count per string
    int64_t* res_start;       // [n] start in res_pool
    uint32_t* res_pool;       // accepted state ids, ascending per string
    int64_t pool_cap;
    unsigned long long* pool_top;  // [1] (zeroed before launch)
    int32_t* overflow;        // [1] set when the pool is exhausted
    const float* mlp;         // learned selector weights or null
    int32_t mlp_every;        // >0: evaluate every k steps; 0: per-byte decision cache
    // step logging (MODE_LOG): per sampled step 16 features + 3 times
    float* log_buf;
    int64_t log_cap;
    unsigned long long* log_top;
    int32_t log_stride;       // sample one step in log_stride
    unsigned long long* method_steps;  // [4] steps per method, or null
    uint64_t* dense;          // hashed storage: two W-word bitmaps per wave of the grid, all zero
    const uint32_t* policy;   // learned selector table (2 bits per entry) or null
    // streaming (LDS-storage engine only): string i continues stream
    // stream_ids[i], whose saved set (W words) and flags are read from that
    // slot; the set and flags after string i are written to row i of the
    // out arrays (committed to the slots once the scan has succeeded)
    const int32_t* stream_ids;
    const uint64_t* stream_state;
    const uint8_t* stream_flags;  // bit 0: a set is saved, bit 1: the last byte so far was a word byte
    uint64_t* stream_out_state;
    uint8_t* stream_out_flags;
    int32_t policy_sparse, policy_dense;  // the table's method for every entry of that list validity, or -1
    int32_t emit_all;         // write every final state, not only the accepting ones (SHOW_RESULTS)
    // device-side serving (Engine::serve_ring): ticket t may start once the
    // host has published more than t requests.  Requests (text, offsets) live
    // in host memory and are copied into the wave's staging slot; results
    // (res_*) and a done flag per request go to host memory.
    int32_t serve;
    const int32_t* serve_published;  // host: requests published so far
    const int32_t* serve_abort;      // host: nonzero stops every wait
    int32_t* serve_mirror;           // device: [0] published as last read, [1] stop seen
    int32_t* serve_done;             // host: per request, 1 once its results are written
    uint8_t* serve_stage;            // device: serve_slot bytes per wave of the grid
    int64_t serve_slot;
    int64_t serve_wait_ticks;        // longest wait for one ticket, 100 MHz wall-clock ticks
};

enum SelectMode : int32_t {
    MODE_INFA = 0,     // engine, iNFA scatter every step
    MODE_TKO = 1,      // engine, TKO every step
    MODE_AS = 2,       // engine, active-state rows every step (iNFA when dense)
    MODE_ADMATCH = 3,  // reference rule: k_c > active ? AS : TKO
    MODE_HEUR = 4,     // AdMatch rule extended with an iNFA cost term
    MODE_LEARNED = 5,  // MLP selector, MODE_HEUR as fallback
    MODE_LOG = 6,      // run all three methods per step and log timings
    MODE_LEARNED_EVERY = 7,  // internal: the network evaluated in the scan every mlp_every steps
    MODE_LEARNED_AS = 8,     // internal: policy table whose reachable list-valid entries all pick AS
};

// Kernels are also built with this bit added to the mode, for rule sets
// without word-boundary transitions and with two-byte rows: the boundary code,
// the path for missing rows and (LDS storage) the stream code are compiled out
// of the step loop, which leaves the loop fewer scalar registers to spill.
// Streams, the logging mode and the every-k-steps network mode use the full
// kernels.
constexpr int32_t kModeLean = 0x100;
// LDS-storage kernels are also built with this bit added to the mode, for
// state vectors of at most 64 * kTkoPassChunks words: a smaller TKO step and
// up to four waves per SIMD.
constexpr int32_t kModeNarrow = 0x200;

enum Method : int32_t { M_INFA = 0, M_TKO = 1, M_AS = 2 };

constexpr int kTkoRecGroups = 60;  // groups held in a byte's TKO record
// TKO row passes: one pass of up to kTkoOnePass 64-word chunks per lane when
// the state vector fits, else passes of kTkoPassChunks chunks; rows padded
// with kTkoPad zero words on each side.
constexpr int kTkoOnePass = 18;
constexpr int kTkoPassChunks = 8;
constexpr int kTkoPad = 64;

constexpr int kFeatures = 16;
constexpr int kPolicyTableWords = 512;  // 16 buckets x 2 list states x 256 bytes x 2 bits
constexpr int kHidden = 64;
constexpr int kOutputs = 3;
// Packed MLP layout (floats): W1[64][16], b1[64], W2[64][64], b2[64],
// W3[3][64], b3[3]; input normalisation folded into W1/b1.
constexpr int kMlpW1 = 0;
constexpr int kMlpB1 = kMlpW1 + kHidden * kFeatures;
constexpr int kMlpW2 = kMlpB1 + kHidden;
constexpr int kMlpB2 = kMlpW2 + kHidden * kHidden;
constexpr int kMlpW3 = kMlpB2 + kHidden;
constexpr int kMlpB3 = kMlpW3 + kOutputs * kHidden;
constexpr int kMlpParams = kMlpB3 + kOutputs;  // 5,443
// log record: features, 3 cycle counts, chosen method, byte, explicit active count
constexpr int kLogRecord = kFeatures + kOutputs + 3;

__device__ __forceinline__ bool word_char(uint32_t c) {
    return (c - 'A' < 26u) || (c - 'a' < 26u) || (c - '0' < 10u) || c == '_';
}
__device__ __forceinline__ int lane_id() { return __lane_id(); }
__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
}
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
// Orders this wave's LDS accesses across lanes (one wave owns a string).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// Cross-lane moves below are DPP modifiers of VALU moves (no LDS round
// trip): quad permutes, row (16-lane) rotates and shifts.  The whole wave
// must be active.
// Sum over the 64 lanes (the same value in every lane).
__device__ __forceinline__ int wave_sum(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xf, 0xf, false);   // quad_perm [1,0,3,2]
    v += __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xf, 0xf, false);   // quad_perm [2,3,0,1]
    v += __builtin_amdgcn_update_dpp(0, v, 0x124, 0xf, 0xf, false);  // row_ror:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x128, 0xf, 0xf, false);  // row_ror:8
    // every lane holds its row's sum
    return __builtin_amdgcn_readlane(v, 0) + __builtin_amdgcn_readlane(v, 16) + __builtin_amdgcn_readlane(v, 32) +
           __builtin_amdgcn_readlane(v, 48);
}
__device__ __forceinline__ float wave_sum_f(float v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// inclusive prefix sum across the 64 lanes
__device__ __forceinline__ int wave_scan(int v) {
    // within each row of 16 (a lane whose source is before its row's start reads 0)
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);  // row_shr:8
    // then the totals of the rows below
    const int r0 = __builtin_amdgcn_readlane(v, 15);
    const int r1 = r0 + __builtin_amdgcn_readlane(v, 31);
    const int r2 = r1 + __builtin_amdgcn_readlane(v, 47);
    const int l = lane_id();
    return v + (l < 16 ? 0 : l < 32 ? r0 : l < 48 ? r1 : r2);
}
__device__ __forceinline__ bool test_bit(const uint64_t* v, uint32_t i) { return (v[i >> 6] >> (i & 63)) & 1ull; }

}  // namespace admatch
