// Host side of the scan engine: device residency of a compiled rule set,
// per-stream scan contexts (buffers, staging, captured graphs), launches and
// result extraction.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "device.hpp"
#include "graph.hpp"

namespace admatch {

enum Status : int {
    ADM_OK = 0,
    ADM_ERR_ARG = -1,
    ADM_ERR_LOAD = -2,
    ADM_ERR_HIP = -3,
    ADM_ERR_STATE = -4,     // call order (e.g. scan before alloc)
    ADM_ERR_CAPACITY = -5,  // output buffer too small
    ADM_ERR_TOO_LARGE = -6, // rule set does not fit the kernel's LDS budget
    ADM_ERR_BUDGET = -7,    // admission budget exhausted
    ADM_ERR_POOL = -8,      // accepted-state pool overflowed (grown; scan again)
};

// Kernel used for a scan: the plain single-kernel iNFA path or the engine
// with a step selector.
enum KernelKind : int32_t { KIND_PLAIN = 0, KIND_ENGINE = 1 };

struct Variant {
    int wpb = 2;        // waves (strings) per workgroup
    int cap = 512;      // active-list capacity per string
    int plain_bs = 64;  // workgroup size of the plain iNFA kernel
    // engine storage for the AS and learned modes: 0 = two S-bit bitmaps per
    // string in LDS; 1 = active lists with LDS hash tables, the bitmaps in
    // HBM (ScanArgs::dense).  The other modes always use LDS bitmaps.
    int store = 0;
};

struct ScanStats {
    float kernel_ms = 0.f;  // last scan: hipEvent time of the launch sequence
    float engine_ms = 0.f;  // last engine scan: hipEvent time of the engine kernel alone
    // phases of the last host-strings scan (the reference's PROFILER_MODE,
    // src/kernel_common.cu:57-95): device buffer growth (host clock; buffers
    // are reused, so usually 0), host-to-device copies and result copies
    // (hipEvents), accept-map lookup on the host (host clock)
    float alloc_ms = 0.f, h2d_ms = 0.f, d2h_ms = 0.f, map_ms = 0.f;
    unsigned long long method_steps[3] = {0, 0, 0};
    long long pool_used = 0;
    // engine kernel instance of the last launch: its mode argument (the mode
    // plus the kModeLean / kModeNarrow bits), -1 for the plain kernel; and
    // its storage (1: hashed lists)
    int32_t kernel_mode = -1, kernel_store = 0;
};

// Everything one in-flight scan needs besides the rule set: a stream,
// result buffers, counters, host/device staging and captured graphs.
// Contexts of one engine may run concurrently on the device.
// Device state of a set of streams (each a string scanned in chunks).
struct StreamSet {
    int32_t cap = 0;               // streams
    int32_t W = 0;                 // u64 words per saved set
    uint64_t* d_state = nullptr;   // [cap][W]
    uint8_t* d_flags = nullptr;    // [cap], 0 = nothing saved
    ~StreamSet();
};

struct Context {
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, done = nullptr;
    hipEvent_t ev_h2d0 = nullptr, ev_h2d1 = nullptr, ev_d2h0 = nullptr, ev_d2h1 = nullptr;
    // events around the engine kernel; it is always launched outside a captured
    // graph, since HIP gives no elapsed time for events recorded in one
    hipEvent_t ev_k0 = nullptr, ev_k1 = nullptr;
    bool engine_timed = false;  // the last launch sequence recorded ev_k0/ev_k1
    // emit every final state, not only the accepting ones; the last scan's
    // final states of string i are fin_states[fin_off[i] .. fin_off[i+1])
    bool emit_all = false;
    std::vector<int32_t> fin_off;
    std::vector<uint32_t> fin_states;
    int32_t res_n_cap = 0;
    int64_t pool_cap = 0;
    int32_t* d_res_count = nullptr;
    int64_t* d_res_start = nullptr;
    uint32_t* d_pool = nullptr;
    int32_t* d_order = nullptr;
    int32_t* d_ctr = nullptr;
    int32_t* h_ctr = nullptr;  // pinned mirrors for asynchronous collection
    int32_t* h_res_count = nullptr;
    int64_t* h_res_start = nullptr;
    int32_t h_res_cap = 0;
    uint8_t* h_text = nullptr;
    uint8_t* d_text = nullptr;
    int64_t text_cap = 0;
    int64_t* h_off = nullptr;
    int64_t* d_off = nullptr;
    int32_t off_cap = 0;
    uint64_t* d_dense = nullptr;  // hashed storage bitmaps (zero between launches)
    size_t dense_words = 0;
    // streaming scans: the stream set of the scan in progress, the stream of
    // each string and the per-string output sets
    const StreamSet* active_stream = nullptr;
    int32_t* d_stream_ids = nullptr;
    uint64_t* d_stream_out = nullptr;
    uint8_t* d_stream_out_flags = nullptr;
    int32_t stream_n_cap = 0;
    int ensure_stream(int32_t n, int32_t W, std::string& err);
    // captured launch prefixes (counter clear, length binning, the plain kernel)
    std::map<std::tuple<const void*, const void*, int32_t, int32_t, int32_t>, hipGraphExec_t> graphs;
    ScanStats stats;

    ~Context();
    int init(std::string& err);
    void release();
    void drop_graphs();
    int ensure_results(int32_t n, std::string& err);
    int ensure_staging(int64_t bytes, int32_t n, std::string& err);
    int grow_pool(int64_t need, std::string& err);
    int ensure_dense(size_t words, std::string& err);
    size_t device_bytes() const;
};

class Engine {
   public:
    Engine();
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    int load(const char* path, std::string& err);  // parse + host tables
    int upload(std::string& err);                 // device tables (allocScratch)
    void release();                               // freeScratch
    bool on_device() const { return dev_ready_; }
    const Graph& graph() const { return g_; }

    // --- default context (the handle's own stream) ---
    int scan_device(const uint8_t* d_text, const int64_t* d_off, int32_t n, KernelKind kind, int32_t mode,
                    bool use_graph, std::string& err) {
        return scan_device(ctx_, d_text, d_off, n, kind, mode, use_graph, err);
    }
    int collect(int32_t n, std::vector<int32_t>& out_off, std::vector<int32_t>& out_ids, std::string& err) {
        return collect(ctx_, n, out_off, out_ids, err);
    }
    // final_states: also keep every string's final states (ctx_.fin_*)
    int scan_host(const uint8_t* const* text, const int32_t* len, int32_t n, KernelKind kind, int32_t mode,
                  std::vector<int32_t>& out_off, std::vector<int32_t>& out_ids, std::string& err,
                  bool final_states = false);
    const ScanStats& stats() const { return ctx_.stats; }
    Context& default_context() { return ctx_; }

    // --- any context ---
    // Scans n strings already in device memory (text + offsets[n+1]) on the
    // context's stream; results stay on the device until collected.
    int scan_device(Context& c, const uint8_t* d_text, const int64_t* d_off, int32_t n, KernelKind kind,
                    int32_t mode, bool use_graph, std::string& err);
    // Copies host strings into the context's pinned staging and enqueues
    // the host-to-device copy.
    int stage(Context& c, const uint8_t* const* text, const int32_t* len, int32_t n, std::string& err);
    // Enqueues the result copies and records c.done (no host wait).
    int begin_collect(Context& c, int32_t n, std::string& err);
    // After c.done: maps accepted states to sorted unique rule ids, string i
    // in out_ids[out_off[i] .. out_off[i+1]).
    int finish_collect(Context& c, int32_t n, std::vector<int32_t>& out_off, std::vector<int32_t>& out_ids,
                       std::string& err);
    int collect(Context& c, int32_t n, std::vector<int32_t>& out_off, std::vector<int32_t>& out_ids,
                std::string& err);
    int new_context(std::unique_ptr<Context>& out, std::string& err);
    void forget_context(Context* c);

    void set_variant(const Variant& v);
    // Grouping threshold of the TKO shift groups: an offset group becomes a
    // shift row when it holds more than this many transitions.  kTkoReference
    // is the reference's state_count/16; kTkoAuto (the default) is W/2, where a
    // row of W words costs about what W/2 scattered edges cost.  Results do
    // not depend on it.  Set before the tables go to the device.
    static constexpr int32_t kTkoReference = -1, kTkoAuto = -2;
    int set_tko_threshold(int32_t thr, std::string& err);
    int32_t tko_threshold_setting() const { return tko_thr_; }
    const Variant& variant() const { return var_; }
    int set_mlp(const float* w, int32_t n, int32_t every, std::string& err);
    // streaming: chunk i continues stream ids[i] (distinct within a call);
    // results are each stream's accepted rules as if it ended after the chunk
    int stream_create(int32_t n, StreamSet& ss, std::string& err);
    int stream_reset(StreamSet& ss, int32_t id, std::string& err);
    int stream_scan(StreamSet& ss, KernelKind kind, int32_t mode, const uint8_t* const* text, const int32_t* len,
                    const int32_t* ids, int32_t n, std::vector<int32_t>& out_off, std::vector<int32_t>& out_ids,
                    std::string& err);
    // allocation-free scan: the launch sequence of scan_device on the caller's
    // stream, its counters and processing order in the file-scope device
    // array of noalloc.hip, results (accepted states per string, as
    // collected by ita_collect before the rule lookup) in caller buffers;
    // status[0] = pool overflow, status[1] = accepted states needed.  LDS
    // storage only; at most kNoAllocMaxStrings strings.
    int scan_noalloc(const uint8_t* d_text, const int64_t* d_off, int32_t n, KernelKind kind, int32_t mode,
                     hipStream_t st, int32_t* d_counts, int64_t* d_starts, uint32_t* d_states, int64_t states_cap,
                     int64_t* d_status, std::string& err);
    // Device-side serving: replays n requests (buf[offsets[i] .. offsets[i+1]),
    // arriving arrival_s[i] seconds after the start, non-decreasing) through
    // one persistent engine launch whose waves take each request as soon as
    // the host publishes its arrival (serve.cpp).
    struct RingResult {
        std::vector<double> latency_s;     // completion seen by the host - arrival
        std::vector<int32_t> offsets, ids;  // rule ids per request
        double span_s = 0;                  // first arrival to last completion
        int64_t rescanned = 0;              // requests rescanned after an output overflow
        int32_t waves = 0;                  // waves of the persistent grid
        unsigned long long method_steps[3] = {0, 0, 0};
    };
    int serve_ring(int32_t mode, const uint8_t* buf, const int64_t* offsets, const double* arrival_s, int32_t n,
                   RingResult& out, std::string& err);
    // HBM bitmaps of the hashed storage for context c (no-op for LDS storage)
    int prepare_dense(Context& c, int32_t mode, std::string& err);
    // the engine kernel a public mode runs (the learned mode's depends on the selector)
    int32_t kernel_mode(int32_t mode) const;
    // the policy table's method for every entry of that list validity that
    // the current list capacity lets a step reach, or -1
    int32_t policy_uniform(bool valid) const;
    int policy_table(uint32_t* out, std::string& err);  // kPolicyTableWords u32
    // replaces the policy table the network was compiled into (a table the
    // host derived from the network, e.g. its expected cost per entry over
    // logged steps); needs a selector set
    int set_policy(const uint32_t* table, int32_t n, std::string& err);
    bool has_mlp() const { return d_mlp_ != nullptr; }
    // step logging for the learned selector (MODE_LOG)
    int set_log(int64_t cap_records, int32_t stride, std::string& err);
    int64_t fetch_log(float* out, int64_t cap_records, std::string& err);
    size_t engine_lds_bytes(const Variant& v, int32_t mode = MODE_LEARNED) const;
    bool hashed(const Variant& v, int32_t mode) const {
        return v.store == 1 && (mode == MODE_AS || mode == MODE_LEARNED || mode == MODE_LEARNED_AS);
    }
    size_t plain_lds_bytes() const;
    int device_cu_count() const { return cu_count_; }
    size_t table_bytes() const { return table_bytes_; }
    int64_t tko_distinct_rows() const { return tko_rows_; }  // after upload

   private:
    int engine_grid(int32_t mode, int32_t n, bool hashed_storage, int* grid, std::string& err, bool stream = false);
    // the kernel instance of a mode: lean kernels for rule sets without
    // boundaries and with two-byte rows (not for streams), narrow LDS
    // kernels for small state vectors
    int32_t engine_kmode(int32_t mode, bool hashed_storage, bool stream = false) const;
    // phases of a launch sequence: the prefix is graph-capturable, the engine
    // kernel runs between ev_k0 and ev_k1
    enum LaunchPhase { kPhaseAll = 0, kPhasePrefix = 1, kPhaseEngine = 2 };
    int launch(Context& c, const uint8_t* d_text, const int64_t* d_off, int32_t n, KernelKind kind, int32_t mode,
               std::string& err, LaunchPhase phase = kPhaseAll);
    ScanArgs make_args(Context& c, const uint8_t* d_text, const int64_t* d_off, int32_t n, int32_t mode);
    void drop_all_graphs();

    Graph g_;
    bool loaded_ = false, dev_ready_ = false;
    DevGraph dev_{};
    std::vector<void*> dev_allocs_;
    size_t table_bytes_ = 0;
    int64_t tko_rows_ = 0;
    int32_t tko_thr_ = kTkoAuto;
    int32_t resolve_tko(const Graph& g) const;
    int cu_count_ = 0;
    size_t max_lds_ = 0;
    Variant var_;
    float* d_mlp_ = nullptr;
    uint32_t* d_policy_ = nullptr;  // learned selector table (kPolicyTableWords)
    std::vector<uint32_t> h_policy_;  // host copy of the table
    int32_t mlp_every_ = 0;  // 0: policy table
    float* d_log_ = nullptr;
    int64_t log_cap_ = 0;
    int32_t log_stride_ = 64;
    Context ctx_;
    std::vector<Context*> contexts_;  // every context handed out by new_context
};

// counters layout in Context::d_ctr (int32 words)
constexpr int kCtrQueue = 0;
constexpr int kCtrOverflow = 1;
constexpr int kCtrPoolTop = 2;    // u64
constexpr int kCtrBins = 4;       // kBins counts
constexpr int kCtrCursors = 44;   // kBins cursors
constexpr int kCtrMethods = 84;   // 4 x u64
constexpr int kCtrLogTop = 92;    // u64
constexpr int kCtrServe = 94;     // serving: published count as last read, stop flag
constexpr int kCtrWords = 96;

// strings per call of the allocation-free scan (its order array holds one
// int32 per string)
constexpr int32_t kNoAllocMaxStrings = 1 << 20;
// looks up (once per device) the current device's copy of the
// allocation-free scan's array
int noalloc_bind_device(uint8_t** out, std::string& err);

}  // namespace admatch
