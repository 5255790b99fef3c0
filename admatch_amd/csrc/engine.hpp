// Host side of the scan engine: device residency of a compiled rule set,
// staging buffers, launches, hipGraph replay and result extraction.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
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
};

// Kernel used for a scan: the plain single-kernel iNFA path or the engine
// with a step selector.
enum KernelKind : int32_t { KIND_PLAIN = 0, KIND_ENGINE = 1 };

struct Variant {
    int wpb = 2;        // waves (strings) per workgroup
    int cap = 512;      // active-list capacity per string
    int plain_bs = 64;  // workgroup size of the plain iNFA kernel
};

struct ScanStats {
    float kernel_ms = 0.f;            // last scan: hipEvent time of the scan kernel
    unsigned long long method_steps[3] = {0, 0, 0};
    long long pool_used = 0;
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

    // Scans n strings already resident in device memory (text + offsets[n+1]).
    // Results stay on the device until collect().
    int scan_device(const uint8_t* d_text, const int64_t* d_off, int32_t n, KernelKind kind, int32_t mode,
                    hipStream_t stream, bool use_graph, std::string& err);
    // Copies the results of the last scan_device to the host and maps
    // accepted states to rule ids: ids of string i are
    // out_ids[out_off[i] .. out_off[i+1]), sorted and unique.
    int collect(int32_t n, std::vector<int32_t>& out_off, std::vector<int32_t>& out_ids, hipStream_t stream,
                std::string& err);
    // Host strings: stage, scan, collect (the BatchedScan path).
    int scan_host(const uint8_t* const* text, const int32_t* len, int32_t n, KernelKind kind, int32_t mode,
                  std::vector<int32_t>& out_off, std::vector<int32_t>& out_ids, std::string& err);

    void set_variant(const Variant& v) {
        var_ = v;
        drop_graphs();
    }
    const Variant& variant() const { return var_; }
    int set_mlp(const float* w, int32_t n, int32_t every, std::string& err);
    bool has_mlp() const { return d_mlp_ != nullptr; }
    // step logging for the learned selector (MODE_LOG)
    int set_log(int64_t cap_records, int32_t stride, std::string& err);
    int64_t fetch_log(float* out, int64_t cap_records, std::string& err);
    const ScanStats& stats() const { return stats_; }
    hipStream_t stream() const { return stream_; }
    size_t engine_lds_bytes(const Variant& v) const;
    size_t plain_lds_bytes() const;
    int device_cu_count() const { return cu_count_; }

   private:
    int ensure_results(int32_t n, std::string& err);
    int launch(const uint8_t* d_text, const int64_t* d_off, int32_t n, KernelKind kind, int32_t mode,
               hipStream_t stream, std::string& err);
    ScanArgs make_args(const uint8_t* d_text, const int64_t* d_off, int32_t n, int32_t mode);

    Graph g_;
    bool loaded_ = false, dev_ready_ = false;
    DevGraph dev_{};
    std::vector<void*> dev_allocs_;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    int cu_count_ = 0;
    size_t max_lds_ = 0;
    Variant var_;
    // result buffers
    int32_t res_n_cap_ = 0;
    int64_t pool_cap_ = 0;
    int32_t* d_res_count_ = nullptr;
    int64_t* d_res_start_ = nullptr;
    uint32_t* d_pool_ = nullptr;
    int32_t* d_order_ = nullptr;
    int32_t* d_ctr_ = nullptr;  // counters (see kCtr*)
    int32_t* h_ctr_ = nullptr;  // pinned mirror
    float* d_mlp_ = nullptr;
    int32_t mlp_every_ = 0;  // 0: per-wave decision cache
    float* d_log_ = nullptr;
    int64_t log_cap_ = 0;
    int32_t log_stride_ = 64;
    // host staging for scan_host
    uint8_t* h_text_ = nullptr;
    int64_t h_text_cap_ = 0;
    uint8_t* d_text_ = nullptr;
    int64_t d_text_cap_ = 0;
    int64_t* h_off_ = nullptr;
    int64_t* d_off_ = nullptr;
    int32_t off_cap_ = 0;
    ScanStats stats_;
    // captured scans keyed by (text, offsets, n, kind, mode, stream)
    std::map<std::tuple<const void*, const void*, int32_t, int32_t, int32_t, void*>, hipGraphExec_t> graphs_;
    void drop_graphs();
};

// counters layout in d_ctr_ (int32 words)
constexpr int kCtrQueue = 0;
constexpr int kCtrOverflow = 1;
constexpr int kCtrPoolTop = 2;    // u64
constexpr int kCtrBins = 4;       // kBins counts
constexpr int kCtrCursors = 44;   // kBins cursors
constexpr int kCtrMethods = 84;   // 4 x u64
constexpr int kCtrLogTop = 92;    // u64
constexpr int kCtrWords = 96;

}  // namespace admatch
