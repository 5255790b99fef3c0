// Host side of the scan engine (see engine.hpp).
#include "engine.hpp"

#include <algorithm>
#include <cstring>

#include "launch.hpp"

namespace admatch {

#define ADM_HIP(call)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            err = std::string(#call) + ": " + hipGetErrorString(e_);                        \
            return ADM_ERR_HIP;                                                             \
        }                                                                                   \
    } while (0)

Engine::Engine() {}

Engine::~Engine() {
    release();
    if (h_ctr_) (void)hipHostFree(h_ctr_);
    if (h_text_) (void)hipHostFree(h_text_);
    if (h_off_) (void)hipHostFree(h_off_);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

int Engine::load(const char* path, std::string& err) {
    if (!path) {
        err = "null path";
        return ADM_ERR_ARG;
    }
    Graph g;
    if (load_nfa(path, g, err) != 0) return ADM_ERR_LOAD;
    build_tko(g);
    build_as(g);
    g_ = std::move(g);
    loaded_ = true;
    return ADM_OK;
}

size_t Engine::engine_lds_bytes(const Variant& v) const {
    return size_t(v.wpb) * (size_t(2) * g_.W + size_t(v.cap) + 32) * 8;
}

size_t Engine::plain_lds_bytes() const { return size_t(3) * g_.W * 8 + 258 * 4; }

void Engine::drop_graphs() {
    for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
    graphs_.clear();
}

void Engine::release() {
    drop_graphs();
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (void* p : dev_allocs_) (void)hipFree(p);
    dev_allocs_.clear();
    auto f = [](auto*& p) {
        if (p) (void)hipFree(p);
        p = nullptr;
    };
    f(d_res_count_);
    f(d_res_start_);
    f(d_pool_);
    f(d_order_);
    f(d_ctr_);
    f(d_mlp_);
    f(d_log_);
    f(d_text_);
    f(d_off_);
    res_n_cap_ = 0;
    pool_cap_ = 0;
    d_text_cap_ = 0;
    off_cap_ = 0;
    log_cap_ = 0;
    dev_ready_ = false;
}

int Engine::upload(std::string& err) {
    if (!loaded_) {
        err = "no rule set loaded";
        return ADM_ERR_STATE;
    }
    if (dev_ready_) return ADM_OK;
    if (!stream_) {
        ADM_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        ADM_HIP(hipEventCreate(&ev0_));
        ADM_HIP(hipEventCreate(&ev1_));
        int dev = 0;
        ADM_HIP(hipGetDevice(&dev));
        ADM_HIP(hipDeviceGetAttribute(&cu_count_, hipDeviceAttributeMultiprocessorCount, dev));
        int lds = 0;
        ADM_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
        max_lds_ = size_t(lds);
    }
    auto put = [&](const void* src, size_t bytes, const void** dst) -> int {
        void* p = nullptr;
        const size_t nb = bytes ? bytes : 8;
        ADM_HIP(hipMalloc(&p, nb));
        dev_allocs_.push_back(p);
        if (bytes) ADM_HIP(hipMemcpy(p, src, bytes, hipMemcpyHostToDevice));
        *dst = p;
        return ADM_OK;
    };
#define ADM_PUT(vec, field)                                                                                   \
    if (int rc_ = put((vec).data(), (vec).size() * sizeof((vec)[0]), reinterpret_cast<const void**>(&dev_.field))) \
        return rc_;
    dev_.S = g_.S;
    dev_.W = g_.W;
    ADM_PUT(g_.init, init)
    ADM_PUT(g_.persis, persis)
    ADM_PUT(g_.accept, accept)
    std::vector<int32_t> soff(g_.sym_off.begin(), g_.sym_off.end());
    ADM_PUT(soff, sym_off)
    ADM_PUT(g_.list, list)
    ADM_PUT(g_.tko_koff, tko_koff)
    ADM_PUT(g_.tko_shift, tko_shift)
    ADM_PUT(g_.tko_lim, tko_lim)
    ADM_PUT(g_.tko_loff, tko_loff)
    ADM_PUT(g_.tko_left, tko_left)
    ADM_PUT(g_.as_off, as_off)
    ADM_PUT(g_.as_dst, as_dst)
    std::vector<int4> st(256);
    for (int c = 0; c < 256; ++c)
        st[c] = make_int4(g_.sym_off[c + 1] - g_.sym_off[c], g_.tko_koff[c + 1] - g_.tko_koff[c],
                          g_.tko_loff[c + 1] - g_.tko_loff[c], 0);
    ADM_PUT(st, sym_stats)
#undef ADM_PUT
    dev_.wb_begin = g_.sym_off[kWordBoundary];
    dev_.wb_count = g_.wb_count();
    ADM_HIP(hipMalloc(&d_ctr_, kCtrWords * sizeof(int32_t)));
    if (!h_ctr_) ADM_HIP(hipHostMalloc(&h_ctr_, kCtrWords * sizeof(int32_t)));
    dev_ready_ = true;
    return ADM_OK;
}

int Engine::ensure_results(int32_t n, std::string& err) {
    if (n > res_n_cap_) {
        drop_graphs();
        const int32_t cap = std::max<int32_t>(n, 1024);
        if (d_res_count_) (void)hipFree(d_res_count_);
        if (d_res_start_) (void)hipFree(d_res_start_);
        if (d_order_) (void)hipFree(d_order_);
        d_res_count_ = nullptr;
        d_res_start_ = nullptr;
        d_order_ = nullptr;
        ADM_HIP(hipMalloc(&d_res_count_, sizeof(int32_t) * cap));
        ADM_HIP(hipMalloc(&d_res_start_, sizeof(int64_t) * cap));
        ADM_HIP(hipMalloc(&d_order_, sizeof(int32_t) * cap));
        res_n_cap_ = cap;
    }
    const int64_t want = std::max<int64_t>(int64_t(8) * n, int64_t(1) << 16);
    if (pool_cap_ < want) {
        drop_graphs();
        if (d_pool_) (void)hipFree(d_pool_);
        d_pool_ = nullptr;
        ADM_HIP(hipMalloc(&d_pool_, sizeof(uint32_t) * want));
        pool_cap_ = want;
    }
    return ADM_OK;
}

ScanArgs Engine::make_args(const uint8_t* d_text, const int64_t* d_off, int32_t n, int32_t mode) {
    ScanArgs a{};
    a.text = d_text;
    a.offsets = d_off;
    a.order = d_order_;
    a.n = n;
    a.mode = mode;
    a.queue = d_ctr_ + kCtrQueue;
    a.res_count = d_res_count_;
    a.res_start = d_res_start_;
    a.res_pool = d_pool_;
    a.pool_cap = pool_cap_;
    a.pool_top = reinterpret_cast<unsigned long long*>(d_ctr_ + kCtrPoolTop);
    a.overflow = d_ctr_ + kCtrOverflow;
    a.mlp = d_mlp_;
    a.mlp_every = mlp_every_ >= 0 ? mlp_every_ : 0;
    a.log_buf = d_log_;
    a.log_cap = log_cap_;
    a.log_top = reinterpret_cast<unsigned long long*>(d_ctr_ + kCtrLogTop);
    a.log_stride = log_stride_ > 0 ? log_stride_ : 1;
    a.method_steps = reinterpret_cast<unsigned long long*>(d_ctr_ + kCtrMethods);
    return a;
}

int Engine::launch(const uint8_t* d_text, const int64_t* d_off, int32_t n, KernelKind kind, int32_t mode,
                   hipStream_t st, std::string& err) {
    ScanArgs a = make_args(d_text, d_off, n, mode);
    ADM_HIP(launch_fill(d_ctr_, 0, kCtrWords, st));
    if (n == 0) return ADM_OK;
    if (kind == KIND_PLAIN) {
        const size_t lds = plain_lds_bytes();
        if (lds > max_lds_) {
            err = "rule set too large for the plain iNFA kernel's LDS budget";
            return ADM_ERR_TOO_LARGE;
        }
        a.order = nullptr;
        ADM_HIP(launch_plain(var_.plain_bs, n, lds, st, dev_, a));
        return ADM_OK;
    }
    const size_t lds = engine_lds_bytes(var_);
    if (lds > max_lds_) {
        err = "rule set too large for the engine's LDS budget (lower the waves per workgroup)";
        return ADM_ERR_TOO_LARGE;
    }
    if (mode == MODE_LOG && (!d_log_ || log_cap_ <= 0)) {
        err = "logging mode needs set_log() first";
        return ADM_ERR_STATE;
    }
    ADM_HIP(launch_order(d_off, n, d_ctr_ + kCtrBins, d_ctr_ + kCtrCursors, d_order_, st));
    int blocks = 0;
    ADM_HIP(engine_occupancy(var_.wpb, var_.cap, lds, &blocks));
    if (blocks <= 0) {
        err = "engine variant does not fit on a CU";
        return ADM_ERR_TOO_LARGE;
    }
    int grid = blocks * cu_count_;
    const int need = (n + var_.wpb - 1) / var_.wpb;
    grid = std::max(1, std::min(grid, need));
    ADM_HIP(launch_engine(var_.wpb, var_.cap, grid, lds, st, dev_, a));
    return ADM_OK;
}

int Engine::scan_device(const uint8_t* d_text, const int64_t* d_off, int32_t n, KernelKind kind, int32_t mode,
                        hipStream_t st, bool use_graph, std::string& err) {
    if (!dev_ready_) {
        err = "rule set not on the device (allocScratch first)";
        return ADM_ERR_STATE;
    }
    if (n < 0 || (n > 0 && (!d_text || !d_off))) {
        err = "bad arguments";
        return ADM_ERR_ARG;
    }
    if (!st) st = stream_;
    if (int rc = ensure_results(n, err)) return rc;
    ADM_HIP(hipEventRecord(ev0_, st));
    if (use_graph && n > 0) {
        auto key = std::make_tuple(static_cast<const void*>(d_text), static_cast<const void*>(d_off), n,
                                   int32_t(kind), mode, static_cast<void*>(st));
        auto it = graphs_.find(key);
        if (it == graphs_.end()) {
            hipGraph_t gr = nullptr;
            ADM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            std::string lerr;
            int rc = launch(d_text, d_off, n, kind, mode, st, lerr);
            hipError_t ec = hipStreamEndCapture(st, &gr);
            if (rc) {
                if (gr) (void)hipGraphDestroy(gr);
                err = lerr;
                return rc;
            }
            if (ec != hipSuccess) {
                err = std::string("hipStreamEndCapture: ") + hipGetErrorString(ec);
                return ADM_ERR_HIP;
            }
            hipGraphExec_t ex = nullptr;
            hipError_t ei = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (ei != hipSuccess) {
                err = std::string("hipGraphInstantiate: ") + hipGetErrorString(ei);
                return ADM_ERR_HIP;
            }
            it = graphs_.emplace(key, ex).first;
        }
        ADM_HIP(hipGraphLaunch(it->second, st));
    } else {
        if (int rc = launch(d_text, d_off, n, kind, mode, st, err)) return rc;
    }
    ADM_HIP(hipEventRecord(ev1_, st));
    return ADM_OK;
}

int Engine::collect(int32_t n, std::vector<int32_t>& out_off, std::vector<int32_t>& out_ids, hipStream_t st,
                    std::string& err) {
    if (!st) st = stream_;
    out_off.assign(size_t(n) + 1, 0);
    out_ids.clear();
    ADM_HIP(hipMemcpyAsync(h_ctr_, d_ctr_, kCtrWords * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    std::vector<int32_t> cnt(static_cast<size_t>(n));
    std::vector<int64_t> start(static_cast<size_t>(n));
    if (n > 0) {
        ADM_HIP(hipMemcpyAsync(cnt.data(), d_res_count_, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
        ADM_HIP(hipMemcpyAsync(start.data(), d_res_start_, sizeof(int64_t) * n, hipMemcpyDeviceToHost, st));
    }
    ADM_HIP(hipStreamSynchronize(st));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ev0_, ev1_) == hipSuccess) stats_.kernel_ms = ms;
    unsigned long long top = 0;
    std::memcpy(&top, h_ctr_ + kCtrPoolTop, 8);
    std::memcpy(stats_.method_steps, h_ctr_ + kCtrMethods, sizeof(stats_.method_steps));
    stats_.pool_used = (long long)top;
    if (h_ctr_[kCtrOverflow]) {
        err = "accepted-state pool overflow";
        return ADM_ERR_CAPACITY;
    }
    std::vector<uint32_t> pool(size_t(std::min<unsigned long long>(top, (unsigned long long)pool_cap_)));
    if (!pool.empty()) {
        ADM_HIP(hipMemcpyAsync(pool.data(), d_pool_, sizeof(uint32_t) * pool.size(), hipMemcpyDeviceToHost, st));
        ADM_HIP(hipStreamSynchronize(st));
    }
    // accept map lookup, sort, unique (src/iNFA_kernel.cu:290-307)
    std::vector<int32_t> tmp;
    for (int32_t i = 0; i < n; ++i) {
        tmp.clear();
        for (int32_t k = 0; k < cnt[i]; ++k) {
            const uint32_t s = pool[size_t(start[i] + k)];
            for (int32_t r = g_.acc_off[s]; r < g_.acc_off[s + 1]; ++r) tmp.push_back(g_.acc_rules[r]);
        }
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        out_ids.insert(out_ids.end(), tmp.begin(), tmp.end());
        out_off[size_t(i) + 1] = int32_t(out_ids.size());
    }
    return ADM_OK;
}

int Engine::scan_host(const uint8_t* const* text, const int32_t* len, int32_t n, KernelKind kind, int32_t mode,
                      std::vector<int32_t>& out_off, std::vector<int32_t>& out_ids, std::string& err) {
    if (n < 0 || (n > 0 && (!text || !len))) {
        err = "bad arguments";
        return ADM_ERR_ARG;
    }
    if (!dev_ready_) {
        err = "rule set not on the device (allocScratch first)";
        return ADM_ERR_STATE;
    }
    int64_t total = 0;
    for (int32_t i = 0; i < n; ++i) {
        if (len[i] < 0 || (len[i] > 0 && !text[i])) {
            err = "bad string";
            return ADM_ERR_ARG;
        }
        total += len[i];
    }
    // pinned staging grows geometrically, device staging likewise
    if (total > h_text_cap_ || total > d_text_cap_) {
        const int64_t cap = std::max<int64_t>(total + (total >> 1), 1 << 20);
        drop_graphs();
        if (h_text_) (void)hipHostFree(h_text_);
        if (d_text_) (void)hipFree(d_text_);
        h_text_ = nullptr;
        d_text_ = nullptr;
        ADM_HIP(hipHostMalloc(&h_text_, size_t(cap)));
        ADM_HIP(hipMalloc(&d_text_, size_t(cap)));
        h_text_cap_ = d_text_cap_ = cap;
    }
    if (n + 1 > off_cap_) {
        const int32_t cap = std::max<int32_t>(n + 1 + (n >> 1), 4096);
        drop_graphs();
        if (h_off_) (void)hipHostFree(h_off_);
        if (d_off_) (void)hipFree(d_off_);
        h_off_ = nullptr;
        d_off_ = nullptr;
        ADM_HIP(hipHostMalloc(&h_off_, sizeof(int64_t) * cap));
        ADM_HIP(hipMalloc(&d_off_, sizeof(int64_t) * cap));
        off_cap_ = cap;
    }
    int64_t o = 0;
    for (int32_t i = 0; i < n; ++i) {
        h_off_[i] = o;
        if (len[i] > 0) std::memcpy(h_text_ + o, text[i], size_t(len[i]));
        o += len[i];
    }
    h_off_[n] = o;
    ADM_HIP(hipMemcpyAsync(d_text_, h_text_, size_t(std::max<int64_t>(total, 1)), hipMemcpyHostToDevice, stream_));
    ADM_HIP(hipMemcpyAsync(d_off_, h_off_, sizeof(int64_t) * (size_t(n) + 1), hipMemcpyHostToDevice, stream_));
    for (int attempt = 0; attempt < 4; ++attempt) {
        if (int rc = scan_device(d_text_, d_off_, n, kind, mode, stream_, false, err)) return rc;
        int rc = collect(n, out_off, out_ids, stream_, err);
        if (rc != ADM_ERR_CAPACITY) return rc;
        // pool overflow: grow and rescan
        const int64_t want = pool_cap_ * 4;
        drop_graphs();
        if (d_pool_) (void)hipFree(d_pool_);
        d_pool_ = nullptr;
        ADM_HIP(hipMalloc(&d_pool_, sizeof(uint32_t) * want));
        pool_cap_ = want;
    }
    err = "accepted-state pool overflow";
    return ADM_ERR_CAPACITY;
}

int Engine::set_mlp(const float* w, int32_t n, int32_t every, std::string& err) {
    if (!dev_ready_) {
        err = "rule set not on the device";
        return ADM_ERR_STATE;
    }
    if (w == nullptr) {
        if (d_mlp_) (void)hipFree(d_mlp_);
        d_mlp_ = nullptr;
        drop_graphs();
        return ADM_OK;
    }
    if (n != kMlpParams) {
        err = "selector weights must hold " + std::to_string(kMlpParams) + " floats";
        return ADM_ERR_ARG;
    }
    if (!d_mlp_) {
        drop_graphs();
        ADM_HIP(hipMalloc(&d_mlp_, sizeof(float) * kMlpParams));
    }
    ADM_HIP(hipMemcpy(d_mlp_, w, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    mlp_every_ = every >= 0 ? every : 0;
    drop_graphs();
    return ADM_OK;
}

int Engine::set_log(int64_t cap_records, int32_t stride, std::string& err) {
    if (!dev_ready_) {
        err = "rule set not on the device";
        return ADM_ERR_STATE;
    }
    drop_graphs();
    if (d_log_) (void)hipFree(d_log_);
    d_log_ = nullptr;
    log_cap_ = 0;
    if (cap_records <= 0) return ADM_OK;
    ADM_HIP(hipMalloc(&d_log_, sizeof(float) * kLogRecord * size_t(cap_records)));
    log_cap_ = cap_records;
    log_stride_ = stride > 0 ? stride : 1;
    return ADM_OK;
}

int64_t Engine::fetch_log(float* out, int64_t cap_records, std::string& err) {
    if (!d_log_) return 0;
    ADM_HIP(hipStreamSynchronize(stream_));
    ADM_HIP(hipMemcpy(h_ctr_, d_ctr_, kCtrWords * sizeof(int32_t), hipMemcpyDeviceToHost));
    unsigned long long top = 0;
    std::memcpy(&top, h_ctr_ + kCtrLogTop, 8);
    int64_t n = std::min<int64_t>(int64_t(top), log_cap_);
    n = std::min<int64_t>(n, cap_records);
    if (n > 0 && out) ADM_HIP(hipMemcpy(out, d_log_, sizeof(float) * kLogRecord * size_t(n), hipMemcpyDeviceToHost));
    return n;
}

}  // namespace admatch
