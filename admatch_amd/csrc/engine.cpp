// Host side of the scan engine (see engine.hpp).
#include "engine.hpp"

#include <algorithm>
#include <chrono>
#include <limits>
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

template <class T>
static void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}
template <class T>
static void hfree(T*& p) {
    if (p) (void)hipHostFree(p);
    p = nullptr;
}

// Elapsed time between two recorded events into *out (left unchanged when HIP
// has none); a failed query leaves no error behind for the next launch check.
static void elapsed_ms(hipEvent_t a, hipEvent_t b, float* out) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) == hipSuccess)
        *out = ms;
    else
        (void)hipGetLastError();
}

// ------------------------------------------------------------------ Context
Context::~Context() {
    release();
    dfree(d_ctr);
    hfree(h_ctr);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (done) (void)hipEventDestroy(done);
    for (hipEvent_t e : {ev_h2d0, ev_h2d1, ev_d2h0, ev_d2h1, ev_k0, ev_k1})
        if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
}

int Context::init(std::string& err) {
    if (stream) return ADM_OK;
    // default priority: every stream of the engine has the same priority
    ADM_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    ADM_HIP(hipEventCreate(&ev0));
    ADM_HIP(hipEventCreate(&ev1));
    ADM_HIP(hipEventCreateWithFlags(&done, hipEventDisableTiming));
    ADM_HIP(hipEventCreate(&ev_h2d0));
    ADM_HIP(hipEventCreate(&ev_h2d1));
    ADM_HIP(hipEventCreate(&ev_d2h0));
    ADM_HIP(hipEventCreate(&ev_d2h1));
    ADM_HIP(hipEventCreate(&ev_k0));
    ADM_HIP(hipEventCreate(&ev_k1));
    ADM_HIP(hipMalloc(&d_ctr, kCtrWords * sizeof(int32_t)));
    ADM_HIP(hipHostMalloc(&h_ctr, kCtrWords * sizeof(int32_t)));
    return ADM_OK;
}

void Context::drop_graphs() {
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
    graphs.clear();
}

void Context::release() {
    drop_graphs();
    if (stream) (void)hipStreamSynchronize(stream);
    dfree(d_res_count);
    dfree(d_res_start);
    dfree(d_pool);
    dfree(d_order);
    dfree(d_text);
    dfree(d_off);
    dfree(d_dense);
    dense_words = 0;
    dfree(d_stream_ids);
    dfree(d_stream_out);
    dfree(d_stream_out_flags);
    stream_n_cap = 0;
    hfree(h_res_count);
    hfree(h_res_start);
    hfree(h_text);
    hfree(h_off);
    res_n_cap = 0;
    h_res_cap = 0;
    pool_cap = 0;
    text_cap = 0;
    off_cap = 0;
}

int Context::ensure_stream(int32_t n, int32_t W, std::string& err) {
    if (n <= stream_n_cap) return ADM_OK;
    if (stream) ADM_HIP(hipStreamSynchronize(stream));
    drop_graphs();
    dfree(d_stream_ids);
    dfree(d_stream_out);
    dfree(d_stream_out_flags);
    const int32_t cap = std::max<int32_t>(n, 256);
    ADM_HIP(hipMalloc(&d_stream_ids, sizeof(int32_t) * cap));
    ADM_HIP(hipMalloc(&d_stream_out, sizeof(uint64_t) * size_t(cap) * size_t(W)));
    ADM_HIP(hipMalloc(&d_stream_out_flags, size_t(cap)));
    stream_n_cap = cap;
    return ADM_OK;
}

StreamSet::~StreamSet() {
    dfree(d_state);
    dfree(d_flags);
}

int Context::ensure_dense(size_t words, std::string& err) {
    if (words <= dense_words) return ADM_OK;
    if (stream) ADM_HIP(hipStreamSynchronize(stream));
    drop_graphs();  // captured launches hold the old buffer
    dfree(d_dense);
    dense_words = 0;
    ADM_HIP(hipMalloc(&d_dense, words * 8));
    ADM_HIP(hipMemset(d_dense, 0, words * 8));  // kernels keep it zero from here on
    ADM_HIP(hipDeviceSynchronize());
    dense_words = words;
    return ADM_OK;
}

size_t Context::device_bytes() const {
    return dense_words * 8 + size_t(res_n_cap) * (4 + 8 + 4) + size_t(pool_cap) * 4 + size_t(text_cap) + size_t(off_cap) * 8 +
           kCtrWords * 4;
}

int Context::ensure_results(int32_t n, std::string& err) {
    if (n > res_n_cap) {
        drop_graphs();
        const int32_t cap = std::max<int32_t>(n, 1024);
        dfree(d_res_count);
        dfree(d_res_start);
        dfree(d_order);
        hfree(h_res_count);
        hfree(h_res_start);
        ADM_HIP(hipMalloc(&d_res_count, sizeof(int32_t) * cap));
        ADM_HIP(hipMalloc(&d_res_start, sizeof(int64_t) * cap));
        ADM_HIP(hipMalloc(&d_order, sizeof(int32_t) * cap));
        ADM_HIP(hipHostMalloc(&h_res_count, sizeof(int32_t) * cap));
        ADM_HIP(hipHostMalloc(&h_res_start, sizeof(int64_t) * cap));
        res_n_cap = h_res_cap = cap;
    }
    const int64_t want = std::max<int64_t>(int64_t(8) * n, int64_t(1) << 16);
    if (pool_cap < want) {
        drop_graphs();
        dfree(d_pool);
        ADM_HIP(hipMalloc(&d_pool, sizeof(uint32_t) * want));
        pool_cap = want;
    }
    return ADM_OK;
}

int Context::grow_pool(int64_t need, std::string& err) {
    const int64_t want = std::max(pool_cap * 2, need + need / 4);
    drop_graphs();
    dfree(d_pool);
    ADM_HIP(hipMalloc(&d_pool, sizeof(uint32_t) * want));
    pool_cap = want;
    return ADM_OK;
}

int Context::ensure_staging(int64_t bytes, int32_t n, std::string& err) {
    if (bytes > text_cap) {
        const int64_t cap = std::max<int64_t>(bytes + (bytes >> 1), 1 << 20);
        drop_graphs();
        hfree(h_text);
        dfree(d_text);
        ADM_HIP(hipHostMalloc(&h_text, size_t(cap)));
        ADM_HIP(hipMalloc(&d_text, size_t(cap)));
        text_cap = cap;
    }
    if (n + 1 > off_cap) {
        const int32_t cap = std::max<int32_t>(n + 1 + (n >> 1), 4096);
        drop_graphs();
        hfree(h_off);
        dfree(d_off);
        ADM_HIP(hipHostMalloc(&h_off, sizeof(int64_t) * cap));
        ADM_HIP(hipMalloc(&d_off, sizeof(int64_t) * cap));
        off_cap = cap;
    }
    return ADM_OK;
}

// ------------------------------------------------------------------ Engine
Engine::Engine() {}

Engine::~Engine() { release(); }

int Engine::load(const char* path, std::string& err) {
    if (!path) {
        err = "null path";
        return ADM_ERR_ARG;
    }
    Graph g;
    if (load_nfa(path, g, err) != 0) return ADM_ERR_LOAD;
    build_tko(g, resolve_tko(g));
    build_as(g);
    build_wb_rows(g);
    g_ = std::move(g);
    loaded_ = true;
    return ADM_OK;
}

int32_t Engine::resolve_tko(const Graph& g) const {
    if (tko_thr_ >= 0) return tko_thr_;
    if (tko_thr_ == kTkoReference) return g.S / 16;
    return g.W / 2;
}

int Engine::set_tko_threshold(int32_t thr, std::string& err) {
    if (thr < kTkoAuto) {
        err = "TKO threshold must be >= 0, -1 (the reference's) or -2 (auto)";
        return ADM_ERR_ARG;
    }
    if (dev_ready_) {
        err = "set the TKO threshold before the tables go to the device";
        return ADM_ERR_STATE;
    }
    tko_thr_ = thr;
    if (loaded_) build_tko(g_, resolve_tko(g_));
    return ADM_OK;
}

size_t Engine::engine_lds_bytes(const Variant& v, int32_t mode) const {
    // per wave: two lists of cap u32 and two W-word bitmaps, or two lists of
    // cap + 64 u32 and two hash tables of 2*cap + 64 u32 (sink entries)
    if (hashed(v, mode)) return size_t(v.wpb) * (size_t(3) * v.cap + 128) * 8;
    return size_t(v.wpb) * (size_t(2) * g_.W + size_t(v.cap)) * 8;
}

size_t Engine::plain_lds_bytes() const { return size_t(3) * g_.W * 8 + 258 * 4; }

void Engine::drop_all_graphs() {
    ctx_.drop_graphs();
    for (Context* c : contexts_) c->drop_graphs();
}

void Engine::set_variant(const Variant& v) {
    var_ = v;
    drop_all_graphs();
}

void Engine::release() {
    drop_all_graphs();
    ctx_.release();
    for (Context* c : contexts_) c->release();
    for (void* p : dev_allocs_) (void)hipFree(p);
    dev_allocs_.clear();
    table_bytes_ = 0;
    dfree(d_mlp_);
    dfree(d_policy_);
    dfree(d_log_);
    log_cap_ = 0;
    dev_ready_ = false;
}

int Engine::new_context(std::unique_ptr<Context>& out, std::string& err) {
    auto c = std::make_unique<Context>();
    if (int rc = c->init(err)) return rc;
    contexts_.push_back(c.get());
    out = std::move(c);
    return ADM_OK;
}

void Engine::forget_context(Context* c) {
    contexts_.erase(std::remove(contexts_.begin(), contexts_.end(), c), contexts_.end());
}

int Engine::upload(std::string& err) {
    if (!loaded_) {
        err = "no rule set loaded";
        return ADM_ERR_STATE;
    }
    if (dev_ready_) return ADM_OK;
    if (int rc = ctx_.init(err)) return rc;
    int dev = 0;
    ADM_HIP(hipGetDevice(&dev));
    ADM_HIP(hipDeviceGetAttribute(&cu_count_, hipDeviceAttributeMultiprocessorCount, dev));
    int lds = 0;
    ADM_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
    max_lds_ = size_t(lds);
    // the allocation-free entry point's array on this device
    if (int rc = noalloc_bind_device(nullptr, err)) return rc;
    auto put = [&](const void* src, size_t bytes, const void** dst) -> int {
        void* p = nullptr;
        const size_t nb = bytes ? bytes : 8;
        ADM_HIP(hipMalloc(&p, nb));
        dev_allocs_.push_back(p);
        table_bytes_ += nb;
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
    {
        // each distinct source-mask row once (groups of different bytes
        // often share one: the persistent states, a class run's chain), so
        // the rows a burst of steps reads stay in cache
        const size_t W = size_t(g_.W);
        const size_t K = g_.tko_shift.size();
        std::vector<uint64_t> rows;
        std::vector<int32_t> row_of(std::max<size_t>(K, 1), 0);
        std::map<uint64_t, std::vector<int32_t>> by_hash;  // content hash -> rows
        for (size_t i = 0; i < K; ++i) {
            const uint64_t* r = g_.tko_lim.data() + i * W;
            uint64_t h = 1469598103934665603ull;
            for (size_t j = 0; j < W; ++j) h = (h ^ r[j]) * 1099511628211ull;
            int32_t found = -1;
            for (int32_t cand : by_hash[h])
                if (std::equal(r, r + W, rows.data() + size_t(cand) * W)) {
                    found = cand;
                    break;
                }
            if (found < 0) {
                found = int32_t(rows.size() / W);
                rows.insert(rows.end(), r, r + W);
                by_hash[h].push_back(found);
            }
            row_of[i] = found;
        }
        tko_rows_ = int64_t(rows.size() / std::max<size_t>(W, 1));
        // each row stored with kTkoPad zero words before it and zeros after
        // it up to the last word a shift pass can read, so the kernels read
        // shifted rows without bounds checks
        const size_t pass = size_t(64 * kTkoPassChunks);
        size_t cover = (W + pass - 1) / pass * pass;  // passes of kTkoPassChunks chunks
        if (W <= size_t(64 * kTkoOnePass)) cover = std::max(cover, size_t(64 * kTkoOnePass));  // one pass
        const size_t stride = size_t(kTkoPad) + cover + size_t(kTkoPad);
        std::vector<uint64_t> padded(std::max<size_t>(size_t(tko_rows_), 1) * stride, 0ull);
        for (int64_t r = 0; r < tko_rows_; ++r)
            std::copy(rows.begin() + r * W, rows.begin() + (r + 1) * W, padded.begin() + r * stride + kTkoPad);
        dev_.tko_stride = int32_t(stride);
        ADM_PUT(padded, tko_lim)
        // (shift, row) per group: a step reads its byte's groups with one load
        std::vector<int2> grp(std::max<size_t>(K, 1), make_int2(0, 0));
        for (size_t i = 0; i < K; ++i) grp[i] = make_int2(g_.tko_shift[i], row_of[i]);
        ADM_PUT(grp, tko_grp)
        // one 512-byte record per byte, read by a wave with one load
        std::vector<int2> rec(256 * 64, make_int2(0, 0));
        for (int c = 0; c < 256; ++c) {
            int2* r = rec.data() + c * 64;
            const int kb = g_.tko_koff[c], ke = g_.tko_koff[c + 1];
            for (int i = 0; i < std::min(ke - kb, kTkoRecGroups); ++i) r[i] = grp[size_t(kb + i)];
            r[kTkoRecGroups] = make_int2(kb, ke);
            r[kTkoRecGroups + 1] = make_int2(g_.tko_loff[c], g_.tko_loff[c + 1]);
        }
        ADM_PUT(rec, tko_rec)
    }
    ADM_PUT(g_.tko_loff, tko_loff)
    ADM_PUT(g_.tko_left, tko_left)
    if (g_.as_ok) {
        ADM_PUT(g_.as_tab, as_tab)
        ADM_PUT(g_.as_multi, as_multi)
    } else {
        dev_.as_tab = nullptr;  // the AS method falls back to iNFA
        dev_.as_multi = nullptr;
    }
    ADM_PUT(g_.alw_dst, alw_dst)
    dev_.alw_total = g_.alw_off[256];
    ADM_PUT(g_.alw_states, alw_states)
    dev_.n_alw_states = int32_t(g_.alw_states.size());
    ADM_PUT(g_.init_explicit, init_explicit)
    if (!g_.pair_off.empty()) {
        ADM_PUT(g_.pair_off, pair_off)
        ADM_PUT(g_.pair_dst, pair_dst)
    } else {
        dev_.pair_off = nullptr;
        dev_.pair_dst = nullptr;
    }
    if (!g_.wbs_off.empty()) {
        ADM_PUT(g_.wbs_off, wbs_off)
        ADM_PUT(g_.wbs_dst, wbs_dst)
        ADM_PUT(g_.wbi_off, wbi_off)
        ADM_PUT(g_.wbi_dst, wbi_dst)
    } else {
        dev_.wbs_off = nullptr;
        dev_.wbs_dst = nullptr;
        dev_.wbi_off = nullptr;
        dev_.wbi_dst = nullptr;
    }
    std::vector<int4> st(256);
    for (int c = 0; c < 256; ++c)
        st[c] = make_int4(g_.sym_off[c + 1] - g_.sym_off[c], g_.tko_koff[c + 1] - g_.tko_koff[c],
                          g_.tko_loff[c + 1] - g_.tko_loff[c], g_.alw_off[c]);
    ADM_PUT(st, sym_stats)
#undef ADM_PUT
    dev_.wb_begin = g_.sym_off[kWordBoundary];
    dev_.wb_count = g_.wb_count();
    dev_.wb_parallel = g_.wb_chain_free() ? 1 : 0;
    {
        // the LDS engine kernel reads its arguments through the kernarg
        // segment pointer: check once that they sit where it looks
        int32_t* d_probe = nullptr;
        ADM_HIP(hipMalloc(&d_probe, sizeof(int32_t)));
        dev_allocs_.push_back(d_probe);
        int32_t h_probe = 0;
        ADM_HIP(launch_kernarg_probe(dev_, d_probe, nullptr));
        ADM_HIP(hipMemcpy(&h_probe, d_probe, sizeof(int32_t), hipMemcpyDeviceToHost));
        if (h_probe != 1) {
            err = "engine kernel arguments are not where the kernel reads them (kernel argument layout)";
            return ADM_ERR_HIP;
        }
    }
    dev_ready_ = true;
    return ADM_OK;
}

ScanArgs Engine::make_args(Context& c, const uint8_t* d_text, const int64_t* d_off, int32_t n, int32_t mode) {
    ScanArgs a{};
    a.text = d_text;
    a.offsets = d_off;
    a.order = c.d_order;
    a.n = n;
    a.mode = mode;
    a.queue = c.d_ctr + kCtrQueue;
    a.res_count = c.d_res_count;
    a.res_start = c.d_res_start;
    a.res_pool = c.d_pool;
    a.pool_cap = c.pool_cap;
    a.pool_top = reinterpret_cast<unsigned long long*>(c.d_ctr + kCtrPoolTop);
    a.overflow = c.d_ctr + kCtrOverflow;
    a.mlp = d_mlp_;
    a.mlp_every = mlp_every_ >= 0 ? mlp_every_ : 0;
    a.policy = (d_mlp_ && a.mlp_every == 0) ? d_policy_ : nullptr;
    a.policy_dense = policy_uniform(false);
    a.policy_sparse = policy_uniform(true);
    if (c.active_stream) {
        a.stream_ids = c.d_stream_ids;
        a.stream_state = c.active_stream->d_state;
        a.stream_flags = c.active_stream->d_flags;
        a.stream_out_state = c.d_stream_out;
        a.stream_out_flags = c.d_stream_out_flags;
    }
    a.log_buf = d_log_;
    a.log_cap = log_cap_;
    a.log_top = reinterpret_cast<unsigned long long*>(c.d_ctr + kCtrLogTop);
    a.log_stride = log_stride_ > 0 ? log_stride_ : 1;
    a.method_steps = reinterpret_cast<unsigned long long*>(c.d_ctr + kCtrMethods);
    a.emit_all = c.emit_all ? 1 : 0;
    return a;
}

int Engine::launch(Context& c, const uint8_t* d_text, const int64_t* d_off, int32_t n, KernelKind kind,
                   int32_t mode, std::string& err, LaunchPhase phase) {
    ScanArgs a = make_args(c, d_text, d_off, n, mode);
    hipStream_t st = c.stream;
    const bool prefix = phase != kPhaseEngine, engine = phase != kPhasePrefix;
    if (prefix) {
        c.engine_timed = false;
        ADM_HIP(launch_fill(c.d_ctr, 0, kCtrWords, st));
    }
    if (n == 0) return ADM_OK;
    if (kind == KIND_PLAIN) {
        const size_t lds = plain_lds_bytes();
        if (lds > max_lds_) {
            err = "rule set too large for the plain iNFA kernel's LDS budget";
            return ADM_ERR_TOO_LARGE;
        }
        a.order = nullptr;
        c.stats.kernel_mode = -1;
        c.stats.kernel_store = 0;
        if (prefix) ADM_HIP(launch_plain(var_.plain_bs, n, lds, st, dev_, a));
        return ADM_OK;
    }
    if (mode == MODE_LOG && (!d_log_ || log_cap_ <= 0)) {
        err = "logging mode needs set_log() first";
        return ADM_ERR_STATE;
    }
    int grid = 0;
    const bool hs = hashed(var_, mode) && !c.active_stream;  // streams use LDS storage
    if (int rc = engine_grid(mode, n, hs, &grid, err, c.active_stream != nullptr)) return rc;
    if (hs) {
        if (c.dense_words < size_t(grid) * var_.wpb * 2 * size_t(g_.W)) {
            err = "dense bitmaps not prepared for this launch";
            return ADM_ERR_STATE;
        }
        a.dense = c.d_dense;
    }
    if (prefix) ADM_HIP(launch_order(d_off, n, c.d_ctr + kCtrBins, c.d_ctr + kCtrCursors, c.d_order, st));
    if (!engine) return ADM_OK;
    const size_t lds = hs ? engine_lds_bytes(var_, mode) : engine_lds_bytes(Variant{var_.wpb, var_.cap, var_.plain_bs, 0});
    const int32_t kmode = engine_kmode(mode, hs, c.active_stream != nullptr);
    c.stats.kernel_mode = kmode;
    c.stats.kernel_store = hs ? 1 : 0;
    ADM_HIP(hipEventRecord(c.ev_k0, st));
    ADM_HIP(launch_engine(var_.wpb, var_.cap, kmode, hs, grid, lds, st, dev_, a));
    ADM_HIP(hipEventRecord(c.ev_k1, st));
    c.engine_timed = true;
    return ADM_OK;
}

int32_t Engine::policy_uniform(bool valid) const {
    // entries a step can reach under the current list capacity: a valid
    // list holds at most cap states, an invalid one more
    if (h_policy_.empty()) return -1;
    int lim = 0;
    while (lim < 15 && (2 << lim) - 1 <= var_.cap) ++lim;  // bucket of cap + 1
    const int b0 = valid ? 0 : lim, b1 = valid ? std::min(lim, 15) : 15;
    int u = -2;
    for (int b = b0; b <= b1; ++b)
        for (int c = 0; c < 256; ++c) {
            const int e = (b * 2 + int(valid)) * 256 + c;
            const int m = int((h_policy_[size_t(e >> 4)] >> ((e & 15) * 2)) & 3u);
            if (u == -2) u = m;
            else if (u != m) return -1;
        }
    return u < 0 ? -1 : u;
}

int32_t Engine::kernel_mode(int32_t mode) const {
    // the learned kernel reads the policy table; without weights the
    // heuristic decides, with every > 0 the network runs in the scan; a
    // table that picks AS for every list-valid step it can reach runs the
    // kernel that decides only the other steps
    if (mode != MODE_LEARNED) return mode;
    if (!d_mlp_) return MODE_HEUR;
    if (mlp_every_ > 0) return MODE_LEARNED_EVERY;
    if (policy_uniform(true) != M_AS) return MODE_LEARNED;
    // AS on every listed set and iNFA on every bitmap-only one is exactly
    // the AS mode's rule: run its kernel (fewer registers than the kernel
    // that reads the table for bitmap-only steps)
    return policy_uniform(false) == M_INFA ? MODE_AS : MODE_LEARNED_AS;
}

int Engine::prepare_dense(Context& c, int32_t mode, std::string& err) {
    mode = kernel_mode(mode);
    if (!hashed(var_, mode)) return ADM_OK;
    // sized for the largest grid of this variant, so that batches of any size
    // reuse it
    int grid = 0;
    if (int rc = engine_grid(mode, std::numeric_limits<int32_t>::max(), true, &grid, err)) return rc;
    return c.ensure_dense(size_t(grid) * var_.wpb * 2 * size_t(g_.W), err);
}

int32_t Engine::engine_kmode(int32_t mode, bool hs, bool stream) const {
    const bool lean_rules = g_.wb_count() == 0 && !g_.pair_off.empty();
    if (hs) return lean_rules ? (mode | kModeLean) : mode;
    int32_t k = g_.W <= 64 * kTkoPassChunks ? (mode | kModeNarrow) : mode;
    if (lean_rules && !stream && mode != MODE_LOG && mode != MODE_LEARNED_EVERY) k |= kModeLean;
    return k;
}

int Engine::engine_grid(int32_t mode, int32_t n, bool hs, int* grid, std::string& err, bool stream) {
    const size_t lds = hs ? engine_lds_bytes(var_, mode) : engine_lds_bytes(Variant{var_.wpb, var_.cap, var_.plain_bs, 0});
    if (lds > max_lds_) {
        err = "rule set too large for the engine's LDS budget (lower the waves per workgroup)";
        return ADM_ERR_TOO_LARGE;
    }
    int blocks = 0;
    ADM_HIP(engine_occupancy(var_.wpb, var_.cap, engine_kmode(mode, hs, stream), hs, lds, &blocks));
    if (blocks <= 0) {
        err = "engine variant does not fit on a CU";
        return ADM_ERR_TOO_LARGE;
    }
    const int64_t need = (int64_t(n) + var_.wpb - 1) / var_.wpb;
    *grid = int(std::max<int64_t>(1, std::min<int64_t>(int64_t(blocks) * cu_count_, need)));
    return ADM_OK;
}

int Engine::scan_device(Context& c, const uint8_t* d_text, const int64_t* d_off, int32_t n, KernelKind kind,
                        int32_t mode, bool use_graph, std::string& err) {
    if (!dev_ready_) {
        err = "rule set not on the device (allocScratch first)";
        return ADM_ERR_STATE;
    }
    if (n < 0 || (n > 0 && (!d_text || !d_off))) {
        err = "bad arguments";
        return ADM_ERR_ARG;
    }
    if (int rc = c.init(err)) return rc;
    const auto ta = std::chrono::steady_clock::now();
    if (int rc = c.ensure_results(n, err)) return rc;
    if (kind == KIND_ENGINE) mode = kernel_mode(mode);
    if (kind == KIND_ENGINE && n > 0 && !c.active_stream)
        if (int rc = prepare_dense(c, mode, err)) return rc;
    c.stats.alloc_ms += std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - ta).count();
    hipStream_t st = c.stream;
    ADM_HIP(hipEventRecord(c.ev0, st));
    if (use_graph && n > 0) {
        auto key = std::make_tuple(static_cast<const void*>(d_text), static_cast<const void*>(d_off), n,
                                   int32_t(kind), mode);
        auto it = c.graphs.find(key);
        if (it == c.graphs.end()) {
            hipGraph_t gr = nullptr;
            ADM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            std::string lerr;
            int rc = launch(c, d_text, d_off, n, kind, mode, lerr, kPhasePrefix);
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
            it = c.graphs.emplace(key, ex).first;
        }
        ADM_HIP(hipGraphLaunch(it->second, st));
        c.engine_timed = false;
        if (int rc = launch(c, d_text, d_off, n, kind, mode, err, kPhaseEngine)) return rc;
    } else {
        if (int rc = launch(c, d_text, d_off, n, kind, mode, err)) return rc;
    }
    ADM_HIP(hipEventRecord(c.ev1, st));
    return ADM_OK;
}

int Engine::stage(Context& c, const uint8_t* const* text, const int32_t* len, int32_t n, std::string& err) {
    if (n < 0 || (n > 0 && (!text || !len))) {
        err = "bad arguments";
        return ADM_ERR_ARG;
    }
    int64_t total = 0;
    for (int32_t i = 0; i < n; ++i) {
        if (len[i] < 0 || (len[i] > 0 && !text[i])) {
            err = "bad string";
            return ADM_ERR_ARG;
        }
        total += len[i];
    }
    if (int rc = c.init(err)) return rc;
    // the previous use of the staging buffers must have finished
    ADM_HIP(hipStreamSynchronize(c.stream));
    const auto ta = std::chrono::steady_clock::now();
    if (int rc = c.ensure_staging(total, n, err)) return rc;
    c.stats.alloc_ms += std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - ta).count();
    int64_t o = 0;
    for (int32_t i = 0; i < n; ++i) {
        c.h_off[i] = o;
        if (len[i] > 0) std::memcpy(c.h_text + o, text[i], size_t(len[i]));
        o += len[i];
    }
    c.h_off[n] = o;
    ADM_HIP(hipEventRecord(c.ev_h2d0, c.stream));
    if (total > 0) ADM_HIP(hipMemcpyAsync(c.d_text, c.h_text, size_t(total), hipMemcpyHostToDevice, c.stream));
    ADM_HIP(hipMemcpyAsync(c.d_off, c.h_off, sizeof(int64_t) * (size_t(n) + 1), hipMemcpyHostToDevice, c.stream));
    ADM_HIP(hipEventRecord(c.ev_h2d1, c.stream));
    return ADM_OK;
}

int Engine::begin_collect(Context& c, int32_t n, std::string& err) {
    if (n > c.res_n_cap) {
        err = "collect: more strings than the last scan";
        return ADM_ERR_ARG;
    }
    ADM_HIP(hipEventRecord(c.ev_d2h0, c.stream));
    ADM_HIP(hipMemcpyAsync(c.h_ctr, c.d_ctr, kCtrWords * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
    if (n > 0) {
        ADM_HIP(hipMemcpyAsync(c.h_res_count, c.d_res_count, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c.stream));
        ADM_HIP(hipMemcpyAsync(c.h_res_start, c.d_res_start, sizeof(int64_t) * n, hipMemcpyDeviceToHost, c.stream));
    }
    ADM_HIP(hipEventRecord(c.done, c.stream));
    return ADM_OK;
}

int Engine::finish_collect(Context& c, int32_t n, std::vector<int32_t>& out_off, std::vector<int32_t>& out_ids,
                           std::string& err) {
    ADM_HIP(hipEventSynchronize(c.done));
    elapsed_ms(c.ev0, c.ev1, &c.stats.kernel_ms);
    c.stats.engine_ms = 0.f;
    if (c.engine_timed) elapsed_ms(c.ev_k0, c.ev_k1, &c.stats.engine_ms);
    unsigned long long top = 0;
    std::memcpy(&top, c.h_ctr + kCtrPoolTop, 8);
    std::memcpy(c.stats.method_steps, c.h_ctr + kCtrMethods, sizeof(c.stats.method_steps));
    c.stats.pool_used = (long long)top;
    if (c.h_ctr[kCtrOverflow]) {
        // pool_top counts every string's accepted states, stored or not
        if (int rc = c.grow_pool(int64_t(top), err)) return rc;
        err = "accepted-state pool overflow (grown to fit; scan again)";
        return ADM_ERR_POOL;
    }
    std::vector<uint32_t> pool(size_t(std::min<unsigned long long>(top, (unsigned long long)c.pool_cap)));
    if (!pool.empty())
        ADM_HIP(hipMemcpyAsync(pool.data(), c.d_pool, sizeof(uint32_t) * pool.size(), hipMemcpyDeviceToHost,
                               c.stream));
    ADM_HIP(hipEventRecord(c.ev_d2h1, c.stream));
    ADM_HIP(hipStreamSynchronize(c.stream));
    elapsed_ms(c.ev_d2h0, c.ev_d2h1, &c.stats.d2h_ms);
    const auto tm = std::chrono::steady_clock::now();
    if (c.emit_all) {
        // every final state, ascending (the reference's get_active_states,
        // src/state_vector.cpp:108-126)
        c.fin_off.assign(size_t(n) + 1, 0);
        c.fin_states.clear();
        for (int32_t i = 0; i < n; ++i) {
            const uint32_t* p = pool.data() + c.h_res_start[i];
            c.fin_states.insert(c.fin_states.end(), p, p + c.h_res_count[i]);
            c.fin_off[size_t(i) + 1] = int32_t(c.fin_states.size());
        }
    }
    // accept map lookup, sort, unique (src/iNFA_kernel.cu:290-307)
    out_off.assign(size_t(n) + 1, 0);
    out_ids.clear();
    std::vector<int32_t> tmp;
    for (int32_t i = 0; i < n; ++i) {
        tmp.clear();
        for (int32_t k = 0; k < c.h_res_count[i]; ++k) {
            const uint32_t s = pool[size_t(c.h_res_start[i] + k)];
            for (int32_t r = g_.acc_off[s]; r < g_.acc_off[s + 1]; ++r) tmp.push_back(g_.acc_rules[r]);
        }
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        out_ids.insert(out_ids.end(), tmp.begin(), tmp.end());
        out_off[size_t(i) + 1] = int32_t(out_ids.size());
    }
    c.stats.map_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - tm).count();
    return ADM_OK;
}

int Engine::collect(Context& c, int32_t n, std::vector<int32_t>& out_off, std::vector<int32_t>& out_ids,
                    std::string& err) {
    if (int rc = begin_collect(c, n, err)) return rc;
    return finish_collect(c, n, out_off, out_ids, err);
}

int Engine::scan_host(const uint8_t* const* text, const int32_t* len, int32_t n, KernelKind kind, int32_t mode,
                      std::vector<int32_t>& out_off, std::vector<int32_t>& out_ids, std::string& err,
                      bool final_states) {
    if (!dev_ready_) {
        err = "rule set not on the device (allocScratch first)";
        return ADM_ERR_STATE;
    }
    ctx_.stats.alloc_ms = 0.f;
    if (int rc = stage(ctx_, text, len, n, err)) return rc;
    // a pool overflow grows the pool to fit, so the second scan succeeds
    ctx_.emit_all = final_states;
    int rc = ADM_ERR_POOL;
    for (int attempt = 0; attempt < 2 && rc == ADM_ERR_POOL; ++attempt) {
        rc = scan_device(ctx_, ctx_.d_text, ctx_.d_off, n, kind, mode, false, err);
        if (rc) break;
        rc = collect(ctx_, n, out_off, out_ids, err);
    }
    ctx_.emit_all = false;
    if (rc == ADM_OK) elapsed_ms(ctx_.ev_h2d0, ctx_.ev_h2d1, &ctx_.stats.h2d_ms);
    return rc;
}

int Engine::set_mlp(const float* w, int32_t n, int32_t every, std::string& err) {
    if (!dev_ready_) {
        err = "rule set not on the device";
        return ADM_ERR_STATE;
    }
    drop_all_graphs();
    if (w == nullptr) {
        dfree(d_mlp_);
        dfree(d_policy_);
        h_policy_.clear();
        return ADM_OK;
    }
    if (n != kMlpParams) {
        err = "selector weights must hold " + std::to_string(kMlpParams) + " floats";
        return ADM_ERR_ARG;
    }
    if (!d_mlp_) ADM_HIP(hipMalloc(&d_mlp_, sizeof(float) * kMlpParams));
    ADM_HIP(hipMemcpy(d_mlp_, w, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    mlp_every_ = every >= 0 ? every : 0;
    // the network's decision per (active-count bucket, list validity, byte)
    if (!d_policy_) ADM_HIP(hipMalloc(&d_policy_, sizeof(uint32_t) * kPolicyTableWords));
    ADM_HIP(launch_policy(dev_, d_mlp_, d_policy_, nullptr));
    ADM_HIP(hipStreamSynchronize(nullptr));
    h_policy_.resize(kPolicyTableWords);
    ADM_HIP(hipMemcpy(h_policy_.data(), d_policy_, sizeof(uint32_t) * h_policy_.size(), hipMemcpyDeviceToHost));
    return ADM_OK;
}

int Engine::stream_create(int32_t n, StreamSet& ss, std::string& err) {
    if (!dev_ready_) {
        err = "rule set not on the device (allocScratch first)";
        return ADM_ERR_STATE;
    }
    if (n <= 0) {
        err = "bad stream count";
        return ADM_ERR_ARG;
    }
    ss.cap = n;
    ss.W = g_.W;
    ADM_HIP(hipMalloc(&ss.d_state, sizeof(uint64_t) * size_t(n) * size_t(g_.W)));
    ADM_HIP(hipMalloc(&ss.d_flags, size_t(n)));
    ADM_HIP(hipMemset(ss.d_flags, 0, size_t(n)));
    ADM_HIP(hipDeviceSynchronize());
    return ADM_OK;
}

int Engine::stream_reset(StreamSet& ss, int32_t id, std::string& err) {
    if (id < 0 || id >= ss.cap) {
        err = "bad stream id";
        return ADM_ERR_ARG;
    }
    if (ctx_.stream) ADM_HIP(hipStreamSynchronize(ctx_.stream));
    ADM_HIP(hipMemset(ss.d_flags + id, 0, 1));
    ADM_HIP(hipDeviceSynchronize());
    return ADM_OK;
}

int Engine::stream_scan(StreamSet& ss, KernelKind kind, int32_t mode, const uint8_t* const* text, const int32_t* len,
                        const int32_t* ids, int32_t n, std::vector<int32_t>& out_off, std::vector<int32_t>& out_ids,
                        std::string& err) {
    if (!dev_ready_) {
        err = "rule set not on the device (allocScratch first)";
        return ADM_ERR_STATE;
    }
    if (ss.W != g_.W || !ss.d_state) {
        err = "stream set of another rule set";
        return ADM_ERR_ARG;
    }
    if (n < 0 || (n > 0 && !ids)) {
        err = "bad arguments";
        return ADM_ERR_ARG;
    }
    std::vector<uint8_t> seen(size_t(ss.cap), 0);
    for (int32_t i = 0; i < n; ++i) {
        if (ids[i] < 0 || ids[i] >= ss.cap) {
            err = "bad stream id";
            return ADM_ERR_ARG;
        }
        if (seen[size_t(ids[i])]++) {
            err = "a stream appears twice in one call";
            return ADM_ERR_ARG;
        }
    }
    if (int rc = stage(ctx_, text, len, n, err)) return rc;
    if (int rc = ctx_.ensure_stream(std::max(n, 1), g_.W, err)) return rc;
    if (n > 0)
        ADM_HIP(hipMemcpyAsync(ctx_.d_stream_ids, ids, sizeof(int32_t) * size_t(n), hipMemcpyHostToDevice,
                               ctx_.stream));
    // the engine kernel with LDS storage; a plain-kernel handle scans with the
    // engine's iNFA mode
    const int32_t m = kind == KIND_PLAIN ? MODE_INFA : mode;
    ctx_.active_stream = &ss;
    int rc = ADM_ERR_POOL;
    for (int attempt = 0; attempt < 2 && rc == ADM_ERR_POOL; ++attempt) {
        rc = scan_device(ctx_, ctx_.d_text, ctx_.d_off, n, KIND_ENGINE, m, false, err);
        if (rc) break;
        rc = collect(ctx_, n, out_off, out_ids, err);
    }
    ctx_.active_stream = nullptr;
    if (rc) return rc;
    // the scan succeeded: its sets become the streams' saved sets
    ADM_HIP(launch_stream_commit(ctx_.d_stream_ids, n, g_.W, ctx_.d_stream_out, ctx_.d_stream_out_flags, ss.d_state,
                                 ss.d_flags, ctx_.stream));
    ADM_HIP(hipStreamSynchronize(ctx_.stream));
    return ADM_OK;
}

int Engine::policy_table(uint32_t* out, std::string& err) {
    if (!d_policy_) {
        err = "no learned selector set";
        return ADM_ERR_STATE;
    }
    ADM_HIP(hipMemcpy(out, d_policy_, sizeof(uint32_t) * kPolicyTableWords, hipMemcpyDeviceToHost));
    return ADM_OK;
}

int Engine::set_policy(const uint32_t* table, int32_t n, std::string& err) {
    if (!d_mlp_ || !d_policy_) {
        err = "no learned selector set";
        return ADM_ERR_STATE;
    }
    if (!table || n != kPolicyTableWords) {
        err = "policy table must hold " + std::to_string(kPolicyTableWords) + " words";
        return ADM_ERR_ARG;
    }
    drop_all_graphs();
    if (ctx_.stream) ADM_HIP(hipStreamSynchronize(ctx_.stream));
    ADM_HIP(hipMemcpy(d_policy_, table, sizeof(uint32_t) * kPolicyTableWords, hipMemcpyHostToDevice));
    h_policy_.assign(table, table + kPolicyTableWords);
    return ADM_OK;
}

int Engine::set_log(int64_t cap_records, int32_t stride, std::string& err) {
    if (!dev_ready_) {
        err = "rule set not on the device";
        return ADM_ERR_STATE;
    }
    drop_all_graphs();
    if (ctx_.stream) ADM_HIP(hipStreamSynchronize(ctx_.stream));
    dfree(d_log_);
    log_cap_ = 0;
    if (cap_records <= 0) return ADM_OK;
    ADM_HIP(hipMalloc(&d_log_, sizeof(float) * kLogRecord * size_t(cap_records)));
    log_cap_ = cap_records;
    log_stride_ = stride > 0 ? stride : 1;
    return ADM_OK;
}

int64_t Engine::fetch_log(float* out, int64_t cap_records, std::string& err) {
    if (!d_log_) return 0;
    ADM_HIP(hipStreamSynchronize(ctx_.stream));
    ADM_HIP(hipMemcpy(ctx_.h_ctr, ctx_.d_ctr, kCtrWords * sizeof(int32_t), hipMemcpyDeviceToHost));
    unsigned long long top = 0;
    std::memcpy(&top, ctx_.h_ctr + kCtrLogTop, 8);
    int64_t n = std::min<int64_t>(int64_t(top), log_cap_);
    n = std::min<int64_t>(n, cap_records);
    if (n > 0 && out) ADM_HIP(hipMemcpy(out, d_log_, sizeof(float) * kLogRecord * size_t(n), hipMemcpyDeviceToHost));
    return n;
}

}  // namespace admatch
