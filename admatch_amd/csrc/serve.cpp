// Continuous-batching request loop over one resident rule set.
//
// Requests (strings with arrival times) are replayed in real time.  The loop
// keeps `lanes` scan contexts, each with its own HIP stream, staging and
// result buffers; a lane that is free takes the requests that have arrived
// and not started yet (up to max_batch strings / max_batch_bytes, and as
// long as the device bytes of all in-flight batches stay within the state
// budget), copies them in, launches the scan and queues the result copies
// without waiting.  Short and long requests (split_bytes) wait in separate
// queues and go to separate batches: a batch lasts as long as its longest
// string's sequential steps, so short requests are not held by long ones.
// Long batches hold at most half of the lanes (at least one), so a short
// request always finds a lane within one short batch's time.  Lanes run
// concurrently on the device, so requests that arrive while others are being
// scanned start as soon as a lane and budget free up.  Completion is detected by polling each lane's event; a request's
// latency is its completion time minus its arrival time.
#include "serve.hpp"

#include "launch.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <limits>
#include <memory>
#include <thread>

namespace admatch {

namespace {
using clk = std::chrono::steady_clock;

struct Lane {
    std::unique_ptr<Context> ctx;
    bool busy = false;
    int cls = 0;               // length class of the batch in flight
    std::vector<int32_t> ids;  // request ids of the batch
    int64_t device_bytes = 0;
};

// device bytes a batch occupies while in flight: input bytes, offsets,
// order/count/start per string, accepted-state pool
int64_t batch_bytes(int64_t text, int32_t n) { return text + 8 * int64_t(n + 1) + 16 * int64_t(n) + 32 * int64_t(n); }
}  // namespace

int serve_trace(Engine& eng, const ServeConfig& cfg, const uint8_t* buf, const int64_t* offsets,
                const double* arrival_s, int32_t n, ServeResult& out, std::string& err) {
    if (n < 0 || (n > 0 && (!buf || !offsets || !arrival_s)) || cfg.lanes < 1 || cfg.max_batch < 1) {
        err = "bad arguments";
        return ADM_ERR_ARG;
    }
    for (int32_t i = 1; i < n; ++i)
        if (arrival_s[i] < arrival_s[i - 1]) {
            err = "arrival times must be non-decreasing";
            return ADM_ERR_ARG;
        }
    int64_t max_req = 0;
    for (int32_t i = 0; i < n; ++i) max_req = std::max<int64_t>(max_req, offsets[i + 1] - offsets[i]);
    if (batch_bytes(max_req, 1) > cfg.state_budget_bytes) {
        err = "a single request exceeds the state budget";
        return ADM_ERR_BUDGET;
    }
    std::vector<Lane> lanes(size_t(cfg.lanes));
    for (auto& l : lanes) {
        if (int rc = eng.new_context(l.ctx, err)) return rc;
        // pre-size buffers so the loop does not allocate in steady state
        if (int rc = l.ctx->ensure_staging(std::max<int64_t>(cfg.max_batch_bytes, max_req), cfg.max_batch, err))
            return rc;
        if (int rc = l.ctx->ensure_results(cfg.max_batch, err)) return rc;
        if (cfg.kind == KIND_ENGINE)
            if (int rc = eng.prepare_dense(*l.ctx, cfg.mode, err)) return rc;
    }
    struct Guard {
        Engine& e;
        std::vector<Lane>& l;
        ~Guard() {
            for (auto& x : l)
                if (x.ctx) e.forget_context(x.ctx.get());
        }
    } guard{eng, lanes};

    out.latency_s.assign(size_t(n), 0.0);
    std::vector<std::vector<int32_t>> res(static_cast<size_t>(n));
    // pending requests by length class (short, long); one class per batch
    std::deque<int32_t> pending_cls[2];
    auto cls_of = [&](int32_t r) {
        return (cfg.split_bytes > 0 && offsets[r + 1] - offsets[r] > cfg.split_bytes) ? 1 : 0;
    };
    const int long_lanes = cfg.split_bytes > 0 ? std::max(1, cfg.lanes / 2) : cfg.lanes;
    int long_busy = 0;
    int32_t next = 0, completed = 0;
    int64_t inflight = 0;
    out.batches = 0;
    out.max_inflight_bytes = 0;
    out.max_lanes_busy = 0;
    double batch_sum = 0;
    std::vector<const uint8_t*> ptrs;
    std::vector<int32_t> lens;
    std::vector<int32_t> boff, bids;
    const auto t0 = clk::now();
    auto now_s = [&]() { return std::chrono::duration<double>(clk::now() - t0).count(); };

    while (completed < n) {
        bool progressed = false;
        const double t = now_s();
        while (next < n && arrival_s[next] <= t) {
            pending_cls[cls_of(next)].push_back(next);
            ++next;
        }
        // completions
        for (auto& l : lanes) {
            if (!l.busy) continue;
            const hipError_t q = hipEventQuery(l.ctx->done);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) {
                err = std::string("hipEventQuery: ") + hipGetErrorString(q);
                return ADM_ERR_HIP;
            }
            const int32_t bn = int32_t(l.ids.size());
            const int frc = eng.finish_collect(*l.ctx, bn, boff, bids, err);
            if (frc == ADM_ERR_POOL) {
                // the lane's pool was grown to fit: scan the batch again
                if (int rc = eng.scan_device(*l.ctx, l.ctx->d_text, l.ctx->d_off, bn, KernelKind(cfg.kind), cfg.mode,
                                             false, err))
                    return rc;
                if (int rc = eng.begin_collect(*l.ctx, bn, err)) return rc;
                continue;
            }
            if (frc) return frc;
            const double tdone = now_s();
            for (int32_t k = 0; k < bn; ++k) {
                const int32_t r = l.ids[size_t(k)];
                res[size_t(r)].assign(bids.begin() + boff[k], bids.begin() + boff[k + 1]);
                out.latency_s[size_t(r)] = tdone - arrival_s[r];
            }
            completed += bn;
            inflight -= l.device_bytes;
            long_busy -= l.cls;
            l.busy = false;
            l.ids.clear();
            progressed = true;
        }
        // admission
        for (auto& l : lanes) {
            if (l.busy || (pending_cls[0].empty() && pending_cls[1].empty())) continue;
            // the class whose oldest request arrived first, long ones only
            // while fewer than long_lanes lanes hold long batches
            const bool long_ok = !pending_cls[1].empty() && long_busy < long_lanes;
            if (pending_cls[0].empty() && !long_ok) break;
            const int k = !long_ok                 ? 0
                          : pending_cls[0].empty() ? 1
                          : (pending_cls[0].front() < pending_cls[1].front() ? 0 : 1);
            std::deque<int32_t>& pending = pending_cls[k];
            int64_t text = 0;
            l.ids.clear();
            while (!pending.empty() && int32_t(l.ids.size()) < cfg.max_batch) {
                const int32_t r = pending.front();
                const int64_t len = offsets[r + 1] - offsets[r];
                if (!l.ids.empty() && text + len > cfg.max_batch_bytes) break;
                if (inflight + batch_bytes(text + len, int32_t(l.ids.size()) + 1) > cfg.state_budget_bytes) break;
                l.ids.push_back(r);
                text += len;
                pending.pop_front();
            }
            if (l.ids.empty()) break;  // budget exhausted until a lane completes
            const int32_t bn = int32_t(l.ids.size());
            ptrs.resize(size_t(bn));
            lens.resize(size_t(bn));
            for (int32_t j = 0; j < bn; ++j) {
                const int32_t r = l.ids[size_t(j)];
                ptrs[size_t(j)] = buf + offsets[r];
                lens[size_t(j)] = int32_t(offsets[r + 1] - offsets[r]);
            }
            if (int rc = eng.stage(*l.ctx, ptrs.data(), lens.data(), bn, err)) return rc;
            if (int rc = eng.scan_device(*l.ctx, l.ctx->d_text, l.ctx->d_off, bn, KernelKind(cfg.kind), cfg.mode,
                                         false, err))
                return rc;
            if (int rc = eng.begin_collect(*l.ctx, bn, err)) return rc;
            l.device_bytes = batch_bytes(text, bn);
            inflight += l.device_bytes;
            l.cls = k;
            long_busy += k;
            l.busy = true;
            out.batches += 1;
            batch_sum += bn;
            out.max_inflight_bytes = std::max(out.max_inflight_bytes, inflight);
            progressed = true;
        }
        int busy = 0;
        for (auto& l : lanes) busy += l.busy;
        out.max_lanes_busy = std::max(out.max_lanes_busy, busy);
        if (!progressed) {
            if (busy == 0 && pending_cls[0].empty() && pending_cls[1].empty() && next < n) {
                // idle until the next arrival
                const double wait = arrival_s[next] - now_s();
                if (wait > 2e-4) std::this_thread::sleep_for(std::chrono::duration<double>(wait - 1e-4));
            } else {
                std::this_thread::yield();
            }
        }
    }
    out.wall_s = now_s();
    out.mean_batch = out.batches ? batch_sum / double(out.batches) : 0.0;
    out.offsets.assign(size_t(n) + 1, 0);
    out.ids.clear();
    for (int32_t i = 0; i < n; ++i) {
        out.ids.insert(out.ids.end(), res[size_t(i)].begin(), res[size_t(i)].end());
        out.offsets[size_t(i) + 1] = int32_t(out.ids.size());
    }
    return ADM_OK;
}

}  // namespace admatch

// ------------------------------------------------------------ device-side serving
namespace admatch {

namespace {
// pinned host memory the GPU reads and writes directly
struct Pinned {
    void* h = nullptr;
    void* d = nullptr;
    ~Pinned() {
        if (h) (void)hipHostFree(h);
    }
    hipError_t alloc(size_t bytes) {
        hipError_t e = hipHostMalloc(&h, bytes ? bytes : 8, hipHostMallocMapped | hipHostMallocCoherent);
        if (e != hipSuccess) return e;
        std::memset(h, 0, bytes ? bytes : 8);
        return hipHostGetDevicePointer(&d, h, 0);
    }
};

#define ADM_RING_HIP(call)                                                  \
    do {                                                                    \
        hipError_t e_ = (call);                                             \
        if (e_ != hipSuccess) {                                             \
            err = std::string(#call) + ": " + hipGetErrorString(e_);        \
            return ADM_ERR_HIP;                                             \
        }                                                                   \
    } while (0)
}  // namespace

int Engine::serve_ring(int32_t mode, const uint8_t* buf, const int64_t* offsets, const double* arrival_s, int32_t n,
                       RingResult& out, std::string& err) {
    if (!dev_ready_) {
        err = "rule set not on the device (allocScratch first)";
        return ADM_ERR_STATE;
    }
    if (n < 0 || (n > 0 && (!buf || !offsets || !arrival_s))) {
        err = "bad arguments";
        return ADM_ERR_ARG;
    }
    for (int32_t i = 1; i < n; ++i)
        if (arrival_s[i] < arrival_s[i - 1]) {
            err = "arrival times must be non-decreasing";
            return ADM_ERR_ARG;
        }
    int64_t max_len = 0;
    for (int32_t i = 0; i < n; ++i) {
        const int64_t len = offsets[i + 1] - offsets[i];
        if (len < 0 || len > (int64_t(1) << 30)) {
            err = "bad request length";
            return ADM_ERR_ARG;
        }
        max_len = std::max(max_len, len);
    }
    out = RingResult();
    out.latency_s.assign(size_t(n), 0.0);
    out.offsets.assign(size_t(n) + 1, 0);
    if (n == 0) return ADM_OK;
    const int64_t total = offsets[n] - offsets[0];
    mode = kernel_mode(mode);
    const bool hs = hashed(var_, mode);
    std::unique_ptr<Context> cp;
    if (int rc = new_context(cp, err)) return rc;
    Context& c = *cp;
    struct Forget {
        Engine& e;
        Context* c;
        ~Forget() { e.forget_context(c); }
    } forget{*this, cp.get()};
    if (hs)
        if (int rc = prepare_dense(c, mode, err)) return rc;
    int grid = 0;
    if (int rc = engine_grid(mode, std::numeric_limits<int32_t>::max(), hs, &grid, err)) return rc;
    const size_t lds = hs ? engine_lds_bytes(var_, mode) : engine_lds_bytes(Variant{var_.wpb, var_.cap, var_.plain_bs, 0});
    const int64_t slot = (max_len + 32 + 15) / 16 * 16;
    const int64_t waves = int64_t(grid) * var_.wpb;
    uint8_t* d_stage = nullptr;
    ADM_RING_HIP(hipMalloc(&d_stage, size_t(waves * slot)));
    struct FreeStage {
        uint8_t* p;
        ~FreeStage() { (void)hipFree(p); }
    } free_stage{d_stage};

    // requests, as a network stack would have left them: in host memory
    Pinned text, offs, flags, done, cnt, start, pool;
    const int64_t pool_cap = std::max<int64_t>(int64_t(8) * n, int64_t(1) << 16);
    ADM_RING_HIP(text.alloc(size_t(total) + 64));
    ADM_RING_HIP(offs.alloc(sizeof(int64_t) * (size_t(n) + 1)));
    ADM_RING_HIP(flags.alloc(64));
    ADM_RING_HIP(done.alloc(sizeof(int32_t) * size_t(n)));
    ADM_RING_HIP(cnt.alloc(sizeof(int32_t) * size_t(n)));
    ADM_RING_HIP(start.alloc(sizeof(int64_t) * size_t(n)));
    ADM_RING_HIP(pool.alloc(sizeof(uint32_t) * size_t(pool_cap)));
    std::memcpy(text.h, buf + offsets[0], size_t(total));
    int64_t* h_off = static_cast<int64_t*>(offs.h);
    for (int32_t i = 0; i <= n; ++i) h_off[i] = offsets[i] - offsets[0];
    int32_t* h_flags = static_cast<int32_t*>(flags.h);  // [0] published, [1] abort
    int32_t* h_done = static_cast<int32_t*>(done.h);

    ScanArgs a = make_args(c, static_cast<const uint8_t*>(text.d), static_cast<const int64_t*>(offs.d), n, mode);
    a.order = nullptr;
    a.res_count = static_cast<int32_t*>(cnt.d);
    a.res_start = static_cast<int64_t*>(start.d);
    a.res_pool = static_cast<uint32_t*>(pool.d);
    a.pool_cap = pool_cap;
    a.serve = 1;
    a.serve_published = static_cast<const int32_t*>(flags.d);
    a.serve_abort = static_cast<const int32_t*>(flags.d) + 1;
    a.serve_mirror = c.d_ctr + kCtrServe;
    a.serve_done = static_cast<int32_t*>(done.d);
    a.serve_stage = d_stage;
    a.serve_slot = slot;
    // a ticket waits at most until the last arrival plus 10 s (100 MHz ticks)
    a.serve_wait_ticks = int64_t((arrival_s[n - 1] - arrival_s[0] + 10.0) * 1e8);
    if (hs) a.dense = c.d_dense;
    const int32_t kmode = engine_kmode(mode, hs);
    ADM_RING_HIP(launch_fill(c.d_ctr, 0, kCtrWords, c.stream));
    ADM_RING_HIP(launch_engine(var_.wpb, var_.cap, kmode, hs, grid, lds, c.stream, dev_, a));
    out.waves = int32_t(waves);

    // replay: publish each request at its arrival, watch the done flags
    auto stop_kernel = [&]() {
        __atomic_store_n(&h_flags[1], 1, __ATOMIC_RELEASE);
        (void)hipStreamSynchronize(c.stream);
    };
    std::vector<uint8_t> seen(size_t(n), 0);
    const auto t0 = clk::now();
    auto now_s = [&]() { return std::chrono::duration<double>(clk::now() - t0).count() + arrival_s[0]; };
    int32_t published = 0, completed = 0, lo = 0;
    double last_done = arrival_s[0];
    while (completed < n) {
        const double t = now_s();
        int32_t p = published;
        while (p < n && arrival_s[p] <= t) ++p;
        if (p != published) {
            published = p;
            __atomic_store_n(&h_flags[0], published, __ATOMIC_RELEASE);
        }
        bool progressed = false;
        for (int32_t i = lo; i < published; ++i) {
            if (seen[size_t(i)] || !__atomic_load_n(&h_done[i], __ATOMIC_ACQUIRE)) continue;
            seen[size_t(i)] = 1;
            last_done = now_s();
            out.latency_s[size_t(i)] = last_done - arrival_s[i];
            ++completed;
            progressed = true;
        }
        while (lo < n && seen[size_t(lo)]) ++lo;
        if (progressed) continue;
        if (t > arrival_s[n - 1] + 60.0) {
            err = "serving: requests still open a minute after the last arrival";
            stop_kernel();
            return ADM_ERR_HIP;
        }
        const hipError_t q = hipStreamQuery(c.stream);
        if (q == hipSuccess) {
            err = "serving kernel ended before every request completed";
            stop_kernel();
            return ADM_ERR_HIP;
        }
        if (q != hipErrorNotReady) {
            err = std::string("serving kernel: ") + hipGetErrorString(q);
            stop_kernel();
            return ADM_ERR_HIP;
        }
        if (published < n && completed == published) {
            const double wait = arrival_s[published] - now_s();
            if (wait > 2e-4) std::this_thread::sleep_for(std::chrono::duration<double>(wait - 1e-4));
        }
    }
    out.span_s = last_done - arrival_s[0];
    // every ticket past the last request exits: the kernel ends
    ADM_RING_HIP(hipStreamSynchronize(c.stream));
    ADM_RING_HIP(hipMemcpy(c.h_ctr, c.d_ctr, kCtrWords * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (c.h_ctr[kCtrServe + 1] != 0) {
        err = "serving kernel stopped a wait (host abort or wait limit)";
        return ADM_ERR_HIP;
    }
    std::memcpy(out.method_steps, c.h_ctr + kCtrMethods, sizeof(out.method_steps));

    // accepted states -> sorted unique rule ids (src/iNFA_kernel.cu:290-307);
    // a request whose states did not fit the pool is scanned again
    const int32_t* h_cnt = static_cast<const int32_t*>(cnt.h);
    const int64_t* h_start = static_cast<const int64_t*>(start.h);
    const uint32_t* h_pool = static_cast<const uint32_t*>(pool.h);
    std::vector<std::vector<int32_t>> res(static_cast<size_t>(n));
    std::vector<int32_t> again;
    std::vector<int32_t> tmp;
    for (int32_t i = 0; i < n; ++i) {
        if (h_cnt[i] < 0) {
            again.push_back(i);
            continue;
        }
        tmp.clear();
        for (int32_t k = 0; k < h_cnt[i]; ++k) {
            const uint32_t s = h_pool[h_start[i] + k];
            for (int32_t r = g_.acc_off[s]; r < g_.acc_off[s + 1]; ++r) tmp.push_back(g_.acc_rules[r]);
        }
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        res[size_t(i)] = tmp;
    }
    if (!again.empty()) {
        std::vector<const uint8_t*> ptr;
        std::vector<int32_t> len;
        for (int32_t i : again) {
            ptr.push_back(buf + offsets[i]);
            len.push_back(int32_t(offsets[i + 1] - offsets[i]));
        }
        std::vector<int32_t> off2, ids2;
        if (int rc = scan_host(ptr.data(), len.data(), int32_t(again.size()), KIND_ENGINE, mode, off2, ids2, err))
            return rc;
        for (size_t k = 0; k < again.size(); ++k)
            res[size_t(again[k])].assign(ids2.begin() + off2[k], ids2.begin() + off2[k + 1]);
        out.rescanned = int64_t(again.size());
    }
    out.ids.clear();
    for (int32_t i = 0; i < n; ++i) {
        out.ids.insert(out.ids.end(), res[size_t(i)].begin(), res[size_t(i)].end());
        out.offsets[size_t(i) + 1] = int32_t(out.ids.size());
    }
    return ADM_OK;
}

}  // namespace admatch
