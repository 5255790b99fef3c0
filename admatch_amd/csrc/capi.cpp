// C ABI (include/admatch.h) over the Engine.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/admatch.h"
#include "engine.hpp"
#include "mlp.hpp"
#include "serve.hpp"

using admatch::Engine;

struct ita_handle {
    uint16_t flags = 0;
    int32_t kind = admatch::KIND_ENGINE;
    int32_t mode = admatch::MODE_ADMATCH;
    Engine eng;
    std::vector<int32_t> off, ids;  // reused result buffers
};

struct ita_streams {
    ita_handle* h = nullptr;
    admatch::StreamSet set;
};

namespace {
thread_local std::string g_err;

int fail(int rc, const std::string& msg) {
    g_err = msg;
    return rc;
}

int check(ita_handle* h) {
    if (!h) return fail(ITA_ERR_ARG, "null handle");
    return ITA_OK;
}

// kernel / selector from the reference's flag bits (Interface.cu:12-40)
int resolve(uint16_t flags, int32_t* kind, int32_t* mode) {
    const int k = ((flags & ITA_INFA_KERNEL) != 0) + ((flags & ITA_AS_KERNEL) != 0) + ((flags & ITA_TKO_KERNEL) != 0);
    if (flags & ITA_AD_MATCHER) {
        if (k != 0) return fail(ITA_ERR_ARG, "AD_MATCHER takes no kernel bit");
        *kind = admatch::KIND_ENGINE;
        if ((flags & ITA_LEARNED_SELECTOR) && (flags & ITA_HEURISTIC_SELECTOR))
            return fail(ITA_ERR_ARG, "choose one selector");
        *mode = (flags & ITA_LEARNED_SELECTOR)     ? admatch::MODE_LEARNED
                : (flags & ITA_HEURISTIC_SELECTOR) ? admatch::MODE_HEUR
                                                   : admatch::MODE_ADMATCH;
        return ITA_OK;
    }
    if (flags & (ITA_LEARNED_SELECTOR | ITA_HEURISTIC_SELECTOR)) return fail(ITA_ERR_ARG, "selector bits need AD_MATCHER");
    if (k != 1) return fail(ITA_ERR_ARG, "exactly one kernel bit required");
    if (flags & ITA_INFA_KERNEL) {
        *kind = admatch::KIND_PLAIN;
        *mode = admatch::MODE_INFA;
    } else {
        *kind = admatch::KIND_ENGINE;
        *mode = (flags & ITA_TKO_KERNEL) ? admatch::MODE_TKO : admatch::MODE_AS;
    }
    return ITA_OK;
}

int status(int rc, const std::string& err) {
    if (rc != 0) g_err = err;
    return rc;
}

// PROFILER_MODE / SHOW_RESULTS output in the reference's format
// (src/kernel_common.cu:36-95).  Device buffers are resident and reused, so
// "Device MemAlloc" is the time spent growing them (usually 0) and "Device
// MemFree" is 0; "Accept Map" is the host's state-to-rule lookup.
void maybe_report(ita_handle* h, int32_t n, double wall_ms) {
    if (h->flags & ITA_SHOW_RESULTS) {
        const auto& c = h->eng.default_context();
        for (int32_t i = 0; i < n; ++i) {
            std::printf("==================\nString [%d] :\nFinal States : ", i + 1);
            if (size_t(i) + 1 < c.fin_off.size())
                for (int32_t k = c.fin_off[size_t(i)]; k < c.fin_off[size_t(i) + 1]; ++k)
                    std::printf("%u ", c.fin_states[size_t(k)]);
            std::printf("\nAccept Rules : ");
            for (int32_t k = h->off[i]; k < h->off[i + 1]; ++k) std::printf("%d ", h->ids[k]);
            std::printf("\n==================\n\n");
        }
    }
    if (h->flags & ITA_PROFILER_MODE) {
        const auto& st = h->eng.stats();
        const double dev = double(st.alloc_ms) + st.h2d_ms + st.kernel_ms + st.d2h_ms;
        auto row = [](const char* what, double ms) { std::printf("%20s%10.4f ms\n", what, ms); };
        row("Total: ", wall_ms);
        row("Per-string: ", n ? wall_ms / n : 0.0);
        row("Device MemAlloc: ", st.alloc_ms);
        row("MemCpy H2D: ", st.h2d_ms);
        row("Kernel Exec: ", st.kernel_ms);
        row("Per-string Kernel: ", n ? double(st.kernel_ms) / n : 0.0);
        row("MemCpy D2H: ", st.d2h_ms);
        row("Device MemFree: ", 0.0);
        row("Accept Map: ", st.map_ms);
        row("Device time: ", dev);
    }
    std::fflush(stdout);
}
}  // namespace

extern "C" {

const char* ita_last_error(void) { return g_err.c_str(); }

int ita_create(uint16_t flags, const char* nfa_path, ita_handle** out) {
    if (!out || !nfa_path) return fail(ITA_ERR_ARG, "null argument");
    *out = nullptr;
    int32_t kind, mode;
    if (int rc = resolve(flags, &kind, &mode)) return rc;
    auto* h = new ita_handle();
    h->flags = flags;
    h->kind = kind;
    h->mode = mode;
    std::string err;
    if (int rc = h->eng.load(nfa_path, err)) {
        delete h;
        return fail(rc, err);
    }
    *out = h;
    return ITA_OK;
}

int ita_alloc(ita_handle* h) {
    if (int rc = check(h)) return rc;
    std::string err;
    return status(h->eng.upload(err), err);
}

int ita_free(ita_handle* h) {
    if (int rc = check(h)) return rc;
    h->eng.release();
    return ITA_OK;
}

int ita_destroy(ita_handle* h) {
    delete h;
    return ITA_OK;
}

int ita_batched_scan(ita_handle* h, const uint8_t* const* text, const int32_t* len, int32_t n,
                     int32_t* out_offsets, int32_t* out_ids, int64_t cap) {
    if (int rc = check(h)) return rc;
    if (n < 0 || (n > 0 && (!out_offsets || (cap > 0 && !out_ids)))) return fail(ITA_ERR_ARG, "bad output buffers");
    std::string err;
    const auto t0 = std::chrono::steady_clock::now();
    int rc = h->eng.scan_host(text, len, n, admatch::KernelKind(h->kind), h->mode, h->off, h->ids, err,
                              (h->flags & ITA_SHOW_RESULTS) != 0);
    if (rc) return fail(rc, err);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (out_offsets) std::memcpy(out_offsets, h->off.data(), sizeof(int32_t) * (size_t(n) + 1));
    const int64_t total = h->off.empty() ? 0 : h->off.back();
    if (total > cap) return fail(ITA_ERR_CAPACITY, "out_ids too small");
    if (total) std::memcpy(out_ids, h->ids.data(), sizeof(int32_t) * size_t(total));
    maybe_report(h, n, ms);
    return ITA_OK;
}

int ita_scan(ita_handle* h, const char* text, int32_t* out_ids, int32_t cap, int32_t* n_out) {
    if (int rc = check(h)) return rc;
    if (!text) return fail(ITA_ERR_ARG, "null text");
    const uint8_t* p = reinterpret_cast<const uint8_t*>(text);
    const int32_t len = int32_t(std::strlen(text));  // Interface.cu:155
    int32_t off[2] = {0, 0};
    int rc = ita_batched_scan(h, &p, &len, 1, off, cap > 0 ? out_ids : nullptr, cap);
    // the count is known on success and when only the output was too small
    if (n_out) *n_out = (rc == ITA_OK || rc == ITA_ERR_CAPACITY) ? off[1] : 0;
    return rc;
}

int ita_scan_device(ita_handle* h, const uint8_t* d_text, const int64_t* d_offsets, int32_t n, int32_t kind,
                    int32_t mode, void* stream, int32_t use_graph) {
    if (int rc = check(h)) return rc;
    if (kind != ITA_KIND_PLAIN && kind != ITA_KIND_ENGINE) return fail(ITA_ERR_ARG, "bad kind");
    if (mode < ITA_MODE_INFA || mode > ITA_MODE_LOG) return fail(ITA_ERR_ARG, "bad mode");
    std::string err;
    admatch::Context& c = h->eng.default_context();
    if (int rc = c.init(err)) return fail(rc, err);
    hipStream_t user = static_cast<hipStream_t>(stream);
    if (user) {  // order after the caller's stream, and make it wait for the scan
        if (hipEventRecord(c.done, user) != hipSuccess || hipStreamWaitEvent(c.stream, c.done, 0) != hipSuccess)
            return fail(ITA_ERR_HIP, "cannot order after the caller's stream");
    }
    int rc = h->eng.scan_device(d_text, d_offsets, n, admatch::KernelKind(kind), mode, use_graph != 0, err);
    if (rc) return fail(rc, err);
    if (user && hipStreamWaitEvent(user, c.ev1, 0) != hipSuccess)
        return fail(ITA_ERR_HIP, "cannot order the caller's stream after the scan");
    return ITA_OK;
}

int ita_scan_device_noalloc(ita_handle* h, const uint8_t* d_text, const int64_t* d_offsets, int32_t n, int32_t kind,
                            int32_t mode, void* stream, int32_t* d_counts, int64_t* d_starts, uint32_t* d_states,
                            int64_t states_cap, int64_t* d_status) {
    if (int rc = check(h)) return rc;
    if (kind != ITA_KIND_PLAIN && kind != ITA_KIND_ENGINE) return fail(ITA_ERR_ARG, "bad kind");
    if (mode < ITA_MODE_INFA || mode > ITA_MODE_LEARNED) return fail(ITA_ERR_ARG, "bad mode");
    std::string err;
    if (int rc = h->eng.scan_noalloc(d_text, d_offsets, n, admatch::KernelKind(kind), mode,
                                     static_cast<hipStream_t>(stream), d_counts, d_starts, d_states, states_cap,
                                     d_status, err))
        return fail(rc, err);
    return ITA_OK;
}

int ita_collect(ita_handle* h, int32_t n, int32_t* out_offsets, int32_t* out_ids, int64_t cap) {
    if (int rc = check(h)) return rc;
    if (n < 0 || !out_offsets) return fail(ITA_ERR_ARG, "bad output buffers");
    std::string err;
    if (int rc = h->eng.collect(n, h->off, h->ids, err)) return fail(rc, err);
    std::memcpy(out_offsets, h->off.data(), sizeof(int32_t) * (size_t(n) + 1));
    const int64_t total = h->off.back();
    if (total > cap) return fail(ITA_ERR_CAPACITY, "out_ids too small");
    if (total) std::memcpy(out_ids, h->ids.data(), sizeof(int32_t) * size_t(total));
    return ITA_OK;
}

int ita_stream_create(ita_handle* h, int32_t n_streams, ita_streams** out) {
    if (int rc = check(h)) return rc;
    if (!out) return fail(ITA_ERR_ARG, "null output");
    auto* s = new ita_streams;
    s->h = h;
    std::string err;
    if (int rc = h->eng.stream_create(n_streams, s->set, err)) {
        delete s;
        return fail(rc, err);
    }
    *out = s;
    return ITA_OK;
}

int ita_stream_reset(ita_streams* s, int32_t stream) {
    if (!s) return fail(ITA_ERR_ARG, "null stream set");
    std::string err;
    if (int rc = s->h->eng.stream_reset(s->set, stream, err)) return fail(rc, err);
    return ITA_OK;
}

int ita_stream_scan(ita_streams* s, const uint8_t* const* chunks, const int32_t* len, const int32_t* stream_ids,
                    int32_t n, int32_t* out_offsets, int32_t* out_ids, int64_t cap) {
    if (!s) return fail(ITA_ERR_ARG, "null stream set");
    ita_handle* h = s->h;
    if (n < 0 || (n > 0 && (!out_offsets || (cap > 0 && !out_ids)))) return fail(ITA_ERR_ARG, "bad output buffers");
    std::string err;
    int rc = h->eng.stream_scan(s->set, admatch::KernelKind(h->kind), h->mode, chunks, len, stream_ids, n, h->off,
                                h->ids, err);
    if (rc) return fail(rc, err);
    if (out_offsets) std::memcpy(out_offsets, h->off.data(), sizeof(int32_t) * (size_t(n) + 1));
    const int64_t total = h->off.empty() ? 0 : h->off.back();
    if (total > cap) return fail(ITA_ERR_CAPACITY, "out_ids too small (the streams have advanced: ita_last_results)");
    if (total) std::memcpy(out_ids, h->ids.data(), sizeof(int32_t) * size_t(total));
    return ITA_OK;
}

int64_t ita_last_results(ita_handle* h, int32_t* out_ids, int64_t cap) {
    if (int rc = check(h)) return rc;
    const int64_t total = h->off.empty() ? 0 : h->off.back();
    if (out_ids && cap > 0) std::memcpy(out_ids, h->ids.data(), sizeof(int32_t) * size_t(std::min(total, cap)));
    return total;
}

int ita_stream_destroy(ita_streams* s) {
    delete s;
    return ITA_OK;
}

int ita_last_stats(ita_handle* h, float* kernel_ms, uint64_t* method_steps3) {
    if (int rc = check(h)) return rc;
    const auto& st = h->eng.stats();
    if (kernel_ms) *kernel_ms = st.kernel_ms;
    if (method_steps3)
        for (int i = 0; i < 3; ++i) method_steps3[i] = st.method_steps[i];
    return ITA_OK;
}

int ita_last_kernel(ita_handle* h, int32_t* out2) {
    if (int rc = check(h)) return rc;
    if (!out2) return fail(ITA_ERR_ARG, "null output");
    const auto& st = h->eng.stats();
    out2[0] = st.kernel_mode;
    out2[1] = st.kernel_store;
    return ITA_OK;
}

int ita_last_times(ita_handle* h, float* out6) {
    if (int rc = check(h)) return rc;
    if (!out6) return fail(ITA_ERR_ARG, "null output");
    const auto& st = h->eng.stats();
    out6[0] = st.kernel_ms;
    out6[1] = st.engine_ms;
    out6[2] = st.alloc_ms;
    out6[3] = st.h2d_ms;
    out6[4] = st.d2h_ms;
    out6[5] = st.map_ms;
    return ITA_OK;
}

int ita_set_variant(ita_handle* h, int32_t wpb, int32_t cap, int32_t plain_bs, int32_t store) {
    if (int rc = check(h)) return rc;
    if ((wpb != 1 && wpb != 2 && wpb != 4) || (cap != 128 && cap != 256 && cap != 512 && cap != 1024) ||
        (plain_bs != 64 && plain_bs != 128 && plain_bs != 256 && plain_bs != 512 && plain_bs != 1024) || (store != 0 && store != 1))
        return fail(ITA_ERR_ARG, "unsupported variant");
    admatch::Variant v;
    v.wpb = wpb;
    v.cap = cap;
    v.plain_bs = plain_bs;
    v.store = store;
    h->eng.set_variant(v);
    return ITA_OK;
}

int ita_set_tko_threshold(ita_handle* h, int32_t threshold) {
    if (int rc = check(h)) return rc;
    std::string err;
    if (int rc = h->eng.set_tko_threshold(threshold, err)) return fail(rc, err);
    return ITA_OK;
}

int ita_get_variant(ita_handle* h, int32_t* wpb, int32_t* cap, int32_t* plain_bs, int32_t* store) {
    if (int rc = check(h)) return rc;
    const auto& v = h->eng.variant();
    if (wpb) *wpb = v.wpb;
    if (cap) *cap = v.cap;
    if (plain_bs) *plain_bs = v.plain_bs;
    if (store) *store = v.store;
    return ITA_OK;
}

int ita_set_selector(ita_handle* h, const float* weights, int32_t n, int32_t every) {
    if (int rc = check(h)) return rc;
    std::string err;
    return status(h->eng.set_mlp(weights, n, every, err), err);
}

int ita_set_policy(ita_handle* h, const uint32_t* table, int32_t n) {
    if (int rc = check(h)) return rc;
    std::string err;
    return status(h->eng.set_policy(table, n, err), err);
}

int ita_set_log(ita_handle* h, int64_t cap_records, int32_t stride) {
    if (int rc = check(h)) return rc;
    std::string err;
    return status(h->eng.set_log(cap_records, stride, err), err);
}

int64_t ita_fetch_log(ita_handle* h, float* out, int64_t cap_records) {
    if (int rc = check(h)) return rc;
    std::string err;
    const int64_t n = h->eng.fetch_log(out, cap_records, err);
    if (n < 0) g_err = err;
    return n;
}

int ita_info(ita_handle* h, int64_t* out) {
    if (int rc = check(h)) return rc;
    if (!out) return fail(ITA_ERR_ARG, "null output");
    const auto& g = h->eng.graph();
    out[0] = g.S;
    out[1] = g.W;
    out[2] = g.sym_off[admatch::kSymbols];
    out[3] = g.wb_count();
    out[4] = int64_t(g.tko_shift.size());
    out[5] = g.tko_threshold;
    out[6] = int64_t(g.as_dst.size());
    out[7] = int64_t(h->eng.engine_lds_bytes(h->eng.variant()));
    out[8] = int64_t(h->eng.plain_lds_bytes());
    out[9] = h->eng.device_cu_count();
    out[10] = int64_t(g.content_hash());
    return ITA_OK;
}

int64_t ita_table(ita_handle* h, int32_t which, void* out, int64_t cap) {
    if (int rc = check(h)) return rc;
    const auto& g = h->eng.graph();
    auto give = [&](const void* src, int64_t count, size_t elem) -> int64_t {
        if (out && cap > 0) std::memcpy(out, src, size_t(count < cap ? count : cap) * elem);
        return count;
    };
    switch (which) {
        case 0: return give(g.sym_off.data(), int64_t(g.sym_off.size()), 4);
        case 1: return give(g.list.data(), int64_t(g.list.size()) * 2, 4);
        case 2: return give(g.tko_koff.data(), int64_t(g.tko_koff.size()), 4);
        case 3: return give(g.tko_shift.data(), int64_t(g.tko_shift.size()), 4);
        case 4: return give(g.tko_lim.data(), int64_t(g.tko_lim.size()), 8);
        case 5: return give(g.tko_loff.data(), int64_t(g.tko_loff.size()), 4);
        case 6: return give(g.tko_left.data(), int64_t(g.tko_left.size()) * 2, 4);
        case 7: return give(g.as_off.data(), int64_t(g.as_off.size()), 4);
        case 8: return give(g.as_dst.data(), int64_t(g.as_dst.size()), 4);
        case 9: {
            std::vector<uint64_t> v;
            v.insert(v.end(), g.init.begin(), g.init.end());
            v.insert(v.end(), g.persis.begin(), g.persis.end());
            v.insert(v.end(), g.accept.begin(), g.accept.end());
            return give(v.data(), int64_t(v.size()), 8);
        }
        case 10: return give(g.as_tab.data(), int64_t(g.as_tab.size()), 4);
        case 11: return give(g.as_multi.data(), int64_t(g.as_multi.size()), 4);
        case 12: return give(g.alw_off.data(), int64_t(g.alw_off.size()), 4);
        case 13: return give(g.alw_dst.data(), int64_t(g.alw_dst.size()), 4);
        case 14: return give(g.pair_off.data(), int64_t(g.pair_off.size()), 4);
        case 15: return give(g.pair_dst.data(), int64_t(g.pair_dst.size()), 4);
        case 16: {
            std::vector<uint32_t> t(admatch::kPolicyTableWords);
            std::string err;
            if (int rc = h->eng.policy_table(t.data(), err)) return fail(rc, err);
            return give(t.data(), int64_t(t.size()), 4);
        }
        case 17: return give(g.wbs_off.data(), int64_t(g.wbs_off.size()), 4);
        case 18: return give(g.wbs_dst.data(), int64_t(g.wbs_dst.size()), 4);
        case 19: return give(g.wbi_off.data(), int64_t(g.wbi_off.size()), 4);
        case 20: return give(g.wbi_dst.data(), int64_t(g.wbi_dst.size()), 4);
        case 21: return give(g.acc_off.data(), int64_t(g.acc_off.size()), 4);
        case 22: return give(g.acc_rules.data(), int64_t(g.acc_rules.size()), 4);
        default: return fail(ITA_ERR_ARG, "bad table id");
    }
}

int ita_selector_train(const float* features, const float* targets, const float* mask, int64_t n, int32_t epochs,
                       int32_t batch, float lr, uint64_t seed, float* weights_inout, float* loss_history) {
    std::string err;
    return status(admatch::mlp_train(features, targets, mask, n, epochs, batch, lr, seed, weights_inout, loss_history,
                                     err),
                  err);
}

int ita_selector_grad(const float* features, const float* targets, const float* mask, int64_t n,
                      const float* weights, float* grad_out, float* loss_out) {
    std::string err;
    return status(admatch::mlp_grad(features, targets, mask, n, weights, grad_out, loss_out, err), err);
}

int ita_selector_forward(const float* features, int64_t n, const float* weights, float* out) {
    std::string err;
    return status(admatch::mlp_forward(features, n, weights, out, err), err);
}

int ita_serve_ring(ita_handle* h, int32_t mode, const uint8_t* buf, const int64_t* offsets, const double* arrival_s,
                   int32_t n, double* latency_s, int32_t* out_offsets, int32_t* out_ids, int64_t cap, double* stats6) {
    if (int rc = check(h)) return rc;
    if (mode < ITA_MODE_INFA || mode > ITA_MODE_LEARNED) return fail(ITA_ERR_ARG, "bad mode");
    if (n < 0 || (n > 0 && (!latency_s || !out_offsets))) return fail(ITA_ERR_ARG, "null output");
    admatch::Engine::RingResult r;
    std::string err;
    if (int rc = h->eng.serve_ring(mode, buf, offsets, arrival_s, n, r, err)) return fail(rc, err);
    if (n > 0) {
        std::memcpy(latency_s, r.latency_s.data(), sizeof(double) * size_t(n));
        std::memcpy(out_offsets, r.offsets.data(), sizeof(int32_t) * (size_t(n) + 1));
    }
    if (stats6) {
        stats6[0] = r.span_s;
        stats6[1] = double(r.rescanned);
        stats6[2] = double(r.waves);
        for (int k = 0; k < 3; ++k) stats6[3 + k] = double(r.method_steps[k]);
    }
    const int64_t total = r.offsets.empty() ? 0 : r.offsets.back();
    h->off = r.offsets;
    h->ids = r.ids;
    if (total > cap) return fail(ITA_ERR_CAPACITY, "out_ids too small (ita_last_results has them)");
    if (total) std::memcpy(out_ids, r.ids.data(), sizeof(int32_t) * size_t(total));
    return ITA_OK;
}

int ita_serve_trace(ita_handle* h, int32_t kind, int32_t mode, int32_t lanes, int32_t max_batch,
                    int64_t max_batch_bytes, int64_t state_budget_bytes, int64_t split_bytes, const uint8_t* buf,
                    const int64_t* offsets,
                    const double* arrival_s, int32_t n, double* latency_s, int32_t* out_offsets, int32_t* out_ids,
                    int64_t cap, double* stats6) {
    if (int rc = check(h)) return rc;
    if (kind != ITA_KIND_PLAIN && kind != ITA_KIND_ENGINE) return fail(ITA_ERR_ARG, "bad kind");
    if (mode < ITA_MODE_INFA || mode > ITA_MODE_LEARNED) return fail(ITA_ERR_ARG, "bad mode");
    if (!latency_s || !out_offsets) return fail(ITA_ERR_ARG, "null output");
    if (!h->eng.on_device()) return fail(ITA_ERR_STATE, "rule set not on the device (allocScratch first)");
    admatch::ServeConfig cfg;
    cfg.kind = kind;
    cfg.mode = mode;
    cfg.lanes = lanes;
    cfg.max_batch = max_batch;
    cfg.max_batch_bytes = max_batch_bytes;
    cfg.state_budget_bytes = state_budget_bytes;
    cfg.split_bytes = split_bytes;
    admatch::ServeResult r;
    std::string err;
    if (int rc = admatch::serve_trace(h->eng, cfg, buf, offsets, arrival_s, n, r, err)) return fail(rc, err);
    std::memcpy(latency_s, r.latency_s.data(), sizeof(double) * size_t(n));
    std::memcpy(out_offsets, r.offsets.data(), sizeof(int32_t) * (size_t(n) + 1));
    if (stats6) {
        stats6[0] = r.wall_s;
        stats6[1] = double(r.batches);
        stats6[2] = r.mean_batch;
        stats6[3] = double(r.max_inflight_bytes);
        stats6[4] = double(r.max_lanes_busy);
        stats6[5] = n ? double(offsets[n] - offsets[0]) : 0.0;
    }
    const int64_t total = r.offsets.back();
    if (total > cap) return fail(ITA_ERR_CAPACITY, "out_ids too small");
    if (total) std::memcpy(out_ids, r.ids.data(), sizeof(int32_t) * size_t(total));
    return ITA_OK;
}

}  // extern "C"
