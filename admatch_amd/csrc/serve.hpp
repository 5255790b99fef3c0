// Continuous-batching request loop (serve.cpp).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "engine.hpp"

namespace admatch {

struct ServeConfig {
    int32_t kind = KIND_ENGINE;
    int32_t mode = MODE_LEARNED;
    int32_t lanes = 4;                          // batches in flight (one stream each)
    int32_t max_batch = 4096;                   // strings per batch
    int64_t max_batch_bytes = int64_t(8) << 20; // input bytes per batch
    int64_t state_budget_bytes = int64_t(256) << 20;  // device bytes of all in-flight batches
    // requests up to split_bytes and longer ones are batched apart and long
    // batches hold at most half of the lanes, so short requests do not wait
    // for long strings (0: one queue)
    int64_t split_bytes = 2048;
};

struct ServeResult {
    std::vector<double> latency_s;  // per request: completion - arrival
    std::vector<int32_t> offsets, ids;
    double wall_s = 0, mean_batch = 0;
    int64_t batches = 0, max_inflight_bytes = 0;
    int max_lanes_busy = 0;
};

// Replays n requests (buf[offsets[i]..offsets[i+1]) arriving at arrival_s[i],
// non-decreasing, seconds from the start) in real time.
int serve_trace(Engine& eng, const ServeConfig& cfg, const uint8_t* buf, const int64_t* offsets,
                const double* arrival_s, int32_t n, ServeResult& out, std::string& err);

}  // namespace admatch
