// Host entry points of the kernels in kernels.hip and engine_c*_*.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "device.hpp"

namespace admatch {

constexpr int kLenBins = 40;

// dynamic LDS the engine may request: the CU's 160 KB
constexpr int kEngineDynLds = 160 * 1024;

hipError_t launch_engine(int wpb, int cap, int mode, bool hashed, int grid, size_t lds, hipStream_t st,
                         const DevGraph& g, const ScanArgs& a);
hipError_t engine_occupancy(int wpb, int cap, int mode, bool hashed, size_t lds, int* blocks);
// per capacity and storage (defined in engine_c<CAP>_<storage>.hip)
template <int CAP, bool HASHED>
hipError_t engine_launch_cap(int wpb, int mode, int grid, size_t lds, hipStream_t st, const DevGraph& g,
                             const ScanArgs& a);
template <int CAP, bool HASHED>
hipError_t engine_occ_cap(int wpb, int mode, size_t lds, int* blocks);

// plain single-kernel iNFA path, workgroup size bs in {64, 128, 256, 512, 1024}
hipError_t launch_plain(int bs, int n, size_t lds, hipStream_t st, const DevGraph& g, const ScanArgs& a);
hipError_t launch_fill(int32_t* p, int32_t v, int64_t n, hipStream_t st);
// streaming: copies row i of out_state/out_flags to slot ids[i] of state/flags
hipError_t launch_stream_commit(const int32_t* ids, int32_t n, int32_t W, const uint64_t* out_state,
                                const uint8_t* out_flags, uint64_t* state, uint8_t* flags, hipStream_t st);
// writes 1 to *out if a kernel taking (DevGraph, ScanArgs) finds them through
// kernel_args() where its parameters are (the LDS engine kernel reads them
// that way), 2 if not
hipError_t launch_kernarg_probe(const DevGraph& g, int32_t* out, hipStream_t st);
// learned selector: evaluates the network into the policy table (kPolicyTableWords u32)
hipError_t launch_policy(const DevGraph& g, const float* mlp, uint32_t* table, hipStream_t st);
hipError_t launch_order(const int64_t* offsets, int32_t n, int32_t* ctr_bins, int32_t* ctr_cursors, int32_t* order,
                        hipStream_t st);

}  // namespace admatch
