// Kernels of the scan path besides the engine's step loop (whose kernels are
// instantiated in engine_c*_*.hip from engine_kernels.hpp):
//
//  * plain_infa_kernel: the single-kernel iNFA path -- one workgroup per
//    string, state bit vectors in LDS, every step scatters the byte's whole
//    transition list (the design of src/iNFA_kernel.cu:12-173, re-laid out
//    for u64 words and LDS ds_or atomics).  It is the throughput baseline.
//  * bin kernels: longest-first processing order by length bins.
//  * policy_kernel: the learned selector's network evaluated once per (byte,
//    active-count bucket, list validity) into a 2-bit table.
#include "engine_kernels.hpp"
#include "launch.hpp"

namespace admatch {

// ------------------------------------------------------------ plain iNFA
// One workgroup of BS threads per string; cur/fut/persistent vectors and the
// 258 symbol offsets in LDS (src/iNFA_kernel.cu:22-52).
template <int BS>
__global__ __launch_bounds__(BS) void plain_infa_kernel(DevGraph g, ScanArgs a) {
    extern __shared__ uint64_t lds[];
    const int sid = blockIdx.x;
    const int W = g.W;
    uint64_t* cur = lds;
    uint64_t* fut = lds + W;
    uint64_t* per = lds + 2 * W;
    int32_t* soff = reinterpret_cast<int32_t*>(lds + 3 * W);
    const int tid = threadIdx.x;
    for (int i = tid; i < W; i += BS) {
        cur[i] = g.init[i];
        per[i] = g.persis[i];
    }
    for (int i = tid; i <= 257; i += BS) soff[i] = g.sym_off[i];
    __syncthreads();
    const int64_t beg = a.offsets[sid];
    const int64_t L = a.offsets[sid + 1] - beg;
    const uint8_t* x = a.text + beg;
    const int wb_b = soff[256], wb_n = soff[257] - soff[256];
    auto wb_pass = [&](uint64_t* v) {
        if (g.wb_parallel) {
            // no destination is a source: the pass equals one parallel pass
            for (int t = tid; t < wb_n; t += BS) {
                const uint2 e = g.list[wb_b + t];
                if (test_bit(v, e.x)) set_bit_atomic(v, e.y);
            }
        } else if (tid == 0) {
            // in-place pass in (src,dst) order on one thread: the
            // thread_count == 1 semantics of src/iNFA_kernel.cu:145-164
            for (int t = 0; t < wb_n; ++t) {
                const uint2 e = g.list[wb_b + t];
                if (test_bit(v, e.x)) v[e.y >> 6] |= 1ull << (e.y & 63);
            }
        }
        __syncthreads();
    };
    if (wb_n > 0 && L > 0 && word_char(x[0])) wb_pass(cur);
    for (int64_t b = 0; b < L; ++b) {
        for (int i = tid; i < W; i += BS) fut[i] = cur[i] & per[i];
        __syncthreads();
        const int c = x[b];
        const int t0 = soff[c], tn = soff[c + 1] - soff[c];
        for (int t = tid; t < tn; t += BS) {
            const uint2 e = g.list[t0 + t];
            if (test_bit(cur, e.x)) set_bit_atomic(fut, e.y);
        }
        uint64_t* tmp = cur;
        cur = fut;
        fut = tmp;
        __syncthreads();
        if (wb_n == 0) continue;
        if ((b < L - 1 && word_char(x[b]) == word_char(x[b + 1])) || (b == L - 1 && !word_char(x[L - 1]))) continue;
        wb_pass(cur);
    }
    // accept extraction by the first wave
    if (tid < 64) emit_accepts(g, a, sid, cur);
}

__global__ __launch_bounds__(64) void policy_kernel(DevGraph g, const float* mlp, uint32_t* table) {
    const int e = blockIdx.x;  // (bucket * 2 + sparse) * 256 + c
    const int c = e & 255;
    const int sp = (e >> 8) & 1;
    const int bucket = e >> 9;
    const StepCtx s = policy_ctx(g, c, bucket, sp != 0);
    const int m = mlp_choice(mlp, feature(g, s, lane_id()), sp != 0);
    if (lane_id() == 0) atomicOr(&table[(e >> 4)], uint32_t(m) << ((c & 15) * 2));
}

// ------------------------------------------------------------ ordering
constexpr int kBins = 40;

__device__ __forceinline__ int len_bin(int64_t len) {
    // descending length bins: bin 0 holds the longest strings
    const int lg = len <= 0 ? 0 : 64 - __clzll(static_cast<unsigned long long>(len));  // 1..40
    return kBins - 1 - min(lg, kBins - 1);
}

__global__ void fill_i32_kernel(int32_t* p, int32_t v, int64_t n) {
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) p[i] = v;
}

// streaming: row i of a scan's output sets becomes stream ids[i]'s saved set
__global__ void stream_commit_kernel(const int32_t* ids, int32_t W, const uint64_t* out_state,
                                     const uint8_t* out_flags, uint64_t* state, uint8_t* flags) {
    const int i = blockIdx.x;
    const int slot = ids[i];
    for (int j = threadIdx.x; j < W; j += blockDim.x)
        state[size_t(slot) * W + j] = out_state[size_t(i) * W + j];
    if (threadIdx.x == 0) flags[slot] = out_flags[i];
}

// counters layout (int32): [0] queue, [1] overflow, [2..3] pool_top (u64),
// [4..4+kBins) bin counts, [4+kBins..4+2kBins) bin cursors
// Strings fall into a handful of length bins, so each workgroup counts in LDS
// and adds one total per bin to the global counters (a global atomic per
// string would queue tens of thousands of them on a few addresses).
__global__ void bin_count_kernel(const int64_t* offsets, int32_t n, int32_t* bins) {
    __shared__ int32_t cnt[kBins];
    for (int t = threadIdx.x; t < kBins; t += blockDim.x) cnt[t] = 0;
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(&cnt[len_bin(offsets[i + 1] - offsets[i])], 1);
    __syncthreads();
    for (int t = threadIdx.x; t < kBins; t += blockDim.x)
        if (cnt[t]) atomicAdd(&bins[t], cnt[t]);
}

// Each workgroup ranks its strings within their bins in LDS, reserves one
// range per bin with a global atomic, and writes its strings there.
__global__ void bin_scatter_kernel(const int64_t* offsets, int32_t n, const int32_t* bins, int32_t* cursors,
                                   int32_t* order) {
    __shared__ int32_t start[kBins];
    __shared__ int32_t cnt[kBins];
    __shared__ int32_t base[kBins];
    if (threadIdx.x == 0) {
        int32_t s = 0;
        for (int b = 0; b < kBins; ++b) {
            start[b] = s;
            s += bins[b];
        }
    }
    for (int t = threadIdx.x; t < kBins; t += blockDim.x) cnt[t] = 0;
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int b = 0, r = 0;
    if (i < n) {
        b = len_bin(offsets[i + 1] - offsets[i]);
        r = atomicAdd(&cnt[b], 1);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kBins; t += blockDim.x)
        if (cnt[t]) base[t] = atomicAdd(&cursors[t], cnt[t]);
    __syncthreads();
    if (i < n) order[start[b] + base[b] + r] = i;
}


// ------------------------------------------------------------ launchers
namespace {
template <int BS>
hipError_t plain_launch_t(int n, size_t lds, hipStream_t st, const DevGraph& g, const ScanArgs& a) {
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&plain_infa_kernel<BS>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipGetLastError();
        attr = true;
    }
    hipLaunchKernelGGL((plain_infa_kernel<BS>), dim3(n), dim3(BS), lds, st, g, a);
    return hipGetLastError();
}
}  // namespace

#define ADM_DISPATCH_CAP(FN, ...)                                  \
    switch (cap * 2 + int(hashed)) {                               \
        case 256: return FN<128, false>(__VA_ARGS__);              \
        case 257: return FN<128, true>(__VA_ARGS__);               \
        case 512: return FN<256, false>(__VA_ARGS__);              \
        case 513: return FN<256, true>(__VA_ARGS__);               \
        case 1024: return FN<512, false>(__VA_ARGS__);             \
        case 1025: return FN<512, true>(__VA_ARGS__);              \
        case 2048: return FN<1024, false>(__VA_ARGS__);            \
        case 2049: return FN<1024, true>(__VA_ARGS__);             \
        default: return hipErrorInvalidValue;                      \
    }

hipError_t launch_engine(int wpb, int cap, int mode, bool hashed, int grid, size_t lds, hipStream_t st,
                         const DevGraph& g, const ScanArgs& a) {
    ADM_DISPATCH_CAP(engine_launch_cap, wpb, mode, grid, lds, st, g, a)
}

hipError_t engine_occupancy(int wpb, int cap, int mode, bool hashed, size_t lds, int* blocks) {
    ADM_DISPATCH_CAP(engine_occ_cap, wpb, mode, lds, blocks)
}

hipError_t launch_plain(int bs, int n, size_t lds, hipStream_t st, const DevGraph& g, const ScanArgs& a) {
    switch (bs) {
        case 64: return plain_launch_t<64>(n, lds, st, g, a);
        case 128: return plain_launch_t<128>(n, lds, st, g, a);
        case 256: return plain_launch_t<256>(n, lds, st, g, a);
        case 512: return plain_launch_t<512>(n, lds, st, g, a);
        case 1024: return plain_launch_t<1024>(n, lds, st, g, a);
        default: return hipErrorInvalidValue;
    }
}
// The engine kernel's (DevGraph, ScanArgs) signature: kernel_args() against
// the parameters, once per rule-set upload.
__global__ void kernarg_probe_kernel(DevGraph g, ScanArgs a) {
    if (threadIdx.x == 0) {
        const KernelArgs k = kernel_args();
        const bool same = k.g->S == g.S && k.g->W == g.W && k.g->as_tab == g.as_tab && k.a->n == a.n &&
                          k.a->overflow == a.overflow && k.a->log_stride == a.log_stride;
        *a.overflow = same ? 1 : 2;
    }
}

hipError_t launch_kernarg_probe(const DevGraph& g, int32_t* out, hipStream_t st) {
    ScanArgs a{};
    a.n = 0x5a5a;
    a.log_stride = 0x3c3c;
    a.overflow = out;
    hipLaunchKernelGGL(kernarg_probe_kernel, dim3(1), dim3(64), 0, st, g, a);
    return hipGetLastError();
}

hipError_t launch_policy(const DevGraph& g, const float* mlp, uint32_t* table, hipStream_t st) {
    hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(256), 0, st, reinterpret_cast<int32_t*>(table), 0,
                       int64_t(kPolicyWords));
    hipLaunchKernelGGL(policy_kernel, dim3(kPolicyBuckets * 2 * 256), dim3(64), 0, st, g, mlp, table);
    return hipGetLastError();
}

hipError_t launch_stream_commit(const int32_t* ids, int32_t n, int32_t W, const uint64_t* out_state,
                                const uint8_t* out_flags, uint64_t* state, uint8_t* flags, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(stream_commit_kernel, dim3(n), dim3(64), 0, st, ids, W, out_state, out_flags, state, flags);
    return hipGetLastError();
}

hipError_t launch_fill(int32_t* p, int32_t v, int64_t n, hipStream_t st) {
    const int grid = int(n < 65536 ? (n + 255) / 256 : 256);
    hipLaunchKernelGGL(fill_i32_kernel, dim3(grid > 0 ? grid : 1), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launch_order(const int64_t* offsets, int32_t n, int32_t* ctr_bins, int32_t* ctr_cursors, int32_t* order,
                        hipStream_t st) {
    const int grid = (n + 255) / 256;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(bin_count_kernel, dim3(grid), dim3(256), 0, st, offsets, n, ctr_bins);
    hipLaunchKernelGGL(bin_scatter_kernel, dim3(grid), dim3(256), 0, st, offsets, n, ctr_bins, ctr_cursors, order);
    return hipGetLastError();
}

}  // namespace admatch
