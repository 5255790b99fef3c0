// Allocation-free scan (ita_scan_device_noalloc): the launch sequence of
// ita_scan_device with every buffer the kernels need besides inputs and
// outputs -- the work queue, overflow flag, pool cursor, length-bin counters
// and the longest-first order -- in one file-scope device array, `scratch`,
// instead of memory the library allocates.  Each GPU holds its own copy of
// the array for the life of the process; its address on a device is looked
// up once (hipGetSymbolAddress, at ita_alloc or on the first call on that
// device) and that address, never the array's name, is what host code hands
// to kernels.  Every call on one GPU shares the array: calls that overlap on
// one GPU (concurrent streams or threads) race and must be serialised by the
// caller.  The counters are cleared by a fill kernel at the start of every
// call (no hipMemsetAsync, so the sequence can be captured into a graph).
#include <mutex>

#include "engine.hpp"
#include "launch.hpp"

namespace admatch {

#define ADM_HIP_ERR(call)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            err = std::string(#call) + ": " + hipGetErrorString(e_);                        \
            return ADM_ERR_HIP;                                                             \
        }                                                                                   \
    } while (0)

namespace {
// counters (kCtrWords int32), then the processing order (one int32 per string)
constexpr size_t kScratchBytes = size_t(kCtrWords) * 4 + size_t(kNoAllocMaxStrings) * 4;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint8_t* g_scratch_ptr[kMaxDevices] = {};

// status[0] = 1 if the accepted states did not fit the caller's buffer,
// status[1] = accepted states of all strings (stored or not)
__global__ void noalloc_status_kernel(const int32_t* ctr, int64_t* status) {
    if (threadIdx.x == 0) {
        status[0] = ctr[kCtrOverflow] != 0 ? 1 : 0;
        status[1] = int64_t(*reinterpret_cast<const unsigned long long*>(ctr + kCtrPoolTop));
    }
}
}  // namespace

int noalloc_bind_device(uint8_t** out, std::string& err) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        err = std::string("hipGetDevice: ") + hipGetErrorString(e);
        return ADM_ERR_HIP;
    }
    if (dev < 0 || dev >= kMaxDevices) {
        err = "device id beyond the allocation-free entry point's table";
        return ADM_ERR_HIP;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return ADM_ERR_HIP;
        }
        g_scratch_ptr[dev] = static_cast<uint8_t*>(p);
    }
    if (out) *out = g_scratch_ptr[dev];
    return ADM_OK;
}

int Engine::scan_noalloc(const uint8_t* d_text, const int64_t* d_off, int32_t n, KernelKind kind, int32_t mode,
                         hipStream_t st, int32_t* d_counts, int64_t* d_starts, uint32_t* d_states, int64_t states_cap,
                         int64_t* d_status, std::string& err) {
    if (!dev_ready_) {
        err = "rule set not on the device (allocScratch first)";
        return ADM_ERR_STATE;
    }
    if (n < 0 || !d_status || states_cap < 0 || (n > 0 && (!d_text || !d_off || !d_counts || !d_starts)) ||
        (states_cap > 0 && !d_states)) {
        err = "bad arguments";
        return ADM_ERR_ARG;
    }
    if (n > kNoAllocMaxStrings) {
        err = "the allocation-free scan takes at most " + std::to_string(kNoAllocMaxStrings) + " strings per call";
        return ADM_ERR_TOO_LARGE;
    }
    if (mode == MODE_LOG) {
        err = "logging mode is not available without allocation";
        return ADM_ERR_ARG;
    }
    uint8_t* base = nullptr;
    if (int rc = noalloc_bind_device(&base, err)) return rc;
    int32_t* ctr = reinterpret_cast<int32_t*>(base);
    int32_t* order = reinterpret_cast<int32_t*>(base + size_t(kCtrWords) * 4);

    ScanArgs a{};
    a.text = d_text;
    a.offsets = d_off;
    a.order = order;
    a.n = n;
    a.queue = ctr + kCtrQueue;
    a.res_count = d_counts;
    a.res_start = d_starts;
    a.res_pool = d_states;
    a.pool_cap = states_cap;
    a.pool_top = reinterpret_cast<unsigned long long*>(ctr + kCtrPoolTop);
    a.overflow = ctr + kCtrOverflow;
    a.mlp = d_mlp_;
    a.mlp_every = mlp_every_ >= 0 ? mlp_every_ : 0;
    a.policy = (d_mlp_ && a.mlp_every == 0) ? d_policy_ : nullptr;
    a.policy_dense = policy_uniform(false);
    a.policy_sparse = policy_uniform(true);
    a.log_stride = 1;
    a.method_steps = reinterpret_cast<unsigned long long*>(ctr + kCtrMethods);

    // every check before the first launch
    const size_t plain_lds = plain_lds_bytes();
    const size_t engine_lds = engine_lds_bytes(Variant{var_.wpb, var_.cap, var_.plain_bs, 0});
    int grid = 0;
    if (kind == KIND_PLAIN) {
        if (plain_lds > max_lds_) {
            err = "rule set too large for the plain iNFA kernel's LDS budget";
            return ADM_ERR_TOO_LARGE;
        }
        a.order = nullptr;
        a.mode = MODE_INFA;
    } else {
        // LDS-bitmap storage: the hashed storage's device bitmaps would be
        // memory beyond the array
        mode = kernel_mode(mode);
        a.mode = mode;
        if (n > 0)
            if (int rc = engine_grid(mode, n, false, &grid, err)) return rc;
    }
    ADM_HIP_ERR(launch_fill(ctr, 0, kCtrWords, st));
    if (n > 0) {
        if (kind == KIND_PLAIN) {
            ADM_HIP_ERR(launch_plain(var_.plain_bs, n, plain_lds, st, dev_, a));
        } else {
            ADM_HIP_ERR(launch_order(d_off, n, ctr + kCtrBins, ctr + kCtrCursors, order, st));
            ADM_HIP_ERR(
                launch_engine(var_.wpb, var_.cap, engine_kmode(mode, false), false, grid, engine_lds, st, dev_, a));
        }
    }
    hipLaunchKernelGGL(noalloc_status_kernel, dim3(1), dim3(64), 0, st, ctr, d_status);
    ADM_HIP_ERR(hipGetLastError());
    return ADM_OK;
}

}  // namespace admatch
