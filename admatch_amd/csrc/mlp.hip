// Training and batch inference of the learned step selector (16 -> 64 -> 64
// -> 3 MLP, FP32) on the GPU.  The same packed weight layout is read by the
// inference fused into the engine's step loop (kernels.hip, mlp_choice):
//   W1 [16 in][64 out], b1[64], W2 [64 in][64 out], b2[64], W3 [3][64], b3[3]
// Targets are per-method log2 step costs; the loss is the masked mean
// squared error over the methods that were timed for a sample.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <random>
#include <string>
#include <vector>

#include "device.hpp"
#include "mlp.hpp"

namespace admatch {

constexpr int kTile = 32;  // samples per workgroup
constexpr int kThreads = 256;

// Forward + backward for a tile of samples; accumulates parameter gradients
// (scaled by 1/n_total) and the loss into global memory with atomics.
__global__ __launch_bounds__(kThreads) void mlp_fwd_bwd_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                                              const float* __restrict__ mask, const int32_t* idx,
                                                              int64_t n, float inv_n, const float* __restrict__ w,
                                                              float* __restrict__ grad, float* __restrict__ loss) {
    __shared__ float sw[kMlpParams];
    __shared__ float sx[kTile][kFeatures];
    __shared__ float h1[kTile][kHidden];
    __shared__ float h2[kTile][kHidden];
    __shared__ float d2[kTile][kHidden];
    __shared__ float dout[kTile][kOutputs];
    const int tid = threadIdx.x;
    const int64_t base = int64_t(blockIdx.x) * kTile;
    for (int i = tid; i < kMlpParams; i += kThreads) sw[i] = w[i];
    for (int i = tid; i < kTile * kFeatures; i += kThreads) {
        const int t = i / kFeatures, f = i % kFeatures;
        const int64_t s = base + t;
        sx[t][f] = s < n ? x[int64_t(idx ? idx[s] : s) * kFeatures + f] : 0.f;
    }
    __syncthreads();
    // layer 1
    for (int i = tid; i < kTile * kHidden; i += kThreads) {
        const int t = i / kHidden, j = i % kHidden;
        float a = sw[kMlpB1 + j];
#pragma unroll
        for (int f = 0; f < kFeatures; ++f) a += sw[kMlpW1 + f * kHidden + j] * sx[t][f];
        h1[t][j] = fmaxf(a, 0.f);
    }
    __syncthreads();
    // layer 2
    for (int i = tid; i < kTile * kHidden; i += kThreads) {
        const int t = i / kHidden, j = i % kHidden;
        float a = sw[kMlpB2 + j];
        for (int k = 0; k < kHidden; ++k) a += sw[kMlpW2 + k * kHidden + j] * h1[t][k];
        h2[t][j] = fmaxf(a, 0.f);
    }
    __syncthreads();
    // output + loss gradient
    if (tid < kTile * kOutputs) {
        const int t = tid / kOutputs, m = tid % kOutputs;
        const int64_t s = base + t;
        float o = sw[kMlpB3 + m];
        for (int j = 0; j < kHidden; ++j) o += sw[kMlpW3 + m * kHidden + j] * h2[t][j];
        float g = 0.f;
        if (s < n) {
            const int64_t r = idx ? idx[s] : s;
            const float mk = mask[r * kOutputs + m];
            const float e = o - y[r * kOutputs + m];
            g = 2.f * mk * e * inv_n;
            atomicAdd(loss, mk * e * e * inv_n);
        }
        dout[t][m] = g;
    }
    __syncthreads();
    // dH2 = dout W3 * relu'
    for (int i = tid; i < kTile * kHidden; i += kThreads) {
        const int t = i / kHidden, j = i % kHidden;
        float a = 0.f;
#pragma unroll
        for (int m = 0; m < kOutputs; ++m) a += dout[t][m] * sw[kMlpW3 + m * kHidden + j];
        d2[t][j] = h2[t][j] > 0.f ? a : 0.f;
    }
    __syncthreads();
    // grads of W3 / b3
    if (tid < kOutputs * kHidden) {
        const int m = tid / kHidden, j = tid % kHidden;
        float a = 0.f;
        for (int t = 0; t < kTile; ++t) a += h2[t][j] * dout[t][m];
        atomicAdd(&grad[kMlpW3 + m * kHidden + j], a);
    }
    if (tid < kOutputs) {
        float a = 0.f;
        for (int t = 0; t < kTile; ++t) a += dout[t][tid];
        atomicAdd(&grad[kMlpB3 + tid], a);
    }
    // grads of W2 / b2
    for (int i = tid; i < kHidden * kHidden; i += kThreads) {
        const int k = i / kHidden, j = i % kHidden;
        float a = 0.f;
        for (int t = 0; t < kTile; ++t) a += h1[t][k] * d2[t][j];
        atomicAdd(&grad[kMlpW2 + k * kHidden + j], a);
    }
    if (tid < kHidden) {
        float a = 0.f;
        for (int t = 0; t < kTile; ++t) a += d2[t][tid];
        atomicAdd(&grad[kMlpB2 + tid], a);
    }
    __syncthreads();
    // dH1 = dH2 W2^T * relu'  (reuse h2 as storage)
    for (int i = tid; i < kTile * kHidden; i += kThreads) {
        const int t = i / kHidden, k = i % kHidden;
        float a = 0.f;
        for (int j = 0; j < kHidden; ++j) a += d2[t][j] * sw[kMlpW2 + k * kHidden + j];
        h2[t][k] = h1[t][k] > 0.f ? a : 0.f;
    }
    __syncthreads();
    // grads of W1 / b1
    for (int i = tid; i < kFeatures * kHidden; i += kThreads) {
        const int f = i / kHidden, j = i % kHidden;
        float a = 0.f;
        for (int t = 0; t < kTile; ++t) a += sx[t][f] * h2[t][j];
        atomicAdd(&grad[kMlpW1 + f * kHidden + j], a);
    }
    if (tid < kHidden) {
        float a = 0.f;
        for (int t = 0; t < kTile; ++t) a += h2[t][tid];
        atomicAdd(&grad[kMlpB1 + tid], a);
    }
}

__global__ void mlp_forward_kernel(const float* __restrict__ x, int64_t n, const float* __restrict__ w,
                                   float* __restrict__ out) {
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float xf[kFeatures], h[kHidden];
    for (int f = 0; f < kFeatures; ++f) xf[f] = x[s * kFeatures + f];
    for (int j = 0; j < kHidden; ++j) {
        float a = w[kMlpB1 + j];
        for (int f = 0; f < kFeatures; ++f) a += w[kMlpW1 + f * kHidden + j] * xf[f];
        h[j] = fmaxf(a, 0.f);
    }
    float o[kOutputs];
    for (int m = 0; m < kOutputs; ++m) o[m] = w[kMlpB3 + m];
    for (int j = 0; j < kHidden; ++j) {
        float a = w[kMlpB2 + j];
        for (int k = 0; k < kHidden; ++k) a += w[kMlpW2 + k * kHidden + j] * h[k];
        a = fmaxf(a, 0.f);
        for (int m = 0; m < kOutputs; ++m) o[m] += w[kMlpW3 + m * kHidden + j] * a;
    }
    for (int m = 0; m < kOutputs; ++m) out[s * kOutputs + m] = o[m];
}

// Adam (Kingma & Ba, 2015) with bias correction.
__global__ void adam_kernel(float* w, float* m1, float* m2, float* grad, float lr, float b1, float b2, float eps,
                            float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kMlpParams) return;
    const float g = grad[i];
    const float a = b1 * m1[i] + (1.f - b1) * g;
    const float b = b2 * m2[i] + (1.f - b2) * g * g;
    m1[i] = a;
    m2[i] = b;
    w[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
    grad[i] = 0.f;
}

__global__ void zero_kernel(float* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// ------------------------------------------------------------ host side
namespace {
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 4); }
    template <class T>
    T* as() {
        return static_cast<T*>(p);
    }
};
}  // namespace

#define MLP_HIP(call)                                                 \
    do {                                                              \
        hipError_t e_ = (call);                                       \
        if (e_ != hipSuccess) {                                       \
            err = std::string(#call) + ": " + hipGetErrorString(e_);  \
            return -3;                                                \
        }                                                             \
    } while (0)

int mlp_grad(const float* x, const float* y, const float* mask, int64_t n, const float* w, float* grad_out,
             float* loss_out, std::string& err) {
    if (n <= 0 || !x || !y || !mask || !w || !grad_out) {
        err = "bad arguments";
        return -1;
    }
    DevBuf dx, dy, dm, dw, dg, dl;
    MLP_HIP(dx.alloc(sizeof(float) * n * kFeatures));
    MLP_HIP(dy.alloc(sizeof(float) * n * kOutputs));
    MLP_HIP(dm.alloc(sizeof(float) * n * kOutputs));
    MLP_HIP(dw.alloc(sizeof(float) * kMlpParams));
    MLP_HIP(dg.alloc(sizeof(float) * kMlpParams));
    MLP_HIP(dl.alloc(sizeof(float)));
    MLP_HIP(hipMemcpy(dx.p, x, sizeof(float) * n * kFeatures, hipMemcpyHostToDevice));
    MLP_HIP(hipMemcpy(dy.p, y, sizeof(float) * n * kOutputs, hipMemcpyHostToDevice));
    MLP_HIP(hipMemcpy(dm.p, mask, sizeof(float) * n * kOutputs, hipMemcpyHostToDevice));
    MLP_HIP(hipMemcpy(dw.p, w, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(zero_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, nullptr, dg.as<float>(), kMlpParams);
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(32), 0, nullptr, dl.as<float>(), 1);
    const int blocks = int((n + kTile - 1) / kTile);
    hipLaunchKernelGGL(mlp_fwd_bwd_kernel, dim3(blocks), dim3(kThreads), 0, nullptr, dx.as<float>(), dy.as<float>(),
                       dm.as<float>(), static_cast<const int32_t*>(nullptr), n, 1.f / float(n), dw.as<float>(),
                       dg.as<float>(), dl.as<float>());
    MLP_HIP(hipGetLastError());
    MLP_HIP(hipMemcpy(grad_out, dg.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
    if (loss_out) MLP_HIP(hipMemcpy(loss_out, dl.p, sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int mlp_forward(const float* x, int64_t n, const float* w, float* out, std::string& err) {
    if (n <= 0 || !x || !w || !out) {
        err = "bad arguments";
        return -1;
    }
    DevBuf dx, dw, dout;
    MLP_HIP(dx.alloc(sizeof(float) * n * kFeatures));
    MLP_HIP(dw.alloc(sizeof(float) * kMlpParams));
    MLP_HIP(dout.alloc(sizeof(float) * n * kOutputs));
    MLP_HIP(hipMemcpy(dx.p, x, sizeof(float) * n * kFeatures, hipMemcpyHostToDevice));
    MLP_HIP(hipMemcpy(dw.p, w, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(int((n + 255) / 256)), dim3(256), 0, nullptr, dx.as<float>(), n,
                       dw.as<float>(), dout.as<float>());
    MLP_HIP(hipGetLastError());
    MLP_HIP(hipMemcpy(out, dout.p, sizeof(float) * n * kOutputs, hipMemcpyDeviceToHost));
    return 0;
}

int mlp_train(const float* x, const float* y, const float* mask, int64_t n, int32_t epochs, int32_t batch, float lr,
              uint64_t seed, float* w_inout, float* loss_hist, std::string& err) {
    if (n <= 0 || epochs <= 0 || batch <= 0 || !x || !y || !mask || !w_inout) {
        err = "bad arguments";
        return -1;
    }
    DevBuf dx, dy, dm, dw, dg, dm1, dm2, dl, didx;
    MLP_HIP(dx.alloc(sizeof(float) * n * kFeatures));
    MLP_HIP(dy.alloc(sizeof(float) * n * kOutputs));
    MLP_HIP(dm.alloc(sizeof(float) * n * kOutputs));
    MLP_HIP(dw.alloc(sizeof(float) * kMlpParams));
    MLP_HIP(dg.alloc(sizeof(float) * kMlpParams));
    MLP_HIP(dm1.alloc(sizeof(float) * kMlpParams));
    MLP_HIP(dm2.alloc(sizeof(float) * kMlpParams));
    MLP_HIP(dl.alloc(sizeof(float) * size_t(epochs)));
    MLP_HIP(didx.alloc(sizeof(int32_t) * n));
    MLP_HIP(hipMemcpy(dx.p, x, sizeof(float) * n * kFeatures, hipMemcpyHostToDevice));
    MLP_HIP(hipMemcpy(dy.p, y, sizeof(float) * n * kOutputs, hipMemcpyHostToDevice));
    MLP_HIP(hipMemcpy(dm.p, mask, sizeof(float) * n * kOutputs, hipMemcpyHostToDevice));
    MLP_HIP(hipMemcpy(dw.p, w_inout, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    const int pb = (kMlpParams + 255) / 256;
    hipLaunchKernelGGL(zero_kernel, dim3(pb), dim3(256), 0, nullptr, dg.as<float>(), kMlpParams);
    hipLaunchKernelGGL(zero_kernel, dim3(pb), dim3(256), 0, nullptr, dm1.as<float>(), kMlpParams);
    hipLaunchKernelGGL(zero_kernel, dim3(pb), dim3(256), 0, nullptr, dm2.as<float>(), kMlpParams);
    hipLaunchKernelGGL(zero_kernel, dim3((epochs + 255) / 256), dim3(256), 0, nullptr, dl.as<float>(), epochs);
    std::vector<int32_t> perm(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) perm[size_t(i)] = int32_t(i);
    std::mt19937_64 rng(seed);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    int64_t t = 0;
    for (int e = 0; e < epochs; ++e) {
        std::shuffle(perm.begin(), perm.end(), rng);
        MLP_HIP(hipMemcpy(didx.p, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
        for (int64_t s0 = 0; s0 < n; s0 += batch) {
            const int64_t bn = std::min<int64_t>(batch, n - s0);
            const int blocks = int((bn + kTile - 1) / kTile);
            // the epoch's loss is the mean of its batch losses
            hipLaunchKernelGGL(mlp_fwd_bwd_kernel, dim3(blocks), dim3(kThreads), 0, nullptr, dx.as<float>(),
                               dy.as<float>(), dm.as<float>(), didx.as<int32_t>() + s0, bn, 1.f / float(bn),
                               dw.as<float>(), dg.as<float>(), dl.as<float>() + e);
            ++t;
            const float c1 = 1.f - std::pow(b1, float(t)), c2 = 1.f - std::pow(b2, float(t));
            hipLaunchKernelGGL(adam_kernel, dim3(pb), dim3(256), 0, nullptr, dw.as<float>(), dm1.as<float>(),
                               dm2.as<float>(), dg.as<float>(), lr, b1, b2, eps, c1, c2);
        }
        MLP_HIP(hipGetLastError());
    }
    MLP_HIP(hipDeviceSynchronize());
    MLP_HIP(hipMemcpy(w_inout, dw.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
    if (loss_hist) {
        std::vector<float> lh(static_cast<size_t>(epochs));
        MLP_HIP(hipMemcpy(lh.data(), dl.p, sizeof(float) * epochs, hipMemcpyDeviceToHost));
        const int64_t nb = (n + batch - 1) / batch;
        for (int e = 0; e < epochs; ++e) loss_hist[e] = lh[size_t(e)] / float(nb);
    }
    return 0;
}

}  // namespace admatch
