// hbm_copy <MiB>: copies a device buffer of the given size once with a
// plain 16-byte-per-lane copy kernel.  Used to calibrate rocprofv3's
// FETCH_SIZE / WRITE_SIZE counters (known bytes read and written).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

int main(int argc, char** argv) {
    const size_t mib = argc > 1 ? size_t(std::atol(argv[1])) : 256;
    const size_t bytes = mib << 20, n = bytes / sizeof(uint4);
    uint4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
    if (hipMemset(a, 1, bytes) != hipSuccess) return 1;
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, nullptr, a, b, n);
    if (hipDeviceSynchronize() != hipSuccess) return 1;
    std::printf("copied %zu bytes (%zu KiB read, %zu KiB written)\n", bytes, bytes >> 10, bytes >> 10);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
