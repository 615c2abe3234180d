// Achievable HBM bandwidth on this GPU, measured with the library's own
// streaming copy kernels (the roofline's practical ceiling next to the
// datasheet peak): 16-byte vector loads and stores over a grid of at least
// eight workgroups per CU, in a plain and a non-temporal form.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "../../include/splash.h"

namespace splash {
namespace bandwidth {

constexpr int kThreads = 256;
constexpr int kUnroll = 4;
typedef float v4f __attribute__((ext_vector_type(4)));  // native 16-byte vector

// Each thread copies kUnroll float4 elements per grid stride, loads first.
template <bool NT>
__global__ __launch_bounds__(kThreads) void k_copy(const v4f* __restrict__ src, v4f* __restrict__ dst, size_t n) {
    const size_t stride = size_t(gridDim.x) * kThreads * kUnroll;
    for (size_t base = size_t(blockIdx.x) * kThreads * kUnroll + threadIdx.x; base < n; base += stride) {
        v4f v[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const size_t i = base + size_t(u) * kThreads;
            if (i < n) v[u] = NT ? __builtin_nontemporal_load(&src[i]) : src[i];
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const size_t i = base + size_t(u) * kThreads;
            if (i < n) {
                if (NT) __builtin_nontemporal_store(v[u], &dst[i]);
                else dst[i] = v[u];
            }
        }
    }
}

}  // namespace bandwidth
}  // namespace splash

extern "C" int splash_copy_bandwidth(int64_t bytes, int reps, double* gbs_out) {
    using namespace splash::bandwidth;
    if (bytes < (int64_t(1) << 20) || reps < 1 || !gbs_out) return SPLASH_EINVAL;
    const size_t n = size_t(bytes) / sizeof(v4f);
    v4f *src = nullptr, *dst = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = SPLASH_OK;
    double best = 0;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipMalloc(&src, n * sizeof(v4f)) != hipSuccess || hipMalloc(&dst, n * sizeof(v4f)) != hipSuccess ||
        hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess || hipMemsetAsync(src, 0, n * sizeof(v4f), s) != hipSuccess) {
        rc = SPLASH_EDEVICE;
    } else {
        const size_t per_block = size_t(kThreads) * kUnroll;
        const unsigned blocks = unsigned(std::min<size_t>((n + per_block - 1) / per_block, size_t(cus) * 16));
        for (int nt = 0; nt < 2 && rc == SPLASH_OK; ++nt) {
            auto kernel = nt ? k_copy<true> : k_copy<false>;
            kernel<<<blocks, kThreads, 0, s>>>(src, dst, n);  // warm-up
            if (hipEventRecord(e0, s) != hipSuccess) rc = SPLASH_EDEVICE;
            for (int r = 0; r < reps; ++r) kernel<<<blocks, kThreads, 0, s>>>(src, dst, n);
            float ms = 0;
            if (hipGetLastError() != hipSuccess || hipEventRecord(e1, s) != hipSuccess ||
                hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0)
                rc = SPLASH_EDEVICE;
            else best = std::max(best, 2.0 * double(n * sizeof(v4f)) * reps / (double(ms) * 1e-3) / 1e9);
        }
    }
    if (s) (void)hipStreamSynchronize(s);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    if (rc == SPLASH_OK) *gbs_out = best;
    return rc;
}
