// Probes for the measurements that report a stage kernel's use of the chip
// against what it can do.  The vector-ALU issue rate: independent fused
// multiply-add chains (no memory traffic in the timed loop), FP32 or FP64, on
// enough waves to fill every SIMD.  The streaming HBM rate: a kernel that
// reads five float streams and writes one, the mix of an FP32 extension stage
// (stage value, start value, three velocity components in; the stage out),
// over buffers larger than the 256 MB last-level cache.  Not part of the
// level-set path.
#include <hip/hip_runtime.h>

#include "../../include/vesicle_ls.h"

namespace vls {

template <typename T, int CH>
__global__ void __launch_bounds__(256) k_valu_rate(T* __restrict__ out, T a, T b, int iters)
{
    T x[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) x[c] = (T)(threadIdx.x + c);
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int c = 0; c < CH; ++c) x[c] = fma(x[c], a, b);
    }
    T s = 0;
#pragma unroll
    for (int c = 0; c < CH; ++c) s += x[c];
    out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

template <typename T>
hipError_t valu_rate(int blocks, int iters, double* insts_per_s, hipStream_t st)
{
    constexpr int CH = 16;
    T* out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc((void**)&out, sizeof(T) * (size_t)blocks * 256);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    float ms = 0;
    if (e == hipSuccess) {
        k_valu_rate<T, CH><<<dim3(blocks), dim3(256), 0, st>>>(out, (T)0.999, (T)0.001, iters);  // warm-up
        (void)hipEventRecord(e0, st);
        k_valu_rate<T, CH><<<dim3(blocks), dim3(256), 0, st>>>(out, (T)0.999, (T)0.001, iters);
        (void)hipEventRecord(e1, st);
        e = hipEventSynchronize(e1);
    }
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess) e = hipGetLastError();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (out) (void)hipFree(out);
    if (e == hipSuccess) *insts_per_s = (double)blocks * 4.0 * iters * CH / (ms * 1e-3);
    return e;
}

__global__ void __launch_bounds__(256) k_stream5(float* __restrict__ out, const float* __restrict__ a, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < 5; ++k) s += a[(size_t)k * n + i];
        out[i] = s;
    }
}

}  // namespace vls

extern "C" int vls_probe_stream_rate(int64_t points, double* bytes_per_s, vls_stream_t stream)
{
    if (!bytes_per_s || points < 1024) return VLS_ERR_INVALID;
    const hipStream_t st = (hipStream_t)stream;
    const size_t n = (size_t)points;
    float *a = nullptr, *o = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc((void**)&a, sizeof(float) * n * 5);
    if (e == hipSuccess) e = hipMalloc((void**)&o, sizeof(float) * n);
    if (e == hipSuccess) e = hipMemsetAsync(a, 0, sizeof(float) * n * 5, st);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    float ms = 0;
    const int reps = 10, blocks = 8192;
    if (e == hipSuccess) {
        vls::k_stream5<<<dim3(blocks), dim3(256), 0, st>>>(o, a, n);  // warm-up
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < reps; ++r) vls::k_stream5<<<dim3(blocks), dim3(256), 0, st>>>(o, a, n);
        (void)hipEventRecord(e1, st);
        e = hipEventSynchronize(e1);
    }
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess) e = hipGetLastError();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (o) (void)hipFree(o);
    if (e != hipSuccess) return VLS_ERR_HIP;
    *bytes_per_s = 6.0 * sizeof(float) * (double)n * reps / (ms * 1e-3);
    return VLS_OK;
}

extern "C" int vls_probe_valu_rate(int fp64, double* insts_per_s, vls_stream_t stream)
{
    if (!insts_per_s) return VLS_ERR_INVALID;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        return VLS_ERR_HIP;
    const int blocks = cus * 8, iters = 4096;  // 32 waves per CU
    const hipError_t e = fp64 ? vls::valu_rate<double>(blocks, iters, insts_per_s, (hipStream_t)stream)
                              : vls::valu_rate<float>(blocks, iters, insts_per_s, (hipStream_t)stream);
    return e == hipSuccess ? VLS_OK : VLS_ERR_HIP;
}
