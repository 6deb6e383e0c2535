// Issue-rate probe of the vector ALU, for the measurements that report a
// stage kernel's VALU use against what the chip can issue: independent
// fused multiply-add chains (no memory traffic in the timed loop), FP32 or
// FP64, on enough waves to fill every SIMD.  Not part of the level-set path.
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

}  // namespace vls

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
