// FP32 VALU issue-rate probe: a measurement tool, not part of the library.
//
// Built as its own shared object (libnsx_probe.so) that only bench.py's
// --valu-probe loads. It times one launch of 32 waves per CU running iters
// x 32 dependent-chain operations per lane of one instruction kind, which
// fixes the VALU peak that the roofline fractions are quoted against.
#include <hip/hip_runtime.h>

namespace {

// VALU throughput probe: 8 independent dependency chains per lane of the
// instruction mix under test; every result feeds the output so nothing is
// folded away. kind 0: v_sub_f32; 1: v_max3_f32 (|.| modifiers);
// 2: v_fma_f32; 3: the KSG inner-loop mix (sub, max3, compare, select).
__global__ __launch_bounds__(256) void valu_probe_kernel(float *out, float a, float b, int iters,
                                                         int kind) {
    float x[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) x[c] = threadIdx.x * 1e-3f + c;
    if (kind == 0) {
        for (int it = 0; it < iters; ++it)
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int c = 0; c < 8; ++c) x[c] = x[c] - a;
    } else if (kind == 1) {
        for (int it = 0; it < iters; ++it)
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int c = 0; c < 8; ++c) x[c] = fmaxf(fmaxf(fabsf(x[c]), fabsf(a)), fabsf(b - x[c]));
    } else if (kind == 2) {
        for (int it = 0; it < iters; ++it)
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int c = 0; c < 8; ++c) x[c] = fmaf(x[c], a, b);
    } else {
        for (int it = 0; it < iters; ++it)
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    const float d = fmaxf(fabsf(x[c] - a), fabsf(x[c] - b));
                    x[c] = (d < b) ? d : x[c] - a;
                }
    }
    float s = 0.f;
#pragma unroll
    for (int c = 0; c < 8; ++c) s += x[c];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace

extern "C" {

// (ms, lanes) of one launch after an untimed one; 0 on success, else the
// HIP error code.
int nsx_valu_probe(int kind, int iters, double *ms, long long *lanes) {
    if (kind < 0 || kind > 3 || iters < 1 || !ms || !lanes) return -1;
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return (int)e;
    const int blocks = cus * 8;  // 8 x 256 threads per CU = 32 waves
    float *out = nullptr;
    if ((e = hipMalloc(&out, (size_t)blocks * 256 * sizeof(float))) != hipSuccess) return (int)e;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(valu_probe_kernel, dim3(blocks), dim3(256), 0, nullptr, out, 1.25f,
                           0.75f, iters, kind);
        e = hipEventRecord(e0, nullptr);
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(valu_probe_kernel, dim3(blocks), dim3(256), 0, nullptr, out, 1.25f,
                           0.75f, iters, kind);
        e = hipEventRecord(e1, nullptr);
    }
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
    if (e != hipSuccess) return (int)e;
    *ms = t;
    *lanes = (long long)blocks * 256;
    return 0;
}

}  // extern "C"
