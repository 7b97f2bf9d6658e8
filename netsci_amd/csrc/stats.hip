// Row mean and population standard deviation on gfx950.
//
// Reference: netcalc/src/math_utils_gpu.cu:23-68 (meanKernel), :70-156
// (standardDeviationKernel) and :158-206 (drivers). The reference fixes a
// float summation order, kept here so results are bit-identical: per row,
// thread t of 1024 adds a[t], a[t + 1024], ... in order, and the 1024
// partial sums are folded as a tree, s[t] += s[t + o] for o = 512, ..., 1;
// u = s[0] / (float)n. The deviation pass adds (a - u)^2 in the same order
// and sigma = sqrt(s[0] / (float)n). The square is fl(d * d) and the sum a
// separate add (no FMA contraction); the reference squares with powf(d, 2).
// HBM-bound (one read of the matrix per pass); one 1024-lane workgroup per
// row, rows dealt round-robin to the resident workgroups.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "../../include/netsci_hip.h"
#include "nsx_internal.h"

namespace nsx {
namespace {

constexpr int kLanes = 1024;

// Tree fold of s[0..1023]; every lane returns s[0].
__device__ __forceinline__ float tree_fold(float *s, int t) {
    for (int o = kLanes / 2; o >= 1; o >>= 1) {
        __syncthreads();
        if (t < o) s[t] = s[t] + s[t + o];
    }
    __syncthreads();
    return s[0];
}

__global__ __launch_bounds__(kLanes) void row_stats_kernel(const float *__restrict__ a,
                                                           long long m, long long n,
                                                           float *__restrict__ u,
                                                           float *__restrict__ sigma) {
#pragma clang fp contract(off)
    __shared__ float s[kLanes];
    const int t = threadIdx.x;
    for (long long r = blockIdx.x; r < m; r += gridDim.x) {
        const float *row = a + r * n;
        float acc = 0.0f;
        for (long long i = t; i < n; i += kLanes) acc += row[i];
        s[t] = acc;
        const float mean = tree_fold(s, t) / (float)n;
        if (sigma) {
            float acc2 = 0.0f;
            for (long long i = t; i < n; i += kLanes) {
                const float d = row[i] - mean;
                acc2 += d * d;
            }
            __syncthreads();  // every lane has read s[0]
            s[t] = acc2;
            const float sq = tree_fold(s, t);
            if (t == 0) sigma[r] = sqrtf(sq / (float)n);
        }
        if (t == 0) u[r] = mean;
        __syncthreads();  // s is reused by the next row
    }
}

int launch_row_stats(const float *d_a, long long m, long long n, float *d_u, float *d_sigma,
                     void *stream) {
    NSX_REQUIRE(d_a && d_u && m >= 0 && n >= 1, "row statistics: invalid arguments");
    if (m == 0) return kOk;
    int rc = ensure_device();
    if (rc) return rc;
    const long long groups = std::min<long long>(m, (long long)cu_count() * 2);
    hipLaunchKernelGGL(row_stats_kernel, dim3((unsigned)groups), dim3(kLanes), 0,
                       (hipStream_t)stream, d_a, m, n, d_u, d_sigma);
    NSX_HIP(hipGetLastError(), "row statistics launch");
    return kOk;
}

}  // namespace
}  // namespace nsx

using namespace nsx;

extern "C" {

int nsx_row_mean(const float *d_a, long long m, long long n, float *d_u, void *stream) {
    TraceRange trace("nsx_row_mean");
    clear_error();
    return launch_row_stats(d_a, m, n, d_u, nullptr, stream);
}

int nsx_row_std(const float *d_a, long long m, long long n, float *d_u, float *d_sigma,
                void *stream) {
    TraceRange trace("nsx_row_std");
    clear_error();
    NSX_REQUIRE(d_sigma, "nsx_row_std: null sigma");
    return launch_row_stats(d_a, m, n, d_u, d_sigma, stream);
}

}  // extern "C"
