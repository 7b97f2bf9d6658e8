// Learned pair screener for all-pairs generalized correlation (no reference
// counterpart; the reference computes KSG for every pair).
//
// 1. Pair features. The centred trajectory D (3N x n, rows (node, coordinate))
//    gives the covariance S = D D^T / n with one plain rocBLAS SGEMM (FP32, a
//    library GEMM). Per node: mean, 3x3 covariance, the inverse of its
//    Cholesky factor, RMSF. Per pair (a, b) one thread forms the 3x3
//    cross-covariance, the whitened T = La^-1 Cab Lb^-T, its singular values
//    (canonical correlations, from the closed-form eigenvalues of T^T T) and
//    16 features:
//      0 |trace-Pearson r|  1-3 canonical correlations  4 Gaussian MI
//      -0.5 sum ln(1 - rho^2)  5 its GC  6 mean distance / 10 A
//      7 min(7.5 / distance, 2)  8-9 RMSF a, b  10 RMSF ratio
//      11 log1p(|a-b|) / log1p(N)  12 ||T||_F / sqrt 3  13 max |coordinate
//      correlation|  14 ln(n) / 10  15 1
//    stored as bf16 (round to nearest even).
// 2. MLP 16 -> 64 -> 64 -> 1 (ReLU) plus a skip of feature 5 to the
//    output (out = net(x) + x5, so the network starts from, and corrects,
//    the Gaussian-MI GC estimate), hand-written with
//    v_mfma_f32_16x16x16_bf16: a wave takes 16 pairs at a time; bf16 operands,
//    FP32 accumulation; FP32 master weights converted to bf16 per launch.
//    Operand layouts of the 16x16x16 instruction (lane l, element e):
//      A[i][k]: i = l % 16, k = 4 (l / 16) + e
//      B[k][j]: j = l % 16, k = 4 (l / 16) + e
//      D[i][j]: j = l % 16, i = 4 (l / 16) + e
//    so a D-layout tile of activations is directly the A operand of the
//    weight-gradient products (h1^T dz2, x^T dz1) and the B operand for the
//    gradient side; the forward and dh1 products read their A operand through
//    a 16 x 64 LDS transpose.
// 3. Training: loss = mean_i w_i (out_i - y_i)^2; every wave writes its
//    partial gradient, a second kernel sums the partials in wave order
//    (deterministic), Adam updates the FP32 master weights.
// 4. Selection: candidates (prediction >= threshold) plus a deterministic
//    hashed audit sample of the rest are compacted into a pair list for the
//    exact KSG path.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <mutex>

#include "../../include/netsci_hip.h"
#include "nsx_internal.h"

namespace nsx {
namespace {

constexpr int kF = 16, kH = 64;
constexpr int oW1 = 0, oB1 = 1024, oW2 = 1088, oB2 = 5184, oW3 = 5248, oB3 = 5312;
constexpr int kParams = 5313;
constexpr int kNodeInfo = 16;
constexpr int kSkipFeature = 5;  // Gaussian-MI GC, added to the network output

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
    bf16x4 r;
    r[0] = (short)f2bf(a);
    r[1] = (short)f2bf(b);
    r[2] = (short)f2bf(c);
    r[3] = (short)f2bf(d);
    return r;
}

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// ---- 1. features -------------------------------------------------------------

// Row means of the (node, coordinate) rows and the centred copy D.
__global__ void centre_kernel(const float *__restrict__ X, long long ld, long long ds, int N,
                              int n, float *__restrict__ D, float *__restrict__ mu) {
    const int r = blockIdx.x;  // row = node * 3 + coordinate
    const int a = r / 3, c = r % 3;
    const float *x = X + (long long)a * ld + (long long)c * ds;
    double s = 0.0;
    for (int t = threadIdx.x; t < n; t += blockDim.x) s += x[t];
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    const float m = (float)(red[0] / n);
    if (threadIdx.x == 0) mu[r] = m;
    float *dr = D + (long long)r * n;
    for (int t = threadIdx.x; t < n; t += blockDim.x) dr[t] = x[t] - m;
    (void)N;
}

// Per node: mean (3), covariance diagonal (3), trace, inverse Cholesky
// factor (w00 w10 w11 w20 w21 w22), RMSF.
__global__ void node_info_kernel(const float *__restrict__ S, const float *__restrict__ mu,
                                 int N, float *__restrict__ info) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= N) return;
    const long long M = 3LL * N;
    const float *s = S + 3LL * a * M + 3LL * a;
    const float c00 = s[0], c01 = s[1], c02 = s[2];
    const float c11 = s[M + 1], c12 = s[M + 2], c22 = s[2 * M + 2];
    const float eps = 1e-12f;
    const float l00 = sqrtf(fmaxf(c00, eps));
    const float l10 = c01 / l00, l20 = c02 / l00;
    const float l11 = sqrtf(fmaxf(c11 - l10 * l10, eps));
    const float l21 = (c12 - l20 * l10) / l11;
    const float l22 = sqrtf(fmaxf(c22 - l20 * l20 - l21 * l21, eps));
    const float w00 = 1.0f / l00, w11 = 1.0f / l11, w22 = 1.0f / l22;
    const float w10 = -l10 * w00 / l11;
    const float w21 = -l21 * w11 / l22;
    const float w20 = -(l20 * w00 + l21 * w10) / l22;
    float *o = info + (long long)a * kNodeInfo;
    o[0] = mu[3 * a];
    o[1] = mu[3 * a + 1];
    o[2] = mu[3 * a + 2];
    o[3] = fmaxf(c00, eps);
    o[4] = fmaxf(c11, eps);
    o[5] = fmaxf(c22, eps);
    o[6] = fmaxf(c00 + c11 + c22, eps);
    o[7] = w00;
    o[8] = w10;
    o[9] = w11;
    o[10] = w20;
    o[11] = w21;
    o[12] = w22;
    o[13] = sqrtf(o[6]);
}

// Eigenvalues (descending) of a symmetric 3x3 matrix (trigonometric method).
__device__ void sym3_eig(float g00, float g01, float g02, float g11, float g12, float g22,
                         float &e1, float &e2, float &e3) {
    const float p1 = g01 * g01 + g02 * g02 + g12 * g12;
    const float q = (g00 + g11 + g22) / 3.0f;
    const float a0 = g00 - q, a1 = g11 - q, a2 = g22 - q;
    const float p2 = a0 * a0 + a1 * a1 + a2 * a2 + 2.0f * p1;
    if (p2 <= 1e-30f) {
        e1 = e2 = e3 = q;
        return;
    }
    const float p = sqrtf(p2 / 6.0f);
    const float ip = 1.0f / p;
    const float b00 = a0 * ip, b11 = a1 * ip, b22 = a2 * ip;
    const float b01 = g01 * ip, b02 = g02 * ip, b12 = g12 * ip;
    const float det = b00 * (b11 * b22 - b12 * b12) - b01 * (b01 * b22 - b12 * b02) +
                      b02 * (b01 * b12 - b11 * b02);
    const float r = fminf(fmaxf(det * 0.5f, -1.0f), 1.0f);
    const float phi = acosf(r) / 3.0f;
    e1 = q + 2.0f * p * cosf(phi);
    e3 = q + 2.0f * p * cosf(phi + 2.0943951023931953f);
    e2 = 3.0f * q - e1 - e3;
}

__global__ void features_kernel(const float *__restrict__ S, const float *__restrict__ info,
                                int N, int n, const int *__restrict__ ab, long long P,
                                unsigned short *__restrict__ feat) {
    const long long M = 3LL * N;
    const float ln_n = logf((float)n) / 10.0f;
    const float ln_N = log1pf((float)N);
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < P;
         p += (long long)gridDim.x * blockDim.x) {
        const int a = ab[2 * p], b = ab[2 * p + 1];
        const float *ia = info + (long long)a * kNodeInfo;
        const float *ib = info + (long long)b * kNodeInfo;
        float C[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) C[i][j] = S[(3LL * a + i) * M + 3LL * b + j];
        const float r = (C[0][0] + C[1][1] + C[2][2]) * rsqrtf(ia[6] * ib[6]);
        // Wa (lower triangular) * C
        const float wa[3][3] = {{ia[7], 0.f, 0.f}, {ia[8], ia[9], 0.f}, {ia[10], ia[11], ia[12]}};
        const float wb[3][3] = {{ib[7], 0.f, 0.f}, {ib[8], ib[9], 0.f}, {ib[10], ib[11], ib[12]}};
        float Mw[3][3], T[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                Mw[i][j] = wa[i][0] * C[0][j] + wa[i][1] * C[1][j] + wa[i][2] * C[2][j];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                T[i][j] = Mw[i][0] * wb[j][0] + Mw[i][1] * wb[j][1] + Mw[i][2] * wb[j][2];
        float G[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) G[i][j] = T[0][i] * T[0][j] + T[1][i] * T[1][j] + T[2][i] * T[2][j];
        float e1, e2, e3;
        sym3_eig(G[0][0], G[0][1], G[0][2], G[1][1], G[1][2], G[2][2], e1, e2, e3);
        const float cap = 1.0f - 1e-6f;
        const float r1 = sqrtf(fminf(fmaxf(e1, 0.f), cap));
        const float r2 = sqrtf(fminf(fmaxf(e2, 0.f), cap));
        const float r3 = sqrtf(fminf(fmaxf(e3, 0.f), cap));
        const float ig = fminf(-0.5f * (logf(1.f - r1 * r1) + logf(1.f - r2 * r2) + logf(1.f - r3 * r3)),
                               10.0f);
        const float gcg = sqrtf(fmaxf(1.0f - expf(-2.0f * ig / 3.0f), 0.f));
        const float dx = ia[0] - ib[0], dy = ia[1] - ib[1], dz = ia[2] - ib[2];
        const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
        const float ra = ia[13], rb = ib[13];
        float cmax = 0.f;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                cmax = fmaxf(cmax, fabsf(C[i][j]) * rsqrtf(ia[3 + i] * ib[3 + j]));
        const float frob = sqrtf(fmaxf(G[0][0] + G[1][1] + G[2][2], 0.f) / 3.0f);
        const float f[kF] = {fabsf(r), r1, r2, r3, ig, gcg, dist / 10.0f,
                             fminf(7.5f / fmaxf(dist, 1.0f), 2.0f), ra, rb,
                             fminf(ra, rb) / fmaxf(ra, rb), log1pf((float)abs(a - b)) / ln_N,
                             frob, cmax, ln_n, 1.0f};
        unsigned short *o = feat + p * kF;
        for (int i = 0; i < kF; ++i) o[i] = f2bf(f[i]);
    }
}

// ---- 2. MLP -------------------------------------------------------------------

struct Weights {
    bf16x4 w1[4];      // B operand of x W1, N tile t
    bf16x4 w2[4][4];   // B operand of h1 W2, K chunk c, N tile t
    float b1[4], b2[4], w3[4];
    float b3;
};

__device__ void load_weights(const float *__restrict__ P, int lane, Weights &W, bf16x4 (*w2t)[4]) {
    const int j = lane & 15, g = lane >> 4;
    for (int t = 0; t < 4; ++t) {
        const int col = 16 * t + j;
        W.w1[t] = pack4(P[oW1 + (4 * g + 0) * kH + col], P[oW1 + (4 * g + 1) * kH + col],
                        P[oW1 + (4 * g + 2) * kH + col], P[oW1 + (4 * g + 3) * kH + col]);
        for (int c = 0; c < 4; ++c) {
            const int k0 = 16 * c + 4 * g;
            W.w2[c][t] = pack4(P[oW2 + (k0 + 0) * kH + col], P[oW2 + (k0 + 1) * kH + col],
                               P[oW2 + (k0 + 2) * kH + col], P[oW2 + (k0 + 3) * kH + col]);
            if (w2t) {
                // B[k = unit2][j = unit1] = W2[unit1][unit2]
                const float *row = P + oW2 + (long long)col * kH + k0;
                w2t[c][t] = pack4(row[0], row[1], row[2], row[3]);
            }
        }
        W.b1[t] = P[oB1 + col];
        W.b2[t] = P[oB2 + col];
        W.w3[t] = P[oW3 + col];
    }
    W.b3 = P[oB3];
}

struct Fwd {
    f32x4 z1[4], z2[4];
    float out[4];
};

// Forward pass of 16 samples starting at s0 (rows >= P give zero features).
// xs: wave LDS [16][16] bf16; hs: wave LDS [16][64] bf16.
__device__ void forward16(const unsigned short *__restrict__ feat, long long P, long long s0,
                          int lane, const Weights &W, unsigned short *xs, unsigned short *hs,
                          Fwd &f) {
    const int j = lane & 15, g = lane >> 4;
    // x tile -> LDS (16 samples x 16 features): each lane stores 4 features.
    {
        const long long s = s0 + j;
        unsigned long long v = 0;
        if (s < P) v = *reinterpret_cast<const unsigned long long *>(feat + s * kF + 4 * g);
        *reinterpret_cast<unsigned long long *>(xs + j * kF + 4 * g) = v;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    bf16x4 a1 = *reinterpret_cast<const bf16x4 *>(xs + j * kF + 4 * g);
    for (int t = 0; t < 4; ++t) {
        f32x4 z = {0.f, 0.f, 0.f, 0.f};
        f.z1[t] = mfma(a1, W.w1[t], z);
        for (int e = 0; e < 4; ++e) {
            f.z1[t][e] += W.b1[t];
            hs[(4 * g + e) * kH + 16 * t + j] = f2bf(fmaxf(f.z1[t][e], 0.f));
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    bf16x4 a2[4];
    for (int c = 0; c < 4; ++c) a2[c] = *reinterpret_cast<const bf16x4 *>(hs + j * kH + 16 * c + 4 * g);
    for (int t = 0; t < 4; ++t) {
        f32x4 z = {0.f, 0.f, 0.f, 0.f};
        for (int c = 0; c < 4; ++c) z = mfma(a2[c], W.w2[c][t], z);
        for (int e = 0; e < 4; ++e) z[e] += W.b2[t];
        f.z2[t] = z;
    }
    // Output head plus the skip connection from feature 5 (the Gaussian-MI
    // GC): the network learns a correction to that estimate.
    for (int e = 0; e < 4; ++e) {
        float s = 0.f;
        for (int t = 0; t < 4; ++t) s += fmaxf(f.z2[t][e], 0.f) * W.w3[t];
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        s += __shfl_xor(s, 4);
        s += __shfl_xor(s, 8);
        f.out[e] = s + W.b3 + bf2f(xs[(4 * g + e) * kF + kSkipFeature]);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__global__ __launch_bounds__(256) void mlp_forward_kernel(const float *__restrict__ params,
                                                          const unsigned short *__restrict__ feat,
                                                          long long P, float *__restrict__ out) {
    __shared__ __attribute__((aligned(16))) unsigned short lds[4][16 * kF + 16 * kH];
    const int lane = threadIdx.x & 63, wib = threadIdx.x >> 6;
    unsigned short *xs = lds[wib], *hs = lds[wib] + 16 * kF;
    Weights W;
    load_weights(params, lane, W, nullptr);
    const long long groups = (P + 15) / 16;
    const long long nw = (long long)gridDim.x * 4;
    for (long long gi = (long long)blockIdx.x * 4 + wib; gi < groups; gi += nw) {
        Fwd f;
        forward16(feat, P, gi * 16, lane, W, xs, hs, f);
        const int j = lane & 15, g = lane >> 4;
        if (j == 0)
            for (int e = 0; e < 4; ++e) {
                const long long s = gi * 16 + 4 * g + e;
                if (s < P) out[s] = f.out[e];
            }
    }
}

// Partial gradients of one wave over its groups of 16 samples.
__global__ __launch_bounds__(256) void mlp_grad_kernel(const float *__restrict__ params,
                                                       const unsigned short *__restrict__ feat,
                                                       const float *__restrict__ y,
                                                       const float *__restrict__ wt, long long B,
                                                       float *__restrict__ partial,
                                                       float *__restrict__ loss_partial) {
    __shared__ __attribute__((aligned(16))) unsigned short lds[4][16 * kF + 16 * kH];
    const int lane = threadIdx.x & 63, wib = threadIdx.x >> 6;
    const int j = lane & 15, g = lane >> 4;
    unsigned short *xs = lds[wib], *hs = lds[wib] + 16 * kF;
    Weights W;
    bf16x4 w2t[4][4];
    load_weights(params, lane, W, w2t);
    f32x4 gW1[4], gW2[4][4];
    float gb1[4] = {0, 0, 0, 0}, gb2[4] = {0, 0, 0, 0}, gw3[4] = {0, 0, 0, 0}, gb3 = 0.f, lsum = 0.f;
    for (int t = 0; t < 4; ++t) {
        gW1[t] = f32x4{0.f, 0.f, 0.f, 0.f};
        for (int c = 0; c < 4; ++c) gW2[c][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    const long long groups = (B + 15) / 16;
    const long long nw = (long long)gridDim.x * 4;
    const long long wave = (long long)blockIdx.x * 4 + wib;
    const float invB = 1.0f / (float)B;
    for (long long gi = wave; gi < groups; gi += nw) {
        const long long s0 = gi * 16;
        Fwd f;
        forward16(feat, B, s0, lane, W, xs, hs, f);
        float gs[4];
        for (int e = 0; e < 4; ++e) {
            const long long s = s0 + 4 * g + e;
            gs[e] = 0.f;
            if (s < B) {
                const float d = f.out[e] - y[s];
                gs[e] = 2.0f * wt[s] * d * invB;
                if (j == 0) lsum += wt[s] * d * d;
            }
            gb3 += (j == 0) ? gs[e] : 0.f;
        }
        // dz2 (D layout), dw3, db2; dz2 bf16 -> LDS for the dh1 product.
        bf16x4 dz2b[4];
        for (int t = 0; t < 4; ++t) {
            float v[4];
            for (int e = 0; e < 4; ++e) {
                const float h2 = fmaxf(f.z2[t][e], 0.f);
                gw3[t] += h2 * gs[e];
                v[e] = (f.z2[t][e] > 0.f) ? gs[e] * W.w3[t] : 0.f;
                gb2[t] += v[e];
                hs[(4 * g + e) * kH + 16 * t + j] = f2bf(v[e]);
            }
            dz2b[t] = pack4(v[0], v[1], v[2], v[3]);
        }
        // dW2[c][t] += h1^T dz2: the D-layout h1 tile c is the A operand.
        for (int c = 0; c < 4; ++c) {
            const bf16x4 h1c = pack4(fmaxf(f.z1[c][0], 0.f), fmaxf(f.z1[c][1], 0.f),
                                     fmaxf(f.z1[c][2], 0.f), fmaxf(f.z1[c][3], 0.f));
            for (int t = 0; t < 4; ++t) gW2[c][t] = mfma(h1c, dz2b[t], gW2[c][t]);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // dh1 = dz2 W2^T, A operand read through the LDS transpose.
        bf16x4 a[4];
        for (int c = 0; c < 4; ++c) a[c] = *reinterpret_cast<const bf16x4 *>(hs + j * kH + 16 * c + 4 * g);
        // x^T as A operand: A[i = feature j][k = sample 4g + e].
        const bf16x4 xt = {(short)xs[(4 * g + 0) * kF + j], (short)xs[(4 * g + 1) * kF + j],
                           (short)xs[(4 * g + 2) * kF + j], (short)xs[(4 * g + 3) * kF + j]};
        for (int t = 0; t < 4; ++t) {
            f32x4 dh = {0.f, 0.f, 0.f, 0.f};
            for (int c = 0; c < 4; ++c) dh = mfma(a[c], w2t[c][t], dh);
            float v[4];
            for (int e = 0; e < 4; ++e) {
                v[e] = (f.z1[t][e] > 0.f) ? dh[e] : 0.f;
                gb1[t] += v[e];
            }
            gW1[t] = mfma(xt, pack4(v[0], v[1], v[2], v[3]), gW1[t]);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    // Sum the per-lane bias/w3 partials over the four lane groups.
    for (int t = 0; t < 4; ++t) {
        gb1[t] += __shfl_xor(gb1[t], 16);
        gb1[t] += __shfl_xor(gb1[t], 32);
        gb2[t] += __shfl_xor(gb2[t], 16);
        gb2[t] += __shfl_xor(gb2[t], 32);
        gw3[t] += __shfl_xor(gw3[t], 16);
        gw3[t] += __shfl_xor(gw3[t], 32);
    }
    gb3 += __shfl_xor(gb3, 16);
    gb3 += __shfl_xor(gb3, 32);
    lsum += __shfl_xor(lsum, 16);
    lsum += __shfl_xor(lsum, 32);
    float *o = partial + wave * kParams;
    for (int t = 0; t < 4; ++t) {
        for (int e = 0; e < 4; ++e) {
            o[oW1 + (4 * g + e) * kH + 16 * t + j] = gW1[t][e];
            for (int c = 0; c < 4; ++c) o[oW2 + (16 * c + 4 * g + e) * kH + 16 * t + j] = gW2[c][t][e];
        }
        if (g == 0) {
            o[oB1 + 16 * t + j] = gb1[t];
            o[oB2 + 16 * t + j] = gb2[t];
            o[oW3 + 16 * t + j] = gw3[t];
        }
    }
    if (lane == 0) {
        o[oB3] = gb3;
        loss_partial[wave] = lsum;
    }
}

__global__ void reduce_partials_kernel(const float *__restrict__ partial, int waves,
                                       float *__restrict__ grad, const float *__restrict__ lp,
                                       float *__restrict__ loss, long long B) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kParams) {
        float s = 0.f;
        for (int w = 0; w < waves; ++w) s += partial[(long long)w * kParams + i];
        grad[i] = s;
    }
    if (i == 0 && loss) {
        float s = 0.f;
        for (int w = 0; w < waves; ++w) s += lp[w];
        *loss = s / (float)B;
    }
}

__global__ void adam_kernel(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, float lr, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kParams) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

// ---- 4. selection ---------------------------------------------------------------

__device__ __forceinline__ unsigned mix32(unsigned long long x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return (unsigned)x;
}

__global__ void select_kernel(const float *__restrict__ pred, long long P, float thr,
                              unsigned audit_cut, unsigned seed, const int *__restrict__ ab,
                              int *__restrict__ ab_out, long long *__restrict__ idx_out,
                              unsigned char *__restrict__ flag,
                              unsigned long long *__restrict__ counts) {
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < P;
         p += (long long)gridDim.x * blockDim.x) {
        unsigned char f = 0;
        if (pred[p] >= thr) f = 1;
        else if (mix32(((unsigned long long)seed << 40) ^ (unsigned long long)p) < audit_cut) f = 2;
        flag[p] = f;
        if (f) {
            const unsigned long long slot = atomicAdd(&counts[0], 1ull);
            if (f == 2) atomicAdd(&counts[1], 1ull);
            ab_out[2 * slot] = ab[2 * p];
            ab_out[2 * slot + 1] = ab[2 * p + 1];
            idx_out[slot] = p;
        }
    }
}

std::mutex g_blas_mutex;
rocblas_handle g_blas = nullptr;

}  // namespace
}  // namespace nsx

using namespace nsx;

extern "C" {

int nsx_mlp_param_count(int *count) {
    if (count) *count = kParams;
    return kOk;
}

int nsx_screen_features(const float *d_X, long long ld, long long dstride, int N, int n,
                        const int *d_ab, long long P, unsigned short *d_feat, void *stream) {
    TraceRange trace("nsx_screen_features");
    clear_error();
    NSX_REQUIRE(d_X && d_ab && d_feat && N >= 1 && n >= 2 && P >= 0,
                "nsx_screen_features: invalid arguments");
    NSX_REQUIRE(2 * dstride + n <= ld && dstride >= n, "nsx_screen_features: layout needs d = 3");
    int rc = ensure_device();
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    const long long M = 3LL * N;
    float *D = (float *)workspace(kWsScreen0, (size_t)(M * n + M * M + M + (long long)N * kNodeInfo) * sizeof(float));
    if (!D) return kHipError;
    float *S = D + M * n;
    float *mu = S + M * M;
    float *info = mu + M;
    hipLaunchKernelGGL(centre_kernel, dim3((unsigned)M), dim3(256), 0, s, d_X, ld, dstride, N, n,
                       D, mu);
    NSX_HIP(hipGetLastError(), "centre launch");
    {
        std::lock_guard<std::mutex> lock(g_blas_mutex);
        if (!g_blas && rocblas_create_handle(&g_blas) != rocblas_status_success) {
            set_error("rocblas_create_handle failed");
            return kInternal;
        }
        rocblas_set_stream(g_blas, s);
        // Column-major view: D is n x M (lda = n); S = D^T D / n (M x M).
        const float alpha = 1.0f / (float)n, beta = 0.0f;
        if (rocblas_sgemm(g_blas, rocblas_operation_transpose, rocblas_operation_none, (int)M,
                          (int)M, n, &alpha, D, n, D, n, &beta, S, (int)M) !=
            rocblas_status_success) {
            set_error("rocblas_sgemm (covariance) failed");
            return kInternal;
        }
    }
    hipLaunchKernelGGL(node_info_kernel, dim3((N + 255) / 256), dim3(256), 0, s, S, mu, N, info);
    NSX_HIP(hipGetLastError(), "node info launch");
    if (P > 0) {
        const unsigned blocks = (unsigned)std::min<long long>((P + 255) / 256, 65536);
        hipLaunchKernelGGL(features_kernel, dim3(blocks), dim3(256), 0, s, S, info, N, n, d_ab, P,
                           d_feat);
        NSX_HIP(hipGetLastError(), "features launch");
    }
    return kOk;
}

int nsx_mlp_forward(const float *d_params, const unsigned short *d_feat, long long P,
                    float *d_out, void *stream) {
    TraceRange trace("nsx_mlp_forward");
    clear_error();
    NSX_REQUIRE(d_params && d_feat && d_out && P >= 0, "nsx_mlp_forward: invalid arguments");
    if (P == 0) return kOk;
    int rc = ensure_device();
    if (rc) return rc;
    const long long groups = (P + 15) / 16;
    const unsigned blocks = (unsigned)std::min<long long>((groups + 3) / 4, 8192);
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                       d_params, d_feat, P, d_out);
    NSX_HIP(hipGetLastError(), "mlp forward launch");
    return kOk;
}

int nsx_mlp_grad(const float *d_params, const unsigned short *d_feat, const float *d_y,
                 const float *d_w, long long B, float *d_grad, float *d_loss, void *stream) {
    TraceRange trace("nsx_mlp_grad");
    clear_error();
    NSX_REQUIRE(d_params && d_feat && d_y && d_w && d_grad && B >= 1,
                "nsx_mlp_grad: invalid arguments");
    int rc = ensure_device();
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    const long long groups = (B + 15) / 16;
    const int blocks = (int)std::min<long long>((groups + 3) / 4, 64);
    const int waves = blocks * 4;
    float *partial = (float *)workspace(kWsScreen1, ((size_t)waves * kParams + waves) * sizeof(float));
    if (!partial) return kHipError;
    float *lp = partial + (size_t)waves * kParams;
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(blocks), dim3(256), 0, s, d_params, d_feat, d_y, d_w,
                       B, partial, lp);
    NSX_HIP(hipGetLastError(), "mlp grad launch");
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((kParams + 255) / 256), dim3(256), 0, s,
                       partial, waves, d_grad, lp, d_loss, B);
    NSX_HIP(hipGetLastError(), "grad reduce launch");
    return kOk;
}

int nsx_mlp_adam(float *d_params, const float *d_grad, float *d_m, float *d_v, int step,
                 float lr, void *stream) {
    TraceRange trace("nsx_mlp_adam");
    clear_error();
    NSX_REQUIRE(d_params && d_grad && d_m && d_v && step >= 1, "nsx_mlp_adam: invalid arguments");
    const float c1 = 1.0f - powf(0.9f, (float)step), c2 = 1.0f - powf(0.999f, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3((kParams + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, d_params, d_grad, d_m, d_v, lr, c1, c2);
    NSX_HIP(hipGetLastError(), "adam launch");
    return kOk;
}

int nsx_screen_select(const float *d_pred, long long P, float threshold, float audit_rate,
                      unsigned seed, const int *d_ab, int *d_ab_out, long long *d_idx_out,
                      unsigned char *d_flag, long long *h_n_selected, long long *h_n_audit,
                      void *stream) {
    TraceRange trace("nsx_screen_select");
    clear_error();
    NSX_REQUIRE(d_pred && d_ab && d_ab_out && d_idx_out && d_flag && P >= 0 && audit_rate >= 0 &&
                    audit_rate <= 1,
                "nsx_screen_select: invalid arguments");
    int rc = ensure_device();
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    unsigned long long *cnt = (unsigned long long *)workspace(kWsCounter, 64);
    if (!cnt) return kHipError;
    rc = fill_u32(cnt, 0u, 4, s);
    if (rc) return rc;
    const unsigned cut = (unsigned)std::min(4294967295.0, (double)audit_rate * 4294967296.0);
    if (P > 0) {
        const unsigned blocks = (unsigned)std::min<long long>((P + 255) / 256, 65536);
        hipLaunchKernelGGL(select_kernel, dim3(blocks), dim3(256), 0, s, d_pred, P, threshold, cut,
                           seed, d_ab, d_ab_out, d_idx_out, d_flag, cnt);
        NSX_HIP(hipGetLastError(), "select launch");
    }
    unsigned long long h[2] = {0, 0};
    NSX_HIP(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, s), "select count readback");
    NSX_HIP(hipStreamSynchronize(s), "select sync");
    if (h_n_selected) *h_n_selected = (long long)h[0];
    if (h_n_audit) *h_n_audit = (long long)h[1];
    return kOk;
}

}  // extern "C"
