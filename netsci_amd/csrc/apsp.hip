// Hedetniemi all-pairs shortest paths on gfx950.
//
// Reference: netcalc/src/hedetniemi_gpu.cu:28-107 (kernels) and :109-261
// (drivers). One iteration is
//     H_{p+1}[i][j] = min_k fl(A[j][k] + H_p[k][i])
// — a min-plus product of H_p^T and A^T. The reference updates H in place
// while other blocks read it (a race); this engine double-buffers (Jacobi),
// so every run gives the same H and the same snapshots Hi[p] = H_p.
//
// min-plus kernel: 64 x 64 output tile per 256-thread workgroup, 4 x 4
// outputs per thread, k streamed through LDS in steps of 16 (both operand
// tiles stored k-major so each thread reads four consecutive values with one
// ds_read_b128). Each output folds k in ascending order. When A holds no
// negative zero, v_min_f32 gives exactly the reference's strict-`<` result
// (sums can only be -0 if both addends are -0); otherwise the kernel takes a
// compare-and-select loop that keeps the first minimum, as the reference
// does. Out-of-range rows/columns are padded with +inf, which never wins.
// Roofline: FP32 VALU, 2 lane-ops (add, min) per (i, j, k).
//
// Sparse step (graphs whose rows of A hold at most kDenseDegree entries
// < +inf on average, e.g. a 7.5 A contact graph): only k with
// A[j][k] < +inf can win the fold (any other k gives +inf or NaN, which the
// strict `<` never takes), so H_{p+1}[i][j] folds just those k, ascending,
// with the compare-and-select — bit-identical to the dense fold. Lanes own
// i and a wave takes one j at a time, so j's neighbour list is wave-uniform
// and each k costs one coalesced 256-byte read of H_p[k][i0..i0+63]. Cost
// per step: N^2 * degree / 64 such reads instead of N^3 lane-ops.
//
// Backtrack (hedetniemi_gpu.cu:70-107): for each (i, j) and p from L-1 down
// to 0, take the first q != j (ascending) with
// fl(fl(Hi[p][i][q] + A[q][k]) - len) < tol, the reference's first-hit rule.
// Only q with A[q][k] finite or -inf can pass that test: with A[q][k] = +inf
// the sum is +inf or NaN and the difference is +inf or NaN, never below tol;
// NaN likewise. So the candidates of node k are the column list
//     cand(k) = { q : A[q][k] < +inf }, ascending,
// built once on the device (count, scan, fill; each list padded to a
// multiple of 4 with (q = 0, +inf), which never passes), and
//   * sparse graphs (mean list length <= kDenseDegree): one LANE per (i, j)
//     walks cand(k) four entries at a time — at the ~11 contacts per node of
//     a 7.5 A contact graph a step costs a few reads instead of a scan of
//     all N nodes; the workgroup's lanes share i, so the row Hi[p][i][.]
//     they read from is staged in LDS once per step;
//   * dense graphs: one WAVE per (i, j) tests 64 q at once against a
//     transposed copy of A (coalesced) and a ballot picks the first hit.
// A device flag written by the scan selects the kernel: the backtrack
// launches both variants and the other one exits at once; for the steps the
// host polls the flag of A's lists (written to pinned host memory) once per
// call and launches only the one that works (L launches that exit at once
// would cost more than the wait).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>

#include "../../include/netsci_hip.h"
#include "nsx_internal.h"

namespace nsx {
namespace {

constexpr int kTile = 64;   // output tile edge
constexpr int kKStep = 16;  // k values per LDS stage
constexpr int kMicro = 4;   // outputs per thread per dimension

// out[i][j] = min_k fl(A[j][k] + Hp[k][i]).
template <bool kExactTies>
__device__ __forceinline__ void minplus_tile(const float *__restrict__ A,
                                             const float *__restrict__ Hp,
                                             float *__restrict__ out, int N, float (*Hs)[kTile],
                                             float (*As)[kTile]) {
    const int tid = threadIdx.x;
    const int tx = tid & 15;   // i micro-tile
    const int ty = tid >> 4;   // j micro-tile
    const int i0 = blockIdx.x * kTile;
    const int j0 = blockIdx.y * kTile;
    float acc[kMicro][kMicro];
#pragma unroll
    for (int a = 0; a < kMicro; ++a)
#pragma unroll
        for (int b = 0; b < kMicro; ++b) acc[a][b] = __builtin_inff();
    const float inf = __builtin_inff();

    for (int k0 = 0; k0 < N; k0 += kKStep) {
        // Hs[kk][ii] = Hp[k0+kk][i0+ii]; 16 x 64 values, 4 per thread.
#pragma unroll
        for (int r = 0; r < (kKStep * kTile) / 256; ++r) {
            const int e = r * 256 + tid;
            const int kk = e / kTile, ii = e % kTile;
            const int k = k0 + kk, i = i0 + ii;
            Hs[kk][ii] = (k < N && i < N) ? Hp[(long long)k * N + i] : inf;
        }
        // As[kk][jj] = A[j0+jj][k0+kk]; read along k, store k-major.
#pragma unroll
        for (int r = 0; r < (kKStep * kTile) / 256; ++r) {
            const int e = r * 256 + tid;
            const int jj = e / kKStep, kk = e % kKStep;
            const int k = k0 + kk, j = j0 + jj;
            As[kk][jj] = (k < N && j < N) ? A[(long long)j * N + k] : inf;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kKStep; ++kk) {
            const float4 hv = *reinterpret_cast<const float4 *>(&Hs[kk][tx * kMicro]);
            const float4 av = *reinterpret_cast<const float4 *>(&As[kk][ty * kMicro]);
            const float h[4] = {hv.x, hv.y, hv.z, hv.w};
            const float v[4] = {av.x, av.y, av.z, av.w};
#pragma unroll
            for (int a = 0; a < kMicro; ++a)
#pragma unroll
                for (int b = 0; b < kMicro; ++b) {
                    const float s = v[b] + h[a];
                    if (kExactTies) acc[a][b] = (s < acc[a][b]) ? s : acc[a][b];
                    else acc[a][b] = __builtin_fminf(acc[a][b], s);
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < kMicro; ++a) {
        const int i = i0 + tx * kMicro + a;
        if (i >= N) continue;
#pragma unroll
        for (int b = 0; b < kMicro; ++b) {
            const int j = j0 + ty * kMicro + b;
            if (j < N) out[(long long)i * N + j] = acc[a][b];
        }
    }
}

// Dense step; runs only when the neighbour lists of A flagged the graph
// dense (otherwise minplus_sparse_kernel computes the same step).
__global__ __launch_bounds__(256) void minplus_kernel(const float *__restrict__ A,
                                                      const float *__restrict__ Hp,
                                                      float *__restrict__ out, int N,
                                                      const unsigned *__restrict__ negzero,
                                                      const unsigned *__restrict__ dense) {
    __shared__ __attribute__((aligned(16))) float Hs[kKStep][kTile];
    __shared__ __attribute__((aligned(16))) float As[kKStep][kTile];
    if (!dense[0]) return;
    if (*negzero) minplus_tile<true>(A, Hp, out, N, Hs, As);
    else minplus_tile<false>(A, Hp, out, N, Hs, As);
}

// Sparse step. out[i][j] = min over k in nbr(j) (ascending) of
// fl(A[j][k] + Hp[k][i]) with the reference's strict-`<` select from +inf,
// where nbr(j) = { k : A[j][k] < +inf }: every other k gives +inf or NaN,
// which the select never takes, so the result equals the dense fold bit for
// bit (signed zeros and NaN included).
// Workgroup = a (64 NIT) (i) x TJ (j) output tile of 16 waves; lanes own i
// (NIT columns 64 apart), each wave takes TJ / 16 of the j (16 waves rather
// than 4 keep enough row reads in flight at small N: 0.77 -> 0.41 ms for 32
// steps at N = 1000). For one j the neighbour list is wave-uniform (scalar
// loads) and every k reads the coalesced row segments Hp[k][i0..i0+63] of
// the wave's NIT tiles. Results go through an LDS tile so the output is
// written as row segments out[i][j0..j0+TJ). Workgroups g, g+8, g+16, ...
// share an XCD; they get the same i-range, so that XCD's L2 keeps the
// column slice of Hp they read.
// Measured alternatives at N = 4000, L = 132 (32 steps lengths only, this
// kernel 10.9 ms; all bit-identical): lists padded to 16 entries so no
// 4-entry tail groups remain 11.5 ms; 16-entry rounds with per-group guards
// 14.3 ms, two j per wave interleaved 19.9 ms; a v_min3_f32 fold (exact when
// A holds no -0.0) beside the compare-and-select one, picked on the device,
// 14.3 ms; buffer loads with the row offset in a scalar register 17.2 ms.
// Counters (N = 4000 step): 5.8 M row reads, L1 hit rate ~93%, 2.5 TB/s from
// HBM, VALU ~55% busy.
// TJ = 16 (one j per wave) for small N, where the 64-wide tiles would give
// each CU only a few waves, each with a chain of four lists to walk.
// Two i-tiles per wave (kStepTiles): every list entry's scalar loads serve
// two row reads and the grid has half the waves (one round at N = 1000):
// N = 1000 lengths 0.374 -> 0.339 ms, N = 4000 10.43 -> 9.66 ms; four tiles
// need 102 VGPRs and halve the resident waves (0.65 and 12.9 ms); two
// adjacent columns per lane by one 8-byte load, 0.41 and 14.3 ms; the tail
// groups as one more 16-entry round that repeats its last group, 0.37 and
// 10.9 ms.
// Write-out: with TJ = 16 a wave stores four 16-column row segments per
// instruction instead of one (N = 1000 lengths 0.343 -> 0.312 ms). kStream
// makes those stores non-temporal, so the step leaves no dirty output in
// L2 for the end-of-kernel write-back: 0.312 -> 0.292 ms for a lengths
// call, whose outputs are read only by the next step. The path pass keeps
// plain stores (its snapshots feed the backtrack: 0.558 ms with paths
// against 0.581 non-temporal), and so does TJ = 64 (N = 4000 lengths 9.61
// ms plain, 10.95 non-temporal).
constexpr int kStepWaves = 16;  // waves per sparse-step workgroup
constexpr int kStepTiles = 2;   // 64-column i-tiles per wave

// NIT: 64-column i-tiles per wave; each list entry's scalar loads then
// serve NIT row reads, and the grid has 1/NIT as many waves.
template <int TJ, int NIT, bool kStream>
__global__ __launch_bounds__(kStepWaves * 64) void minplus_sparse_kernel(
    const float *__restrict__ Hp, const long long *__restrict__ rows, const int *__restrict__ lst,
    const unsigned *__restrict__ dense, float *__restrict__ out, int N, int nI, int nJ) {
    __shared__ float T[64 * NIT][TJ + 1];
    (void)dense;  // launched only for sparse lists (the host polls the flag)
    const int g = blockIdx.x;
    const int s = g >> 3;
    const int iI = (g & 7) + 8 * (s / nJ);  // group of NIT i-tiles
    const int jJ = s % nJ;
    if (iI >= nI) return;
    const int i0 = iI * 64 * NIT, j0 = jJ * TJ;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const float *col[NIT];  // col[t][k * N] = Hp[k][i0 + 64 t + lane]
#pragma unroll
    for (int t = 0; t < NIT; ++t) {
        const int i = i0 + 64 * t + lane;
        col[t] = Hp + (i < N ? i : 0);
    }
    for (int jj = w; jj < TJ; jj += kStepWaves) {
        const int j = j0 + jj;
        float c[NIT];
#pragma unroll
        for (int t = 0; t < NIT; ++t) c[t] = __builtin_inff();
        if (j < N) {
            const long long end = rows[j + 1];
            long long e = rows[j];
            // 16 entries per round: the list loads and the row reads are
            // all in flight before the first fold.
            for (; e + 16 <= end; e += 16) {
                int kq[16];
                float ka[16], v[NIT][16];
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    kq[u] = lst[2 * e + 8 * (u >> 2) + (u & 3)];
                    ka[u] = __int_as_float(lst[2 * e + 8 * (u >> 2) + 4 + (u & 3)]);
                }
#pragma unroll
                for (int u = 0; u < 16; ++u)
#pragma unroll
                    for (int t = 0; t < NIT; ++t) v[t][u] = col[t][(long long)kq[u] * N];
#pragma unroll
                for (int u = 0; u < 16; ++u)
#pragma unroll
                    for (int t = 0; t < NIT; ++t) {
                        const float x = ka[u] + v[t][u];
                        c[t] = (x < c[t]) ? x : c[t];
                    }
            }
            for (; e < end; e += 4) {
                const int4 k4 = *reinterpret_cast<const int4 *>(lst + 2 * e);
                const float4 a4 = *reinterpret_cast<const float4 *>(lst + 2 * e + 4);
                const int ks[4] = {k4.x, k4.y, k4.z, k4.w};
                const float as[4] = {a4.x, a4.y, a4.z, a4.w};
                float v[NIT][4];
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int t = 0; t < NIT; ++t) v[t][u] = col[t][(long long)ks[u] * N];
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int t = 0; t < NIT; ++t) {
                        const float x = as[u] + v[t][u];
                        c[t] = (x < c[t]) ? x : c[t];
                    }
            }
        }
#pragma unroll
        for (int t = 0; t < NIT; ++t) T[64 * t + lane][jj] = c[t];
    }
    __syncthreads();
    constexpr int kRowsPerStore = 64 / TJ;
    const int rr = lane / TJ, cc = lane % TJ;
    for (int r = w * kRowsPerStore + rr; r < 64 * NIT; r += kStepWaves * kRowsPerStore) {
        const int ii = i0 + r, j = j0 + cc;
        if (ii < N && j < N) {
            if constexpr (kStream) __builtin_nontemporal_store(T[r][cc], out + (long long)ii * N + j);
            else out[(long long)ii * N + j] = T[r][cc];
        }
    }
}

// out = in^T; skipped entirely when `skip` is non-null and set.
__global__ void transpose_kernel(const float *__restrict__ in, float *__restrict__ out, int N,
                                 const unsigned *__restrict__ skip) {
    __shared__ float t[32][33];
    if (skip && skip[0]) return;
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    for (int r = threadIdx.y; r < 32; r += blockDim.y) {
        const int row = by + r, col = bx + threadIdx.x;
        if (row < N && col < N) t[r][threadIdx.x] = in[(long long)row * N + col];
    }
    __syncthreads();
    for (int r = threadIdx.y; r < 32; r += blockDim.y) {
        const int row = bx + r, col = by + threadIdx.x;
        if (row < N && col < N) out[(long long)row * N + col] = t[threadIdx.x][r];
    }
}

// Mean candidate-list length above which the wave-per-(i, j) kernel runs.
constexpr long long kDenseDegree = 128;

// a < +inf (finite or -inf; NaN and +inf are not), tested on the bits. As
// a float compare handed straight to a ballot, `a < +inf` (fcmp one) has
// been emitted as the unordered v_cmp_neq_f32, which takes NaN.
__device__ __forceinline__ bool is_candidate(float a) {
    const unsigned b = __float_as_uint(a);
    return (b & 0x7fffffffu) < 0x7f800000u || b == 0xff800000u;
}

// Candidate counts: one wave per row k of M stores the list length rounded
// up to a multiple of 4, and in negz[k] whether the row holds a -0.0 (the
// dense step's choice of fold; the scan ORs the rows).
__global__ __launch_bounds__(256) void cand_count_kernel(const float *__restrict__ M,
                                                         long long *__restrict__ len4,
                                                         unsigned *__restrict__ negz, int N) {
    const int lane = threadIdx.x & 63;
    const int k = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (k >= N) return;
    const float *row = M + (long long)k * N;
    long long cnt = 0;
    bool nz = false;
    // Four 64-column chunks' loads in flight before their ballots.
    for (int q0 = 0; q0 < N; q0 += 256) {
        float a[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int q = q0 + 64 * u + lane;
            a[u] = q < N ? row[q] : __builtin_inff();
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            cnt += __popcll(__ballot(is_candidate(a[u])));
            nz |= __float_as_uint(a[u]) == 0x80000000u;
        }
    }
    const bool any_nz = __ballot(nz) != 0ull;
    if (lane == 0) {
        len4[k] = (cnt + 3) & ~3LL;
        negz[k] = any_nz ? 1u : 0u;
    }
}

// Exclusive scan of the padded lengths into rows[0..N] (one workgroup of
// 1024 threads, 1024 nodes per round) and the dense/sparse flag.
// dense[1]: the zero-diagonal flag that cand_fill_kernel clears (see there);
// dense[2]: whether M holds a -0.0. With `hostf` (coherent pinned host
// memory) the flags also go to hostf[1..2] and then `seq` to hostf[0], so
// the host can poll for them instead of synchronising the stream.
__global__ __launch_bounds__(1024) void cand_scan_kernel(const long long *__restrict__ len4,
                                                         const unsigned *__restrict__ negz,
                                                         long long *__restrict__ rows,
                                                         unsigned *__restrict__ dense, int N,
                                                         int allow_sparse, unsigned *hostf,
                                                         unsigned seq) {
    __shared__ long long s[1024];
    __shared__ unsigned any_nz;
    const int t = threadIdx.x;
    if (t == 0) any_nz = 0u;
    __syncthreads();
    long long carry = 0;
    for (int base = 0; base < N; base += 1024) {
        if (base + t < N && negz[base + t]) any_nz = 1u;
        s[t] = (base + t < N) ? len4[base + t] : 0;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const long long v = (t >= off) ? s[t - off] : 0;
            __syncthreads();
            s[t] += v;
            __syncthreads();
        }
        if (base + t < N) rows[base + t + 1] = carry + s[t];
        carry += s[1023];
        __syncthreads();
    }
    if (t == 0) {
        rows[0] = 0;
        const unsigned d = (!allow_sparse || carry > kDenseDegree * (long long)N) ? 1u : 0u;
        dense[0] = d;
        dense[1] = 1u;
        dense[2] = any_nz;
        if (hostf) {
            hostf[1] = d;
            hostf[2] = any_nz;
            __threadfence_system();
            __hip_atomic_store(hostf, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

// Candidate lists: one wave per node k (row k of M) writes the columns c
// with M[k][c] < +inf in ascending order (ballot + lane rank), then the
// (0, +inf) padding. Entries are stored in groups of four as 32 bytes,
// {c0, c1, c2, c3, a0, a1, a2, a3}, so entry e's column is lst[2e - e%4 ...]:
// group e/4 starts at int offset 8 * (e / 4) = 2e - 2 (e % 4).
__device__ __forceinline__ void put_entry(int *__restrict__ lst, long long e, int c, float a) {
    const long long g = (e >> 2) << 3;
    const int slot = (int)(e & 3);
    lst[g + slot] = c;
    lst[g + 4 + slot] = __float_as_int(a);
}

// With `self` set (the backtrack's lists, M = A^T), row k also records the
// index of its own entry q = k in self[k] (-1 when M[k][k] is not a
// candidate) and clears dense[1] unless M[k][k] is +0.0: the backtrack's
// gate needs every diagonal weight to be exactly +0.0.
__global__ __launch_bounds__(256) void cand_fill_kernel(const float *__restrict__ M,
                                                        const long long *__restrict__ rows,
                                                        unsigned *__restrict__ dense,
                                                        int *__restrict__ lst,
                                                        int *__restrict__ self, int N) {
    if (dense[0]) return;
    const int lane = threadIdx.x & 63;
    const int k = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (k >= N) return;
    const float *row = M + (long long)k * N;
    const long long end = rows[k + 1];
    long long at = rows[k];
    long long own = -1;  // index of the entry q = k (wave-uniform)
    // Four 64-column chunks' loads in flight before their ballots.
    for (int c0 = 0; c0 < N; c0 += 256) {
        float av[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int q = c0 + 64 * u + lane;
            av[u] = q < N ? row[q] : __builtin_inff();
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int q0 = c0 + 64 * u;
            const int q = q0 + lane;
            const float a = av[u];
            const bool ok = is_candidate(a);  // +inf past N
            const unsigned long long m = __ballot(ok);
            const int rank = __builtin_amdgcn_mbcnt_hi(
                (unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
            if (ok) put_entry(lst, at + rank, q, a);
            if (k >= q0 && k < q0 + 64 && ((m >> (k - q0)) & 1ull))
                own = at + __popcll(m & ((1ull << (k - q0)) - 1ull));
            at += __popcll(m);
        }
    }
    if (self && lane == 0) {
        self[k] = (int)own;
        if (__float_as_uint(row[k]) != 0u) atomicAnd(dense + 1, 0u);
    }
    for (long long e = at + lane; e < end; e += 64) put_entry(lst, e, 0, __builtin_inff());
}

// A node's list span [rows[k], rows[k + 1]) in one 16-byte load.
struct __attribute__((aligned(8))) Span {
    long long beg, end;
};

// Sparse backtrack: one lane per (i, j); a workgroup of 1024 lanes takes
// one i and 1024 consecutive j. Every step p all its lanes read from the row
// Hi[p][i][.], so the workgroup stages that row in LDS (kLdsRow; N <=
// kLdsRowMax) and the per-candidate reads become LDS reads. Two row buffers:
// while the lanes scan step p from one, row p-1 streams into the other by
// LDS DMA (16-byte rows; otherwise it is copied after the scan), so one
// barrier per step separates the scans and the row fetch is hidden behind
// them. Workgroups g, g+8, ... share an XCD and get the same i, so the row
// comes from that XCD's L2 for all but the first of them. Writes
// paths[i][j][0..L-1] (diagonal: i).
// Measured at N = 4000, L = 132 (7.5 A contact graph): one dword store per
// step 58.7 ms; non-temporal dword stores 109 ms; 16-byte non-temporal
// blocks 44.5 ms (plain 16-byte stores 47.1); list starts in LDS 44.0;
// two list groups in flight per iteration 37.8 ms; four need 77 VGPRs,
// which halves the resident waves: 54.8 ms. Tests of a group evaluated
// without branches (a select chain keeps the first passing slot; the
// per-entry early exits compiled to nested exec-mask branches) 35.7 ms;
// with that, three groups 36.6 ms and four 44.0 ms. Later, with the stay
// gate (24.2 -> 22.6 ms): 32-byte blocks 15.7 ms (the 16-byte stores were
// written to memory as 32-byte sectors: 16.5 GB for 8.45 GB of entries;
// 10.2 GB with 32-byte blocks); 64-byte blocks and one list group in flight
// 15.2 ms; the next row by LDS DMA instead of through registers 13.5 ms.
constexpr int kBtThreads = 1024;
constexpr int kMaxNodes = 92680;  // N * (N + 3) / 4 list groups fit an int
constexpr int kLdsRowMax = 8192;
constexpr int kRowPerLane = kLdsRowMax / kBtThreads;

typedef int nsx_v4i __attribute__((ext_vector_type(4)));

// First entry of one list group {q0..q3 | a0..a3} that passes the test, in
// slot order; -1 if none (or if `valid` is false). Slots below `smin` are
// skipped (a scan that resumes inside the group). Branch-free: the four
// tests are evaluated and a select chain keeps the lowest passing slot.
template <bool kLdsRow>
__device__ __forceinline__ int first_pass(const int4 q, const float4 a, const float *cur,
                                          const float *hrow, int j, float len, float tol,
                                          bool valid, int smin, float &af) {
    const float h0 = kLdsRow ? cur[q.x] : hrow[q.x];
    const float h1 = kLdsRow ? cur[q.y] : hrow[q.y];
    const float h2 = kLdsRow ? cur[q.z] : hrow[q.z];
    const float h3 = kLdsRow ? cur[q.w] : hrow[q.w];
    const bool p0 = valid & (smin <= 0) & (q.x != j) & ((h0 + a.x) - len < tol);
    const bool p1 = valid & (smin <= 1) & (q.y != j) & ((h1 + a.y) - len < tol);
    const bool p2 = valid & (smin <= 2) & (q.z != j) & ((h2 + a.z) - len < tol);
    const bool p3 = valid & (q.w != j) & ((h3 + a.w) - len < tol);
    int f = p3 ? q.w : -1;
    float fa = a.w;
    f = p2 ? q.z : f;
    fa = p2 ? a.z : fa;
    f = p1 ? q.y : f;
    fa = p1 ? a.y : fa;
    f = p0 ? q.x : f;
    fa = p0 ? a.x : fa;
    af = fa;
    return f;
}

// kLdsRow: Hi[p][i][.] staged in LDS (two row buffers, N <= kLdsRowMax).
// kLdsSpan: the list starts (in groups of 4 entries) of every node follow
// the row buffers in LDS, so a step's list lookup is an LDS read.
// kQuad: steps whose paths entries fill an aligned 64-byte block (16
// entries) inside the lane's own paths[i][j][0..L) are collected in
// registers and written as four 16-byte non-temporal stores, the partial
// blocks at either end as dwords (needs a 16-byte aligned paths); otherwise
// one dword store per step.
// kGroups: list groups whose loads are issued together before their tests
// (loads past the list's end are skipped); the tests keep the slot order.
// kGate (active when every A[k][k] is +0.0, dense[1]; instantiated with
// kLdsRow, the bound also reads global rows without it): the stay gate. With a zero diagonal, k = j is a candidate of every Jacobi
// fold with weight 0, so H_{p+1}[i][j] <= H_p[j][i] and H_{p+2} <= H_p entry
// by entry (entries of H_0 = A that are NaN pass no test and need no
// bound): LB[q] = min(Hi[L-1][i][q], Hi[L-2][i][q]) is below Hi[p][i][q] for
// every p. Rounding is monotone, so with
//     bound[k] = min over q < k in cand(k) of fl(LB[q] + A[q][k])
// (-inf if any such sum is NaN), fl(bound[k] - len) >= tol means that no
// q < k passes the test at any step. The lane then tests q = k alone
// (weight +0.0: len is unchanged when it passes) and, if that fails (or
// k = j, which is never taken), resumes the scan after k's own entry
// self[k]; otherwise it scans cand(k) from the start. Most steps of a path
// that has reached its last hop count are such stays, so the gate replaces
// their scan with two LDS reads. bound[] is computed per workgroup (one i)
// into LDS after the row buffers and the span table (those present).
template <bool kLdsRow, bool kLdsSpan, bool kQuad, int kGroups, bool kGate>
__global__ __launch_bounds__(kBtThreads) __attribute__((amdgpu_waves_per_eu(8))) void
backtrack_sparse_kernel(
    const float *__restrict__ H, const float *__restrict__ Hi, const long long *__restrict__ rows,
    const int *__restrict__ lst, const unsigned *__restrict__ dense,
    const int *__restrict__ self, int *__restrict__ paths, float tol, int N, int L, int nC) {
    // kLdsRow: 2 x N floats; kLdsSpan: then N + 1 ints; kGate: then N floats
    extern __shared__ float hrow_lds[];
    if (dense[0]) return;
    const bool gate = kGate && dense[1] != 0u;
    const int g = blockIdx.x;
    const int s = g >> 3;
    const int i = (g & 7) + 8 * (s / nC);
    if (i >= N) return;
    const int t0 = threadIdx.x;
    const int j = (s % nC) * kBtThreads + t0;
    const long long NN = (long long)N * N;
    const bool active = j < N && j != i;
    const long long base = ((long long)i * N + (j < N ? j : 0)) * L;
    int *pp = paths + base;
    if (j < N && j == i)
        for (int p = 0; p < L; ++p) pp[p] = i;
    float len = active ? H[(long long)i * N + j] : 0.0f;
    int k = j;
    // Row staging: 16-byte loads and LDS writes when rows are 16-byte aligned.
    const bool v4 = (N & 3) == 0 && ((uintptr_t)Hi & 15) == 0;
    int *span_lds = reinterpret_cast<int *>(hrow_lds + (kLdsRow ? 2 * N : 0));
    float *bound_lds = hrow_lds + (kLdsRow ? 2 * N : 0) + (kLdsSpan ? N + 1 : 0);
    if (kLdsSpan)
        for (int t = t0; t <= N; t += kBtThreads) span_lds[t] = (int)(rows[t] >> 2);
    if (kLdsRow) {
        const float *row = Hi + (long long)(L - 1) * NN + (long long)i * N;
        for (int t = t0; t < N; t += kBtThreads) hrow_lds[t] = row[t];
        // The gate's bound reads Hi[L-2][i][.] too; step L-1 stages the same
        // row into this buffer again.
        if (gate && L >= 2)
            for (int t = t0; t < N; t += kBtThreads) hrow_lds[N + t] = row[t - NN];
    }
    if (kLdsRow || kLdsSpan) __syncthreads();
    if (gate) {
        const float inf = __builtin_inff();
        const float *last = Hi + (long long)(L - 1) * NN + (long long)i * N;  // Hi[L-1][i][.]
        const float *r0 = kLdsRow ? hrow_lds : last;
        const float *r1 = kLdsRow ? hrow_lds + N : (L >= 2 ? last - NN : last);
        for (int kk = t0; kk < N; kk += kBtThreads) {
            float m = inf;
            const int own = self[kk];  // >= 0: the zero diagonal makes kk a candidate
            int e = kLdsSpan ? 4 * span_lds[kk] : (int)rows[kk];
            for (; e < own; e += 4) {
                const int *grp = lst + 2 * (size_t)(unsigned)e;  // group e / 4 (e % 4 == 0)
                const int4 q = *reinterpret_cast<const int4 *>(grp);
                const float4 a = *reinterpret_cast<const float4 *>(grp + 4);
                const int qs[4] = {q.x, q.y, q.z, q.w};
                const float as[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (e + u >= own) break;
                    const float lb =
                        L >= 2 ? __builtin_fminf(r0[qs[u]], r1[qs[u]]) : r0[qs[u]];
                    const float t = lb + as[u];
                    m = (t != t) ? -inf : __builtin_fminf(m, t);
                }
            }
            bound_lds[kk] = m;
        }
        __syncthreads();
    }
    nsx_v4i quad[4] = {};
    for (int p = L - 1; p >= 0; --p) {
        const float *hrow = Hi + (long long)p * NN + (long long)i * N;
        const float *cur = hrow_lds + ((L - 1 - p) & 1) * N;
        if (kLdsRow && p > 0 && v4) {
            // Hi[p-1][i][.] goes straight into the other row buffer while the
            // lanes scan this step (global_load_lds_dwordx4: 16 bytes per lane
            // land at the wave's LDS base + 16 x lane), so no registers are
            // held across the scan. The wait before the step's barrier
            // completes it.
            const float *row = hrow - NN;
            float *nxt = hrow_lds + ((L - p) & 1) * N;
            const int wv = t0 >> 6;
#pragma unroll
            for (int r = 0; r < kRowPerLane / 4; ++r) {
                const int tb = 4 * (64 * wv + r * kBtThreads);  // the wave's first float
                const int t = 4 * (t0 + r * kBtThreads);        // this lane's (t < N: t + 3 < N)
                if (tb < N && t < N)
                    __builtin_amdgcn_global_load_lds(
                        (const __attribute__((address_space(1))) void *)(row + t),
                        (__attribute__((address_space(3))) void *)(nxt + tb), 16, 0, 0);
            }
        }
        if (active) {
            int gb, ge;  // groups [gb, ge) of k's list (N <= kMaxNodes: fits int)
            if (kLdsSpan) {
                gb = span_lds[k];
                ge = span_lds[k + 1];
            } else {
                const Span sp = *reinterpret_cast<const Span *>(rows + k);
                gb = (int)(sp.beg >> 2);
                ge = (int)(sp.end >> 2);
            }
            int found = -1;
            float af = 0.0f;
            int est = 4 * gb;  // first entry the scan tests
            bool scan = true;
            if (gate && bound_lds[k] - len >= tol) {
                const float hk = kLdsRow ? cur[k] : hrow[k];
                if (k != j && (hk + 0.0f) - len < tol) {
                    found = k;  // stay: weight A[k][k] = +0.0
                    scan = false;
                } else {
                    est = self[k] + 1;
                }
            }
            for (int gi = est >> 2; scan && gi < ge && found < 0; gi += kGroups) {
                // Groups past the list's end are read at the last group of
                // the list (in bounds) and masked out.
                int4 q[kGroups];
                float4 a[kGroups];
#pragma unroll
                for (int u = 0; u < kGroups; ++u) {
                    const int gu = gi + u < ge ? gi + u : ge - 1;
                    const int *grp = lst + 8 * (size_t)(unsigned)gu;
                    q[u] = *reinterpret_cast<const int4 *>(grp);
                    a[u] = *reinterpret_cast<const float4 *>(grp + 4);
                }
#pragma unroll
                for (int u = 0; u < kGroups; ++u) {
                    float au;
                    const int fu = first_pass<kLdsRow>(q[u], a[u], cur, hrow, j, len, tol,
                                                       gi + u < ge, est - 4 * (gi + u), au);
                    const bool take = found < 0 && fu >= 0;
                    af = take ? au : af;
                    found = take ? fu : found;
                }
            }
            if (found >= 0) {
                len -= af;
                k = found;
            }
            if (kQuad) {
                // a = index of paths[i][j][p] in the int array; its aligned
                // 16-entry block [a & ~15, (a & ~15) + 15] lies inside this
                // lane's entries [base, base + L) or is written entry by entry.
                const long long a = base + p;
                const long long b0 = a & ~15LL;
                const int slot = (int)(a & 15);
                if (b0 >= base && b0 + 15 < base + L) {
                    int *blk = reinterpret_cast<int *>(&quad);
#pragma unroll
                    for (int u = 0; u < 16; ++u) blk[u] = slot == u ? k : blk[u];
                    if (slot == 0) {
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            __builtin_nontemporal_store(
                                quad[u], reinterpret_cast<nsx_v4i *>(paths + b0 + 4 * u));
                    }
                } else {
                    pp[p] = k;
                }
            } else {
                pp[p] = k;
            }
        }
        if (kLdsRow && p > 0) {
            if (v4) {
                __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this lane's row DMA landed
            } else {
                const float *row = hrow - NN;  // Hi[p-1][i][.]
                float *nxt = hrow_lds + ((L - p) & 1) * N;
                for (int t = t0; t < N; t += kBtThreads) nxt[t] = row[t];
            }
            __syncthreads();  // scans of `cur` done, `nxt` complete
        }
    }
}

// Dense backtrack (hedetniemi_gpu.cu:70-107 with the diagonal written as
// paths[i][i][*] = i): one wave per (i, j), 64 candidates per probe.
__global__ __launch_bounds__(256) void backtrack_kernel(const float *__restrict__ H,
                                                        const float *__restrict__ Hi,
                                                        const float *__restrict__ AT,
                                                        const unsigned *__restrict__ dense,
                                                        int *__restrict__ paths, float tol,
                                                        int N, int L) {
    if (!dense[0]) return;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int i = blockIdx.x;
    const long long NN = (long long)N * N;
    for (int j = wave; j < N; j += 4) {
        int *pp = paths + (long long)i * N * L + (long long)j * L;
        if (i == j) {
            for (int p = lane; p < L; p += 64) pp[p] = i;
            continue;
        }
        float len = H[(long long)i * N + j];
        int k = j;
        for (int p = L - 1; p >= 0; --p) {
            const float *hrow = Hi + (long long)p * NN + (long long)i * N;
            const float *arow = AT + (long long)k * N;  // arow[q] = A[q][k]
            int found = -1;
            for (int q0 = 0; q0 < N; q0 += 64) {
                const int q = q0 + lane;
                bool hit = false;
                if (q < N && q != j) {
                    const float t = hrow[q] + arow[q];
                    hit = (t - len) < tol;
                }
                const unsigned long long m = __ballot(hit);
                if (m) {
                    found = __builtin_amdgcn_readfirstlane(q0 + __ffsll((long long)m) - 1);
                    break;
                }
            }
            if (found >= 0) {
                len -= arow[found];
                k = found;
            }
            if (lane == 0) pp[p] = k;
        }
    }
}

__global__ void corr_to_adj_kernel(const float *__restrict__ C, float *__restrict__ A, int N) {
    const long long count = (long long)N * N;
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (; e < count; e += stride) {
        const int i = (int)(e / N), j = (int)(e % N);
        const float c = C[e];
        A[e] = (i == j) ? 0.0f : (c == 0.0f ? __builtin_inff() : 1.0f / c);
    }
}

// Per-row candidate lists of M (row r: the columns c with M[r][c] < +inf,
// ascending, padded to a multiple of 4) and the dense flag, in the given
// workspace slots. allow_sparse = false forces the dense flag.
struct Lists {
    long long *rows = nullptr;
    int *lst = nullptr;  // groups of 4 entries: 4 node ids, then 4 weights
    unsigned *dense = nullptr;  // [0] dense flag, [1] zero-diagonal flag (backtrack lists)
    int *self = nullptr;        // backtrack lists: index of node k's own entry
    bool host_dense = false;  // dense[0], when read back (step lists only)
};

int build_lists(const float *M, int N, int slot_rows, int slot_cand, bool allow_sparse,
                bool want_self, unsigned *hostf, unsigned seq, hipStream_t s, Lists &l) {
    // rows[0..N], len4[0..N), 4 flag words, self[0..N), negz[0..N)
    l.rows = (long long *)workspace(
        slot_rows, (size_t)(2 * N + 1) * sizeof(long long) + 16 + (size_t)N * 2 * sizeof(int));
    if (!l.rows) return kHipError;
    long long *len4 = l.rows + N + 1;
    l.dense = (unsigned *)(len4 + N);
    int *self = (int *)(l.dense + 4);
    unsigned *negz = (unsigned *)(self + N);
    l.self = want_self ? self : nullptr;
    const size_t cap = ((size_t)N * (size_t)(N + 3) + 3) & ~(size_t)3;
    l.lst = (int *)workspace(slot_cand, cap * (sizeof(int) + sizeof(float)));
    if (!l.lst) return kHipError;
    const int wb = (N + 3) / 4;
    hipLaunchKernelGGL(cand_count_kernel, dim3(wb), dim3(256), 0, s, M, len4, negz, N);
    NSX_HIP(hipGetLastError(), "candidate count launch");
    hipLaunchKernelGGL(cand_scan_kernel, dim3(1), dim3(1024), 0, s, len4, negz, l.rows,
                       l.dense, N, allow_sparse ? 1 : 0, hostf, seq);
    NSX_HIP(hipGetLastError(), "candidate scan launch");
    hipLaunchKernelGGL(cand_fill_kernel, dim3(wb), dim3(256), 0, s, M, l.rows, l.dense, l.lst,
                       l.self, N);
    NSX_HIP(hipGetLastError(), "candidate fill launch");
    return kOk;
}

// One Jacobi step out = step(Hp): the sparse or the dense kernel, as the
// dense flag of A's neighbour lists (read back once per call) says.
// `stream_out`: out is read only by the next step (lengths-only calls).
int minplus_step(const float *A, const float *Hp, float *out, int N, const Lists &la,
                 bool stream_out, hipStream_t s) {
    if (!la.host_dense) {
        const int nI = (N + 63) / 64;
        // 64-wide j tiles when they give every CU at least four workgroups.
        if ((long long)nI * nI >= 4LL * cu_count()) {
            const int nIg = (nI + kStepTiles - 1) / kStepTiles;
            const int groups = 8 * ((nIg + 7) / 8) * nI;
            hipLaunchKernelGGL((minplus_sparse_kernel<64, kStepTiles, false>), dim3(groups),
                               dim3(kStepWaves * 64), 0, s, Hp, la.rows, la.lst, la.dense, out, N,
                               nIg, nI);
        } else {
            const int nJ = (N + 15) / 16;
            const int nIg = (nI + kStepTiles - 1) / kStepTiles;
            const int groups = 8 * ((nIg + 7) / 8) * nJ;
            if (stream_out)
                hipLaunchKernelGGL((minplus_sparse_kernel<16, kStepTiles, true>), dim3(groups),
                                   dim3(kStepWaves * 64), 0, s, Hp, la.rows, la.lst, la.dense,
                                   out, N, nIg, nJ);
            else
                hipLaunchKernelGGL((minplus_sparse_kernel<16, kStepTiles, false>), dim3(groups),
                                   dim3(kStepWaves * 64), 0, s, Hp, la.rows, la.lst, la.dense,
                                   out, N, nIg, nJ);
        }
        NSX_HIP(hipGetLastError(), "sparse min-plus launch");
        return kOk;
    }
    const int tiles = (N + kTile - 1) / kTile;
    hipLaunchKernelGGL(minplus_kernel, dim3(tiles, tiles), dim3(256), 0, s, A, Hp, out, N,
                       la.dense + 2, (const unsigned *)la.dense);
    NSX_HIP(hipGetLastError(), "min-plus launch");
    return kOk;
}


// Shared set-up of both entry points: A's neighbour lists with the dense
// flag (dense[0]) and the negative-zero flag (dense[2]). The host needs
// dense[0] to launch only the step kernel that does the work (a launch that
// exits at once costs a few microseconds of a step of about ten at
// N = 1000): the scan kernel writes the flags to coherent pinned host memory
// and the host polls them, which returns tens of microseconds sooner than a
// stream synchronisation; if the stream completes without them (or the
// poll budget runs out) they are read back with a copy.
int step_setup(const float *A, int N, hipStream_t s, Lists &la) {
    static std::atomic<unsigned> seq{0};
    unsigned *hostf = nullptr, *devf = nullptr;
    int rc = host_flags(&hostf, &devf);
    if (rc) return rc;
    unsigned my = seq.fetch_add(1u) + 1u;
    if (my == 0u) my = seq.fetch_add(1u) + 1u;  // 0 is the block's initial value
    rc = build_lists(A, N, kWsApspRowsA, kWsApspCandA, true, false, devf, my, s, la);
    if (rc) return rc;
    volatile unsigned *hf = hostf;
    for (long long it = 0;; ++it) {
        if (__atomic_load_n(hostf, __ATOMIC_ACQUIRE) == my) {
            la.host_dense = hf[1] != 0;
            return kOk;
        }
        if ((it & 255) == 255) {
            const hipError_t e = hipStreamQuery(s);
            if (e == hipErrorNotReady && it < (1LL << 26)) continue;
            if (e != hipSuccess && e != hipErrorNotReady) NSX_HIP(e, "list set-up");
            if (__atomic_load_n(hostf, __ATOMIC_ACQUIRE) == my) continue;
            unsigned h_flags[4] = {0, 0, 0, 0};
            NSX_HIP(hipMemcpyAsync(h_flags, la.dense, sizeof(h_flags), hipMemcpyDeviceToHost, s),
                    "dense flag readback");
            NSX_HIP(hipStreamSynchronize(s), "dense flag sync");
            la.host_dense = h_flags[0] != 0;
            return kOk;
        }
    }
}

}  // namespace
}  // namespace nsx

using namespace nsx;

extern "C" {

int nsx_hedetniemi_lengths(const float *d_A, int N, int L, float *d_H, void *stream) {
    TraceRange trace("nsx_hedetniemi_lengths");
    clear_error();
    NSX_REQUIRE(d_A && d_H && N >= 1 && L >= 0, "nsx_hedetniemi_lengths: invalid arguments");
    int rc = ensure_device();
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    const size_t bytes = (size_t)N * N * sizeof(float);
    Lists la;
    rc = step_setup(d_A, N, s, la);
    if (rc) return rc;
    float *tmp = (float *)workspace(kWsApsp0, bytes);
    if (!tmp) return kHipError;
    // Step p writes d_H when L-1-p is even and tmp otherwise, so the last
    // step lands in d_H; step 0 reads A itself. If A and H overlap, A is
    // copied into the buffer step 0 does not write first.
    const bool overlap = d_A < d_H + (size_t)N * N && d_H < d_A + (size_t)N * N;
    const float *in = d_A;
    if (overlap || L == 0) {
        float *h0 = (L % 2 == 0) ? d_H : tmp;
        if (h0 != d_A)
            NSX_HIP(hipMemcpyAsync(h0, d_A, bytes, hipMemcpyDeviceToDevice, s), "H0 = A");
        in = h0;
    }
    for (int p = 0; p < L; ++p) {
        float *out = ((L - 1 - p) % 2 == 0) ? d_H : tmp;
        rc = minplus_step(d_A, in, out, N, la, true, s);
        if (rc) return rc;
        in = out;
    }
    return kOk;
}

int nsx_hedetniemi(const float *d_A, int N, int L, float tol, float *d_H, float *d_Hi,
                   int *d_paths, void *stream) {
    TraceRange trace("nsx_hedetniemi");
    clear_error();
    NSX_REQUIRE(d_A && d_H && N >= 1 && L >= 1, "nsx_hedetniemi: invalid arguments");
    NSX_REQUIRE(!d_paths || d_Hi, "nsx_hedetniemi: paths need the Hi snapshot buffer");
    NSX_REQUIRE(!d_paths || N <= kMaxNodes, "nsx_hedetniemi: paths need N <= 92680");
    if (!d_Hi) return nsx_hedetniemi_lengths(d_A, N, L, d_H, stream);
    int rc = ensure_device();
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    const long long NN = (long long)N * N;
    const size_t bytes = (size_t)NN * sizeof(float);
    Lists la;
    rc = step_setup(d_A, N, s, la);
    if (rc) return rc;
    NSX_HIP(hipMemcpyAsync(d_Hi, d_A, bytes, hipMemcpyDeviceToDevice, s), "Hi[0] = A");
    for (int p = 0; p < L; ++p) {
        float *out = (p + 1 < L) ? d_Hi + (size_t)(p + 1) * NN : d_H;
        rc = minplus_step(d_A, d_Hi + (size_t)p * NN, out, N, la, false, s);
        if (rc) return rc;
    }
    if (!d_paths) return kOk;
    // Backtrack candidates of node k are the q with A[q][k] < +inf: the
    // per-row lists of A^T.
    float *AT = (float *)workspace(kWsApsp0, bytes);
    if (!AT) return kHipError;
    const int tb = (N + 31) / 32;
    hipLaunchKernelGGL(transpose_kernel, dim3(tb, tb), dim3(32, 8), 0, s, d_A, AT, N,
                       (const unsigned *)nullptr);
    NSX_HIP(hipGetLastError(), "transpose launch");
    Lists lt;
    rc = build_lists(AT, N, kWsApspRows, kWsApspCand, true, true, nullptr, 0u, s, lt);
    if (rc) return rc;
    const int nC = (N + kBtThreads - 1) / kBtThreads;
    const int bt_groups = 8 * ((N + 7) / 8) * nC;
    const size_t row_lds = 2 * (size_t)N * sizeof(float);
    const size_t span_lds = row_lds + (size_t)(N + 1) * sizeof(int);
    const size_t bound_bytes = (size_t)N * sizeof(float);
    const bool quad = ((uintptr_t)d_paths & 15) == 0;
#define NSX_BT(R, S, W, G, T, lds)                                                       \
    hipLaunchKernelGGL((backtrack_sparse_kernel<R, S, W, G, T>), dim3(bt_groups),        \
                       dim3(kBtThreads), lds, s, d_H, d_Hi, lt.rows, lt.lst, lt.dense, \
                       lt.self, d_paths, tol, N, L, nC)
    // LDS: up to 80 KB keeps two workgroups per CU.
    if (N > kLdsRowMax && quad) NSX_BT(false, false, true, 2, false, 0);
    else if (N > kLdsRowMax) NSX_BT(false, false, false, 2, false, 0);
    else if (!quad) NSX_BT(true, false, false, 2, false, row_lds);
    else if (span_lds + bound_bytes <= 80 * 1024)
        NSX_BT(true, true, true, 1, true, span_lds + bound_bytes);
    else NSX_BT(true, false, true, 2, true, row_lds + bound_bytes);
#undef NSX_BT
    NSX_HIP(hipGetLastError(), "sparse backtrack launch");
    hipLaunchKernelGGL(backtrack_kernel, dim3(N), dim3(256), 0, s, d_H, d_Hi, AT, lt.dense,
                       d_paths, tol, N, L);
    NSX_HIP(hipGetLastError(), "dense backtrack launch");
    return kOk;
}

int nsx_corr_to_adj(const float *d_C, int N, float *d_A, void *stream) {
    TraceRange trace("nsx_corr_to_adj");
    clear_error();
    NSX_REQUIRE(d_C && d_A && N >= 0, "nsx_corr_to_adj: invalid arguments");
    if (N == 0) return kOk;
    int rc = ensure_device();
    if (rc) return rc;
    const long long count = (long long)N * N;
    const int blocks = (int)std::min<long long>((count + 255) / 256, 8192);
    hipLaunchKernelGGL(corr_to_adj_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, d_C,
                       d_A, N);
    NSX_HIP(hipGetLastError(), "adjacency launch");
    return kOk;
}

}  // extern "C"
