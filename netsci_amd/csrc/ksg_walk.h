// KSG through per-node sorted neighbour lists ("list path") for gfx950.
//
// Same results as ksg_kernel.h (and the reference CPU path,
// netcalc/src/mutual_information_cpu.cpp:6-121) bit for bit, with far fewer
// distance evaluations. Two facts make it possible:
//
//   * a node's marginal distances D_m(i, j) = max_d |x_d(i) - x_d(j)| are the
//     same for every pair the node takes part in, so each node's frames can
//     be sorted once per call, per query frame i, by (D_m(i, j), j) — the
//     sorted frame list ord_m[i][.] — together with the upper-bound rank
//     rank_m[i][j] = #{j' : D_m(i, j') <= D_m(i, j)};
//   * the joint distance is z(i, j) = max(D_a(i, j), D_b(i, j)), so walking
//     ord_a[i] and ord_b[i] in step (one entry of each per step) and keeping
//     the k+1 smallest z of the distinct frames seen so far, every frame not
//     yet seen in either list has z >= max(last D_a walked, last D_b walked).
//     Once that bound exceeds the current (k+1)-th smallest z, the k+1
//     smallest are final (the threshold-algorithm stopping rule).
//
// The marginal counts then need no second pass: with eps_x = D_a(i, j*) for
// the selected frame j* of largest D_a, n_x + 1 = #{j : D_a(i, j) <= eps_x}
// = rank_a[i][j*] (self included, as the reference counts it before
// subtracting one), and likewise n_y.
//
// A frame seen in one list is recognised in the other without memory: ord_m
// is sorted by the pair (D_m, j), so frame j has already been walked in list
// m exactly when (D_m(i, j), j) <= the last walked (D_m, j) in that order.
//
// Exact ties at the k-th / (k+1)-th joint distance take the libstdc++
// std::sort-order selection of sort_select.h, as in ksg_kernel.h.
//
// Mapping:
//   * list build: one wavefront per (node, query frame) row: the n keys are
//     packed as (float bits << 32 | j) and bitonic-sorted in registers
//     (32 per lane, cross-lane stages by ds_bpermute); ranks come from a
//     suffix minimum over the ends of equal-key runs;
//   * walk: a workgroup owns one node pair at a time with both nodes'
//     coordinates in LDS, one 2d-float record per frame (gathered at random
//     frames with 8-byte loads); its lanes take query
//     frames from an LDS counter, each lane walking its own query in bodies
//     of 8 steps (queries start at a body boundary, so every body reads one
//     aligned 8-entry block per list at fixed register positions, and the
//     block two ahead is requested one body before it is needed). Lanes
//     whose walk ended wait until `batch` of them have (or no lane walks),
//     then the wave processes them together (eps, ties, rank lookups, whose
//     counts are stored one round later) and hands out new queries, so that
//     per-query work runs densely;
//   * the workgroup's counts go to a per-workgroup buffer and one wave sums
//     psi[n_x] + psi[n_y] in frame order, exactly as the reference's epilogue.
// Coordinates must be finite (the host checks and otherwise takes
// ksg_kernel.h, whose NaN behaviour is documented there).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ksg_kernel.h"
#include "sort_select.h"

namespace nsx {
namespace ksgw {

constexpr int kWave = 64;
constexpr int kMaxFrames = 2048;  // 32 sort keys per lane
constexpr int kBuildWaves = 4;
constexpr int kMaxWalkWaves = 8;

struct BuildArgs {
    const float *X;
    long long ld;       // row stride of X (floats)
    long long dstride;  // stride between coordinate dimensions (floats)
    int n;              // frames
    long long rs;       // row stride of the tables (entries, a multiple of 8)
    const int *node_of; // table slot -> node
    int groups;         // workgroups per node
    uint16_t *ord;      // [slot][i][rs] frames sorted by (D, j)
    uint16_t *rank;     // [slot][i][rs] upper-bound ranks, indexed by frame
    unsigned *bad;      // set to 1 when a coordinate is not finite
};

struct WalkArgs {
    const float *X;
    long long ld, dstride;
    const int *ab;
    long long P;
    int n, k;
    const float *psi;
    float inv_k;
    float *mi_out;
    int *counts_out;  // nullable: [p][frame][2]
    const uint16_t *ord, *rank;
    long long rs;
    const int *slot_of;       // node -> table slot
    KeyIdx *scratch;          // per wave: n records (tie rows)
    unsigned *cnt;            // per workgroup: n packed counts n_x | n_y << 16
    unsigned long long *stats;  // [0] tie frames, [1] list steps (entries per list)
    int wpb;
    int batch;
};

using BuildFn = void (*)(BuildArgs);
using WalkFn = void (*)(WalkArgs);

// LDS bytes of the two kernels.
inline size_t build_lds_bytes(int d, int n, long long rs) {
    return (((size_t)d * n * 4 + 15) & ~(size_t)15) + (size_t)kBuildWaves * rs * 2;
}
inline size_t walk_lds_bytes(int d, int n) { return (size_t)2 * d * n * 4 + 16; }

// Sort keys per lane for n frames (8, 16 or 32).
inline int build_width(int n) { return n <= 512 ? 8 : n <= 1024 ? 16 : 32; }

// Kernel instances, one translation unit per coordinate dimension.
BuildFn build_d1(int width);
BuildFn build_d2(int width);
BuildFn build_d3(int width);
WalkFn walk_d1(int k);
WalkFn walk_d2(int k);
WalkFn walk_d3(int k);

#if defined(__HIPCC__)

using ksg::marg;

__device__ __forceinline__ unsigned long long shfl_xor_u64(unsigned long long v, int m) {
    const unsigned lo = (unsigned)__shfl_xor((int)(unsigned)v, m);
    const unsigned hi = (unsigned)__shfl_xor((int)(unsigned)(v >> 32), m);
    return ((unsigned long long)hi << 32) | lo;
}

// Ascending bitonic sort of 64 * R distinct keys; element e = lane * R + r.
template <int R>
__device__ __forceinline__ void bitonic_sort(unsigned long long (&v)[R], int lane) {
#pragma unroll
    for (int size = 2; size <= kWave * R; size <<= 1) {
#pragma unroll
        for (int stride = size / 2; stride > 0; stride >>= 1) {
            if (stride >= R) {
                const int lx = stride / R;
                const bool lower = (lane & lx) == 0;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const bool asc = ((lane * R + r) & size) == 0;
                    const unsigned long long o = shfl_xor_u64(v[r], lx);
                    const bool keep = (v[r] < o) == (lower == asc);
                    v[r] = keep ? v[r] : o;
                }
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (r & stride) continue;
                    const int r2 = r | stride;
                    const bool asc = ((lane * R + r) & size) == 0;
                    const unsigned long long x = v[r], y = v[r2];
                    const bool sw = asc ? (x > y) : (x < y);
                    v[r] = sw ? y : x;
                    v[r2] = sw ? x : y;
                }
            }
        }
    }
}

// One workgroup per (node slot, row group); each wave sorts whole rows.
template <int D, int R>
__global__ __launch_bounds__(kWave *kBuildWaves) void list_build_kernel(BuildArgs a) {
    extern __shared__ __attribute__((aligned(16))) float lds[];
    const int n = a.n;
    const int slot = blockIdx.x / a.groups;
    const int g = blockIdx.x - slot * a.groups;
    const float *x = a.X + (long long)a.node_of[slot] * a.ld;
    const int lane = threadIdx.x & (kWave - 1);
    const int wib = threadIdx.x / kWave;
    bool bad = false;
    for (int t = threadIdx.x; t < D * n; t += blockDim.x) {
        const int d = t / n, f = t - d * n;
        const float v = x[d * a.dstride + f];
        bad |= !__builtin_isfinite(v);
        lds[t] = v;
    }
    if (__ballot(bad) && lane == 0) atomicOr(a.bad, 1u);
    __syncthreads();
    uint16_t *st = reinterpret_cast<uint16_t *>(
                       reinterpret_cast<char *>(lds) + ((D * n * 4 + 15) & ~15)) +
                   (long long)wib * a.rs;
    for (int i = g * kBuildWaves + wib; i < n; i += a.groups * kBuildWaves) {
        float q[D];
#pragma unroll
        for (int d = 0; d < D; ++d) q[d] = lds[d * n + i];
        unsigned long long v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int e = lane * R + r;
            if (e < n) {
                float c[D];
#pragma unroll
                for (int d = 0; d < D; ++d) c[d] = lds[d * n + e];
                const float key = marg<D>(q, c);
                v[r] = ((unsigned long long)__float_as_uint(key) << 32) | (unsigned)e;
            } else {
                v[r] = ~0ull;
            }
        }
        bitonic_sort<R>(v, lane);

        // ord row: this lane's R consecutive positions, 16 bytes at a time.
        uint16_t *orow = a.ord + ((long long)slot * n + i) * a.rs;
#pragma unroll
        for (int r0 = 0; r0 < R; r0 += 8) {
            const long long e0 = (long long)lane * R + r0;
            if (e0 < a.rs) {
                uint4 w;
                w.x = (unsigned)(v[r0 + 0] & 0xffff) | ((unsigned)(v[r0 + 1] & 0xffff) << 16);
                w.y = (unsigned)(v[r0 + 2] & 0xffff) | ((unsigned)(v[r0 + 3] & 0xffff) << 16);
                w.z = (unsigned)(v[r0 + 4] & 0xffff) | ((unsigned)(v[r0 + 5] & 0xffff) << 16);
                w.w = (unsigned)(v[r0 + 6] & 0xffff) | ((unsigned)(v[r0 + 7] & 0xffff) << 16);
                *reinterpret_cast<uint4 *>(orow + e0) = w;
            }
        }

        // Upper-bound rank of position e: the first end of an equal-key run
        // at or after e (a run ends where the next key differs; the padding
        // keys 0xffffffff differ from every finite key).
        constexpr unsigned kBig = 0x7fffffffu;
        const unsigned next0 = (unsigned)__shfl_down((int)(unsigned)(v[0] >> 32), 1);
        unsigned s[R];
        unsigned m = kBig;
#pragma unroll
        for (int r = R - 1; r >= 0; --r) {
            const unsigned kr = (unsigned)(v[r] >> 32);
            const unsigned kn = (r == R - 1) ? (lane == kWave - 1 ? 0xffffffffu : next0)
                                             : (unsigned)(v[r + 1] >> 32);
            const unsigned end = kn != kr ? (unsigned)(lane * R + r + 1) : kBig;
            m = min(m, end);
            s[r] = m;
        }
        unsigned t = m;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const unsigned o = (unsigned)__shfl_down((int)t, off);
            t = lane + off < kWave ? min(t, o) : t;
        }
        unsigned after = (unsigned)__shfl_down((int)t, 1);
        after = lane == kWave - 1 ? kBig : after;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int e = lane * R + r;
            if (e < n) st[(unsigned)(v[r] & 0xffff)] = (uint16_t)min(s[r], after);
        }
        ksg::wave_sync();
        uint16_t *rrow = a.rank + ((long long)slot * n + i) * a.rs;
        for (long long c = (long long)lane * 8; c < a.rs; c += kWave * 8)
            *reinterpret_cast<uint4 *>(rrow + c) = *reinterpret_cast<const uint4 *>(st + c);
        ksg::wave_sync();
    }
}

// Insert (c, j) into the ascending h[0..K] without branches: every slot
// is two compares, a med3 and two selects. c >= h[K] (+inf for a rejected
// candidate) leaves everything unchanged.
template <int K>
__device__ __forceinline__ void insert_flat(float (&h)[K + 1], int (&id)[K + 1], float c, int j) {
#pragma unroll
    for (int m = K; m >= 1; --m) {
        const int t = c < h[m] ? j : id[m];
        id[m] = c < h[m - 1] ? id[m - 1] : t;
        h[m] = __builtin_amdgcn_fmed3f(h[m - 1], c, h[m]);
    }
    id[0] = c < h[0] ? j : id[0];
    h[0] = fminf(c, h[0]);
}

// (D, j) as one signed 64-bit value ordered like the list's (D, j) order:
// D >= 0 is finite, so its bits compare like the float; -1 sorts first.
__device__ __forceinline__ long long lex(float key, int j) {
    return (long long)(((unsigned long long)__float_as_uint(key) << 32) | (unsigned)j);
}

__device__ __forceinline__ float lex_key(long long v) {
    return __uint_as_float((unsigned)((unsigned long long)v >> 32));
}

// Frame j of the pair's LDS record array: 2D floats per frame (node a's D
// coordinates, then node b's), read as D 8-byte LDS loads.
template <int D>
__device__ __forceinline__ void fetch(const float *rec, int j, float (&va)[D], float (&vb)[D]) {
    const float2 *r = reinterpret_cast<const float2 *>(rec + 2 * D * j);
    float v[2 * D];
#pragma unroll
    for (int t = 0; t < D; ++t) {
        const float2 w = r[t];
        v[2 * t] = w.x;
        v[2 * t + 1] = w.y;
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
        va[d] = v[d];
        vb[d] = v[D + d];
    }
}

__device__ __forceinline__ uint4 load_block(const uint16_t *p) {
    return *reinterpret_cast<const uint4 *>(p);
}

// Entry q (0..7) of an 8-entry block (q is a compile-time index in the walk).
__device__ __forceinline__ int block_entry(const uint4 &b, int q) {
    const unsigned w = (q & 4) ? ((q & 2) ? b.w : b.z) : ((q & 2) ? b.y : b.x);
    return (int)((w >> ((q & 1) * 16)) & 0xffffu);
}

template <int D, int K>
__global__ __launch_bounds__(kWave *kMaxWalkWaves) void walk_kernel(WalkArgs a) {
    extern __shared__ __attribute__((aligned(16))) float lds[];
    const int n = a.n;
    const float *rec = lds;  // [n][2D]: frame j's node-a then node-b coordinates
    int *counter = reinterpret_cast<int *>(lds + 2 * D * n);
    const int lane = threadIdx.x & (kWave - 1);
    const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    KeyIdx *scratch = a.scratch + ((long long)blockIdx.x * a.wpb + wib) * n;
    unsigned *cnt = a.cnt + (long long)blockIdx.x * n;
    const long long rs = a.rs;
    const long long table = (long long)n * rs;
    unsigned long long ties = 0, steps = 0;

    for (long long p = blockIdx.x; p < a.P; p += gridDim.x) {
        const int na = a.ab[2 * p], nb = a.ab[2 * p + 1];
        const float *xa = a.X + (long long)na * a.ld;
        const float *xb = a.X + (long long)nb * a.ld;
        const uint16_t *ordA = a.ord + (long long)a.slot_of[na] * table;
        const uint16_t *ordB = a.ord + (long long)a.slot_of[nb] * table;
        const uint16_t *rankA = a.rank + (long long)a.slot_of[na] * table;
        const uint16_t *rankB = a.rank + (long long)a.slot_of[nb] * table;
        __syncthreads();  // the previous pair's readers are done
        for (int t = threadIdx.x; t < D * n; t += blockDim.x) {
            const int d = t / n, f = t - d * n;
            lds[2 * D * f + d] = xa[d * a.dstride + f];
            lds[2 * D * f + D + d] = xb[d * a.dstride + f];
        }
        if (threadIdx.x == 0) *counter = 0;
        __syncthreads();

        // Lane state: 0 wants a query, 1 walks, 2 no queries left.
        int state = 0;
        bool res = false;  // a finished walk waits for its counts
        int i = 0, pos = 0;
        // Last walked entry of each list as lex(D, j); -1: none yet (its key
        // reads as NaN, so the stop test below is false).
        long long lastA = -1, lastB = -1;
        float qa[D], qb[D], h[K + 1];
        int id[K + 1];
        // List blocks: cur holds entries [pos, pos + 8) of each list, nxt
        // the 8 after them (pos is a multiple of 8 at every body start).
        uint4 curA = {0, 0, 0, 0}, curB = {0, 0, 0, 0}, nxtA = {0, 0, 0, 0},
              nxtB = {0, 0, 0, 0};
        const uint16_t *rowA = ordA, *rowB = ordB;
        // Counts of the previous refill round, stored one round later so
        // that the rank lookups' latency is not waited for.
        bool pend = false;
        int pend_i = 0;
        unsigned pend_x = 1u, pend_y = 1u;
#pragma unroll
        for (int d = 0; d < D; ++d) qa[d] = qb[d] = 0.0f;
#pragma unroll
        for (int m = 0; m <= K; ++m) {
            h[m] = __builtin_inff();
            id[m] = 0;
        }

        while (true) {
            const unsigned long long want = __ballot(state == 0);
            const unsigned long long walking = __ballot(state == 1);
            if (want && (__popcll(want) >= (unsigned)a.batch || walking == 0)) {
                // (0) counts looked up in the previous round.
                if (pend) cnt[pend_i] = (pend_x - 1u) | ((pend_y - 1u) << 16);
                pend = false;
                // (1) eps of the finished walks.
                const bool tie = res && (h[K - 1] == h[K]);
                int jx = 0, jy = 0;
                if (res && !tie) {
                    float ex = -1.0f, ey = -1.0f;
#pragma unroll
                    for (int m = 0; m < K; ++m) {
                        float va[D], vb[D];
                        fetch<D>(rec, id[m], va, vb);
                        const float da = marg<D>(qa, va), db = marg<D>(qb, vb);
                        jx = da > ex ? id[m] : jx;
                        ex = fmaxf(ex, da);
                        jy = db > ey ? id[m] : jy;
                        ey = fmaxf(ey, db);
                    }
                }
                unsigned long long tm = __ballot(tie);
                ties += (unsigned)__popcll(tm);
                while (tm) {
                    // The reference order: std::sort of the full joint row,
                    // eps over positions 1..k (mutual_information_cpu.cpp:48-77).
                    const int t = __builtin_amdgcn_readfirstlane(__ffsll((long long)tm) - 1);
                    float tqa[D], tqb[D];
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        tqa[d] = __shfl(qa[d], t);
                        tqb[d] = __shfl(qb[d], t);
                    }
                    for (int j = lane; j < n; j += kWave) {
                        float va[D], vb[D];
                        fetch<D>(rec, j, va, vb);
                        KeyIdx r;
                        r.key = fmaxf(marg<D>(tqa, va), marg<D>(tqb, vb));
                        r.idx = j;
                        scratch[j] = r;
                    }
                    ksg::wave_sync_global();
                    int rjx = 0, rjy = 0;
                    if (lane == 0) {
                        sort_prefix(scratch, n, K + 1);
                        float ex = -1.0f, ey = -1.0f;
                        for (int m = 1; m <= K; ++m) {
                            const int j = scratch[m].idx;
                            float va[D], vb[D];
                            fetch<D>(rec, j, va, vb);
                            const float da = marg<D>(tqa, va), db = marg<D>(tqb, vb);
                            rjx = da > ex ? j : rjx;
                            ex = fmaxf(ex, da);
                            rjy = db > ey ? j : rjy;
                            ey = fmaxf(ey, db);
                        }
                    }
                    rjx = __shfl(rjx, 0);
                    rjy = __shfl(rjy, 0);
                    jx = lane == t ? rjx : jx;
                    jy = lane == t ? rjy : jy;
                    ksg::wave_sync_global();
                    tm &= tm - 1;
                }
                if (res) {
                    pend_x = rankA[(long long)i * rs + jx];
                    pend_y = rankB[(long long)i * rs + jy];
                    pend_i = i;
                    pend = true;
                    steps += (unsigned)pos;
                }
                res = false;
                // (2) new queries for every lane that wants one.
                const int leader = __builtin_amdgcn_readfirstlane(__ffsll((long long)want) - 1);
                int base = 0;
                if (lane == leader) base = atomicAdd(counter, __popcll(want));
                base = __shfl(base, leader);
                const int below = __builtin_amdgcn_mbcnt_hi(
                    (unsigned)(want >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)want, 0u));
                if (state == 0) {
                    const int qi = base + below;
                    if (qi < n) {
                        i = qi;
                        fetch<D>(rec, i, qa, qb);
#pragma unroll
                        for (int m = 0; m <= K; ++m) {
                            h[m] = __builtin_inff();
                            id[m] = 0;
                        }
                        pos = 0;
                        lastA = lastB = -1;
                        rowA = ordA + (long long)i * rs;
                        rowB = ordB + (long long)i * rs;
                        curA = load_block(rowA);
                        curB = load_block(rowB);
                        if (rs > 8) {
                            nxtA = load_block(rowA + 8);
                            nxtB = load_block(rowB + 8);
                        }
                        state = 1;
                    } else {
                        state = 2;
                    }
                }
            }
            if (__ballot(state == 1) == 0) break;
            // A body: 8 steps, one entry of each list per step. The block
            // two ahead is requested first and used one body later.
            uint4 tA = nxtA, tB = nxtB;
            if (state == 1 && pos + 16 < rs) {
                tA = load_block(rowA + pos + 16);
                tB = load_block(rowB + pos + 16);
            }
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                if (state == 1) {
                    const int ja = block_entry(curA, s), jb = block_entry(curB, s);
                    ++pos;
                    float ua[D], ub[D], va[D], vb[D];
                    fetch<D>(rec, ja, ua, ub);
                    fetch<D>(rec, jb, va, vb);
                    const float daA = marg<D>(qa, ua), dbA = marg<D>(qb, ub);
                    const float daB = marg<D>(qa, va), dbB = marg<D>(qb, vb);
                    // Entry ja of list a: already seen if list b has walked it.
                    const bool seenA = lex(dbA, ja) <= lastB;
                    lastA = lex(daA, ja);
                    const float zA = (seenA || ja == i) ? __builtin_inff() : fmaxf(daA, dbA);
                    insert_flat<K>(h, id, zA, ja);
                    // Entry jb of list b: already seen if list a has walked
                    // it (including ja just now).
                    const bool seenB = lex(daB, jb) <= lastA;
                    lastB = lex(dbB, jb);
                    const float zB = (seenB || jb == i) ? __builtin_inff() : fmaxf(daB, dbB);
                    insert_flat<K>(h, id, zB, jb);
                    if (fmaxf(lex_key(lastA), lex_key(lastB)) > h[K] || pos >= n) {
                        state = 0;
                        res = true;
                    }
                }
            }
            if (state == 1) {
                curA = nxtA;
                curB = nxtB;
                nxtA = tA;
                nxtB = tB;
            }
        }
        if (pend) cnt[pend_i] = (pend_x - 1u) | ((pend_y - 1u) << 16);
        __syncthreads();  // every count of this pair is written
        if (wib == 0) {
            // psi terms summed sequentially in frame order
            // (mutual_information_cpu.cpp:110-118).
            float acc = 0.0f;
            for (int f0 = 0; f0 < n; f0 += kWave) {
                const int f = f0 + lane;
                float term = 0.0f;
                if (f < n) {
                    const unsigned c = cnt[f];
                    const int nx = (int)(c & 0xffffu), ny = (int)(c >> 16);
                    term = a.psi[nx] + a.psi[ny];
                    if (a.counts_out) {
                        a.counts_out[(p * n + f) * 2 + 0] = nx;
                        a.counts_out[(p * n + f) * 2 + 1] = ny;
                    }
                }
                const int lim = min(kWave, n - f0);
                for (int l = 0; l < lim; ++l)
                    acc += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(term), l));
            }
            const float avg = acc / (float)n;
            const float mi = a.psi[n] + a.psi[a.k] - a.inv_k - avg;
            if (lane == 0) a.mi_out[p] = mi;
        }
    }
    // Per-wave totals.
    unsigned long long s = steps;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) {
        if (ties) atomicAdd(a.stats, ties);
        atomicAdd(a.stats + 1, s);
    }
}

template <int D, int K>
struct WalkInstances {
    static void fill(WalkFn *t) {
        t[K] = walk_kernel<D, K>;
        WalkInstances<D, K - 1>::fill(t);
    }
};
template <int D>
struct WalkInstances<D, 0> {
    static void fill(WalkFn *) {}
};

template <int D>
struct WalkTable {
    WalkFn t[ksg::kMaxK + 1] = {};
    WalkTable() { WalkInstances<D, ksg::kMaxK>::fill(t); }
};

template <int D>
WalkFn walk_lookup(int k) {
    static const WalkTable<D> tab;
    return (k >= 1 && k <= ksg::kMaxK) ? tab.t[k] : nullptr;
}

template <int D>
BuildFn build_lookup(int width) {
    switch (width) {
        case 8: return list_build_kernel<D, 8>;
        case 16: return list_build_kernel<D, 16>;
        case 32: return list_build_kernel<D, 32>;
        default: return nullptr;
    }
}

#endif  // __HIPCC__

}  // namespace ksgw
}  // namespace nsx
