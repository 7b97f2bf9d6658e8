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
//   * the joint distance is z(i, j) = max(D_a(i, j), D_b(i, j)) >= D_a(i, j),
//     so walking ord_a[i] alone and keeping the k+1 smallest z of the frames
//     walked, every frame not walked yet has z >= the last D_a walked. Once
//     that bound exceeds the current (k+1)-th smallest z, the k+1 smallest
//     are final (the threshold-algorithm stopping rule, with one list). The
//     k+1 nearest lie within either list's walked prefix, so either list
//     gives the same result; the walk takes the one whose keys grow faster
//     around the query (the larger "prior": the marginal distance at list
//     position 64, which the table build records per node and frame); on a
//     C-alpha trajectory (N = 1000, n = 2000) that list needs no more
//     8-entry blocks than the other for 98% of the queries.
//
// The marginal counts then need no second pass: with eps_x = D_a(i, j*) for
// the selected frame j* of largest D_a, n_x + 1 = #{j : D_a(i, j) <= eps_x}
// = rank_a[i][j*] (self included, as the reference counts it before
// subtracting one), and likewise n_y.
//
// Exact ties at the k-th / (k+1)-th joint distance take the libstdc++
// std::sort-order selection of sort_select.h, as in ksg_kernel.h.
//
// Mapping:
//   * list build: one wavefront per (node, query frame) row (up to four
//     for rows of more than 2048 frames): the n keys are packed as 32-bit
//     (distance code << id bits | j) and bitonic-sorted in registers, an
//     odd-even pass on exact keys restores (D, j) order inside runs of
//     equal codes; ranks come from a suffix minimum over the ends of
//     equal-key runs, and the key at position 64 is the frame's prior;
//   * walk: a workgroup owns a few node pairs at a time with both nodes'
//     coordinates in LDS, one 2d-float record per frame (gathered at random
//     frames with 8-byte loads) and one list-choice bit per frame; its lanes
//     take query frames from an LDS counter, each lane walking its own query
//     in bodies of 8 steps of its one list (queries start at a body
//     boundary, so every body reads one aligned 8-entry block at fixed
//     register positions, and the block two ahead is requested one body
//     before it is needed). Lanes
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
constexpr int kMaxFrames = 8192;  // 32 sort keys per lane, up to 4 waves per row
constexpr int kBuildWaves = 4;
constexpr int kMaxWalkWaves = 16;
constexpr int kTieMaxWaves = 8;  // waves per workgroup of the exact-tie kernels

// Waves per walk workgroup a k allows: with k+1 packed 32-bit heap
// entries the walk fits 128 VGPRs (16 waves, 4 per SIMD) for every k up to
// ksg::kMaxK (97 at k = 16).
constexpr int walk_max_waves(int) { return 16; }

struct BuildArgs {
    const float *X;
    long long ld;       // row stride of X (floats)
    long long dstride;  // stride between coordinate dimensions (floats)
    int n;              // frames
    long long rs;       // row stride of the rank tables (entries, a multiple of 8)
    long long ors;      // row stride of the ord tables: the entries kept, ord_stride(n)
    const int *node_of; // table slot -> node
    int groups;         // workgroups per node (rows of <= 2048 frames)
    int nodes;          // nodes (table slots) of the launch
    uint16_t *ord;      // [slot][i][rs] frames sorted by (D, j)
    uint16_t *rank;     // [slot][i][rs] upper-bound ranks, indexed by frame
    const float *span;  // per node: max over dimensions of (max - min) of the
                        // coordinate, a bound of every marginal distance
    const unsigned *abort;  // nullable: return at once when *abort != 0 (a
                            // non-finite coordinate was found beforehand)
    float *prior;       // [slot][table_rows(n)]: the marginal distance at list
                        // position prior_pos(n), the walk's choice of list
};

// Packed counts of a tied query the tie list had no room for: the counts
// of a real query are below n - 1 <= 8191 each.
constexpr unsigned kTieMark = 0xffffffffu;

struct WalkArgs {
    const float *X;
    long long ld, dstride;
    const int *ab;
    long long p0;  // first pair of this launch
    long long P;   // pairs of this launch
    int n, k;
    const uint16_t *ord, *rank;
    long long rs;               // rank row stride
    long long ors;              // ord row stride (entries kept per list, ord_stride(n))
    const int *slot_of;         // node -> table slot
    unsigned *cnt;              // [P][n] packed counts n_x | n_y << 16
    unsigned long long *stats;  // [0] tie frames, [1] list entries walked (both lists)
    unsigned *tie_list;         // tied queries (local pair * n + frame), resolved later
    unsigned *tie_count;        // entries appended (may exceed tie_cap; the rest are
                                // marked kTieMark in cnt)
    unsigned tie_cap;
    unsigned *exact_list;       // tie-list entries whose k-th and (k+1)-th joint
    unsigned *exact_count;      // distances are equal (tie_cap capacity)
    int nbuf;                   // pair buffers per workgroup (LDS ring)
    int batch;                  // lanes that wait before a refill round
    int checkpoint;             // loop passes between workgroup checkpoints
    const float *span;          // per node: max over d of (max - min) of coordinate d
    int id_bits;                // frame bits of a heap entry (2^id_bits >= n)
    int code_shift;             // float-bit units per distance code step: 2^code_shift
    const unsigned *abort;      // nullable: the kernels return at once when *abort != 0
                                // (a non-finite coordinate; the all-frames kernel runs)
    const float *prior;         // [slot][table_rows(n)] (BuildArgs::prior)
};

// Frame-order psi sums of the packed counts (one lane per pair). Local pair
// p of the launch is pair pbase + p of the pair list, whose output index is
// perm[pbase + p] (perm nullable: the index itself).
struct SumArgs {
    const unsigned *cnt;
    long long P;
    int n, k;
    const float *psi;
    float inv_k;
    float *mi_out;    // [pairs]
    int *counts_out;  // nullable: [pairs][n][2]
    const int *perm;
    long long pbase;
    const unsigned *abort;  // nullable, as in WalkArgs
};

using BuildFn = void (*)(BuildArgs);
using WalkFn = void (*)(WalkArgs);
using TieFn = void (*)(WalkArgs);

// Sort elements per wave-lane for n frames: 8, 16 or 32 for one wave per
// row; 64, 96 and 128 mean 2, 3 and 4 waves per row (32 each).
inline int build_width(int n) {
    return n <= 512 ? 8 : n <= 1024 ? 16 : n <= 2048 ? 32 : n <= 4096 ? 64 : n <= 6144 ? 96 : 128;
}
inline int build_row_waves(int width) { return width <= 32 ? 1 : width / 32; }
// Waves per build workgroup: 4 (4, 2 or 1 rows), and 3 for rows of 3 waves
// (a workgroup of 6 waves could not share a CU with a second one: its
// waves fill the SIMDs unevenly, 2-2-1-1, and 3 per SIMD is the limit).
constexpr int build_block_waves(int width) { return width == 96 ? 3 : kBuildWaves; }

// LDS bytes of the two kernels.
inline size_t build_lds_bytes(int d, int n, long long rs) {
    if (n > 2048) return (size_t)build_block_waves(build_width(n)) * 2048 * 4;  // exchange rows
    return (((size_t)d * n * 4 + 15) & ~(size_t)15) + (size_t)kBuildWaves * rs * 2;
}
// LDS bytes of one staged pair of the walk: both nodes' coordinates per
// frame and one list-choice bit per frame.
inline size_t walk_buffer_bytes(int d, int n) {
    return (size_t)2 * d * n * 4 + (size_t)4 * ((n + 31) / 32);
}
inline size_t walk_lds_bytes(int d, int n, int nbuf) {
    return (size_t)nbuf * walk_buffer_bytes(d, n) + (size_t)(8 + 5 * nbuf) * sizeof(int);
}

// Table layout: the rows of 8 consecutive frames of a node are stored
// block-interleaved, [slot][frame / 8][position / 8][frame % 8][8], so the
// lanes that walk consecutive frames of one pair side by side read the same
// block of their rows from one 128-byte line (groups of 16, 32 and 64 rows
// walk as fast; 8 keeps the build's writes densest). table_row gives the
// offset of (slot, frame i)'s position 0, table_at that of position e.
constexpr int kRowGroup = 8;
NSX_HD inline long long table_rows(int n) {
    return ((long long)n + kRowGroup - 1) / kRowGroup * kRowGroup;
}
NSX_HD inline long long table_row(int slot, int i, int n, long long rs) {
    const long long g = (long long)slot * (table_rows(n) / kRowGroup) + i / kRowGroup;
    return g * rs * kRowGroup + (long long)(i % kRowGroup) * 8;
}
NSX_HD inline long long table_at(long long row, long long e) {
    return row + (e >> 3) * (8 * kRowGroup) + (e & 7);
}

// Entries kept per ord row. A walk reads a prefix of its list (on C-alpha
// trajectories the longest walks are a few hundred entries at n = 5000), so
// rows of more than 512 frames keep max(512, n / 8) entries, rounded up to
// 8; a walk that reaches the end of a cut row without meeting the stopping
// rule goes to the tie kernels, which work on the whole row. The rank
// tables keep every frame (a count can reach far beyond the walk).
NSX_HD inline long long ord_stride(int n) {
    const long long rs = ((long long)n + 7) & ~7LL;
    long long c = ((long long)n / 8 + 7) & ~7LL;
    c = c < 512 ? 512 : c;
    return rs < c ? rs : c;
}

// The list position whose marginal distance (the "prior" of a node's frame)
// tells how fast the frame's list grows: the walk takes, per query, the list
// whose prior is larger, i.e. whose prefix below the query's k+1 nearest is
// expected to be shorter. Position 64, or for rows of at most 64 frames the
// last multiple of the build's per-lane width R below n, so that it is
// register 0 of lane prior_pos / R in the builds.
constexpr int kPriorPos = 64;
NSX_HD inline int prior_pos(int n, int R) {
    return n > kPriorPos ? kPriorPos : ((n - 1) / R) * R;
}

// LDS bytes per wave of the tie kernels: the row (n records), the two
// stopper position lists of a partition (n 16-bit entries each) and the
// sub-range stack (64 x 3 ints).
NSX_HD inline size_t tie_wave_lds_bytes(int n) {
    return (size_t)n * sizeof(KeyIdx) + (size_t)4 * n + 64 * 3 * sizeof(int);
}

// Heap entries of the walk: (distance code << id_bits) | frame in 32 bits,
// id_bits the frame bits for n frames. The code covers the 2^27 float-bit
// values (16 binades) below the pair's coordinate span at 2^code_shift per
// step: id_bits + (32 - id_bits) code bits - code_shift = 27.
inline int heap_id_bits(int n) {
    int b = 1;
    while ((1 << b) < n) ++b;
    return b;
}
inline int heap_code_shift(int n) { return heap_id_bits(n) > 5 ? heap_id_bits(n) - 5 : 0; }


// Kernel instances, one translation unit per coordinate dimension.
BuildFn build_d1(int width);
BuildFn build_d2(int width);
BuildFn build_d3(int width);
WalkFn walk_d1(int k);
WalkFn walk_d2(int k);
WalkFn walk_d3(int k);
// Tie-resolution kernels by kind: kTieSelect (every deferred query, no
// LDS), kTieExact (the exact ties among them, row in LDS), kTieOverflow
// (queries the tie list had no room for).
enum TieKind : int { kTieSelect = 0, kTieExact = 1, kTieOverflow = 2 };
TieFn tie_d1(int k, int kind);
TieFn tie_d2(int k, int kind);
TieFn tie_d3(int k, int kind);
// The one-wave-per-query kernel for k > ksg::kMaxK (any k).
WalkFn bigk_d1();
WalkFn bigk_d2();
WalkFn bigk_d3();

#if defined(__HIPCC__)

using ksg::marg;

// The value of lane (lane ^ M): DPP for M = 1, 2 (quad permutations) and 8
// (row rotation by 8 within 16 lanes), all VALU; ds_swizzle in bitmask mode
// for M = 4, 16, ds_bpermute for M = 32.
template <int M>
__device__ __forceinline__ unsigned xor_lane(unsigned v) {
    if constexpr (M == 1) {
        return (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);
    } else if constexpr (M == 2) {
        return (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false);
    } else if constexpr (M == 8) {
        return (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x128, 0xF, 0xF, false);  // row_ror:8
    } else if constexpr (M < 32) {
        return (unsigned)__builtin_amdgcn_ds_swizzle((int)v, 0x1F | (M << 10));
    } else {
        return (unsigned)__shfl_xor((int)v, M);
    }
}

// Sort keys of the table builds: a marginal distance as a code of CB bits
// relative to the node's span (no marginal distance of the node exceeds
// it), 2^(27 - CB) float-bit values per step, so the codes cover the 2^27
// float-bit values (16 binades) below the span; smaller keys share code 0.
// The code is monotone in the key. Its largest value is 2^CB - 2, so no
// packed (code << (32 - CB)) | j equals the padding 0xffffffff.
template <int CB>
__device__ __forceinline__ unsigned build_code_base(float span) {
    constexpr unsigned kMaxOff = (((1u << CB) - 2u) << (27 - CB)) | ((1u << (27 - CB)) - 1u);
    const unsigned top = __float_as_uint(span);
    return top > kMaxOff ? top - kMaxOff : 0u;
}
template <int CB>
__device__ __forceinline__ unsigned build_code(float key, unsigned base) {
    return __builtin_elementwise_sub_sat(__float_as_uint(key), base) >> (27 - CB);
}

// ---- rows of n <= 2048 frames: one wave per row, 32-bit sort keys ----------
//
// A row's elements are packed as build_code<21>(key) << 11 | j: equal codes
// only for keys within 2^-17 relative of each other (or both below the
// node's span by more than 16 binades), and j < 2048 fits the low bits.
// These 32-bit values are sorted with every merge ascending (the
// first stage of a merge compares mirrored positions), so a compare-exchange
// is one v_min_u32 and one v_max_u32. The exact (key, j) order inside runs of
// equal codes is then restored by odd-even transposition on the exact
// keys, repeated until no pair swaps.

// Partner value of the lane (lane ^ M), for any M < 64.
template <int M>
__device__ __forceinline__ unsigned xor_lane_any(unsigned v) {
    if constexpr (M == 1 || M == 2 || M == 4 || M == 8 || M == 16 || M == 32) {
        return xor_lane<M>(v);
    } else if constexpr (M == 3) {
        return (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x1B, 0xF, 0xF, false);
    } else if constexpr (M == 7) {
        return (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x141, 0xF, 0xF, false);  // row_half_mirror
    } else if constexpr (M == 15) {
        return (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x140, 0xF, 0xF, false);  // row_mirror
    } else if constexpr (M < 32) {
        return (unsigned)__builtin_amdgcn_ds_swizzle((int)v, 0x1F | (M << 10));
    } else {
        return (unsigned)__shfl_xor((int)v, M);
    }
}

// DPP control of the partner lane for lane ^ M (M = 1, 2, 8) or the
// mirrored lane (M = 3, 7, 15: quad, half-row and row mirrors); 0 where
// none exists.
template <int M>
constexpr int dpp_ctrl() {
    return M == 1 ? 0xB1 : M == 2 ? 0x4E : M == 3 ? 0x1B : M == 7 ? 0x141 : M == 8 ? 0x128
         : M == 15 ? 0x140 : 0;
}

// DPP read of src from the lane CTRL names, with bound_ctrl (lanes whose
// source is outside the row read 0). Each call is its own instruction, so
// the compiler folds it into the one v_min_u32 / v_max_u32 that uses it
// (v_min_u32_dpp): a compare-exchange across lanes is three instructions
// (min, max, select) instead of four with a separate v_mov_b32_dpp.
template <int CTRL>
__device__ __forceinline__ unsigned dpp_read(unsigned src) {
    return (unsigned)__builtin_amdgcn_update_dpp(0, (int)src, CTRL, 0xF, 0xF, true);
}

// min and max of v with DPP reads (CTRL_LO, CTRL_HI) of src, both
// computed with every lane active: a DPP read of a lane that is switched
// off returns 0, so the compiler must not move them into the two arms of
// a branch on `lower` (the empty asm makes both values needed before it).
template <int CTRL_LO, int CTRL_HI>
__device__ __forceinline__ unsigned dpp_cmpx(unsigned v, unsigned src, bool lower) {
    unsigned lo = min(v, dpp_read<CTRL_LO>(src));
    unsigned hi = max(v, dpp_read<CTRL_HI>(src));
    asm volatile("" : "+v"(lo), "+v"(hi));
    return lower ? lo : hi;
}

// lower ? min(v, p) : max(v, p), p = src of the partner lane of M.
template <int M>
__device__ __forceinline__ unsigned lane_cmpx(unsigned v, unsigned src, bool lower) {
    constexpr int C = dpp_ctrl<M>();
    if constexpr (C != 0) {
        return dpp_cmpx<C, C>(v, src, lower);
    } else {
        const unsigned o = xor_lane_any<M>(src);
        return lower ? min(v, o) : max(v, o);
    }
}

// Half-cleaner stages STRIDE, STRIDE/2, .., 1 (all ascending). Partner
// lanes at distance 16 or 32: v_permlane16/32_swap of two registers puts
// both members of each pair of either register in one lane (the lower
// member in the first result), a min and a max order them, and the same
// swap puts them back. At distance 4: DPP row_shl:4 (lane l + 4) for the
// lower lanes, row_shr:4 (lane l - 4) for the upper ones. Both stay off the
// LDS unit, which the other cross-lane exchanges share with the row loads.
template <int R, int STRIDE>
__device__ __forceinline__ void asc_stages(unsigned (&v)[R], int lane) {
    if constexpr (STRIDE >= R) {
        constexpr int LX = STRIDE / R;
        if constexpr (LX == 16 || LX == 32) {
#pragma unroll
            for (int r = 0; r < R; r += 2) {
                unsigned lo, hi;
                if constexpr (LX == 16) {
                    const auto t = __builtin_amdgcn_permlane16_swap(v[r], v[r + 1], false, false);
                    lo = min((unsigned)t[0], (unsigned)t[1]);
                    hi = max((unsigned)t[0], (unsigned)t[1]);
                    const auto u = __builtin_amdgcn_permlane16_swap(lo, hi, false, false);
                    v[r] = u[0];
                    v[r + 1] = u[1];
                } else {
                    const auto t = __builtin_amdgcn_permlane32_swap(v[r], v[r + 1], false, false);
                    lo = min((unsigned)t[0], (unsigned)t[1]);
                    hi = max((unsigned)t[0], (unsigned)t[1]);
                    const auto u = __builtin_amdgcn_permlane32_swap(lo, hi, false, false);
                    v[r] = u[0];
                    v[r + 1] = u[1];
                }
            }
        } else if constexpr (LX == 4) {
            // row_shl:4 (lane + 4) for the lower lanes, row_shr:4 (lane - 4)
            // for the upper ones; the lanes whose read leaves the row take
            // the other branch.
            const bool lower = (lane & 4) == 0;
#pragma unroll
            for (int r = 0; r < R; ++r) v[r] = dpp_cmpx<0x104, 0x114>(v[r], v[r], lower);
        } else {
            const bool lower = (lane & LX) == 0;
#pragma unroll
            for (int r = 0; r < R; ++r) v[r] = lane_cmpx<LX>(v[r], v[r], lower);
        }
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (r & STRIDE) continue;
            const unsigned x = v[r], y = v[r | STRIDE];
            v[r] = min(x, y);
            v[r | STRIDE] = max(x, y);
        }
    }
    if constexpr (STRIDE > 1) asc_stages<R, STRIDE / 2>(v, lane);
}

// Merges of size SIZE .. 64 R: a mirror stage, then the half-cleaners.
template <int R, int SIZE>
__device__ __forceinline__ void mirror_merges(unsigned (&v)[R], int lane) {
    if constexpr (SIZE <= R) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int r2 = r ^ (SIZE - 1);
            if (r2 <= r) continue;
            const unsigned x = v[r], y = v[r2];
            v[r] = min(x, y);
            v[r2] = max(x, y);
        }
    } else {
        // Element (lane, r) pairs with (lane ^ M, R - 1 - r); registers r
        // and R - 1 - r are updated together from their old values.
        constexpr int M = SIZE / R - 1;
        const bool lower = (lane & ((M + 1) >> 1)) == 0;
#pragma unroll
        for (int r = 0; r < R / 2; ++r) {
            const unsigned x = v[r], y = v[R - 1 - r];
            v[r] = lane_cmpx<M>(x, y, lower);
            v[R - 1 - r] = lane_cmpx<M>(y, x, lower);
        }
    }
    if constexpr (SIZE >= 4) asc_stages<R, SIZE / 4>(v, lane);
    if constexpr (SIZE < kWave * R) mirror_merges<R, SIZE * 2>(v, lane);
}

// One workgroup per (node slot, row group); each wave sorts whole rows.
// Three waves per SIMD: at R = 32 that costs 21 VGPRs of spills and
// measures 3% faster than two waves without.
template <int D, int R>
__global__ __launch_bounds__(kWave *kBuildWaves, 3) void list_build_kernel(BuildArgs a) {
    extern __shared__ __attribute__((aligned(16))) float lds[];
    const int n = a.n;
    const int slot = blockIdx.x / a.groups;
    const int g = blockIdx.x - slot * a.groups;
    const float *x = a.X + (long long)a.node_of[slot] * a.ld;
    const int wib = threadIdx.x / kWave;
    if (a.abort && *a.abort) return;
    const unsigned cbase = build_code_base<21>(a.span[a.node_of[slot]]);
    for (int t = threadIdx.x; t < D * n; t += blockDim.x) {
        const int d = t / n, f = t - d * n;
        lds[t] = x[d * a.dstride + f];
    }
    __syncthreads();
    uint16_t *st = reinterpret_cast<uint16_t *>(
                       reinterpret_cast<char *>(lds) + ((D * n * 4 + 15) & ~15)) +
                   (long long)wib * a.rs;
    constexpr unsigned kPad = 0xffffffffu;  // above every packed finite key
    for (int i = g * kBuildWaves + wib; i < n; i += a.groups * kBuildWaves) {
        // The lane index is re-derived every row from a value the compiler
        // cannot see through, so per-lane addresses and bounds are computed
        // where they are used instead of being hoisted out of the loop and
        // held in registers across it.
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int lane = tid & (kWave - 1);
        float q[D];
#pragma unroll
        for (int d = 0; d < D; ++d) q[d] = lds[d * n + i];
        // The sort takes any initial placement: frame r * 64 + lane starts in
        // register r of this lane, so each LDS read covers 64 consecutive
        // words (no bank conflicts).
        unsigned v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int e = r * kWave + lane;
            if (e < n) {
                float c[D];
#pragma unroll
                for (int d = 0; d < D; ++d) c[d] = lds[d * n + e];
                v[r] = (build_code<21>(marg<D>(q, c), cbase) << 11) | (unsigned)e;
            } else {
                v[r] = kPad;
            }
        }
        mirror_merges<R, 2>(v, lane);

        // Exact order: the elements as (exact key bits << 32 | j), then
        // odd-even transposition until no pair swaps. Pairs whose codes
        // differ are already in exact order, so only equal-code
        // runs (typically of 2 or 3) move; the loop ends after as many passes
        // as the longest run.
        unsigned long long ex[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (v[r] != kPad) {
                float c[D];
                const int j = (int)(v[r] & 0x7ffu);
#pragma unroll
                for (int d = 0; d < D; ++d) c[d] = lds[d * n + j];
                ex[r] = ((unsigned long long)__float_as_uint(marg<D>(q, c)) << 32) | (unsigned)j;
            } else {
                ex[r] = ~0ull;
            }
        }
        for (;;) {
            bool swapped = false;
#pragma unroll
            for (int r = 0; r + 1 < R; r += 2) {
                const unsigned long long x = ex[r], y = ex[r + 1];
                swapped |= x > y;
                ex[r] = x < y ? x : y;
                ex[r + 1] = x < y ? y : x;
            }
#pragma unroll
            for (int r = 1; r + 1 < R; r += 2) {
                const unsigned long long x = ex[r], y = ex[r + 1];
                swapped |= x > y;
                ex[r] = x < y ? x : y;
                ex[r + 1] = x < y ? y : x;
            }
            // The pair across lanes: (lane, R-1) and (lane+1, 0).
            const unsigned long long nx = ((unsigned long long)(unsigned)__shfl_down((int)(ex[0] >> 32), 1) << 32) |
                                          (unsigned)__shfl_down((int)(unsigned)ex[0], 1);
            const unsigned long long px = ((unsigned long long)(unsigned)__shfl_up((int)(ex[R - 1] >> 32), 1) << 32) |
                                          (unsigned)__shfl_up((int)(unsigned)ex[R - 1], 1);
            const bool lsw = lane < kWave - 1 && ex[R - 1] > nx;
            const bool rsw = lane > 0 && px > ex[0];
            ex[R - 1] = lsw ? nx : ex[R - 1];
            ex[0] = rsw ? px : ex[0];
            swapped |= lsw;
            if (__ballot(swapped) == 0) break;
        }
        float kx[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            v[r] = (unsigned)ex[r] & 0x7ffu;
            kx[r] = ex[r] == ~0ull ? __builtin_inff() : __uint_as_float((unsigned)(ex[r] >> 32));
        }
        if (lane == prior_pos(n, R) / R) a.prior[(long long)slot * table_rows(n) + i] = kx[0];

        // ord row (its first a.ors positions): this lane's R consecutive
        // positions, 16 bytes at a time.
        const long long orow = table_row(slot, i, n, a.ors);
#pragma unroll
        for (int r0 = 0; r0 < R; r0 += 8) {
            const long long e0 = (long long)lane * R + r0;
            if (e0 < a.ors) {
                // Padding positions (past n) hold the row's own frame i: the
                // walk reads whole blocks and never inserts frame i.
                unsigned e8[8];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    e8[u] = e0 + u < n ? (v[r0 + u] & 0x7ffu) : (unsigned)i;
                uint4 w;
                w.x = e8[0] | (e8[1] << 16);
                w.y = e8[2] | (e8[3] << 16);
                w.z = e8[4] | (e8[5] << 16);
                w.w = e8[6] | (e8[7] << 16);
                *reinterpret_cast<uint4 *>(a.ord + table_at(orow, e0)) = w;
            }
        }

        // Upper-bound rank of position e: the first end of an equal-key run
        // at or after e (a run ends where the next exact key differs; the
        // padding's +inf differs from every finite key).
        constexpr unsigned kBig = 0x7fffffffu;
        const float next0 = __shfl_down(kx[0], 1);
        unsigned sfx[R];
        unsigned m = kBig;
#pragma unroll
        for (int r = R - 1; r >= 0; --r) {
            const float kn = (r == R - 1) ? (lane == kWave - 1 ? __builtin_inff() : next0) : kx[r + 1];
            const bool end = (kn != kx[r]) || (r == R - 1 && lane == kWave - 1);
            m = end ? min(m, (unsigned)(lane * R + r + 1)) : m;
            sfx[r] = m;
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
            if (e < n) st[v[r]] = (uint16_t)min(sfx[r], after);
        }
        ksg::wave_sync();
        const long long rrow = table_row(slot, i, n, a.rs);
        for (long long c = (long long)lane * 8; c < a.rs; c += kWave * 8)
            *reinterpret_cast<uint4 *>(a.rank + table_at(rrow, c)) =
                *reinterpret_cast<const uint4 *>(st + c);
        ksg::wave_sync();
    }
}

// Cross-wave stages of the all-ascending merges of a row sorted by W waves
// (2048 32-bit values each, element e = lane * 32 + r of wave wl). The LDS
// copy of a wave's block is stored as [r][lane], so both the writes and the
// mirrored reads cover 64 consecutive words.
// A partner wave pw >= W is the all-padding block of a row of W = 3 waves
// (sorted as 4): an exchange with it leaves this wave's values as they are.
__device__ __forceinline__ void wave_exchange(unsigned (&v)[32], int lane, int wl, int pw,
                                              bool mirror, unsigned *xr, int W) {
#pragma unroll
    for (int r = 0; r < 32; ++r) xr[wl * 2048 + r * kWave + lane] = v[r];
    __syncthreads();
    if (pw < W) {
        const bool lower = wl < pw;
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            const unsigned o = mirror ? xr[pw * 2048 + (31 - r) * kWave + (kWave - 1 - lane)]
                                      : xr[pw * 2048 + r * kWave + lane];
            v[r] = lower ? min(v[r], o) : max(v[r], o);
        }
    }
    __syncthreads();
}

// Merges of sizes 4096 and (W = 3, 4) 8192 after each wave sorted its
// block: the mirror stage pairs local e of one wave with local 2047 - e of
// its partner, the stride-2048 stage pairs equal positions, and the strides
// 1024 .. 1 stay within the wave. A row of 3 waves is sorted as 4 whose
// fourth block is all padding (+inf): it stays +inf through every stage,
// so the exchanges with it change nothing and are skipped.
template <int W>
__device__ __forceinline__ void wide_merges32(unsigned (&v)[32], int lane, int wl, unsigned *xr) {
    wave_exchange(v, lane, wl, wl ^ 1, true, xr, W);
    asc_stages<32, 1024>(v, lane);
    if constexpr (W >= 3) {
        wave_exchange(v, lane, wl, 3 - wl, true, xr, W);
        wave_exchange(v, lane, wl, wl ^ 1, false, xr, W);
        asc_stages<32, 1024>(v, lane);
    }
}

// Rows of 2048 < n <= 2048 W frames: W waves sort one row together (each a
// 2048-element block, 32 per lane) as 32-bit values build_code<19>(key) <<
// 13 | j: every in-wave compare-exchange is a v_min_u32 / v_max_u32 with
// DPP or permlane partners, as in list_build_kernel, and the cross-wave
// stages go through LDS. The code t = value >> 13 is monotone in the key
// (equal codes only within 2^-15 relative, or both more than 16 binades
// below the node's span) and j < 8192 fits the low bits. Only elements that
// share t with a neighbour need their exact key bits (ek): odd-even
// transposition on (ek, j) inside those runs, across the row's waves,
// restores exact (key, j) order, and the upper-bound ranks come from the
// ends of runs of equal (t, ek). A workgroup of build_block_waves(32 W)
// waves holds that many / W rows at a time, taken in turn from all the
// launch's rows; coordinates are read from global memory (a node's rows
// are built by many workgroups at once, so its coordinates stay in L2).
template <int D, int W>
__global__ __launch_bounds__(kWave *build_block_waves(32 * W), 3) void list_build_wide_kernel(
    BuildArgs a) {
    constexpr int kBW = build_block_waves(32 * W);
    extern __shared__ __attribute__((aligned(16))) unsigned xch[];
    __shared__ unsigned edge_u[2][kBW], edge_k[2][kBW];  // [0] first, [1] last
    __shared__ unsigned wave_min[kBW];
    __shared__ int any_swap;
    constexpr int kRows = kBW / W;
    constexpr unsigned kPad = 0xffffffffu;  // above every packed finite key
    const int n = a.n;
    const int wib = threadIdx.x / kWave;
    const int rg = wib / W, wl = wib - rg * W;  // row in the workgroup, block in the row
    const int ebase = wl * 2048;
    if (a.abort && *a.abort) return;
    unsigned *xr = xch + rg * (W * 2048);
    uint16_t *st = reinterpret_cast<uint16_t *>(xr);  // rank staging after the sort
    // Rows t = slot * n + i of all the launch's nodes, dealt to the
    // workgroups in turn, so every workgroup gets the same number of rows
    // (within one) whatever the node count.
    const int total = a.nodes * n;  // below 2^31: nodes <= table slots, n <= 8192
    for (int base = blockIdx.x * kRows; base < total; base += gridDim.x * kRows) {
        // The lane index is re-derived every row from a value the compiler
        // cannot see through, so per-lane addresses and bounds are computed
        // where they are used instead of being hoisted out of the loop and
        // held in registers across it.
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int lane = tid & (kWave - 1);
        const int tr = base + rg;
        const bool row = tr < total;
        const int slot = row ? tr / n : 0;
        const int i = row ? tr - slot * n : 0;
        const float *x = a.X + (long long)a.node_of[slot] * a.ld;
        const unsigned cbase = build_code_base<19>(a.span[a.node_of[slot]]);
        float q[D];
#pragma unroll
        for (int d = 0; d < D; ++d) q[d] = row ? x[d * a.dstride + i] : 0.0f;
        // Any initial placement sorts the same: frame ebase + r * 64 + lane
        // starts in register r of this lane, so each load is coalesced.
        unsigned u[32];
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            const int e = ebase + r * kWave + lane;
            if (row && e < n) {
                float c[D];
#pragma unroll
                for (int d = 0; d < D; ++d) c[d] = x[d * a.dstride + e];
                u[r] = (build_code<19>(marg<D>(q, c), cbase) << 13) | (unsigned)e;
            } else {
                u[r] = kPad;
            }
            // At most 8 elements' loads in flight (the compiler would issue
            // all 96 first and run out of registers).
            if ((r & 7) == 7) asm volatile("" ::: "memory");
        }
        mirror_merges<32, 2>(u, lane);
        wide_merges32<W>(u, lane, wl, xr);

        // Neighbours across lanes and waves: (lane, 31) -> (lane + 1, 0),
        // the wave's last -> the next wave's first (kPad past the row).
        if (lane == 0) edge_u[0][wib] = u[0];
        if (lane == kWave - 1) edge_u[1][wib] = u[31];
        __syncthreads();
        unsigned un = (unsigned)__shfl_down((int)u[0], 1);
        unsigned up = (unsigned)__shfl_up((int)u[31], 1);
        if (lane == kWave - 1) un = wl + 1 < W ? edge_u[0][wib + 1] : kPad;
        if (lane == 0) up = wl > 0 ? edge_u[1][wib - 1] : kPad;
        __syncthreads();
        // Exact key bits of the elements in runs of equal code.
        unsigned ek[32];
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            const unsigned t = u[r] >> 13;
            const unsigned tp = (r == 0 ? up : u[r - 1]) >> 13;
            const unsigned tn = (r == 31 ? un : u[r + 1]) >> 13;
            ek[r] = 0u;
            if (u[r] != kPad && (t == tp || t == tn)) {
                const int j = (int)(u[r] & 0x1fffu);
                float c[D];
#pragma unroll
                for (int d = 0; d < D; ++d) c[d] = x[d * a.dstride + j];
                ek[r] = __float_as_uint(marg<D>(q, c));
            }
            if ((r & 7) == 7) asm volatile("" ::: "memory");
        }
        // Odd-even transposition within runs, until no pair of the
        // workgroup's rows swaps. An element is (ek << 32) | u as one 64-bit
        // value: inside a run (equal codes) u orders like j, so (ek, u)
        // orders like the exact (key, j), and one 64-bit compare decides a
        // swap; a pair with different codes is in order already.
        unsigned long long K[32];
#pragma unroll
        for (int r = 0; r < 32; ++r) K[r] = ((unsigned long long)ek[r] << 32) | u[r];
        auto cx = [](unsigned long long &x, unsigned long long &y) {
            const bool sw = ((unsigned)x ^ (unsigned)y) < 0x2000u && x > y;
            const unsigned long long t = x;
            x = sw ? y : x;
            y = sw ? t : y;
            return sw;
        };
        for (;;) {
            bool swapped = false;
#pragma unroll
            for (int r = 0; r + 1 < 32; r += 2) swapped |= cx(K[r], K[r + 1]);
#pragma unroll
            for (int r = 1; r + 1 < 32; r += 2) swapped |= cx(K[r], K[r + 1]);
            if (threadIdx.x == 0) any_swap = 0;
            if (lane == 0) {
                edge_u[0][wib] = (unsigned)K[0];
                edge_k[0][wib] = (unsigned)(K[0] >> 32);
            }
            if (lane == kWave - 1) {
                edge_u[1][wib] = (unsigned)K[31];
                edge_k[1][wib] = (unsigned)(K[31] >> 32);
            }
            unsigned nu = (unsigned)__shfl_down((int)(unsigned)K[0], 1);
            unsigned nk = (unsigned)__shfl_down((int)(unsigned)(K[0] >> 32), 1);
            unsigned pu = (unsigned)__shfl_up((int)(unsigned)K[31], 1);
            unsigned pk = (unsigned)__shfl_up((int)(unsigned)(K[31] >> 32), 1);
            __syncthreads();
            if (lane == kWave - 1) {
                nu = wl + 1 < W ? edge_u[0][wib + 1] : kPad;
                nk = wl + 1 < W ? edge_k[0][wib + 1] : 0u;
            }
            if (lane == 0) {
                pu = wl > 0 ? edge_u[1][wib - 1] : 0u;
                pk = wl > 0 ? edge_k[1][wib - 1] : 0u;
            }
            // (lane, 31) keeps the smaller of itself and its successor,
            // (lane, 0) the larger of its predecessor and itself.
            {
                unsigned long long nx = ((unsigned long long)nk << 32) | nu;
                swapped |= cx(K[31], nx);
            }
            {
                unsigned long long px = ((unsigned long long)pk << 32) | pu;
                cx(px, K[0]);
            }
            if (__ballot(swapped) && lane == 0) atomicOr(&any_swap, 1);
            __syncthreads();
            const int again = any_swap;
            __syncthreads();
            if (!again) break;
        }
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            u[r] = (unsigned)K[r];
            ek[r] = (unsigned)(K[r] >> 32);
        }
        // The prior (position 64 = register 0 of lane 2 of the row's first
        // wave), from its code: within 2^-15 relative of the key.
        if (row && wl == 0 && lane == kPriorPos / 32)
            a.prior[(long long)slot * table_rows(n) + i] = __uint_as_float(((u[0] >> 13) << 8) + cbase);

        // ord row (its first a.ors positions).
        const long long orow = table_row(slot, row ? i : 0, n, a.ors);
#pragma unroll
        for (int r0 = 0; r0 < 32; r0 += 8) {
            const long long e0 = (long long)ebase + lane * 32 + r0;
            if (row && e0 < a.ors) {
                // Padding positions (past n) hold the row's own frame i: the
                // walk reads whole blocks and never inserts frame i.
                unsigned e8[8];
#pragma unroll
                for (int v = 0; v < 8; ++v)
                    e8[v] = e0 + v < n ? (u[r0 + v] & 0x1fffu) : (unsigned)i;
                uint4 w;
                w.x = e8[0] | (e8[1] << 16);
                w.y = e8[2] | (e8[3] << 16);
                w.z = e8[4] | (e8[5] << 16);
                w.w = e8[6] | (e8[7] << 16);
                *reinterpret_cast<uint4 *>(a.ord + table_at(orow, e0)) = w;
            }
        }

        // Upper-bound ranks: the first end of a run of equal exact keys at or
        // after each position, across the row's waves. Position e ends a run
        // when its successor has another code or exact key.
        if (lane == 0) {
            edge_u[0][wib] = u[0];
            edge_k[0][wib] = ek[0];
        }
        __syncthreads();
        constexpr unsigned kBig = 0x7fffffffu;
        unsigned su = (unsigned)__shfl_down((int)u[0], 1);
        unsigned sk = (unsigned)__shfl_down((int)ek[0], 1);
        if (lane == kWave - 1) {
            su = wl + 1 < W ? edge_u[0][wib + 1] : kPad;
            sk = wl + 1 < W ? edge_k[0][wib + 1] : 0u;
        }
        // The lane's first run end, then the same scan again below while
        // the ranks are written (no per-position array).
        auto ends_run = [&](int r) {
            const unsigned nu2 = r == 31 ? su : u[r + 1];
            const unsigned nk2 = r == 31 ? sk : ek[r + 1];
            return (nu2 >> 13) != (u[r] >> 13) || nk2 != ek[r];
        };
        unsigned m = kBig;
#pragma unroll
        for (int r = 31; r >= 0; --r)
            m = ends_run(r) ? (unsigned)(ebase + lane * 32 + r + 1) : m;
        unsigned t = m;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const unsigned o = (unsigned)__shfl_down((int)t, off);
            t = lane + off < kWave ? min(t, o) : t;
        }
        if (lane == 0) wave_min[wib] = t;
        __syncthreads();
        unsigned later = kBig;  // minimum over the row's later waves
        for (int w2 = wl + 1; w2 < W; ++w2) later = min(later, wave_min[rg * W + w2]);
        unsigned after = (unsigned)__shfl_down((int)t, 1);
        after = lane == kWave - 1 ? later : min(after, later);
        if (row) {
            unsigned cur = after;
#pragma unroll
            for (int r = 31; r >= 0; --r) {
                const int e = ebase + lane * 32 + r;
                cur = ends_run(r) ? (unsigned)(e + 1) : cur;
                if (e < n) st[u[r] & 0x1fffu] = (uint16_t)cur;
            }
        }
        __syncthreads();
        if (row) {
            const long long rrow = table_row(slot, i, n, a.rs);
            for (long long c = ((long long)wl * kWave + lane) * 8; c < a.rs; c += W * kWave * 8)
                *reinterpret_cast<uint4 *>(a.rank + table_at(rrow, c)) =
                    *reinterpret_cast<const uint4 *>(st + c);
        }
        __syncthreads();
    }
}

// med3 of unsigned 32-bit values (v_med3_u32; the compiler forms it only
// for constant bounds).
__device__ __forceinline__ unsigned umed3(unsigned x, unsigned y, unsigned z) {
    unsigned r;
    asm("v_med3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(x), "v"(y), "v"(z));
    return r;
}

// Insert c into the ascending h[0..K]: one med3 per slot. c >= h[K]
// (0xffffffff for a rejected candidate) leaves everything unchanged.
template <int K>
__device__ __forceinline__ void insert_packed(unsigned (&h)[K + 1], unsigned c) {
#pragma unroll
    for (int m = K; m >= 1; --m) h[m] = umed3(h[m - 1], c, h[m]);
    h[0] = min(h[0], c);
}

// Distance code of a key for a pair whose code window starts at float bits
// `base`: monotone non-decreasing in the key, so code(x) < code(y) implies
// x < y. Keys never exceed the pair's span (both coordinates lie within
// each dimension's [min, max], and rounding is monotone), so the code fits
// the 32 - id_bits code bits.
__device__ __forceinline__ unsigned key_code(float key, unsigned base, int shift) {
    return __builtin_elementwise_sub_sat(__float_as_uint(key), base) >> shift;
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

// Ring buffer of local pair s (the ring holds at most 3 pairs; constant
// divisors compile to multiplications).
__device__ __forceinline__ int ring_slot(int s, int nbuf) {
    return nbuf == 3 ? s % 3 : (nbuf == 2 ? (s & 1) : 0);
}

// Stage local pair s of this workgroup into ring buffer s % nbuf: both rows'
// coordinates as per-frame records, one list-choice bit per frame (1: node
// a's list, whose prior is at least node b's), the two nodes' table slots and
// the pair's code base: the larger span's float bits minus the code window's
// width less one (at least 0), so that every key's code fits 32 - id_bits
// bits.
template <int D>
__device__ __forceinline__ void stage_pair(const WalkArgs &a, float *rec0, unsigned *choice0,
                                           int *meta, long long s, int n) {
    const long long p = a.p0 + (long long)blockIdx.x + s * (long long)gridDim.x;
    const int b = ring_slot((int)s, a.nbuf);
    const int na = a.ab[2 * p], nb = a.ab[2 * p + 1];
    const float *xa = a.X + (long long)na * a.ld;
    const float *xb = a.X + (long long)nb * a.ld;
    float *rec = rec0 + (long long)b * 2 * D * n;
    for (int t = threadIdx.x; t < D * n; t += blockDim.x) {
        const int d = t / n, f = t - d * n;
        rec[2 * D * f + d] = xa[d * a.dstride + f];
        rec[2 * D * f + D + d] = xb[d * a.dstride + f];
    }
    const long long rows = table_rows(n);
    const float *pa = a.prior + (long long)a.slot_of[na] * rows;
    const float *pb = a.prior + (long long)a.slot_of[nb] * rows;
    const int cwords = (n + 31) / 32;
    unsigned *choice = choice0 + (long long)b * cwords;
    // blockDim.x is a multiple of 64: each wave's ballot covers 64 frames
    // from a multiple of 64, i.e. two choice words.
    for (int f0 = 0; f0 < n; f0 += blockDim.x) {
        const int f = f0 + (int)threadIdx.x;
        const unsigned long long m = __ballot(f < n && pa[f] >= pb[f]);
        const int w = (f0 + ((int)threadIdx.x & ~(kWave - 1))) / 32;
        if ((threadIdx.x & (kWave - 1)) == 0) {
            if (w < cwords) choice[w] = (unsigned)m;
            if (w + 1 < cwords) choice[w + 1] = (unsigned)(m >> 32);
        }
    }
    if (threadIdx.x == 0) {
        meta[3 * b] = a.slot_of[na];
        meta[3 * b + 1] = a.slot_of[nb];
        const long long window = 1LL << (32 - a.id_bits + a.code_shift);
        const long long top = (long long)__float_as_uint(fmaxf(a.span[na], a.span[nb]));
        meta[3 * b + 2] = (int)(unsigned)(top - (window - 1) > 0 ? top - (window - 1) : 0);
    }
}

// One persistent workgroup per CU walks the queries of its pairs
// blockIdx.x, blockIdx.x + gridDim.x, ... (local pairs 0, 1, ...). Query q
// is frame q % n of local pair q / n; lanes reserve query numbers from one
// counter, so when a pair's queries run out its lanes go on to the next pair
// while the pair's longest walks finish. Up to nbuf pairs are staged at once
// (an LDS ring); every `checkpoint` loop passes all waves meet, the pairs
// whose queries are all done retire in order and free their buffers for the
// next pairs.
template <int D, int K>
__global__ __launch_bounds__(kWave *walk_max_waves(K)) void walk_kernel(WalkArgs a) {
    extern __shared__ __attribute__((aligned(16))) float lds[];
    if (a.abort && *a.abort) return;
    const int n = a.n;
    const int nbuf = a.nbuf;
    const int rec_floats = 2 * D * n;
    int *ctrl = reinterpret_cast<int *>(lds + nbuf * rec_floats);
    // ctrl[0] next query, [1] staged pairs, [2] retired pairs, [3] pairs to
    // stage now, [8..8+nbuf) done queries per buffer, then per buffer the 2
    // table slots and the code base, then the pairs to stage, then per
    // buffer the list-choice bits.
    int *done = ctrl + 8;
    int *meta = done + nbuf;
    int *to_stage = meta + 3 * nbuf;
    unsigned *choice = reinterpret_cast<unsigned *>(to_stage + nbuf);
    const int cwords = (n + 31) / 32;
    const int idb = a.id_bits, shift = a.code_shift;
    const unsigned id_mask = (1u << idb) - 1u;
    const int lane = threadIdx.x & (kWave - 1);
    const long long rs = a.rs, ors = a.ors;
    // Entries of a list: n, or the ord rows' length when they are cut.
    const int nl = (long long)n < ors ? n : (int)ors;
    const bool cut_rows = nl < n;
    const long long G = gridDim.x;
    const long long npairs =
        (long long)blockIdx.x < a.P ? (a.P - (long long)blockIdx.x + G - 1) / G : 0;
    const int nq = (int)(npairs * n);  // the host keeps P * n below 2^31
    const double inv_n = 1.0 / (double)n;
    unsigned long long ties = 0, steps = 0;

    if (threadIdx.x == 0) {
        ctrl[0] = ctrl[1] = ctrl[2] = ctrl[3] = 0;
        for (int b = 0; b < nbuf; ++b) done[b] = 0;
    }
    const int first = (int)(npairs < nbuf ? npairs : nbuf);
    for (int s = 0; s < first; ++s) stage_pair<D>(a, lds, choice, meta, s, n);
    __syncthreads();
    if (threadIdx.x == 0) ctrl[1] = first;

    // Lane state: 0 wants a query, 3 holds a reserved query whose pair is
    // not staged yet, 1 walks, 2 no queries left.
    int state = 0;
    bool res = false;  // a finished walk waits for its counts
    bool inc = false;  // the walk reached the end of a cut row without stopping
    int q = 0;   // the lane's query number: frame i = q - qs * n of local pair qs
    int qs = 0;
    int i = 0, pos = 0;  // entries walked in the lane's list
    int rec = 0;  // LDS float offset of the lane's pair records (an offset, not a
                  // pointer, so every access stays an LDS instruction)
    // Rank table row offsets of (node a, frame i) and (node b, frame i).
    long long rA = 0, rB = 0;
    bool useA = true;  // the walked list is node a's
    unsigned base = 0; // the pair's code base
    float qa[D], qb[D];
    // The k+1 smallest (code(z) << idb | j) of the frames evaluated, in
    // ascending order. Codes are monotone in z, so when the k-th and
    // (k+1)-th codes differ the first k entries are the k nearest frames
    // exactly; equal codes (a tie of z, or two z too close for the code)
    // go to the tie list.
    unsigned h[K + 1];
    // List blocks: cur holds entries [pos, pos + 8), nxt the 8 after them
    // (pos is a multiple of 8 at every body start until the list ends).
    // nxt is requested at the end of the body before the one that reads
    // it, so its load has a whole body to arrive.
    uint4 cur = {0, 0, 0, 0}, nxt = {0, 0, 0, 0};
    // The lane's list at entry pos (a block of its list, also once the
    // walk has ended, so that every block request stays in its row).
    const uint16_t *lpos = a.ord;
    // Counts of the previous refill round, stored one round later so that
    // the rank lookups' latency is not waited for.
    bool pend = false;
    long long pend_at = 0;
    unsigned pend_x = 1u, pend_y = 1u;
#pragma unroll
    for (int d = 0; d < D; ++d) qa[d] = qb[d] = 0.0f;
#pragma unroll
    for (int m = 0; m <= K; ++m) h[m] = 0xffffffffu;

    for (int pass = 0;; ++pass) {
        if (pass % a.checkpoint == 0) {
            __syncthreads();
            if (threadIdx.x == 0) {
                int staged = ctrl[1], retired = ctrl[2], ns = 0;
                while (retired < staged && done[ring_slot(retired, nbuf)] == n) {
                    done[ring_slot(retired, nbuf)] = 0;
                    ++retired;
                    if (staged + ns < npairs) {
                        to_stage[ns] = staged + ns;
                        ++ns;
                    }
                }
                ctrl[2] = retired;
                ctrl[3] = ns;
            }
            __syncthreads();
            const int ns = ctrl[3];
            if (ns > 0) {
                for (int x = 0; x < ns; ++x) stage_pair<D>(a, lds, choice, meta, to_stage[x], n);
                __syncthreads();
                // Published without a barrier: a wave that reads the old
                // count starts the new pair's queries one pass later, and
                // the next checkpoint's barrier orders this write.
                if (threadIdx.x == 0)
                    __hip_atomic_fetch_add(ctrl + 1, ns, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            if (ctrl[2] >= npairs) break;
        }
        // Queries of staged pairs (the count grows at checkpoints).
        const int avail =
            __hip_atomic_load(ctrl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) * n;
        const unsigned long long want = __ballot(state == 0);
        const unsigned long long ready = __ballot(state == 3 && q < avail);
        const unsigned long long walking = __ballot(state == 1);
        if ((want | ready) &&
            (__popcll(want) + __popcll(ready) >= (unsigned)a.batch || walking == 0)) {
            // (0) counts looked up in the previous round.
            if (pend)
                __builtin_nontemporal_store((pend_x - 1u) | ((pend_y - 1u) << 16), a.cnt + pend_at);
            pend = false;
            // (1) eps of the finished walks. Walks that ended at the end of a
            // cut row without meeting the stopping rule go to the tie list
            // with the ties (the tie kernels work on the whole row).
            const bool tie = res && (inc || ((h[K - 1] ^ h[K]) >> idb) == 0u);
            int jx = 0, jy = 0;
            if (res && !tie) {
                float ex = -1.0f, ey = -1.0f;
#pragma unroll
                for (int m = 0; m < K; ++m) {
                    const int j = (int)(h[m] & id_mask & (kMaxFrames - 1));  // < 8192: 24-bit address math
                    float va[D], vb[D];
                    fetch<D>(lds + rec, j, va, vb);
                    const float da = marg<D>(qa, va), db = marg<D>(qb, vb);
                    jx = da > ex ? j : jx;
                    ex = fmaxf(ex, da);
                    jy = db > ey ? j : jy;
                    ey = fmaxf(ey, db);
                }
            }
            // Ties go to the tie list, resolved by tie_kernel after the walk;
            // those beyond its capacity are marked in cnt for
            // tie_overflow_kernel.
            unsigned long long tm = __ballot(tie);
            ties += (unsigned)__popcll(tm);
            bool deferred = false;
            if (tm) {
                const int leader = __builtin_amdgcn_readfirstlane(__ffsll((long long)tm) - 1);
                unsigned tbase = 0;
                if (lane == leader) tbase = atomicAdd(a.tie_count, (unsigned)__popcll(tm));
                tbase = __shfl(tbase, leader);
                const unsigned slot = tbase + __builtin_amdgcn_mbcnt_hi(
                                                  (unsigned)(tm >> 32),
                                                  __builtin_amdgcn_mbcnt_lo((unsigned)tm, 0u));
                if (tie && slot < a.tie_cap) {
                    a.tie_list[slot] = (unsigned)(((long long)blockIdx.x + qs * G) * n + i);
                    deferred = true;
                }
            }
            if (res) {
                if (!deferred) {
                    // (0 - 1) | (0 - 1) << 16 is kTieMark.
                    // Non-temporal: each rank line serves one lookup, and
                    // keeping it out of L2 leaves room for the list rows.
                    pend_x = tie ? 0u : (unsigned)__builtin_nontemporal_load(a.rank + table_at(rA, jx));
                    pend_y = tie ? 0u : (unsigned)__builtin_nontemporal_load(a.rank + table_at(rB, jy));
                    pend_at = ((long long)blockIdx.x + qs * G) * n + i;
                    pend = true;
                }
                steps += (unsigned)pos;
                atomicAdd(done + ring_slot(qs, nbuf), 1);  // its LDS records are no longer read
            }
            res = false;
            inc = false;
            // (2) reserve query numbers for the lanes that want one.
            if (want) {
                const int leader = __builtin_amdgcn_readfirstlane(__ffsll((long long)want) - 1);
                int base = 0;
                if (lane == leader) base = atomicAdd(ctrl, __popcll(want));
                base = __shfl(base, leader);
                const int below = __builtin_amdgcn_mbcnt_hi(
                    (unsigned)(want >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)want, 0u));
                if (state == 0) {
                    q = base + below;
                    state = q < nq ? 3 : 2;
                    // qs = q / n, by a double product corrected to the exact quotient.
                    int d = (int)((double)q * inv_n);
                    const int r = q - d * n;
                    d = r < 0 ? d - 1 : (r >= n ? d + 1 : d);
                    qs = d;
                }
            }
            // (3) start the reserved queries whose pairs are staged: the
            // walk takes the list the staged choice bit names.
            if (state == 3 && q < avail) {
                const int b = ring_slot(qs, nbuf);
                i = q - qs * n;
                rec = b * rec_floats;
                fetch<D>(lds + rec, i, qa, qb);
#pragma unroll
                for (int m = 0; m <= K; ++m) h[m] = 0xffffffffu;
                pos = 0;
                useA = ((choice[b * cwords + (i >> 5)] >> (i & 31)) & 1u) != 0u;
                rA = table_row(meta[3 * b], i, n, rs);
                rB = table_row(meta[3 * b + 1], i, n, rs);
                base = (unsigned)meta[3 * b + 2];
                // Entries [0, 8) and [8, 16) of the list (table_at(row, 8) =
                // row + 64).
                lpos = a.ord + table_row(meta[3 * b + (useA ? 0 : 1)], i, n, ors);
                cur = load_block(lpos);
                nxt = load_block(ors > 8 ? lpos + 64 : lpos);
                state = 1;
            }
        }
        // No lane walks: the body below changes nothing (one path into the
        // next pass keeps the loop's registers in place), and the next
        // pass is a checkpoint.
        if (__ballot(state == 1) == 0) pass = (pass / a.checkpoint + 1) * a.checkpoint - 1;
        // A body: the next 8 entries of the lane's list.
        const bool walks = state == 1;
        const int rem = nl - pos;  // entries left in the list
        float key7 = 0.0f;        // the list's key of the body's last entry
        // The body's 8 records are requested at once (their gathers overlap
        // instead of each step waiting for its own). Block entries are
        // frames (the tables' padding past a list's end holds the query
        // frame i, which is never inserted; blocks of no walk are 0 or a
        // previous walk's), so every read stays within the lane's records.
        float u[8][D], v[8][D];
#pragma unroll
        for (int st = 0; st < 8; ++st) fetch<D>(lds + rec, block_entry(cur, st), u[st], v[st]);
#pragma unroll
        for (int st = 0; st < 8; ++st) {
            // Every lane runs every step (no branches: the loop's registers
            // stay in place); a lane that is not walking inserts nothing.
            const int j = block_entry(cur, st);
            const float da = marg<D>(qa, u[st]), db = marg<D>(qb, v[st]);
            const unsigned c = (key_code(fmaxf(da, db), base, shift) << idb) | (unsigned)j;
            const bool take = walks & (j != i);
            insert_packed<K>(h, take ? c : 0xffffffffu);
            if (st == 7) key7 = useA ? da : db;
        }
        // The stopping rule, once per body: every frame not walked yet has a
        // key of at least key7 in this list, so a joint distance of at least
        // key7; once code(key7) exceeds h[K]'s code none of them can enter.
        // The rule is monotone (keys only grow, h[K] only falls), so a walk
        // that meets it inside the body meets it here too, and the entries
        // after that point change nothing. A list walked to its end has
        // evaluated every frame; a cut row walked to its end has not, and
        // the walk is marked incomplete.
        pos += walks ? min(rem, 8) : 0;
        const unsigned stop = key_code(key7, base, shift) << idb;
        const bool met = stop > h[K];
        const bool fin = walks & ((rem <= 8) | met);
        res = res | fin;
        inc = inc | (fin & !met & cut_rows);
        state = fin ? 0 : state;
        // The next body's block, and the request of the one after it
        // (table_at(row, pos + 8) = lpos + 64), by every lane: a lane that
        // will not walk, or whose list ends before that block, reads its
        // current block again.
        const bool go = walks & !fin;
        lpos += go ? 64 : 0;
        const uint16_t *lnext = go && pos + 8 < ors ? lpos + 64 : lpos;
        cur = nxt;
        // The copy is made before the request is issued, so the request can
        // land in nxt's own registers: issued first, it would land in
        // temporaries whose copy into nxt waits for it at once.
        asm volatile("" : "+v"(cur.x), "+v"(cur.y), "+v"(cur.z), "+v"(cur.w)::"memory");
        nxt = load_block(lnext);
    }
    if (pend) __builtin_nontemporal_store((pend_x - 1u) | ((pend_y - 1u) << 16), a.cnt + pend_at);
    // Per-wave totals.
    unsigned long long sum = steps;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) sum += __shfl_xor(sum, off);
    if (lane == 0) {
        if (ties) atomicAdd(a.stats, ties);
        atomicAdd(a.stats + 1, sum);
    }
}

// libstdc++'s unguarded partition of a[f1, l) around the pivot record a[f1 - 1]
// (partition_unguarded in sort_select.h), by one wave. The sequential scans
// stop at the "left stoppers" (positions p >= f1 with !(a[p] < pivot)) from
// the left and at the "right stoppers" (p >= f1 - 1 with !(pivot < a[p]); the
// pivot is one) from the right. Until the scans cross, the t-th swap
// exchanges the t-th left stopper Ls[t] with the t-th right stopper Rs[t]
// counted from the right, all in unswapped territory. The loop ends at the
// first t with Ls[t] >= Rs[t] (or no such stopper) and returns min(Ls[t],
// Rs[t - 1]), with Rs[0] = l. So: list both kinds of stoppers by prefix
// counts, find that t, and do the t - 1 swaps at once (their positions are
// distinct). Returns the cut, wave-uniform.
__device__ int partition_wave(KeyIdx *a, int f1, int l, uint16_t *lasc, uint16_t *rasc, int lane) {
    const float pivot = a[f1 - 1].key;
    int nl = 0, nr = 0;
    for (int c0 = f1 - 1; c0 < l; c0 += kWave) {
        const int p = c0 + lane;
        const bool in = p < l;
        const float v = in ? a[p].key : 0.0f;
        const bool isl = in && p >= f1 && !(v < pivot);
        const bool isr = in && !(pivot < v);
        const unsigned long long ml = __ballot(isl), mr = __ballot(isr);
        const unsigned bl = __builtin_amdgcn_mbcnt_hi((unsigned)(ml >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((unsigned)ml, 0u));
        const unsigned br = __builtin_amdgcn_mbcnt_hi((unsigned)(mr >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((unsigned)mr, 0u));
        if (isl) lasc[nl + bl] = (uint16_t)p;
        if (isr) rasc[nr + br] = (uint16_t)p;
        nl += __popcll(ml);
        nr += __popcll(mr);
    }
    ksg::wave_sync();
    // Ls[t] = lasc[t - 1], Rs[t] = rasc[nr - t].
    int T = 0;
    for (int t0 = 1; T == 0; t0 += kWave) {
        const int t = t0 + lane;
        const bool stop = t > nl || t > nr || lasc[t - 1] >= rasc[nr - t];
        const unsigned long long ms = __ballot(stop);
        if (ms) T = t0 + __ffsll((long long)ms) - 1;
    }
    for (int t0 = 1; t0 < T; t0 += kWave) {
        const int t = t0 + lane;
        if (t < T) {
            const int pl = lasc[t - 1], pr = rasc[nr - t];
            const KeyIdx x = a[pl], y = a[pr];
            a[pl] = y;
            a[pr] = x;
        }
    }
    const int ls = T <= nl ? (int)lasc[T - 1] : 0x7fffffff;
    const int rs_prev = T >= 2 ? (int)rasc[nr - (T - 1)] : l;
    ksg::wave_sync();
    return min(ls, rs_prev);
}

// sort_prefix (sort_select.h) by one wave: the same sub-range steps in the
// same order, with each partition of more than 128 records done by
// partition_wave and the small steps (median of three, short partitions,
// the heap-sort fallback, the final insertion sort) by lane 0. a[0, m) ends
// as std::sort leaves it.
__device__ int sort_prefix_wave(KeyIdx *a, int n, int m, uint16_t *lasc, uint16_t *rasc,
                                int *stk, int lane) {
    if (n <= 0) return 0;
    if (m > n) m = n;
    constexpr int kThreshold = 16, kStack = 64, kWaveMin = 128;
    int bound = n;
    if (lane == 0) {
        stk[0] = 0;
        stk[1] = n;
        stk[2] = 2 * floor_log2(n);
    }
    ksg::wave_sync();
    int sp = 1;
    while (sp > 0) {
        --sp;
        int f = __builtin_amdgcn_readfirstlane(stk[3 * sp]);
        int l = __builtin_amdgcn_readfirstlane(stk[3 * sp + 1]);
        int depth = __builtin_amdgcn_readfirstlane(stk[3 * sp + 2]);
        while (l - f > kThreshold) {
            if (depth == 0) {
                if (lane == 0) heap_sort_range(a + f, l - f);
                ksg::wave_sync();
                break;
            }
            --depth;
            const int mid = f + (l - f) / 2;
            int cut = 0;
            if (lane == 0) median_to_first(a, f, f + 1, mid, l - 1);
            ksg::wave_sync();
            if (l - f > kWaveMin) {
                cut = partition_wave(a, f + 1, l, lasc, rasc, lane);
            } else {
                if (lane == 0) cut = (int)partition_unguarded(a, f + 1, l, f);
                cut = __shfl(cut, 0);
                ksg::wave_sync();
            }
            cut = __builtin_amdgcn_readfirstlane(cut);
            if (cut >= m) {
                bound = min(bound, cut);
            } else if (sp < kStack) {
                // The right part still reaches [0, m): process it too.
                if (lane == 0) {
                    stk[3 * sp] = cut;
                    stk[3 * sp + 1] = l;
                    stk[3 * sp + 2] = depth;
                }
                ksg::wave_sync();
                ++sp;
            }
            l = cut;
        }
    }
    if (lane == 0) insertion_sort(a, bound);
    ksg::wave_sync();
    return bound;
}

// Coordinates of frames j0 + u * 64 (u < U) of both nodes, every load issued
// before any is used, so a row pass waits once per U frames per lane (the
// tie kernels run few waves per CU). Frames past n read frame j0 (< n).
template <int D, int U>
__device__ __forceinline__ void load_frames(const float *xa, const float *xb, long long ds,
                                            int j0, int n, float (&va)[U][D], float (&vb)[U][D]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int j = j0 + u * kWave;
        const int jc = j < n ? j : j0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            va[u][d] = xa[d * ds + jc];
            vb[u][d] = xb[d * ds + jc];
        }
    }
}

// Counts of one deferred query (local pair pl, frame i; equal k-th and
// (k+1)-th heap codes), by one wave with an n-record LDS row: the wave
// rebuilds the query's joint distance row. When the k-th and (k+1)-th
// smallest distances over the other frames differ (the codes only
// collided), the k nearest frames are those with distance <= the k-th, as
// in the walk. Otherwise the row's positions 0..k are selected in libstdc++
// std::sort order (mutual_information_cpu.cpp:48-77) and eps taken over
// positions 1..k. The counts come from the rank tables.
template <int D, int K>
__device__ void resolve_tie(const WalkArgs &a, long long pl, int i, char *wave_lds, int lane) {
    KeyIdx *row = reinterpret_cast<KeyIdx *>(wave_lds);
    uint16_t *lasc = reinterpret_cast<uint16_t *>(wave_lds + (size_t)a.n * sizeof(KeyIdx));
    uint16_t *rasc = lasc + a.n;
    int *stk = reinterpret_cast<int *>(rasc + a.n);
    const int n = a.n;
    const long long p = a.p0 + pl;
    const int na = a.ab[2 * p], nb = a.ab[2 * p + 1];
    const float *xa = a.X + (long long)na * a.ld;
    const float *xb = a.X + (long long)nb * a.ld;
    float qa[D], qb[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        qa[d] = xa[d * a.dstride + i];
        qb[d] = xb[d * a.dstride + i];
    }
    // The row (self included, as the reference sorts it) and each lane's
    // k+1 smallest distances over the other frames.
    float h[K + 1];
#pragma unroll
    for (int m = 0; m <= K; ++m) h[m] = __builtin_inff();
    for (int j0 = lane; j0 < n; j0 += 4 * kWave) {
        float va[4][D], vb[4][D];
        load_frames<D, 4>(xa, xb, a.dstride, j0, n, va, vb);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int j = j0 + u * kWave;
            if (j < n) {
                KeyIdx r;
                r.key = fmaxf(marg<D>(qa, va[u]), marg<D>(qb, vb[u]));
                r.idx = j;
                row[j] = r;
                const float c = j == i ? __builtin_inff() : r.key;
#pragma unroll
                for (int m = K; m >= 1; --m) h[m] = __builtin_amdgcn_fmed3f(h[m - 1], c, h[m]);
                h[0] = fminf(h[0], c);
            }
        }
    }
    // Merge across the wave: afterwards every lane holds the row's k+1
    // smallest.
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        float o[K + 1];
#pragma unroll
        for (int m = 0; m <= K; ++m) o[m] = __shfl_xor(h[m], off);
#pragma unroll
        for (int t = 0; t <= K; ++t) {
#pragma unroll
            for (int m = K; m >= 1; --m) h[m] = __builtin_amdgcn_fmed3f(h[m - 1], o[t], h[m]);
            h[0] = fminf(h[0], o[t]);
        }
    }
    ksg::wave_sync();
    int jx = 0, jy = 0;
    if (h[K - 1] < h[K]) {
        float ex = -1.0f, ey = -1.0f;
        for (int j = lane; j < n; j += kWave) {
            if (j == i || !(row[j].key <= h[K - 1])) continue;
            float va[D], vb[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                va[d] = xa[d * a.dstride + j];
                vb[d] = xb[d * a.dstride + j];
            }
            const float da = marg<D>(qa, va), db = marg<D>(qb, vb);
            jx = da > ex ? j : jx;
            ex = fmaxf(ex, da);
            jy = db > ey ? j : jy;
            ey = fmaxf(ey, db);
        }
        // Any frame of the largest marginal distance gives the same rank.
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const float oex = __shfl_xor(ex, off), oey = __shfl_xor(ey, off);
            const int ojx = __shfl_xor(jx, off), ojy = __shfl_xor(jy, off);
            jx = oex > ex ? ojx : jx;
            ex = fmaxf(ex, oex);
            jy = oey > ey ? ojy : jy;
            ey = fmaxf(ey, oey);
        }
    } else {
        sort_prefix_wave(row, n, K + 1, lasc, rasc, stk, lane);
        float ex = -1.0f, ey = -1.0f;
        for (int m = 1; m <= K; ++m) {
            const int j = row[m].idx;
            float va[D], vb[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                va[d] = xa[d * a.dstride + j];
                vb[d] = xb[d * a.dstride + j];
            }
            const float da = marg<D>(qa, va), db = marg<D>(qb, vb);
            jx = da > ex ? j : jx;
            ex = fmaxf(ex, da);
            jy = db > ey ? j : jy;
            ey = fmaxf(ey, db);
        }
    }
    if (lane == 0) {
        const long long offA = table_row(a.slot_of[na], i, n, a.rs);
        const long long offB = table_row(a.slot_of[nb], i, n, a.rs);
        const unsigned rx = a.rank[table_at(offA, jx)], ry = a.rank[table_at(offB, jy)];
        a.cnt[pl * n + i] = (rx - 1u) | ((ry - 1u) << 16);
    }
    ksg::wave_sync();
}

// The k+1 smallest joint distances of query frame i over the frames j != i,
// exact (each lane its frames, one med3 per slot, then a butterfly merge:
// every lane ends with the row's k+1 smallest). When the k-th and (k+1)-th
// differ, the k nearest are the frames at or below the k-th: eps is their
// largest marginal distance and the counts come from the rank tables, as in
// resolve_tie; the function writes them and returns true. An exact tie
// returns false (its selection needs the whole row, in libstdc++ order).
template <int D, int K>
__device__ bool select_nearest(const WalkArgs &a, long long pl, int i, int lane) {
    const int n = a.n;
    const long long p = a.p0 + pl;
    const int na = a.ab[2 * p], nb = a.ab[2 * p + 1];
    const float *xa = a.X + (long long)na * a.ld;
    const float *xb = a.X + (long long)nb * a.ld;
    float qa[D], qb[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        qa[d] = xa[d * a.dstride + i];
        qb[d] = xb[d * a.dstride + i];
    }
    float h[K + 1];
#pragma unroll
    for (int m = 0; m <= K; ++m) h[m] = __builtin_inff();
    for (int j0 = lane; j0 < n; j0 += 4 * kWave) {
        float va[4][D], vb[4][D];
        load_frames<D, 4>(xa, xb, a.dstride, j0, n, va, vb);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int j = j0 + u * kWave;
            const float z = fmaxf(marg<D>(qa, va[u]), marg<D>(qb, vb[u]));
            const float c = (j == i || j >= n) ? __builtin_inff() : z;
#pragma unroll
            for (int m = K; m >= 1; --m) h[m] = __builtin_amdgcn_fmed3f(h[m - 1], c, h[m]);
            h[0] = fminf(h[0], c);
        }
    }
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        float o[K + 1];
#pragma unroll
        for (int m = 0; m <= K; ++m) o[m] = __shfl_xor(h[m], off);
#pragma unroll
        for (int t = 0; t <= K; ++t) {
#pragma unroll
            for (int m = K; m >= 1; --m) h[m] = __builtin_amdgcn_fmed3f(h[m - 1], o[t], h[m]);
            h[0] = fminf(h[0], o[t]);
        }
    }
    if (!(h[K - 1] < h[K])) return false;
    float ex = -1.0f, ey = -1.0f;
    int jx = 0, jy = 0;
    for (int j0 = lane; j0 < n; j0 += 4 * kWave) {
        float va[4][D], vb[4][D];
        load_frames<D, 4>(xa, xb, a.dstride, j0, n, va, vb);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int j = j0 + u * kWave;
            const float da = marg<D>(qa, va[u]), db = marg<D>(qb, vb[u]);
            const bool sel = j < n && j != i && fmaxf(da, db) <= h[K - 1];
            jx = sel && da > ex ? j : jx;
            ex = sel ? fmaxf(ex, da) : ex;
            jy = sel && db > ey ? j : jy;
            ey = sel ? fmaxf(ey, db) : ey;
        }
    }
    // Any frame of the largest marginal distance gives the same rank.
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const float oex = __shfl_xor(ex, off), oey = __shfl_xor(ey, off);
        const int ojx = __shfl_xor(jx, off), ojy = __shfl_xor(jy, off);
        jx = oex > ex ? ojx : jx;
        ex = fmaxf(ex, oex);
        jy = oey > ey ? ojy : jy;
        ey = fmaxf(ey, oey);
    }
    if (lane == 0) {
        const long long offA = table_row(a.slot_of[na], i, n, a.rs);
        const long long offB = table_row(a.slot_of[nb], i, n, a.rs);
        const unsigned rx = a.rank[table_at(offA, jx)], ry = a.rank[table_at(offB, jy)];
        a.cnt[pl * n + i] = (rx - 1u) | ((ry - 1u) << 16);
    }
    return true;
}

// Every entry of the tie list, one wave each and no LDS (so the CUs hold
// many waves): most entries are distance-code collisions, which
// select_nearest settles; the exact ties go to the exact list for
// tie_kernel.
template <int D, int K>
__global__ __launch_bounds__(256) void tie_select_kernel(WalkArgs a) {
    if (a.abort && *a.abort) return;
    const int lane = threadIdx.x & (kWave - 1);
    const long long wave = (long long)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    const long long nwaves = (long long)gridDim.x * (blockDim.x / kWave);
    const long long count = min((long long)*a.tie_count, (long long)a.tie_cap);
    for (long long e = wave; e < count; e += nwaves) {
        const unsigned v = a.tie_list[e];
        const long long pl = v / (unsigned)a.n;
        const bool done = select_nearest<D, K>(a, pl, (int)(v - pl * a.n), lane);
        if (!done && lane == 0) a.exact_list[atomicAdd(a.exact_count, 1u)] = v;
    }
}

// The exact ties of the tie list (kept by tie_select_kernel), one wave each,
// with tie_wave_lds_bytes(n) of LDS per wave (the row of n records and the
// partition's lists).
template <int D, int K>
__global__ __launch_bounds__(kWave * kTieMaxWaves) void tie_kernel(WalkArgs a) {
    extern __shared__ __attribute__((aligned(16))) char tie_lds[];
    if (a.abort && *a.abort) return;
    const int lane = threadIdx.x & (kWave - 1);
    const long long wave = (long long)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    const long long nwaves = (long long)gridDim.x * (blockDim.x / kWave);
    char *wl = tie_lds + (threadIdx.x / kWave) * tie_wave_lds_bytes(a.n);
    const long long count = (long long)*a.exact_count;
    for (long long e = wave; e < count; e += nwaves) {
        const unsigned v = a.exact_list[e];
        const long long pl = v / (unsigned)a.n;
        resolve_tie<D, K>(a, pl, (int)(v - pl * a.n), wl, lane);
    }
}

// The tied queries the tie list had no room for (kTieMark in cnt): nothing
// to do unless the list overflowed. Each wave scans 64-entry spans of cnt.
template <int D, int K>
__global__ __launch_bounds__(kWave * kTieMaxWaves) void tie_overflow_kernel(WalkArgs a) {
    if ((a.abort && *a.abort) || *a.tie_count <= a.tie_cap) return;
    extern __shared__ __attribute__((aligned(16))) char tie_lds[];
    const int lane = threadIdx.x & (kWave - 1);
    const long long wave = (long long)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    const long long nwaves = (long long)gridDim.x * (blockDim.x / kWave);
    char *wl = tie_lds + (threadIdx.x / kWave) * tie_wave_lds_bytes(a.n);
    const long long total = a.P * a.n;
    for (long long s0 = wave * kWave; s0 < total; s0 += nwaves * kWave) {
        const long long e = s0 + lane;
        unsigned long long marked = __ballot(e < total && a.cnt[e] == kTieMark);
        while (marked) {
            const int t = __ffsll((long long)marked) - 1;
            const long long q = s0 + t;
            const long long pl = q / a.n;
            resolve_tie<D, K>(a, pl, (int)(q - pl * a.n), wl, lane);
            marked &= marked - 1;
        }
    }
}

// ---- k above ksg::kMaxK: one wave per query frame ------------------------------
//
// The register heaps of the walk and of ksg_kernel.h hold at most kMaxK + 1
// entries; the reference selects any k (k + 1 argmin sweeps,
// mutual_information_gpu.cu:62-687). Here a wave takes one query (pair pl,
// frame i) at a time: it writes the joint distances z_j of the row to LDS,
// finds t = the float bits of the k-th smallest z_j over the other frames
// (j != i) by bisection on the bits (each step one pass of compares and a
// wave sum: z >= 0, so its bits order like its value), and counts the
// others at or below t. When that count is k, the k nearest frames are
// exactly those at or below t; when it exceeds k the k-th and (k+1)-th
// distances tie, and the row's positions 0..k are selected in libstdc++
// std::sort order (sort_prefix_wave, mutual_information_cpu.cpp:48-77).
// eps_x / eps_y are the largest marginal distances of the selected frames.
// The counts come from the rank tables when the tables exist (a.rank, the
// list path), else from one more pass over the row (the all-frames path).
// Coordinates are read from global memory; consecutive waves take
// consecutive frames of one pair, so its rows stay in cache.
NSX_HD inline size_t bigk_lds_bytes(int n) {
    return (((size_t)n * 4 + 15) & ~(size_t)15) + tie_wave_lds_bytes(n);
}

__device__ __forceinline__ unsigned wave_sum_u32(unsigned v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += (unsigned)__shfl_xor((int)v, off);
    return v;
}

template <int D>
__global__ __launch_bounds__(kWave) void bigk_kernel(WalkArgs a) {
    extern __shared__ __attribute__((aligned(16))) char bigk_lds[];
    if (a.abort && *a.abort) return;
    const int lane = threadIdx.x;
    const int n = a.n, k = a.k;
    float *zk = reinterpret_cast<float *>(bigk_lds);
    char *wl = bigk_lds + (((size_t)n * 4 + 15) & ~(size_t)15);
    KeyIdx *row = reinterpret_cast<KeyIdx *>(wl);
    uint16_t *lasc = reinterpret_cast<uint16_t *>(wl + (size_t)n * sizeof(KeyIdx));
    uint16_t *rasc = lasc + n;
    int *stk = reinterpret_cast<int *>(rasc + n);
    const long long total = a.P * n;
    unsigned long long ties = 0;
    for (long long w = blockIdx.x; w < total; w += gridDim.x) {
        const long long pl = w / n;
        const int i = (int)(w - pl * n);
        const long long p = a.p0 + pl;
        const int na = a.ab[2 * p], nb = a.ab[2 * p + 1];
        const float *xa = a.X + (long long)na * a.ld;
        const float *xb = a.X + (long long)nb * a.ld;
        float qa[D], qb[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            qa[d] = xa[d * a.dstride + i];
            qb[d] = xb[d * a.dstride + i];
        }
        for (int j = lane; j < n; j += kWave) {
            float va[D], vb[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                va[d] = xa[d * a.dstride + j];
                vb[d] = xb[d * a.dstride + j];
            }
            zk[j] = fmaxf(marg<D>(qa, va), marg<D>(qb, vb));
        }
        ksg::wave_sync();
        // Smallest t with #{j != i : bits(z_j) <= t} >= k (k <= n - 1 others).
        unsigned lo = 0u, hi = 0xffffffffu;
        while (lo < hi) {
            const unsigned mid = lo + ((hi - lo) >> 1);
            unsigned c = 0;
            for (int j = lane; j < n; j += kWave)
                c += (j != i && __float_as_uint(zk[j]) <= mid) ? 1u : 0u;
            c = wave_sum_u32(c);
            if (c >= (unsigned)k) hi = mid;
            else lo = mid + 1u;
        }
        const unsigned t = lo;
        unsigned c_le = 0;
        for (int j = lane; j < n; j += kWave)
            c_le += (j != i && __float_as_uint(zk[j]) <= t) ? 1u : 0u;
        c_le = wave_sum_u32(c_le);
        const bool tie = c_le > (unsigned)k;
        ties += tie ? 1u : 0u;
        float ex = -1.0f, ey = -1.0f;
        int jx = 0, jy = 0;
        auto take = [&](int j) {
            float va[D], vb[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                va[d] = xa[d * a.dstride + j];
                vb[d] = xb[d * a.dstride + j];
            }
            const float da = marg<D>(qa, va), db = marg<D>(qb, vb);
            jx = da > ex ? j : jx;
            ex = fmaxf(ex, da);
            jy = db > ey ? j : jy;
            ey = fmaxf(ey, db);
        };
        if (!tie) {
            for (int j = lane; j < n; j += kWave)
                if (j != i && __float_as_uint(zk[j]) <= t) take(j);
        } else {
            // The row as the reference sorts it (self included).
            for (int j = lane; j < n; j += kWave) {
                KeyIdx r;
                r.key = zk[j];
                r.idx = j;
                row[j] = r;
            }
            ksg::wave_sync();
            sort_prefix_wave(row, n, k + 1, lasc, rasc, stk, lane);
            for (int m = 1 + lane; m <= k; m += kWave) take(row[m].idx);
        }
        // Largest marginal distances over the wave (any frame of the largest
        // distance gives the same rank).
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const float oex = __shfl_xor(ex, off), oey = __shfl_xor(ey, off);
            const int ojx = __shfl_xor(jx, off), ojy = __shfl_xor(jy, off);
            jx = oex > ex ? ojx : jx;
            ex = fmaxf(ex, oex);
            jy = oey > ey ? ojy : jy;
            ey = fmaxf(ey, oey);
        }
        unsigned nx, ny;
        if (a.rank) {
            const long long offA = table_row(a.slot_of[na], i, n, a.rs);
            const long long offB = table_row(a.slot_of[nb], i, n, a.rs);
            nx = (unsigned)a.rank[table_at(offA, jx)] - 1u;
            ny = (unsigned)a.rank[table_at(offB, jy)] - 1u;
        } else {
            // <= eps over every frame, self included, minus one
            // (mutual_information_cpu.cpp:94-109).
            unsigned cx = 0, cy = 0;
            for (int j = lane; j < n; j += kWave) {
                float va[D], vb[D];
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    va[d] = xa[d * a.dstride + j];
                    vb[d] = xb[d * a.dstride + j];
                }
                cx += marg<D>(qa, va) <= ex ? 1u : 0u;
                cy += marg<D>(qb, vb) <= ey ? 1u : 0u;
            }
            nx = wave_sum_u32(cx) - 1u;
            ny = wave_sum_u32(cy) - 1u;
        }
        if (lane == 0) a.cnt[pl * n + i] = nx | (ny << 16);
        ksg::wave_sync();
    }
    if (lane == 0) atomicAdd(a.stats, ties);
}

template <int D, int K>
struct WalkInstances {
    static void fill(WalkFn *t, TieFn *s, TieFn *u, TieFn *o) {
        t[K] = walk_kernel<D, K>;
        s[K] = tie_select_kernel<D, K>;
        u[K] = tie_kernel<D, K>;
        o[K] = tie_overflow_kernel<D, K>;
        WalkInstances<D, K - 1>::fill(t, s, u, o);
    }
};
template <int D>
struct WalkInstances<D, 0> {
    static void fill(WalkFn *, TieFn *, TieFn *, TieFn *) {}
};

template <int D>
struct WalkTable {
    WalkFn t[ksg::kMaxK + 1] = {};
    TieFn s[ksg::kMaxK + 1] = {};
    TieFn u[ksg::kMaxK + 1] = {};
    TieFn o[ksg::kMaxK + 1] = {};
    WalkTable() { WalkInstances<D, ksg::kMaxK>::fill(t, s, u, o); }
};

template <int D>
const WalkTable<D> &walk_table() {
    static const WalkTable<D> tab;
    return tab;
}

template <int D>
WalkFn walk_lookup(int k) {
    return (k >= 1 && k <= ksg::kMaxK) ? walk_table<D>().t[k] : nullptr;
}

template <int D>
TieFn tie_lookup(int k, int kind) {
    if (k < 1 || k > ksg::kMaxK) return nullptr;
    return kind == kTieSelect ? walk_table<D>().s[k]
                              : (kind == kTieOverflow ? walk_table<D>().o[k] : walk_table<D>().u[k]);
}

template <int D>
BuildFn build_lookup(int width) {
    switch (width) {
        case 8: return list_build_kernel<D, 8>;
        case 16: return list_build_kernel<D, 16>;
        case 32: return list_build_kernel<D, 32>;
        case 64: return list_build_wide_kernel<D, 2>;
        case 96: return list_build_wide_kernel<D, 3>;
        case 128: return list_build_wide_kernel<D, 4>;
        default: return nullptr;
    }
}

#endif  // __HIPCC__

}  // namespace ksgw
}  // namespace nsx
