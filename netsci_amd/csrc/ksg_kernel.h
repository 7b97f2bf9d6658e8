// KSG (Kraskov algorithm 2) mutual-information kernel for gfx950.
//
// Reference semantics: netcalc/src/mutual_information_cpu.cpp:6-121 (the CPU
// path, which fixes tie order through std::sort); the GPU driver it replaces
// is netcalc/src/mutual_information_gpu.cu:1007-1079. Output equals the
// reference CPU path bit-for-bit: the same float32 distances and `<=`
// counts, the k-th/(k+1)-th tie resolved in libstdc++ order, and the psi
// epilogue summed sequentially in frame order.
//
// Mapping (one persistent launch for every (window, pair) item):
//   * a wavefront owns one item at a time (items are dealt round-robin to
//     the resident waves, longest windows first) and walks its query frames
//     in blocks of 64*QPL (lane l owns queries q0 + l and, for QPL = 2,
//     q0 + 64 + l);
//   * the item's frames stream through a wave-private LDS tile of `chunk`
//     frames (SoA per coordinate, NaN-padded) and are read back as 16-byte
//     broadcasts, so every lane compares its queries against the same j;
//   * pass 1 keeps, per query, the k+1 smallest joint max-norm distances to
//     frames j != i in registers (sorted; insertion is one v_med3_f32 plus
//     index selects per slot, taken only when some lane's distance beats its
//     current (k+1)-th);
//   * a query whose k-th and (k+1)-th distances are equal rebuilds its
//     distance row in scratch and selects the k nearest in libstdc++
//     std::sort order (sort_select.h);
//   * eps_x / eps_y come from the k selected frames; pass 2 counts
//     dx <= eps_x and dy <= eps_y over all frames (self included, then minus
//     one, as the reference does);
//   * the wave adds psi[n_x] + psi[n_y] over its queries in frame order.
// No MFMA: the work is compare/max/select, not a contraction. The roofline
// is FP32 VALU.
#pragma once

#include <hip/hip_runtime.h>

#include "sort_select.h"

namespace nsx {
namespace ksg {

constexpr int kWave = 64;
constexpr int kMaxK = 16;
constexpr int kMaxChunk = 256;
constexpr int kMaxWavesPerBlock = 8;

struct Args {
    const float *X;
    long long ld;        // row stride (floats)
    long long dstride;   // stride between coordinate dimensions (floats)
    const int *ab;       // P x 2 node indices
    long long P;
    const int *win_off;  // per processing slot: first frame
    const int *win_n;    // per processing slot: frame count
    const int *win_rep;  // per processing slot: output replicate index
    int R;
    int k;
    const float *psi;    // psi[0..max n]
    float inv_k;         // (float)(1.0 / (float)k)
    float *mi_out;       // [rep][P]
    int *counts_out;     // nullable: [p][frame][2] (single window only)
    KeyIdx *scratch;     // per wave: scratch_stride records
    long long scratch_stride;
    long long scratch_n; // records per scratch region (max window length)
    unsigned long long *stats;  // [0] tie frames
    int chunk;           // frames per LDS tile: 64, 128 or 256
    int wpb;             // waves per workgroup
    const unsigned *run_if;  // nullable: the launch does nothing unless *run_if != 0
                             // (the fallback a list-path graph carries for
                             // non-finite coordinates)
};

using KernelFn = void (*)(Args);

// LDS floats of one wave: the tile (2d rows of kMaxChunk), the psi terms
// and, for the symmetric pass 2, the tile frames' eps_x and eps_y.
constexpr int lds_floats_per_wave(int d, int qpl, bool sym) {
    return 2 * d * kMaxChunk + kWave * qpl + (sym ? 2 * kMaxChunk : 0);
}

// Dynamic LDS bytes a launch needs.
inline size_t lds_bytes(int d, int qpl, int wpb, bool sym) {
    return (size_t)wpb * (size_t)lds_floats_per_wave(d, qpl, sym) * sizeof(float);
}

// Scratch records (KeyIdx) per wave: the tie row, plus eps and counts for
// the symmetric pass 2.
inline long long scratch_records(long long nmax, bool sym) { return sym ? 3 * nmax : nmax; }

// Kernel instances, one table per coordinate dimension (ksg_d{1,2,3}.hip).
KernelFn kernel_d1(int k, int qpl, bool sym);
KernelFn kernel_d2(int k, int qpl, bool sym);
KernelFn kernel_d3(int k, int qpl, bool sym);

#if defined(__HIPCC__)

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void wave_sync_global() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Marginal max-norm distance |q - x|_inf over D coordinates.
template <int D>
__device__ __forceinline__ float marg(const float (&q)[D], const float (&x)[D]) {
    float r = fabsf(q[0] - x[0]);
#pragma unroll
    for (int d = 1; d < D; ++d) r = fmaxf(r, fabsf(q[d] - x[d]));
    return r;
}

// Marginal distance from query q to frame j read from global memory.
template <int D>
__device__ __forceinline__ float marg_global(const float (&q)[D], const float *x,
                                             long long ds, int j) {
    float v[D];
#pragma unroll
    for (int d = 0; d < D; ++d) v[d] = x[d * ds + j];
    return marg<D>(q, v);
}

// Insert (c, j) into the ascending heap h[0..K]. With c >= h[K] (+inf in
// particular) it changes nothing: no compare succeeds and every med3 returns
// the slot's own value.
template <int K>
__device__ __forceinline__ void heap_insert(float (&h)[K + 1], int (&id)[K + 1], float c,
                                            int j) {
#pragma unroll
    for (int m = K; m >= 1; --m) {
        const bool lp = c < h[m - 1];
        const bool lc = c < h[m];
        id[m] = lp ? id[m - 1] : (lc ? j : id[m]);
        h[m] = __builtin_amdgcn_fmed3f(h[m - 1], c, h[m]);
    }
    const bool l0 = c < h[0];
    id[0] = l0 ? j : id[0];
    h[0] = l0 ? c : h[0];
}

// A wave's LDS tile: 2D rows (a-coordinates then b-coordinates) of a fixed
// stride kMaxChunk, so every row address is the tile base plus an immediate
// offset; `chunk` (64, 128 or 256) frames are staged per pass.
template <int D>
struct Tile {
    float *s;
    int chunk;
    __device__ __forceinline__ float4 get(int dim, int jj) const {
        return *reinterpret_cast<const float4 *>(s + dim * kMaxChunk + jj);
    }
};

// Stage frames [c0, c0 + chunk) of both rows into the wave's LDS tile.
// Frames at or beyond n are NaN so that every comparison with them fails.
template <int D>
__device__ __forceinline__ void stage_tile(const Tile<D> &t, const float *xa, const float *xb,
                                           long long ds, int c0, int n, int lane) {
    const float qnan = __builtin_nanf("");
    for (int h = 0; h < t.chunk; h += kWave) {
        const int f = c0 + h + lane;
        const bool ok = f < n;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const float va = ok ? xa[d * ds + f] : qnan;
            const float vb = ok ? xb[d * ds + f] : qnan;
            t.s[d * kMaxChunk + h + lane] = va;
            t.s[(D + d) * kMaxChunk + h + lane] = vb;
        }
    }
    wave_sync();
}

template <int D>
__device__ __forceinline__ void unpack(const float4 (&A)[D], const float4 (&B)[D], int u,
                                       float (&xa)[D], float (&xb)[D]) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
        xa[d] = u == 0 ? A[d].x : u == 1 ? A[d].y : u == 2 ? A[d].z : A[d].w;
        xb[d] = u == 0 ? B[d].x : u == 1 ? B[d].y : u == 2 ? B[d].z : B[d].w;
    }
}

// Pass 1 over one staged tile: keep the k+1 nearest frames j != i per query.
template <int D, int K, int QPL, bool kSelf>
__device__ __forceinline__ void pass1_tile(const Tile<D> &t, int cnt4, int c0,
                                           const int (&i)[QPL], const float (&qa)[QPL][D],
                                           const float (&qb)[QPL][D], float (&h)[QPL][K + 1],
                                           int (&id)[QPL][K + 1]) {
    for (int jj = 0; jj < cnt4; jj += 4) {
        float4 A[D], B[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            A[d] = t.get(d, jj);
            B[d] = t.get(D + d, jj);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            float xa[D], xb[D];
            unpack<D>(A, B, u, xa, xb);
            const int j = c0 + jj + u;
#pragma unroll
            for (int q = 0; q < QPL; ++q) {
                float dz = fmaxf(marg<D>(qa[q], xa), marg<D>(qb[q], xb));
                if (kSelf) dz = (j == i[q]) ? __builtin_nanf("") : dz;
                if (dz < h[q][K]) heap_insert<K>(h[q], id[q], dz, j);
            }
        }
    }
}

// mask * 2 + cond in one v_addc_co_u32: the compare's lane mask (a ballot,
// which is the compare's own SGPR result) is the carry-in. Written out
// because the compiler otherwise regroups the shift-ins into a select, a
// shift and a three-input or per frame (`mask + mask + (c ? 1 : 0)` and
// __builtin_addc both end there): 770 -> 709 VALU instructions per 32-frame
// block of two queries, C1 launch 1,352 -> 1,311 ms, identical results.
__device__ __forceinline__ unsigned shift_in(unsigned mask, bool cond) {
    const unsigned long long lanes = __builtin_amdgcn_ballot_w64(cond);
    unsigned out;
    unsigned long long carry;
    asm("v_addc_co_u32_e64 %0, %1, %2, %2, %3" : "=v"(out), "=s"(carry) : "v"(mask), "s"(lanes));
    return out;
}

// Pass 1 with deferred insertion. For each block of 32 frames every lane
// first records, in a bit mask, the frames whose joint distance beats its
// (k+1)-th nearest at the start of the block (one compare and one
// add-with-carry per frame); then the wave inserts the flagged frames in
// rounds, each lane taking its next flagged frame, re-reading that frame
// from LDS and inserting it if it still beats the heap. A lane's flagged
// frames are a superset of those the per-frame scan would insert, and the
// heap's content does not depend on insertion order except for which of
// several frames tied at the (k+1)-th distance stays — a case the tie check
// after pass 1 hands to the exact path anyway.
template <int D, int K, int QPL, bool kSelf>
__device__ __forceinline__ void pass1_tile_deferred(const Tile<D> &t, int cnt, int c0,
                                                    const int (&i)[QPL],
                                                    const float (&qa)[QPL][D],
                                                    const float (&qb)[QPL][D],
                                                    float (&h)[QPL][K + 1],
                                                    int (&id)[QPL][K + 1]) {
    for (int jb = 0; jb < cnt; jb += 32) {
        unsigned mask[QPL];
        float thr[QPL];
#pragma unroll
        for (int q = 0; q < QPL; ++q) {
            mask[q] = 0u;
            thr[q] = h[q][K];
        }
#pragma unroll
        for (int j4 = 0; j4 < 32; j4 += 4) {
            float4 A[D], B[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                A[d] = t.get(d, jb + j4);
                B[d] = t.get(D + d, jb + j4);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                float xa[D], xb[D];
                unpack<D>(A, B, u, xa, xb);
#pragma unroll
                for (int q = 0; q < QPL; ++q) {
                    float dz = fmaxf(marg<D>(qa[q], xa), marg<D>(qb[q], xb));
                    if (kSelf) dz = (c0 + jb + j4 + u == i[q]) ? __builtin_nanf("") : dz;
                    // bit 31 - (j4 + u) marks frame jb + j4 + u
                    mask[q] = shift_in(mask[q], dz < thr[q]);
                }
            }
        }
        // (Measured on MI355X: this divergent form beats a straight-line round
        // that inserts +inf for idle lanes, 1,351 vs 1,365 ms at N=1000,
        // n=2000, and one that skips the insertion by ballot, 1,407 ms.)
#pragma unroll
        for (int q = 0; q < QPL; ++q) {
            while (__ballot(mask[q] != 0u)) {
                if (mask[q] != 0u) {
                    const int u = __clz(mask[q]);  // lowest pending frame first
                    mask[q] &= ~(0x80000000u >> u);
                    float xa[D], xb[D];
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        xa[d] = t.s[d * kMaxChunk + jb + u];
                        xb[d] = t.s[(D + d) * kMaxChunk + jb + u];
                    }
                    const float dz = fmaxf(marg<D>(qa[q], xa), marg<D>(qb[q], xb));
                    if (dz < h[q][K]) heap_insert<K>(h[q], id[q], dz, c0 + jb + u);
                }
            }
        }
    }
}

// Pass 2 over the tile frames [lo, hi): marginal counts within eps (self
// included).
template <int D, int QPL>
__device__ __forceinline__ void pass2_tile(const Tile<D> &t, int lo, int hi,
                                           const float (&qa)[QPL][D], const float (&qb)[QPL][D],
                                           const float (&ex)[QPL], const float (&ey)[QPL],
                                           int (&cx)[QPL], int (&cy)[QPL]) {
    for (int jj = lo; jj < hi; jj += 4) {
        float4 A[D], B[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            A[d] = t.get(d, jj);
            B[d] = t.get(D + d, jj);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            float xa[D], xb[D];
            unpack<D>(A, B, u, xa, xb);
#pragma unroll
            for (int q = 0; q < QPL; ++q) {
                cx[q] += marg<D>(qa[q], xa) <= ex[q] ? 1 : 0;
                cy[q] += marg<D>(qb[q], xb) <= ey[q] ? 1 : 0;
            }
        }
    }
}

// Symmetric pass 2 over the tile frames [lo, hi) (a multiple of 64 apart,
// all after the wave's query block): for query i (lane) and frame j
// (wave-uniform) one evaluation of dx = |x_i - x_j|_inf serves both
//   n_x(i) += [dx <= eps_x(i)]  (lane counter), and
//   n_x(j) += [dx <= eps_x(j)]  summed over the wave's valid queries by
//                               ballot + popcount,
// because |fl(a - b)| = |fl(b - a)|. Frame j's wave sum goes into lane
// j % 64 of vx / vy (a lane-compare select) and every 64 frames one
// coalesced atomic adds them to the item's count arrays Cx / Cy. NaN-padded frames (>= n)
// fail every compare; their lanes are not flushed.
template <int D, int QPL>
__device__ __forceinline__ void pass2_sym(const Tile<D> &t, const float *tex, const float *tey,
                                          int lo, int hi, int s, int n, int lane,
                                          const bool (&valid)[QPL], const float (&qa)[QPL][D],
                                          const float (&qb)[QPL][D], const float (&ex)[QPL],
                                          const float (&ey)[QPL], int (&cx)[QPL],
                                          int (&cy)[QPL], int *Cx, int *Cy) {
    for (int g = lo; g < hi; g += kWave) {
        int vx = 0, vy = 0;
        for (int jj = g; jj < g + kWave; jj += 4) {
            float4 A[D], B[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                A[d] = t.get(d, jj);
                B[d] = t.get(D + d, jj);
            }
            const float4 EX = *reinterpret_cast<const float4 *>(tex + jj);
            const float4 EY = *reinterpret_cast<const float4 *>(tey + jj);
            const float exj[4] = {EX.x, EX.y, EX.z, EX.w};
            const float eyj[4] = {EY.x, EY.y, EY.z, EY.w};
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                float xa[D], xb[D];
                unpack<D>(A, B, u, xa, xb);
                int nx = 0, ny = 0;
#pragma unroll
                for (int q = 0; q < QPL; ++q) {
                    const float da = marg<D>(qa[q], xa), db = marg<D>(qb[q], xb);
                    cx[q] += da <= ex[q] ? 1 : 0;
                    cy[q] += db <= ey[q] ? 1 : 0;
                    nx += __popcll(__ballot(valid[q] && da <= exj[u]));
                    ny += __popcll(__ballot(valid[q] && db <= eyj[u]));
                }
                const bool mine = lane == ((jj + u) & (kWave - 1));
                vx = mine ? nx : vx;
                vy = mine ? ny : vy;
            }
        }
        const int f = s + g + lane;
        if (f < n) {
            if (vx) atomicAdd(Cx + f, vx);
            if (vy) atomicAdd(Cy + f, vy);
        }
    }
}

// Resolve the eps of every lane flagged `tie`: rebuild the frame's joint
// distance row into scratch and select the k+1 first entries in libstdc++
// std::sort order, then take eps over positions 1..k
// (mutual_information_cpu.cpp:48-77). Returns the number of tie frames.
template <int D>
__device__ unsigned resolve_ties(bool tie, const float (&qa)[D], const float (&qb)[D],
                                 const float *xa, const float *xb, long long ds, int n, int k,
                                 KeyIdx *scratch, int lane, float &ex, float &ey) {
    unsigned long long tmask = __ballot(tie);
    const unsigned count = (unsigned)__popcll(tmask);
    while (tmask) {
        const int t = __builtin_amdgcn_readfirstlane(__ffsll((long long)tmask) - 1);
        float tqa[D], tqb[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            tqa[d] = __shfl(qa[d], t);
            tqb[d] = __shfl(qb[d], t);
        }
        for (int j = lane; j < n; j += kWave) {
            float va[D], vb[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                va[d] = xa[d * ds + j];
                vb[d] = xb[d * ds + j];
            }
            KeyIdx r;
            r.key = fmaxf(marg<D>(tqa, va), marg<D>(tqb, vb));
            r.idx = j;
            scratch[j] = r;
        }
        wave_sync_global();
        float rea = 0.0f, reb = 0.0f;
        if (lane == 0) {
            sort_prefix(scratch, n, k + 1);
            for (int m = 1; m <= k; ++m) {
                const int j = scratch[m].idx;
                rea = fmaxf(rea, marg_global<D>(tqa, xa, ds, j));
                reb = fmaxf(reb, marg_global<D>(tqb, xb, ds, j));
            }
        }
        rea = __shfl(rea, 0);
        reb = __shfl(reb, 0);
        ex = (lane == t) ? rea : ex;
        ey = (lane == t) ? reb : ey;
        wave_sync_global();
        tmask &= tmask - 1;
    }
    return count;
}

// Stage the eps of tile frames [c0, c0 + chunk) (NaN beyond n) next to the
// coordinates, for the symmetric pass 2.
__device__ __forceinline__ void stage_eps(float *tex, float *tey, const float *Ex,
                                          const float *Ey, int c0, int chunk, int n, int lane) {
    const float qnan = __builtin_nanf("");
    for (int h = 0; h < chunk; h += kWave) {
        const int f = c0 + h + lane;
        tex[h + lane] = f < n ? Ex[f] : qnan;
        tey[h + lane] = f < n ? Ey[f] : qnan;
    }
}

__device__ __forceinline__ void fence_agent() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
    __builtin_amdgcn_wave_barrier();
}

// Every wave walks items w = wave_gid, wave_gid + nwaves, ... All loop
// bounds and item fields are wave-uniform scalars; the only lane-divergent
// regions are the heap insertion, the tie rebuild's single-lane selection
// and the per-lane loads/stores.
//
// kSym = false: per query block, pass 1 then pass 2 over all frames, and the
// block's psi terms are summed at once.
// kSym = true: pass 1 and eps for every block first (eps to the item's
// scratch, counts zeroed), then pass 2 per block over its own frames and
// the frames after it only, with the symmetric count of pass2_sym, then the
// psi terms of all frames in frame order. Per item the wave's scratch holds
// [tie records: n | eps_x, eps_y: n floats each | n_x, n_y: n ints each].
template <int D, int K, int QPL, bool kSym>
__global__ __launch_bounds__(kWave *kMaxWavesPerBlock) void ksg_kernel(Args a) {
    extern __shared__ __attribute__((aligned(16))) float dyn_lds[];
    if (a.run_if && *a.run_if == 0u) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int chunk = a.chunk;
    float *wave_lds = dyn_lds + wib * lds_floats_per_wave(D, QPL, kSym);
    const Tile<D> tile{wave_lds, chunk};
    float *terms = wave_lds + 2 * D * kMaxChunk;
    float *tex = terms + kWave * QPL;
    float *tey = tex + kMaxChunk;
    const long long nwaves = (long long)gridDim.x * a.wpb;
    const long long wave_gid = (long long)blockIdx.x * a.wpb + wib;
    KeyIdx *scratch = a.scratch + wave_gid * a.scratch_stride;
    const long long total = (long long)a.R * a.P;
    const long long ds = a.dstride;
    constexpr int kBlock = kWave * QPL;
    unsigned long long ties = 0;

    for (long long w = wave_gid; w < total; w += nwaves) {
        const int slot = (int)(w / a.P);
        const long long p = w - (long long)slot * a.P;
        const int off = a.win_off[slot];
        const int n = a.win_n[slot];
        const float *xa = a.X + (long long)a.ab[2 * p] * a.ld + off;
        const float *xb = a.X + (long long)a.ab[2 * p + 1] * a.ld + off;
        float *Ex = reinterpret_cast<float *>(scratch + a.scratch_n);
        float *Ey = Ex + a.scratch_n;
        int *Cx = reinterpret_cast<int *>(scratch + 2 * a.scratch_n);
        int *Cy = Cx + a.scratch_n;

        float acc = 0.0f;
        for (int q0 = 0; q0 < n; q0 += kBlock) {
            int i[QPL];
            bool valid[QPL];
            float qa[QPL][D], qb[QPL][D];
            float h[QPL][K + 1];
            int id[QPL][K + 1];
#pragma unroll
            for (int q = 0; q < QPL; ++q) {
                i[q] = q0 + q * kWave + lane;
                valid[q] = i[q] < n;
                const int iq = valid[q] ? i[q] : 0;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    qa[q][d] = xa[d * ds + iq];
                    qb[q][d] = xb[d * ds + iq];
                }
#pragma unroll
                for (int m = 0; m <= K; ++m) {
                    h[q][m] = __builtin_inff();
                    id[q][m] = 0;
                }
            }
            // Pass 1: k+1 nearest by joint distance.
            for (int c0 = 0; c0 < n; c0 += chunk) {
                stage_tile<D>(tile, xa, xb, ds, c0, n, lane);
                const bool self = (c0 < q0 + kBlock) && (q0 < c0 + chunk);
                const int cnt = min(chunk, n - c0);
                if (self) pass1_tile_deferred<D, K, QPL, true>(tile, cnt, c0, i, qa, qb, h, id);
                else pass1_tile_deferred<D, K, QPL, false>(tile, cnt, c0, i, qa, qb, h, id);
                wave_sync();
            }
            // eps over the k nearest; boundary ties take the exact path.
            float ex[QPL], ey[QPL];
#pragma unroll
            for (int q = 0; q < QPL; ++q) {
                ex[q] = 0.0f;
                ey[q] = 0.0f;
#pragma unroll
                for (int m = 0; m < K; ++m) {
                    ex[q] = fmaxf(ex[q], marg_global<D>(qa[q], xa, ds, id[q][m]));
                    ey[q] = fmaxf(ey[q], marg_global<D>(qb[q], xb, ds, id[q][m]));
                }
                const bool tie = valid[q] && (h[q][K - 1] == h[q][K]);
                ties += resolve_ties<D>(tie, qa[q], qb[q], xa, xb, ds, n, K, scratch, lane,
                                        ex[q], ey[q]);
            }
            if (kSym) {
#pragma unroll
                for (int q = 0; q < QPL; ++q)
                    if (valid[q]) {
                        Ex[i[q]] = ex[q];
                        Ey[i[q]] = ey[q];
                        Cx[i[q]] = 0;
                        Cy[i[q]] = 0;
                    }
                continue;
            }
            // Pass 2: marginal counts.
            int cx[QPL], cy[QPL];
#pragma unroll
            for (int q = 0; q < QPL; ++q) cx[q] = cy[q] = 0;
            for (int c0 = 0; c0 < n; c0 += chunk) {
                stage_tile<D>(tile, xa, xb, ds, c0, n, lane);
                const int cnt4 = (min(chunk, n - c0) + 3) & ~3;
                pass2_tile<D, QPL>(tile, 0, cnt4, qa, qb, ex, ey, cx, cy);
                wave_sync();
            }
#pragma unroll
            for (int q = 0; q < QPL; ++q) {
                const int nx = cx[q] - 1, ny = cy[q] - 1;
                terms[q * kWave + lane] = valid[q] ? a.psi[nx] + a.psi[ny] : 0.0f;
                if (a.counts_out && valid[q]) {
                    a.counts_out[(p * n + i[q]) * 2 + 0] = nx;
                    a.counts_out[(p * n + i[q]) * 2 + 1] = ny;
                }
            }
            wave_sync();
            // Sequential float32 sum in frame order (every lane runs the same
            // sum over LDS broadcasts).
            const int lim = min(kBlock, n - q0);
            for (int l = 0; l < lim; ++l) acc += terms[l];
            wave_sync();
        }
        if (kSym) {
            fence_agent();  // eps and zeroed counts visible to every lane
            for (int q0 = 0; q0 < n; q0 += kBlock) {
                int i[QPL];
                bool valid[QPL];
                float qa[QPL][D], qb[QPL][D], ex[QPL], ey[QPL];
                int cx[QPL], cy[QPL];
#pragma unroll
                for (int q = 0; q < QPL; ++q) {
                    i[q] = q0 + q * kWave + lane;
                    valid[q] = i[q] < n;
                    const int iq = valid[q] ? i[q] : 0;
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        qa[q][d] = xa[d * ds + iq];
                        qb[q][d] = xb[d * ds + iq];
                    }
                    ex[q] = Ex[iq];
                    ey[q] = Ey[iq];
                    cx[q] = cy[q] = 0;
                }
                // Frames of the own block count for the lane only; frames
                // after it count both ways; frames before it were counted
                // by earlier blocks.
                const int own_end = min(q0 + kBlock, n);
                for (int c0 = q0; c0 < n; c0 += chunk) {
                    stage_tile<D>(tile, xa, xb, ds, c0, n, lane);
                    const int real = min(chunk, n - c0);            // frames < n in the tile
                    const int own = max(0, min(own_end - c0, chunk));  // own-block frames
                    // Own frames: the lane's count only. Rounded up to 4 into
                    // the NaN padding, which counts nothing (own < 4-multiple
                    // only when the own block ends at n).
                    if (own > 0)
                        pass2_tile<D, QPL>(tile, 0, (own + 3) & ~3, qa, qb, ex, ey, cx, cy);
                    // Frames after the own block: both ways. Here own_end < n,
                    // so own_end, c0 and hence `own` are multiples of 64, and
                    // the 64-frame groups end inside the tile.
                    if (real > own) {
                        stage_eps(tex, tey, Ex, Ey, c0, chunk, n, lane);
                        wave_sync();
                        pass2_sym<D, QPL>(tile, tex, tey, own, (real + kWave - 1) & ~(kWave - 1),
                                          c0, n, lane, valid, qa, qb, ex, ey, cx, cy, Cx, Cy);
                    }
                    wave_sync();
                }
#pragma unroll
                for (int q = 0; q < QPL; ++q)
                    if (valid[q]) {
                        atomicAdd(Cx + i[q], cx[q]);
                        atomicAdd(Cy + i[q], cy[q]);
                    }
            }
            fence_agent();  // every count complete
            for (int f0 = 0; f0 < n; f0 += kWave) {
                const int f = f0 + lane;
                float term = 0.0f;
                if (f < n) {
                    const int nx = Cx[f] - 1, ny = Cy[f] - 1;
                    term = a.psi[nx] + a.psi[ny];
                    if (a.counts_out) {
                        a.counts_out[(p * n + f) * 2 + 0] = nx;
                        a.counts_out[(p * n + f) * 2 + 1] = ny;
                    }
                }
                terms[lane] = term;
                wave_sync();
                const int lim = min(kWave, n - f0);
                for (int l = 0; l < lim; ++l) acc += terms[l];
                wave_sync();
            }
        }
        const float avg = acc / (float)n;
        const float mi = a.psi[n] + a.psi[a.k] - a.inv_k - avg;
        if (lane == 0) a.mi_out[(long long)a.win_rep[slot] * a.P + p] = mi;
    }
    if (lane == 0 && ties) atomicAdd(a.stats, ties);
}

template <int D, int K>
struct Instances {
    static void fill(KernelFn (*t)[3][2]) {
        t[K][1][0] = ksg_kernel<D, K, 1, false>;
        t[K][2][0] = ksg_kernel<D, K, 2, false>;
        t[K][1][1] = ksg_kernel<D, K, 1, true>;
        t[K][2][1] = ksg_kernel<D, K, 2, true>;
        Instances<D, K - 1>::fill(t);
    }
};
template <int D>
struct Instances<D, 0> {
    static void fill(KernelFn (*)[3][2]) {}
};

// Table lookup shared by the per-dimension translation units.
template <int D>
struct InstanceTable {
    KernelFn t[kMaxK + 1][3][2] = {};
    InstanceTable() { Instances<D, kMaxK>::fill(t); }
};

template <int D>
KernelFn kernel_lookup(int k, int qpl, bool sym) {
    static const InstanceTable<D> tab;  // thread-safe static initialisation
    if (k < 1 || k > kMaxK || qpl < 1 || qpl > 2) return nullptr;
    return tab.t[k][qpl][sym ? 1 : 0];
}

#endif  // __HIPCC__

}  // namespace ksg
}  // namespace nsx
