// KSG (Kraskov algorithm 2) mutual information over node pairs on gfx950.
//
// Reference semantics: netcalc/src/mutual_information_cpu.cpp:6-121 (the
// CPU path, which fixes tie order through std::sort) and the GPU driver
// netcalc/src/mutual_information_gpu.cu:1007-1079. Output equals the
// reference CPU path bit-for-bit: the same float32 distances and `<=`
// counts, the k-th/(k+1)-th tie resolved in libstdc++ order, and the psi
// epilogue summed sequentially in frame order.
//
// Mapping (one persistent launch for every (window, pair) item):
//   * a wavefront owns one item at a time (items are dealt round-robin to
//     the resident waves, longest windows first) and walks its query frames
//     in blocks of 64 (lane = query frame);
//   * the item's frames stream through a wave-private LDS chunk (SoA per
//     coordinate, NaN-padded) and are read back as 16-byte broadcasts, so
//     every lane compares its query against the same frame j;
//   * pass 1 keeps, per lane, the k+1 smallest joint max-norm distances to
//     frames j != i in registers (sorted; insertion is one v_med3_f32 plus
//     index selects per slot, taken only when some lane's distance beats its
//     current (k+1)-th);
//   * a lane whose k-th and (k+1)-th distances are equal rebuilds the
//     frame's distance row in scratch and selects the k nearest in
//     libstdc++ std::sort order (sort_select.h);
//   * eps_x / eps_y come from the k selected frames; pass 2 counts
//     dx <= eps_x and dy <= eps_y over all frames (self included, then minus
//     one, as the reference does);
//   * the wave adds psi[n_x] + psi[n_y] over its 64 lanes in lane order,
//     block after block, so the float32 sum runs in frame order.
// No MFMA: the work is compare/max/select, not a contraction. The roofline
// is FP32 VALU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "../../include/netsci_hip.h"
#include "nsx_internal.h"
#include "sort_select.h"

namespace nsx {
namespace {

constexpr int kWave = 64;
constexpr int kChunk = 128;         // frames per LDS chunk
constexpr int kWavesPerBlock = 4;   // independent waves per workgroup
constexpr int kMaxK = 16;

struct KsgArgs {
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
    int *counts_out;     // nullable: [p][frame][2]
    KeyIdx *scratch;     // per wave: scratch_stride records
    long long scratch_stride;
    unsigned long long *stats;  // [0] tie frames
};

// Make a wave-uniform 64-bit value provably uniform (SGPR) for the compiler.
__device__ __forceinline__ unsigned long long uniform_u64(unsigned long long v) {
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

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

// Insert (c, j) into the ascending heap h[0..K] (caller checked c < h[K]).
template <int K>
__device__ __forceinline__ void heap_insert(float (&h)[K + 1], int (&id)[K + 1],
                                            float c, int j) {
#pragma unroll
    for (int m = K; m >= 1; --m) {
        const bool lp = c < h[m - 1];
        const bool lc = c < h[m];
        id[m] = lp ? id[m - 1] : (lc ? j : id[m]);
        h[m] = __builtin_amdgcn_fmed3f(h[m - 1], c, h[m]);
    }
    const bool l0 = c < h[0];
    id[0] = l0 ? j : id[0];
    h[0] = fminf(h[0], c);
}

template <int D>
struct Chunk {
    float *s;  // [2D][kChunk]: a-coordinates then b-coordinates
    __device__ __forceinline__ float4 get(int dim, int jj) const {
        return *reinterpret_cast<const float4 *>(s + dim * kChunk + jj);
    }
};

// Stage frames [c0, c0 + kChunk) of both rows into the wave's LDS chunk.
// Frames at or beyond n are NaN so that every comparison with them fails.
template <int D>
__device__ __forceinline__ void stage_chunk(const Chunk<D> &ch, const float *xa,
                                            const float *xb, long long ds, int c0,
                                            int n, int lane) {
    const float qnan = __builtin_nanf("");
#pragma unroll
    for (int h = 0; h < kChunk / kWave; ++h) {
        const int f = c0 + h * kWave + lane;
        const bool ok = f < n;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const float va = ok ? xa[d * ds + f] : qnan;
            const float vb = ok ? xb[d * ds + f] : qnan;
            ch.s[d * kChunk + h * kWave + lane] = va;
            ch.s[(D + d) * kChunk + h * kWave + lane] = vb;
        }
    }
    wave_sync();
}

// Pass 1 over one staged chunk: keep the k+1 nearest frames j != i.
template <int D, int K, bool kSelf>
__device__ __forceinline__ void pass1_chunk(const Chunk<D> &ch, int cnt4, int c0, int i,
                                            const float (&qa)[D], const float (&qb)[D],
                                            float (&h)[K + 1], int (&id)[K + 1]) {
    for (int jj = 0; jj < cnt4; jj += 4) {
        float4 A[D], B[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            A[d] = ch.get(d, jj);
            B[d] = ch.get(D + d, jj);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            float xa[D], xb[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                xa[d] = u == 0 ? A[d].x : u == 1 ? A[d].y : u == 2 ? A[d].z : A[d].w;
                xb[d] = u == 0 ? B[d].x : u == 1 ? B[d].y : u == 2 ? B[d].z : B[d].w;
            }
            float dz = fmaxf(marg<D>(qa, xa), marg<D>(qb, xb));
            const int j = c0 + jj + u;
            if (kSelf) dz = (j == i) ? __builtin_nanf("") : dz;
            if (dz < h[K]) heap_insert<K>(h, id, dz, j);
        }
    }
}

// Pass 2 over one staged chunk: marginal counts within eps (self included).
template <int D>
__device__ __forceinline__ void pass2_chunk(const Chunk<D> &ch, int cnt4,
                                            const float (&qa)[D], const float (&qb)[D],
                                            float ex, float ey, int &cx, int &cy) {
    for (int jj = 0; jj < cnt4; jj += 4) {
        float4 A[D], B[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            A[d] = ch.get(d, jj);
            B[d] = ch.get(D + d, jj);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            float xa[D], xb[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                xa[d] = u == 0 ? A[d].x : u == 1 ? A[d].y : u == 2 ? A[d].z : A[d].w;
                xb[d] = u == 0 ? B[d].x : u == 1 ? B[d].y : u == 2 ? B[d].z : B[d].w;
            }
            cx += marg<D>(qa, xa) <= ex ? 1 : 0;
            cy += marg<D>(qb, xb) <= ey ? 1 : 0;
        }
    }
}

// Resolve the eps of every lane flagged `tie`: rebuild the frame's joint
// distance row into scratch and select the k+1 first entries in libstdc++
// std::sort order, then take eps over positions 1..k
// (mutual_information_cpu.cpp:48-77).
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

// Every wave walks items w = wave_gid, wave_gid + nwaves, ... All loop
// bounds and item fields are wave-uniform scalars; the only lane-divergent
// regions are the heap insertion, the tie rebuild's single-lane selection
// and the per-lane loads/stores.
template <int D, int K>
__global__ __launch_bounds__(kWave *kWavesPerBlock) void ksg_kernel(KsgArgs a) {
    __shared__ __attribute__((aligned(16))) float lds[kWavesPerBlock][2 * D * kChunk];
    __shared__ float sterm[kWavesPerBlock][kWave];
    const int lane = threadIdx.x & (kWave - 1);
    const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    Chunk<D> ch{lds[wib]};
    float *terms = sterm[wib];
    const long long nwaves = (long long)gridDim.x * kWavesPerBlock;
    const long long wave_gid = (long long)blockIdx.x * kWavesPerBlock + wib;
    KeyIdx *scratch = a.scratch + wave_gid * a.scratch_stride;
    const long long total = (long long)a.R * a.P;
    const long long ds = a.dstride;
    unsigned long long ties = 0;

    for (long long w = wave_gid; w < total; w += nwaves) {
        const int slot = (int)(w / a.P);
        const long long p = w - (long long)slot * a.P;
        const int off = a.win_off[slot];
        const int n = a.win_n[slot];
        const float *xa = a.X + (long long)a.ab[2 * p] * a.ld + off;
        const float *xb = a.X + (long long)a.ab[2 * p + 1] * a.ld + off;

        float acc = 0.0f;
        for (int q0 = 0; q0 < n; q0 += kWave) {
            const int i = q0 + lane;
            const bool valid = i < n;
            const int iq = valid ? i : 0;
            float qa[D], qb[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                qa[d] = xa[d * ds + iq];
                qb[d] = xb[d * ds + iq];
            }
            float h[K + 1];
            int id[K + 1];
#pragma unroll
            for (int m = 0; m <= K; ++m) {
                h[m] = __builtin_inff();
                id[m] = 0;
            }
            // Pass 1: k+1 nearest by joint distance.
            for (int c0 = 0; c0 < n; c0 += kChunk) {
                stage_chunk<D>(ch, xa, xb, ds, c0, n, lane);
                const int cnt4 = (min(kChunk, n - c0) + 3) & ~3;
                const bool self = (c0 < q0 + kWave) && (q0 < c0 + kChunk);
                if (self) pass1_chunk<D, K, true>(ch, cnt4, c0, i, qa, qb, h, id);
                else pass1_chunk<D, K, false>(ch, cnt4, c0, i, qa, qb, h, id);
                wave_sync();
            }
            // eps over the k nearest; boundary ties take the exact path.
            const bool tie = valid && (h[K - 1] == h[K]);
            float ex = 0.0f, ey = 0.0f;
#pragma unroll
            for (int m = 0; m < K; ++m) {
                ex = fmaxf(ex, marg_global<D>(qa, xa, ds, id[m]));
                ey = fmaxf(ey, marg_global<D>(qb, xb, ds, id[m]));
            }
            ties += resolve_ties<D>(tie, qa, qb, xa, xb, ds, n, K, scratch, lane, ex, ey);
            // Pass 2: marginal counts.
            int cx = 0, cy = 0;
            for (int c0 = 0; c0 < n; c0 += kChunk) {
                stage_chunk<D>(ch, xa, xb, ds, c0, n, lane);
                const int cnt4 = (min(kChunk, n - c0) + 3) & ~3;
                pass2_chunk<D>(ch, cnt4, qa, qb, ex, ey, cx, cy);
                wave_sync();
            }
            const int nx = cx - 1, ny = cy - 1;
            terms[lane] = valid ? a.psi[nx] + a.psi[ny] : 0.0f;
            if (a.counts_out && valid) {
                a.counts_out[(p * n + i) * 2 + 0] = nx;
                a.counts_out[(p * n + i) * 2 + 1] = ny;
            }
            wave_sync();
            // Sequential float32 sum in frame order (every lane runs the same
            // sum over LDS broadcasts).
            const int lim = min(kWave, n - q0);
            for (int l = 0; l < lim; ++l) acc += terms[l];
            wave_sync();
        }
        const float avg = acc / (float)n;
        const float mi = a.psi[n] + a.psi[a.k] - a.inv_k - avg;
        if (lane == 0) a.mi_out[(long long)a.win_rep[slot] * a.P + p] = mi;
    }
    if (lane == 0 && ties) atomicAdd(a.stats, ties);
}

using KernelFn = void (*)(KsgArgs);

template <int D, int K>
struct KTable {
    static void fill(KernelFn *t) {
        t[K] = ksg_kernel<D, K>;
        KTable<D, K - 1>::fill(t);
    }
};
template <int D>
struct KTable<D, 0> {
    static void fill(KernelFn *) {}
};

KernelFn kernel_for(int d, int k) {
    static KernelFn tab[4][kMaxK + 1] = {};
    static bool init = false;
    if (!init) {
        KTable<1, kMaxK>::fill(tab[1]);
        KTable<2, kMaxK>::fill(tab[2]);
        KTable<3, kMaxK>::fill(tab[3]);
        init = true;
    }
    if (d < 1 || d > 3 || k < 1 || k > kMaxK) return nullptr;
    return tab[d][k];
}

thread_local long long g_last_ties = 0;

int run_ksg(const float *X, long long ld, long long dstride, int d,
            const std::vector<int> &off, const std::vector<int> &nfr,
            const int *ab, long long P, int k, float *mi_out, int *counts_out,
            hipStream_t stream) {
    const int R = (int)off.size();
    NSX_REQUIRE(X && ab && mi_out, "ksg: null pointer argument");
    NSX_REQUIRE(d >= 1 && d <= 3, "ksg: the GPU path supports d = 1, 2 or 3");
    NSX_REQUIRE(k >= 1 && k <= kMaxK, "ksg: the GPU path supports 1 <= k <= 16");
    NSX_REQUIRE(P >= 0 && R >= 1, "ksg: bad pair or window count");
    NSX_REQUIRE(ld > 0 && dstride > 0, "ksg: bad strides");
    int nmax = 0;
    for (int r = 0; r < R; ++r) {
        NSX_REQUIRE(nfr[r] > k, "ksg: every window needs more than k frames");
        NSX_REQUIRE(off[r] >= 0 && (long long)off[r] + nfr[r] <= dstride,
                    "ksg: window exceeds the frames of a row");
        nmax = std::max(nmax, nfr[r]);
    }
    NSX_REQUIRE((long long)(d - 1) * dstride + dstride <= ld,
                "ksg: d * dstride exceeds the row length");
    int rc = ensure_device();
    if (rc) return rc;
    g_last_ties = 0;
    if (P == 0) return kOk;

    // Processing order: longest windows first (the biggest items start first).
    std::vector<int> order(R);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return nfr[x] > nfr[y]; });
    std::vector<int> host_win(3 * R);
    for (int s = 0; s < R; ++s) {
        host_win[s] = off[order[s]];
        host_win[R + s] = nfr[order[s]];
        host_win[2 * R + s] = order[s];
    }

    KernelFn fn = kernel_for(d, k);
    NSX_REQUIRE(fn, "ksg: no kernel for (d, k)");
    int per_cu = 0;
    NSX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)fn,
                                                         kWave * kWavesPerBlock, 0),
            "occupancy query");
    per_cu = std::max(per_cu, 1);
    const unsigned long long items = (unsigned long long)R * (unsigned long long)P;
    long long blocks = (long long)cu_count() * per_cu;
    blocks = std::min<long long>(blocks, (long long)((items + kWavesPerBlock - 1) / kWavesPerBlock));
    blocks = std::max<long long>(blocks, 1);
    const long long waves = blocks * kWavesPerBlock;

    std::vector<float> psi(nmax + 1);
    psi_table(psi.data(), nmax);

    float *d_psi = (float *)workspace(kWsPsi, psi.size() * sizeof(float));
    unsigned long long *d_ctr = (unsigned long long *)workspace(kWsCounter, 64);
    KeyIdx *d_scr = (KeyIdx *)workspace(kWsTieScratch, (size_t)waves * nmax * sizeof(KeyIdx));
    int *d_win = (int *)workspace(kWsWindows, host_win.size() * sizeof(int));
    if (!d_psi || !d_ctr || !d_scr || !d_win) return kHipError;
    NSX_HIP(hipMemcpyAsync(d_psi, psi.data(), psi.size() * sizeof(float),
                           hipMemcpyHostToDevice, stream), "psi upload");
    NSX_HIP(hipMemcpyAsync(d_win, host_win.data(), host_win.size() * sizeof(int),
                           hipMemcpyHostToDevice, stream), "window upload");
    rc = fill_u32(d_ctr, 0u, 16, stream);  // [0..1]: tie-frame count
    if (rc) return rc;

    KsgArgs args;
    args.X = X;
    args.ld = ld;
    args.dstride = dstride;
    args.ab = ab;
    args.P = P;
    args.win_off = d_win;
    args.win_n = d_win + R;
    args.win_rep = d_win + 2 * R;
    args.R = R;
    args.k = k;
    args.psi = d_psi;
    args.inv_k = (float)(1.0 / (float)k);
    args.mi_out = mi_out;
    args.counts_out = counts_out;
    args.scratch = d_scr;
    args.scratch_stride = nmax;
    args.stats = d_ctr;
    hipLaunchKernelGGL(fn, dim3((unsigned)blocks), dim3(kWave * kWavesPerBlock), 0,
                       stream, args);
    NSX_HIP(hipGetLastError(), "ksg launch");
    // Tie statistics for the run record (synchronises the stream).
    unsigned long long ties = 0;
    NSX_HIP(hipMemcpyAsync(&ties, d_ctr, sizeof(ties), hipMemcpyDeviceToHost, stream),
            "tie count readback");
    NSX_HIP(hipStreamSynchronize(stream), "ksg sync");
    g_last_ties = (long long)ties;
    return kOk;
}

}  // namespace
}  // namespace nsx

using namespace nsx;

extern "C" {

int nsx_ksg_mi(const float *d_X, long long ld, long long dstride, int d, int n,
               const int *d_ab, long long P, int k, float *d_mi_out, int *d_counts_out,
               void *stream) {
    clear_error();
    std::vector<int> off{0}, nfr{n};
    return run_ksg(d_X, ld, dstride, d, off, nfr, d_ab, P, k, d_mi_out, d_counts_out,
                   (hipStream_t)stream);
}

int nsx_ksg_mi_windows(const float *d_X, long long ld, long long dstride, int d,
                       const int *h_win_off, const int *h_win_n, int R, const int *d_ab,
                       long long P, int k, float *d_mi_out, void *stream) {
    clear_error();
    NSX_REQUIRE(h_win_off && h_win_n && R >= 1, "ksg windows: bad window arrays");
    std::vector<int> off(h_win_off, h_win_off + R), nfr(h_win_n, h_win_n + R);
    return run_ksg(d_X, ld, dstride, d, off, nfr, d_ab, P, k, d_mi_out, nullptr,
                   (hipStream_t)stream);
}

int nsx_ksg_last_tie_frames(long long *count) {
    if (count) *count = g_last_ties;
    return kOk;
}

}  // extern "C"
