// Host side of the KSG engine: argument checks, the (window, pair) item
// order, launch configuration, the eager entry points, hipGraph-captured
// plans for repeated batched launches, and the launch-shape autotuner.
//
// Reference: the pair loops netcalc/src/mutual_information.cpp:103-172 and
// generalized_correlation.cpp:104-173 call mutualInformationGpu
// (mutual_information_gpu.cu:1007-1079) once per pair — a map rebuild, psi
// table, four allocations, four blocking copies and one launch each. Here all
// P pairs (times R windows) are one persistent launch of ksg_kernel.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/netsci_hip.h"
#include "ksg_kernel.h"
#include "ksg_walk.h"
#include "nsx_internal.h"

namespace nsx {
namespace {

using ksg::Args;
using ksg::KernelFn;
using ksg::kWave;

// Launch shape. The autotuner chooses among these fields.
struct Config {
    int qpl = 2;     // queries per lane (1 or 2): 64*qpl query frames per sweep
    int chunk = 128; // frames per LDS tile (64, 128, 256)
    int wpb = 4;     // waves per workgroup (2, 4, 8)
    int occ = 0;     // workgroups per CU (0: as many as fit)
    int sym = 0;     // 1: symmetric pass 2 (each unordered frame pair once)
};

KernelFn kernel_for(int d, int k, int qpl, bool sym) {
    switch (d) {
        case 1: return ksg::kernel_d1(k, qpl, sym);
        case 2: return ksg::kernel_d2(k, qpl, sym);
        case 3: return ksg::kernel_d3(k, qpl, sym);
        default: return nullptr;
    }
}

// Validated problem description with the processing order of the windows.
struct Problem {
    const float *X = nullptr;
    long long ld = 0, dstride = 0;
    int d = 0, k = 0, R = 0, nmax = 0;
    const int *ab = nullptr;
    long long P = 0;
    double work = 0;            // sum over items of n(n-1), for reporting
    std::vector<int> host_win;  // [off(R) | n(R) | rep(R)] in processing order
};

int make_problem(Problem &pb, const float *X, long long ld, long long dstride, int d,
                 const int *off, const int *nfr, int R, const int *ab, long long P, int k) {
    NSX_REQUIRE(X && ab, "ksg: null pointer argument");
    NSX_REQUIRE(d >= 1 && d <= 3, "ksg: the GPU path supports d = 1, 2 or 3");
    NSX_REQUIRE(k >= 1, "ksg: k must be >= 1");
    NSX_REQUIRE(P >= 0 && R >= 1, "ksg: bad pair or window count");
    NSX_REQUIRE(ld > 0 && dstride > 0, "ksg: bad strides");
    NSX_REQUIRE((long long)(d - 1) * dstride + dstride <= ld,
                "ksg: d * dstride exceeds the row length");
    int nmax = 0;
    double work = 0;
    for (int r = 0; r < R; ++r) {
        NSX_REQUIRE(nfr[r] > k, "ksg: every window needs more than k frames");
        NSX_REQUIRE(off[r] >= 0 && (long long)off[r] + nfr[r] <= dstride,
                    "ksg: window exceeds the frames of a row");
        nmax = std::max(nmax, nfr[r]);
        work += (double)nfr[r] * (nfr[r] - 1) * (double)P;
    }
    NSX_REQUIRE(k <= ksg::kMaxK || nmax <= ksgw::kMaxFrames,
                "ksg: k > 16 needs windows of at most 8192 frames");
    // Longest windows first, so the biggest items start first.
    std::vector<int> order(R);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return nfr[x] > nfr[y]; });
    pb.host_win.assign(3 * R, 0);
    for (int s = 0; s < R; ++s) {
        pb.host_win[s] = off[order[s]];
        pb.host_win[R + s] = nfr[order[s]];
        pb.host_win[2 * R + s] = order[s];
    }
    pb.X = X;
    pb.ld = ld;
    pb.dstride = dstride;
    pb.d = d;
    pb.k = k;
    pb.R = R;
    pb.nmax = nmax;
    pb.ab = ab;
    pb.P = P;
    pb.work = work;
    return kOk;
}

// Resident waves a launch may use (8 per SIMD, 4 SIMDs per CU): sizes the
// tie scratch of plans and the tuner.
long long max_waves() { return (long long)cu_count() * 32; }

struct Shape {
    KernelFn fn = nullptr;
    long long blocks = 0;
    size_t lds = 0;
};

int shape_for(const Problem &pb, const Config &c, long long items, Shape &s) {
    s.fn = kernel_for(pb.d, pb.k, c.qpl, c.sym != 0);
    NSX_REQUIRE(s.fn, "ksg: no kernel for (d, k, queries per lane)");
    NSX_REQUIRE((c.chunk == 64 || c.chunk == 128 || c.chunk == 256) && c.wpb >= 1 &&
                    c.wpb <= ksg::kMaxWavesPerBlock,
                "ksg: bad launch configuration");
    s.lds = ksg::lds_bytes(pb.d, c.qpl, c.wpb, c.sym != 0);
    int per_cu = 0;
    NSX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)s.fn,
                                                         kWave * c.wpb, s.lds),
            "occupancy query");
    per_cu = std::max(per_cu, 1);
    if (c.occ > 0) per_cu = std::min(per_cu, c.occ);
    long long blocks = (long long)cu_count() * per_cu;
    blocks = std::min<long long>(blocks, (items + c.wpb - 1) / c.wpb);
    blocks = std::min<long long>(blocks, max_waves() / c.wpb);
    s.blocks = std::max<long long>(blocks, 1);
    return kOk;
}

// Device buffers one launch needs besides its inputs and outputs.
struct Buffers {
    float *psi = nullptr;
    int *win = nullptr;
    KeyIdx *scratch = nullptr;
    unsigned long long *stats = nullptr;
};

Args make_args(const Problem &pb, const Config &c, const Buffers &b, float *mi_out,
               int *counts_out, long long P_used) {
    Args a;
    a.X = pb.X;
    a.ld = pb.ld;
    a.dstride = pb.dstride;
    a.ab = pb.ab;
    a.P = P_used;
    a.win_off = b.win;
    a.win_n = b.win + pb.R;
    a.win_rep = b.win + 2 * pb.R;
    a.R = pb.R;
    a.k = pb.k;
    a.psi = b.psi;
    a.inv_k = (float)(1.0 / (float)pb.k);
    a.mi_out = mi_out;
    a.counts_out = counts_out;
    a.scratch = b.scratch;
    a.scratch_stride = ksg::scratch_records(pb.nmax, c.sym != 0);
    a.scratch_n = pb.nmax;
    a.stats = b.stats;
    a.chunk = c.chunk;
    a.wpb = c.wpb;
    a.run_if = nullptr;
    return a;
}

int launch(const Shape &s, const Config &c, const Args &a, hipStream_t stream) {
    hipLaunchKernelGGL(s.fn, dim3((unsigned)s.blocks), dim3(kWave * c.wpb), s.lds, stream, a);
    NSX_HIP(hipGetLastError(), "ksg launch");
    return kOk;
}

// ---- tuning cache ---------------------------------------------------------

std::mutex g_tune_mutex;
std::map<std::string, Config> g_tuned;
bool g_tune_loaded = false;
Config g_override;  // chunk == 0: no override
bool g_has_override = false;

std::string tune_key(const Problem &pb) {
    // Bucket the mean window length by powers of two; batched and single
    // windows tune separately.
    long long nsum = 0;
    for (int s = 0; s < pb.R; ++s) nsum += pb.host_win[pb.R + s];
    const long long nmean = nsum / pb.R;
    int bucket = 0;
    while ((1LL << (bucket + 1)) <= nmean) ++bucket;
    std::ostringstream os;
    os << "d" << pb.d << "_k" << pb.k << "_n2^" << bucket << (pb.R > 1 ? "_batched" : "_plain");
    return os.str();
}

void load_tune_cache() {
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    const char *path = std::getenv("NSX_TUNE_CACHE");
    if (!path) return;
    // One line per key: qpl chunk wpb occ sym.
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream is(line);
        std::string key;
        Config c;
        if (is >> key >> c.qpl >> c.chunk >> c.wpb >> c.occ >> c.sym) g_tuned[key] = c;
    }
}

void store_tune(const std::string &key, const Config &c) {
    g_tuned[key] = c;
    const char *path = std::getenv("NSX_TUNE_CACHE");
    if (!path) return;
    std::ofstream f(path, std::ios::app);
    f << key << " " << c.qpl << " " << c.chunk << " " << c.wpb << " " << c.occ << " " << c.sym
      << "\n";
}

Config config_for(const Problem &pb) {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    if (g_has_override) return g_override;
    load_tune_cache();
    auto it = g_tuned.find(tune_key(pb));
    if (it != g_tuned.end()) return it->second;
    return Config{};
}

// Time each candidate on a sample of the problem's items (every window, the
// first Ps pairs — enough items to give every resident wave six, so the
// uneven tail of the last round weighs little) and keep the fastest.
// Candidates: pass 2 plain or symmetric x queries per lane {1, 2} x LDS
// tile {64, 128, 256} frames x waves per workgroup {4, 8}.
int autotune(const Problem &pb, const Buffers &b, hipStream_t stream, Config &best,
             double *best_ms) {
    const long long target_items = 6 * max_waves();
    long long Ps = std::min<long long>(pb.P, std::max<long long>(1, target_items / pb.R));
    float *d_out = (float *)workspace(kWsScreen1, (size_t)pb.R * Ps * sizeof(float));
    if (!d_out) return kHipError;
    hipEvent_t e0, e1;
    NSX_HIP(hipEventCreate(&e0), "tune event");
    NSX_HIP(hipEventCreate(&e1), "tune event");
    double best_t = 1e300;
    for (int sym : {0, 1})
    for (int qpl : {1, 2})
        for (int chunk : {64, 128, 256})
            for (int wpb : {4, 8}) {
                Config c;
                c.qpl = qpl;
                c.chunk = chunk;
                c.wpb = wpb;
                c.sym = sym;
                Shape s;
                int rc = shape_for(pb, c, (long long)pb.R * Ps, s);
                if (rc) return rc;
                const Args a = make_args(pb, c, b, d_out, nullptr, Ps);
                rc = launch(s, c, a, stream);  // warm-up (code object load, caches)
                if (rc) return rc;
                NSX_HIP(hipEventRecord(e0, stream), "tune record");
                rc = launch(s, c, a, stream);
                if (rc) return rc;
                NSX_HIP(hipEventRecord(e1, stream), "tune record");
                NSX_HIP(hipEventSynchronize(e1), "tune sync");
                float ms = 0;
                NSX_HIP(hipEventElapsedTime(&ms, e0, e1), "tune elapsed");
                if (ms < best_t) {
                    best_t = ms;
                    best = c;
                }
            }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (best_ms) *best_ms = best_t;
    return kOk;
}

// ---- list path (ksg_walk.h) -------------------------------------------------

// Method: 0 picks the list path whenever it applies, 1 always runs the
// all-frames kernel, 2 requires the list path.
int g_method = 0;
int g_walk_wpb = 16;   // waves per walk workgroup (one workgroup per CU)
int g_walk_batch = 16; // finished lanes a wave waits for before a refill round
thread_local long long g_last_ties = 0;
thread_local int g_last_method = 0;         // 1 all-frames, 2 lists
thread_local long long g_last_entries = 0;  // list entries walked (both lists)
// Device time of the last list-path call's phases (hipEvents on its stream):
// [0] node map + table build, [1] walk, [2] tie resolution, [3] psi sums.
thread_local double g_last_phase_ms[4] = {0, 0, 0, 0};
thread_local double g_last_call_ms = 0;  // host wall time of the last call

// Event pairs for phase timing, created once per thread.
struct PhaseEvents {
    std::vector<hipEvent_t> ev;
    hipEvent_t get(size_t i) {
        while (ev.size() <= i) {
            hipEvent_t e = nullptr;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            ev.push_back(e);
        }
        return ev[i];
    }
};
thread_local PhaseEvents g_phase_events;

ksgw::BuildFn build_for(int d, int width) {
    switch (d) {
        case 1: return ksgw::build_d1(width);
        case 2: return ksgw::build_d2(width);
        case 3: return ksgw::build_d3(width);
        default: return nullptr;
    }
}

ksgw::WalkFn walk_for(int d, int k) {
    switch (d) {
        case 1: return ksgw::walk_d1(k);
        case 2: return ksgw::walk_d2(k);
        case 3: return ksgw::walk_d3(k);
        default: return nullptr;
    }
}

ksgw::TieFn tie_for(int d, int k, int kind) {
    switch (d) {
        case 1: return ksgw::tie_d1(k, kind);
        case 2: return ksgw::tie_d2(k, kind);
        case 3: return ksgw::tie_d3(k, kind);
        default: return nullptr;
    }
}

// mm[0] = min, mm[1] = max over the 2P node indices.
__global__ void ab_range_kernel(const int *ab, long long m, int *mm) {
    int lo = 0x7fffffff, hi = -1;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < m;
         t += (long long)gridDim.x * blockDim.x) {
        lo = min(lo, ab[t]);
        hi = max(hi, ab[t]);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        lo = min(lo, __shfl_xor(lo, off));
        hi = max(hi, __shfl_xor(hi, off));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(mm, lo);
        atomicMax(mm + 1, hi);
    }
}

__global__ void mark_nodes_kernel(const int *ab, long long m, int *used) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < m;
         t += (long long)gridDim.x * blockDim.x)
        used[ab[t]] = 1;
}

// One workgroup of 1024 lanes: slot_of[node] = rank of node among the used
// ones (-1 if unused), node_of[slot] = node, *count = number of used nodes.
__global__ __launch_bounds__(1024) void slot_scan_kernel(const int *used, int N, int *slot_of,
                                                         int *node_of, int *count) {
    __shared__ int wave_tot[16];
    __shared__ int carry;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int c0 = 0; c0 < N; c0 += 1024) {
        const int v = c0 + (int)threadIdx.x;
        const bool u = v < N && used[v] != 0;
        const unsigned long long b = __ballot(u);
        const int below = __builtin_amdgcn_mbcnt_hi((unsigned)(b >> 32),
                                                     __builtin_amdgcn_mbcnt_lo((unsigned)b, 0u));
        if (lane == 0) wave_tot[w] = __popcll(b);
        __syncthreads();
        int before = carry;
        for (int x = 0; x < w; ++x) before += wave_tot[x];
        if (v < N) {
            const int s = before + below;
            slot_of[v] = u ? s : -1;
            if (u) node_of[s] = v;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int tot = 0;
            for (int x = 0; x < 16; ++x) tot += wave_tot[x];
            carry += tot;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *count = carry;
}

// One wave per used node (blocks past *count exit): span[node] = max over
// dimensions of (max - min) of the coordinate over the window's n frames,
// an upper bound of every marginal distance of the node (rounding of the
// differences is monotone); *bad |= 1 if any of those coordinates is not
// finite.
__global__ __launch_bounds__(64) void node_stats_kernel(const float *X, long long ld,
                                                        long long dstride, int d, int n,
                                                        const int *node_of, const int *count,
                                                        float *span, unsigned *bad) {
    if ((int)blockIdx.x >= *count) return;
    const int node = node_of[blockIdx.x];
    const float *x = X + (long long)node * ld;
    float best = 0.0f;
    bool nonfinite = false;
    for (int t = 0; t < d; ++t) {
        float lo = __builtin_inff(), hi = -__builtin_inff();
        for (int f = threadIdx.x; f < n; f += 64) {
            const float v = x[t * dstride + f];
            nonfinite |= !__builtin_isfinite(v);
            lo = fminf(lo, v);
            hi = fmaxf(hi, v);
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            lo = fminf(lo, __shfl_xor(lo, off));
            hi = fmaxf(hi, __shfl_xor(hi, off));
        }
        best = fmaxf(best, hi - lo);
    }
    if (threadIdx.x == 0) span[node] = best;
    if (__ballot(nonfinite) && threadIdx.x == 0) atomicOr(bad, 1u);
}

// Node-blocked tables: the block pair of every pair (blocks of bn slots,
// (lo, hi) with lo <= hi), as its position in the schedule, and a histogram.
__global__ void pair_key_kernel(const int *ab, long long P, const int *slot_of, int bn, int nb,
                                const int *sched_of, int *key, unsigned *hist) {
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < P;
         p += (long long)gridDim.x * blockDim.x) {
        const int x = slot_of[ab[2 * p]] / bn, y = slot_of[ab[2 * p + 1]] / bn;
        const int kk = sched_of[min(x, y) * nb + max(x, y)];
        key[p] = kk;
        atomicAdd(hist + kk, 1u);
    }
}

// Scatter the pairs into schedule order (the order within one block pair is
// arbitrary: every pair's result is independent of it).
__global__ void pair_scatter_kernel(const int *ab, long long P, const int *key, unsigned *cursor,
                                    int *ab_perm, int *idx_perm) {
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < P;
         p += (long long)gridDim.x * blockDim.x) {
        const unsigned at = atomicAdd(cursor + key[p], 1u);
        ab_perm[2 * (long long)at] = ab[2 * p];
        ab_perm[2 * (long long)at + 1] = ab[2 * p + 1];
        idx_perm[at] = (int)p;
    }
}

// Resident table slot of the nodes of one block: rslot_of[node_of[s]] = base + s.
__global__ void set_slots_kernel(const int *node_of, int count, int base, int *rslot_of) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < count) rslot_of[node_of[s]] = base + s;
}

// psi terms of every frame summed sequentially in frame order, then the MI
// (mutual_information_cpu.cpp:110-118). A wave owns 64 consecutive pairs,
// one lane each: tiles of 64 pairs x 64 frames are read row by row
// (coalesced), turned into psi terms and transposed through LDS, and every
// lane adds its own pair's terms in frame order. kPsiLds: the psi table
// (n + 1 <= 2049 entries) is staged in LDS, so its two lookups per frame are
// LDS reads instead of scattered global loads (C1: 2.08 -> 1.33 ms).
constexpr int kSumWaves = 4;
constexpr int kSumPsiLds = 2048;  // psi entries staged in LDS (n + 1 <= this)
template <bool kPsiLds>
__global__ __launch_bounds__(64 * kSumWaves) void psi_sum_kernel(ksgw::SumArgs a) {
    __shared__ float tile[kSumWaves][64][65];
    __shared__ float psi_lds[kPsiLds ? kSumPsiLds + 1 : 1];
    if (a.abort && *a.abort) return;
    if (kPsiLds) {
        for (int t = threadIdx.x; t <= a.n; t += blockDim.x) psi_lds[t] = a.psi[t];
        __syncthreads();
    }
    const float *psi = kPsiLds ? psi_lds : a.psi;
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x / 64;
    const long long p0 = ((long long)blockIdx.x * kSumWaves + w) * 64;
    if (p0 >= a.P) return;
    const int rows = (int)min(64LL, a.P - p0);
    // Output index of this lane's pair, and of row r's pair (for counts).
    const long long mine = a.pbase + p0 + lane;
    const long long out_idx = lane < rows ? (a.perm ? (long long)a.perm[mine] : mine) : 0;
    float acc = 0.0f;
    for (int f0 = 0; f0 < a.n; f0 += 64) {
        const int f = f0 + lane;
        const int cols = min(64, a.n - f0);
        // 16 rows' counts in flight at a time, then their psi terms.
        for (int r0 = 0; r0 < rows; r0 += 16) {
            unsigned v[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int r = r0 + u;
                v[u] = (f < a.n && r < rows) ? a.cnt[(p0 + r) * a.n + f] : 0u;
            }
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int r = r0 + u;
                if (r < rows) {
                    // Row r's output index (all lanes take part in the shuffle).
                    long long pr = 0;
                    if (a.counts_out)
                        pr = (long long)(unsigned)__shfl((int)(unsigned)out_idx, r) |
                             ((long long)__shfl((int)(out_idx >> 32), r) << 32);
                    float term = 0.0f;
                    if (f < a.n) {
                        const int nx = (int)(v[u] & 0xffffu), ny = (int)(v[u] >> 16);
                        term = psi[nx] + psi[ny];
                        if (a.counts_out) {
                            const long long at = pr * a.n + f;
                            a.counts_out[at * 2 + 0] = nx;
                            a.counts_out[at * 2 + 1] = ny;
                        }
                    }
                    tile[w][r][lane] = term;
                }
            }
        }
        ksg::wave_sync();
        if (lane < rows)
            for (int c = 0; c < cols; ++c) acc += tile[w][lane][c];
        ksg::wave_sync();
    }
    if (lane < rows) {
        const float avg = acc / (float)a.n;
        a.mi_out[out_idx] = a.psi[a.n] + a.psi[a.k] - a.inv_k - avg;
    }
}

// ---- list path: layout, schedule and launches ------------------------------
//
// Tables hold S node slots. When every used node fits (S >= U) one build
// covers them all and one walk covers all pairs in their own order. Otherwise
// the used nodes (in node order) form nb blocks of bn = S / 2 slots, the
// pairs are grouped by block pair on the device, and the schedule keeps one
// block resident as the "row" while the blocks after it stream through the
// other half of the tables, last to first, so that the next row block is
// already resident when its turn comes: 1 + nb (nb - 1) / 2 block builds.

struct ListOp {
    bool build;       // build block `block` into half `half`; else walk
    int block, half;
    long long p0, pc; // walk: pairs [p0, p0 + pc) of the (permuted) pair list
};

struct ListLayout {
    int N = 0, U = 0;
    long long S = 0;       // resident table slots
    int nb = 1, bn = 0;    // blocks and slots per block
    bool tables = true;    // false: k > kMaxK on the all-frames method (counts by a pass)
    std::vector<ListOp> ops;
    long long node_builds = 0;  // per window
    // Device buffers.
    int *used = nullptr, *slot_of = nullptr, *node_of = nullptr, *rslot_of = nullptr,
        *count = nullptr;
    float *span = nullptr;  // [R][N]
    unsigned *flags = nullptr;  // [0] non-finite flag, [8..11] walk stats (2 x u64),
                                // [16] tie-list count
    const int *ab = nullptr;    // pair list the walks index (pb.ab or ab_perm)
    const int *perm = nullptr;  // nullable: output index of each listed pair
    uint16_t *ord = nullptr, *rank = nullptr;
    float *prior = nullptr;  // [slot][table_rows(n)] list-choice priors
    unsigned *cnt = nullptr, *tie_list = nullptr;
    long long chunk = 0;   // pairs per walk launch (packed counts below 4 GiB)
    unsigned tie_cap = 0;
};

// Fixed launch shape of the walk for this problem.
struct WalkShape {
    ksgw::WalkFn walk = nullptr;
    ksgw::TieFn ties = nullptr, tiek = nullptr, tieo = nullptr;
    int wpb = 16;
    long long grid_max = 0;
};

ksgw::WalkFn bigk_for(int d) {
    switch (d) {
        case 1: return ksgw::bigk_d1();
        case 2: return ksgw::bigk_d2();
        case 3: return ksgw::bigk_d3();
        default: return nullptr;
    }
}

int walk_shape(const Problem &pb, WalkShape &ws) {
    if (pb.k > ksg::kMaxK) {
        // One wave per query (ksgw::bigk_kernel); as many waves per CU as
        // their LDS rows allow, a few rounds of them.
        ws.walk = bigk_for(pb.d);
        NSX_REQUIRE(ws.walk, "ksg: no kernel for (d, k)");
        ws.wpb = 1;
        const size_t per = ksgw::bigk_lds_bytes(pb.nmax);
        ws.grid_max = (long long)cu_count() * std::max<long long>(1, std::min<long long>(16, (160 * 1024) / per));
        return kOk;
    }
    ws.walk = walk_for(pb.d, pb.k);
    ws.ties = tie_for(pb.d, pb.k, ksgw::kTieSelect);
    ws.tiek = tie_for(pb.d, pb.k, ksgw::kTieExact);
    ws.tieo = tie_for(pb.d, pb.k, ksgw::kTieOverflow);
    NSX_REQUIRE(ws.walk && ws.ties && ws.tiek && ws.tieo, "ksg: no list-path kernel for (d, k)");
    const size_t lds_cap = 160 * 1024;
    const size_t rec_bytes = ksgw::walk_buffer_bytes(pb.d, pb.nmax);
    ws.wpb = std::min(g_walk_wpb, ksgw::walk_max_waves(pb.k));
    int per_cu = 0;
    const int nbuf_max = (int)std::min<size_t>(3, (lds_cap - 256) / rec_bytes);
    NSX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, (const void *)ws.walk, ksgw::kWave * ws.wpb,
                ksgw::walk_lds_bytes(pb.d, pb.nmax, nbuf_max)),
            "walk occupancy");
    ws.grid_max = (long long)cu_count() * std::max(per_cu, 1);
    return kOk;
}

// Entries kept per ord row for windows of n frames: ksgw::ord_stride(n), or
// the cap nsx_ksg_set_list_cap set (rounded up to 8, at most the row).
long long g_list_cap = 0;
long long list_stride(int n) {
    const long long rs = ((long long)n + 7) & ~7LL;
    if (g_list_cap <= 0) return ksgw::ord_stride(n);
    return std::min(rs, std::max<long long>(8, (g_list_cap + 7) & ~7LL));
}

// Entries of one node slot's ord and rank tables, and bytes of the whole
// slot (ord, rank and the per-frame priors), for windows of up to nmax
// frames (a shorter window's slots are smaller).
long long ord_entries_for(int n) { return ksgw::table_rows(n) * list_stride(n); }
long long rank_entries_for(int n) {
    return ksgw::table_rows(n) * (((long long)n + 7) & ~7LL);
}
size_t slot_bytes_for(int nmax) {
    return (size_t)(ord_entries_for(nmax) + rank_entries_for(nmax)) * sizeof(uint16_t) +
           (size_t)ksgw::table_rows(nmax) * sizeof(float);
}

// The schedule for U used nodes in S slots (see above). sched_of[x * nb + y]
// (x <= y) is the position of block pair (x, y) among the walks.
void make_schedule(ListLayout &L, std::vector<int> &sched_of) {
    L.ops.clear();
    L.node_builds = 0;
    if (L.S >= L.U) {
        L.nb = 1;
        L.bn = L.U;
        L.ops.push_back({true, 0, 0, 0, 0});
        L.ops.push_back({false, 0, 0, 0, 0});
        L.node_builds = L.U;
        sched_of.assign(1, 0);
        return;
    }
    L.bn = (int)(L.S / 2);
    L.nb = (L.U + L.bn - 1) / L.bn;
    const int nb = L.nb;
    auto bsize = [&](int b) { return std::min(L.bn, L.U - b * L.bn); };
    sched_of.assign((size_t)nb * nb, -1);
    int walks = 0, h = 0;
    L.ops.push_back({true, 0, h, 0, 0});
    L.node_builds += bsize(0);
    for (int I = 0; I < nb; ++I) {
        sched_of[I * nb + I] = walks++;
        L.ops.push_back({false, I * nb + I, h, 0, 0});
        for (int J = nb - 1; J > I; --J) {
            L.ops.push_back({true, J, 1 - h, 0, 0});
            L.node_builds += bsize(J);
            sched_of[I * nb + J] = walks++;
            L.ops.push_back({false, I * nb + J, h, 0, 0});
        }
        h = 1 - h;
    }
}

// Where list_prepare takes its device buffers: the library's grow-only
// workspaces and table buffer (eager calls), or allocations owned by a plan
// (a captured graph keeps its pointers, so nothing it uses may move).
struct ListMem {
    std::vector<void *> *owned = nullptr;  // null: workspaces

    void *get(int slot, size_t bytes) {
        if (!owned) return workspace(slot, bytes);
        void *p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(bytes, 256)) != hipSuccess) {
            (void)hipGetLastError();
            set_error("list path: device allocation failed");
            return nullptr;
        }
        owned->push_back(p);
        return p;
    }

    // Table slots: at most want, at least min (else *slots < min).
    int tables(size_t slot_bytes, long long want, long long min, void **ptr, long long *slots) {
        if (!owned) return table_space(slot_bytes, want, min, ptr, slots);
        *ptr = nullptr;
        *slots = 0;
        size_t free_b = 0, total_b = 0;
        NSX_HIP(hipMemGetInfo(&free_b, &total_b), "memory info");
        double budget = 0.85 * (double)free_b - (double)(1ull << 30);
        if (table_budget()) budget = std::min(budget, (double)table_budget());
        long long s = budget > 0 ? std::min<long long>(want, (long long)(budget / (double)slot_bytes)) : 0;
        while (s >= min && s > 0) {
            void *p = nullptr;
            if (hipMalloc(&p, (size_t)s * slot_bytes + 256) == hipSuccess) {
                owned->push_back(p);
                *ptr = p;
                *slots = s;
                return kOk;
            }
            (void)hipGetLastError();
            s = s * 3 / 4;
        }
        return kOk;
    }
};

// Every window's node spans and the non-finite flag (flags[0]) of the used
// nodes; flags must be cleared first.
int enqueue_node_stats(const Problem &pb, const ListLayout &L, hipStream_t stream) {
    for (int w = 0; w < pb.R; ++w) {
        hipLaunchKernelGGL(node_stats_kernel, dim3((unsigned)L.N), dim3(64), 0, stream,
                           pb.X + pb.host_win[w], pb.ld, pb.dstride, pb.d, pb.host_win[pb.R + w],
                           L.node_of, L.count, L.span + (long long)w * L.N, L.flags);
        NSX_HIP(hipGetLastError(), "node stats launch");
    }
    return kOk;
}

// Host-side sizing with device readbacks (not capturable): the node map,
// every window's spans and the finite check, the table slots, the
// schedule, and for blocked tables the pair permutation. *ok = false: the
// list path does not apply (non-finite coordinates or no room for tables).
int list_prepare(const Problem &pb, ListLayout &L, hipStream_t stream, bool *ok, ListMem mem) {
    *ok = false;
    const long long m2 = 2 * pb.P;
    // flags: [0] min node, [1] max node (first), then reused as described
    // in ListLayout.
    unsigned *flags = (unsigned *)mem.get(kWsListStats, 256);
    if (!flags) return kHipError;
    int *mm = (int *)flags;
    int rc = fill_u32(mm, 0x7fffffffu, 1, stream);
    if (!rc) rc = fill_u32(mm + 1, 0xffffffffu, 1, stream);
    if (rc) return rc;
    const long long rblocks = std::min<long long>((m2 + 255) / 256, 1024);
    hipLaunchKernelGGL(ab_range_kernel, dim3((unsigned)rblocks), dim3(256), 0, stream, pb.ab, m2,
                       mm);
    NSX_HIP(hipGetLastError(), "ab range launch");
    int h_mm[2] = {0, 0};
    NSX_HIP(hipMemcpyAsync(h_mm, mm, sizeof(h_mm), hipMemcpyDeviceToHost, stream),
            "ab range readback");
    NSX_HIP(hipStreamSynchronize(stream), "ab range sync");
    NSX_REQUIRE(h_mm[0] >= 0, "ksg: negative node index in ab");
    const int N = h_mm[1] + 1;
    L.N = N;
    int *map = (int *)mem.get(kWsListMap, ((size_t)4 * N + 16) * sizeof(int));
    L.span = (float *)mem.get(kWsListSpan, (size_t)pb.R * N * sizeof(float));
    if (!map || !L.span) return kHipError;
    L.used = map;
    L.slot_of = map + N;
    L.node_of = map + 2 * N;
    L.rslot_of = map + 3 * N;
    L.count = map + 4 * N;
    L.flags = flags;
    rc = fill_u32(flags, 0u, 32, stream);
    if (!rc) rc = fill_u32(L.used, 0u, (size_t)N, stream);
    if (rc) return rc;
    hipLaunchKernelGGL(mark_nodes_kernel, dim3((unsigned)rblocks), dim3(256), 0, stream, pb.ab,
                       m2, L.used);
    NSX_HIP(hipGetLastError(), "mark nodes launch");
    hipLaunchKernelGGL(slot_scan_kernel, dim3(1), dim3(1024), 0, stream, L.used, N, L.slot_of,
                       L.node_of, L.count);
    NSX_HIP(hipGetLastError(), "slot scan launch");
    rc = enqueue_node_stats(pb, L, stream);
    if (rc) return rc;
    int h_u[2] = {0, 0};
    NSX_HIP(hipMemcpyAsync(&h_u[0], L.count, sizeof(int), hipMemcpyDeviceToHost, stream),
            "slot count readback");
    NSX_HIP(hipMemcpyAsync(&h_u[1], flags, sizeof(int), hipMemcpyDeviceToHost, stream),
            "finite check readback");
    NSX_HIP(hipStreamSynchronize(stream), "slot scan sync");
    L.U = h_u[0];
    if (h_u[1]) {
        NSX_REQUIRE(g_method != 2, "ksg: the list path needs finite coordinates");
        NSX_REQUIRE(pb.k <= ksg::kMaxK, "ksg: k > 16 needs finite coordinates");
        return kOk;
    }

    std::vector<int> sched_of;
    if (pb.k > ksg::kMaxK && g_method == 1) {
        // No tables: one walk over every pair, counts by a pass over the row.
        L.tables = false;
        L.S = L.U;
        L.nb = 1;
        L.bn = L.U;
        L.ops.assign(1, ListOp{false, 0, 0, 0, 0});
        sched_of.assign(1, 0);
    } else {
        // Table slots: all used nodes when they fit, else at least two.
        void *tab = nullptr;
        const size_t sb = slot_bytes_for(pb.nmax);
        rc = mem.tables(sb, L.U, std::min(L.U, 2), &tab, &L.S);
        if (rc) return rc;
        if (L.S < std::min(L.U, 2)) {
            NSX_REQUIRE(g_method != 2 && pb.k <= ksg::kMaxK,
                        "ksg: no device memory for two nodes' list-path tables");
            return kOk;
        }
        L.ord = (uint16_t *)tab;
        L.rank = L.ord + L.S * ord_entries_for(pb.nmax);
        L.prior = reinterpret_cast<float *>(L.rank + L.S * rank_entries_for(pb.nmax));
        make_schedule(L, sched_of);
    }

    // Pair order.
    L.ab = pb.ab;
    L.perm = nullptr;
    if (L.nb == 1) {
        L.ops.back().p0 = 0;
        L.ops.back().pc = pb.P;
    } else {
        const int nb = L.nb;
        const int nwalks = nb * (nb + 1) / 2;
        const size_t perm_words = (size_t)pb.P * 4 + (size_t)nb * nb + 2 * (size_t)nwalks + 64;
        int *pw = (int *)mem.get(kWsListPerm, perm_words * sizeof(int));
        if (!pw) return kHipError;
        int *ab_perm = pw, *idx_perm = pw + 2 * pb.P, *key = pw + 3 * pb.P;
        int *d_sched = pw + 4 * pb.P;
        unsigned *hist = (unsigned *)(d_sched + (size_t)nb * nb);
        unsigned *cursor = hist + nwalks;
        NSX_HIP(hipMemcpyAsync(d_sched, sched_of.data(), sched_of.size() * sizeof(int),
                               hipMemcpyHostToDevice, stream),
                "schedule upload");
        rc = fill_u32(hist, 0u, (size_t)nwalks, stream);
        if (rc) return rc;
        const long long pblocks = std::min<long long>((pb.P + 255) / 256, 4096);
        hipLaunchKernelGGL(pair_key_kernel, dim3((unsigned)pblocks), dim3(256), 0, stream, pb.ab,
                           pb.P, L.slot_of, L.bn, nb, d_sched, key, hist);
        NSX_HIP(hipGetLastError(), "pair key launch");
        std::vector<unsigned> h_hist(nwalks), h_cur(nwalks);
        NSX_HIP(hipMemcpyAsync(h_hist.data(), hist, nwalks * sizeof(unsigned),
                               hipMemcpyDeviceToHost, stream),
                "histogram readback");
        NSX_HIP(hipStreamSynchronize(stream), "histogram sync");
        unsigned acc = 0;
        for (int t = 0; t < nwalks; ++t) {
            h_cur[t] = acc;
            acc += h_hist[t];
        }
        NSX_HIP(hipMemcpyAsync(cursor, h_cur.data(), nwalks * sizeof(unsigned),
                               hipMemcpyHostToDevice, stream),
                "cursor upload");
        hipLaunchKernelGGL(pair_scatter_kernel, dim3((unsigned)pblocks), dim3(256), 0, stream,
                           pb.ab, pb.P, key, cursor, ab_perm, idx_perm);
        NSX_HIP(hipGetLastError(), "pair scatter launch");
        // The cursor upload reads h_cur: wait before it goes out of scope.
        NSX_HIP(hipStreamSynchronize(stream), "pair scatter sync");
        for (ListOp &op : L.ops) {
            if (op.build) continue;
            const int x = op.block / nb, y = op.block % nb;
            const int t = sched_of[x * nb + y];
            op.p0 = h_cur[t];
            op.pc = h_hist[t];
        }
        // Drop the walks without pairs, and the column builds they needed
        // (every column build is followed by its walk); the build of block
        // I + 1 stays, as the next row block.
        std::vector<ListOp> kept;
        L.node_builds = 0;
        for (size_t t = 0; t < L.ops.size(); ++t) {
            const ListOp &op = L.ops[t];
            if (op.build) {
                const ListOp &w = L.ops[t + 1];
                if (t > 0 && op.block != w.block / nb + 1 && w.pc == 0) {
                    ++t;
                    continue;
                }
                kept.push_back(op);
                L.node_builds += std::min(L.bn, L.U - op.block * L.bn);
            } else if (op.pc > 0) {
                kept.push_back(op);
            }
        }
        L.ops.swap(kept);
        L.ab = ab_perm;
        L.perm = idx_perm;
    }

    // Packed counts of one walk launch: at most 4 GiB, so local (pair,
    // frame) indices stay below 2^30; the tie list holds 1/32 of a launch's
    // queries (ties beyond it are marked in the counts and resolved by
    // tie_overflow_kernel). One row of margin: a shorter window's chunk x n
    // can pass chunk x nmax by less than nmax.
    L.chunk = std::max<long long>(
        1, std::min<long long>(pb.P, (1LL << 32) / ((long long)pb.nmax * (long long)sizeof(unsigned))));
    L.tie_cap = (unsigned)std::max<long long>(1024, L.chunk * (long long)pb.nmax / 32);
    const size_t cnt_words = (size_t)L.chunk * pb.nmax + pb.nmax;
    L.cnt = (unsigned *)mem.get(kWsListCnt, (cnt_words + 2 * (size_t)L.tie_cap) * sizeof(unsigned));
    if (!L.cnt) return kHipError;
    L.tie_list = L.cnt + cnt_words;
    *ok = true;
    return kOk;
}

// Event pairs recorded around the phases of a list-path run.
struct PhaseMarks {
    struct Mark {
        int phase;
        hipEvent_t a, b;
    };
    std::vector<Mark> marks;
    size_t next = 1;  // event 0 marks the call's start
    bool on = false;
    hipEvent_t ev() { return on ? g_phase_events.get(next++) : nullptr; }
    void add(int phase, hipEvent_t a, hipEvent_t b) {
        if (a && b) marks.push_back({phase, a, b});
    }
};

// Enqueues every window's builds, walks, tie resolution and psi sums (no
// host synchronisation; capturable once list_prepare has run).
int list_enqueue(const Problem &pb, const ListLayout &L, const WalkShape &ws, const float *d_psi,
                 float *mi_out, int *counts_out, hipStream_t stream, PhaseMarks *pm,
                 long long q_begin = 0, long long q_end = -1, bool builds = true) {
    // q_begin / q_end: only the pairs [q_begin, q_end) of the list (a single
    // block's layout); builds = false skips the table builds (done earlier).
    if (q_end < 0) q_end = pb.P;
    const size_t lds_cap = 160 * 1024;
    unsigned long long *stats = (unsigned long long *)(L.flags + 8);
    unsigned *tie_count = L.flags + 16;
    const unsigned *abort_flag = L.flags;
    const long long tie_waves = (long long)cu_count() * 8;
    for (int w = 0; w < pb.R; ++w) {
        const int off = pb.host_win[w], n = pb.host_win[pb.R + w], rep = pb.host_win[2 * pb.R + w];
        const float *X = pb.X + off;
        const long long rs = ((long long)n + 7) & ~7LL;
        const long long ors = list_stride(n);
        float *out = mi_out + (long long)rep * pb.P;
        const float *span = L.span + (long long)w * L.N;
        const int width = ksgw::build_width(n);
        const ksgw::BuildFn buildk = build_for(pb.d, width);
        NSX_REQUIRE(buildk, "ksg: no list-path table kernel for this frame count");
        const size_t wrec = ksgw::walk_buffer_bytes(pb.d, n);
        const int nbuf = (int)std::min<size_t>(3, (lds_cap - 256) / wrec);
        const size_t wlds = ksgw::walk_lds_bytes(pb.d, n, nbuf);
        const long long wchunk = std::max<long long>(
            1, std::min<long long>(L.chunk, (1LL << 32) / ((long long)n * (long long)sizeof(unsigned))));
        const unsigned wtie_cap = (unsigned)std::min<long long>(
            L.tie_cap, std::max<long long>(1024, wchunk * (long long)n / 32));
        for (const ListOp &op0 : L.ops) {
            ListOp op = op0;
            if (!op.build && L.nb == 1) {
                op.p0 = q_begin;
                op.pc = q_end - q_begin;
            }
            if (op.build && !builds) continue;
            if (op.build) {
                const int first = op.block * L.bn;
                const int count = std::min(L.bn, L.U - first);
                const int base = op.half * L.bn;
                hipEvent_t b0 = pm ? pm->ev() : nullptr, b1 = pm ? pm->ev() : nullptr;
                if (b0) (void)hipEventRecord(b0, stream);
                if (L.nb > 1) {
                    hipLaunchKernelGGL(set_slots_kernel, dim3((unsigned)((count + 255) / 256)),
                                       dim3(256), 0, stream, L.node_of + first, count, base,
                                       L.rslot_of);
                    NSX_HIP(hipGetLastError(), "slot update launch");
                }
                ksgw::BuildArgs ba;
                ba.X = X;
                ba.ld = pb.ld;
                ba.dstride = pb.dstride;
                ba.n = n;
                ba.rs = rs;
                ba.ors = ors;
                ba.node_of = L.node_of + first;
                const int block_waves = ksgw::build_block_waves(width);
                const int rows_per_pass = block_waves / ksgw::build_row_waves(width);
                // Rows of up to 2048 frames: groups of rows per node, about 32
                // workgroups per CU in all (each stages its node's coordinates
                // in LDS), so the last round of workgroups is a small part.
                ba.groups = (int)std::max<long long>(
                    1, std::min<long long>((n + rows_per_pass - 1) / rows_per_pass,
                                           (32LL * cu_count() + count - 1) / std::max(count, 1)));
                ba.nodes = count;
                // Rows of more than 2048 frames are dealt from one range over
                // all the launch's nodes: 64 workgroups per CU in turn.
                const long long grid =
                    width <= 32 ? (long long)count * ba.groups
                                : std::max<long long>(1, std::min<long long>(
                                      ((long long)count * n + rows_per_pass - 1) / rows_per_pass,
                                      64LL * cu_count()));
                ba.ord = L.ord + (long long)base * ord_entries_for(n);
                ba.rank = L.rank + (long long)base * rank_entries_for(n);
                ba.prior = L.prior + (long long)base * ksgw::table_rows(n);
                ba.span = span;
                ba.abort = abort_flag;
                hipLaunchKernelGGL(buildk, dim3((unsigned)grid), dim3(ksgw::kWave * block_waves),
                                   ksgw::build_lds_bytes(pb.d, n, rs), stream, ba);
                NSX_HIP(hipGetLastError(), "list build launch");
                if (b1) (void)hipEventRecord(b1, stream);
                if (pm) pm->add(0, b0, b1);
                continue;
            }
            for (long long q0 = 0; q0 < op.pc; q0 += wchunk) {
                const long long Pc = std::min(wchunk, op.pc - q0);
                const long long p0 = op.p0 + q0;
                hipEvent_t e0 = pm ? pm->ev() : nullptr, e1 = pm ? pm->ev() : nullptr,
                           e2 = pm ? pm->ev() : nullptr, e3 = pm ? pm->ev() : nullptr;
                ksgw::WalkArgs wa;
                wa.X = X;
                wa.ld = pb.ld;
                wa.dstride = pb.dstride;
                wa.ab = L.ab;
                wa.p0 = p0;
                wa.P = Pc;
                wa.n = n;
                wa.k = pb.k;
                wa.ord = L.ord;
                wa.rank = L.rank;
                wa.prior = L.prior;
                wa.rs = rs;
                wa.ors = ors;
                wa.slot_of = L.nb > 1 ? L.rslot_of : L.slot_of;
                wa.cnt = L.cnt;
                wa.stats = stats;
                wa.tie_list = L.tie_list;
                wa.tie_count = tie_count;
                wa.tie_cap = wtie_cap;
                wa.exact_list = L.tie_list + L.tie_cap;
                wa.exact_count = tie_count + 1;
                wa.nbuf = nbuf;
                wa.batch = g_walk_batch;
                wa.checkpoint = 24;
                wa.span = span;
                wa.id_bits = ksgw::heap_id_bits(n);
                wa.code_shift = ksgw::heap_code_shift(n);
                wa.abort = abort_flag;
                if (!L.tables) wa.rank = nullptr;
                int rc = fill_u32(tie_count, 0u, 2, stream);
                if (rc) return rc;
                if (e0) (void)hipEventRecord(e0, stream);
                if (pb.k > ksg::kMaxK) {
                    // One wave per query frame; no tie list (ties are resolved
                    // in place).
                    const long long grid = std::max<long long>(
                        1, std::min<long long>(Pc * n, ws.grid_max * 8));
                    hipLaunchKernelGGL(ws.walk, dim3((unsigned)grid), dim3(ksgw::kWave),
                                       ksgw::bigk_lds_bytes(n), stream, wa);
                    NSX_HIP(hipGetLastError(), "large-k launch");
                    if (e1) (void)hipEventRecord(e1, stream);
                    if (e2) (void)hipEventRecord(e2, stream);
                } else {
                const long long grid = std::max<long long>(1, std::min<long long>(Pc, ws.grid_max));
                hipLaunchKernelGGL(ws.walk, dim3((unsigned)grid), dim3(ksgw::kWave * ws.wpb), wlds,
                                   stream, wa);
                NSX_HIP(hipGetLastError(), "walk launch");
                if (e1) (void)hipEventRecord(e1, stream);
                // Every deferred query (no LDS: 8 workgroups of 4 waves per
                // CU), then the exact ties among them: as many waves per CU
                // as their LDS rows and lists allow (up to 8, one workgroup
                // per CU; the waves take the entries in a grid stride).
                hipLaunchKernelGGL(ws.ties, dim3((unsigned)tie_waves), dim3(ksgw::kWave * 4), 0,
                                   stream, wa);
                NSX_HIP(hipGetLastError(), "tie select launch");
                const size_t tie_wave_bytes = ksgw::tie_wave_lds_bytes(n);
                const int tie_wpb = (int)std::max<size_t>(
                    1, std::min<size_t>(ksgw::kTieMaxWaves, lds_cap / tie_wave_bytes));
                const unsigned tie_groups = (unsigned)cu_count();
                hipLaunchKernelGGL(ws.tiek, dim3(tie_groups), dim3(ksgw::kWave * tie_wpb),
                                   (size_t)tie_wpb * tie_wave_bytes, stream, wa);
                NSX_HIP(hipGetLastError(), "tie launch");
                hipLaunchKernelGGL(ws.tieo, dim3(tie_groups), dim3(ksgw::kWave * tie_wpb),
                                   (size_t)tie_wpb * tie_wave_bytes, stream, wa);
                NSX_HIP(hipGetLastError(), "tie overflow launch");
                if (e2) (void)hipEventRecord(e2, stream);
                }
                ksgw::SumArgs sa;
                sa.cnt = L.cnt;
                sa.P = Pc;
                sa.n = n;
                sa.k = pb.k;
                sa.psi = d_psi;
                sa.inv_k = (float)(1.0 / (float)pb.k);
                sa.mi_out = out;
                sa.counts_out = counts_out;
                sa.perm = L.perm;
                sa.pbase = p0;
                sa.abort = abort_flag;
                const long long sum_blocks = (Pc + 64 * kSumWaves - 1) / (64 * kSumWaves);
                if (n + 1 <= kSumPsiLds + 1)
                    hipLaunchKernelGGL(psi_sum_kernel<true>, dim3((unsigned)sum_blocks),
                                       dim3(64 * kSumWaves), 0, stream, sa);
                else
                    hipLaunchKernelGGL(psi_sum_kernel<false>, dim3((unsigned)sum_blocks),
                                       dim3(64 * kSumWaves), 0, stream, sa);
                NSX_HIP(hipGetLastError(), "psi sum launch");
                if (e3) (void)hipEventRecord(e3, stream);
                if (pm) {
                    pm->add(1, e0, e1);
                    pm->add(2, e1, e2);
                    pm->add(3, e2, e3);
                }
            }
        }
    }
    return kOk;
}

thread_local int g_last_blocks = 0;
thread_local long long g_last_slots = 0, g_last_node_builds = 0;

// Runs the whole problem on the list path when it applies (windows of at
// most ksgw::kMaxFrames frames whose pair records fit LDS, room for at least
// two nodes' tables, finite coordinates); otherwise returns kOk with
// *ran == false. Windows run one after another, each with its own tables
// (the sorted lists depend on the frame set), into mi_out[rep * P + p].
int run_lists(const Problem &pb, const float *d_psi, float *mi_out, int *counts_out,
              hipStream_t stream, bool *ran) {
    *ran = false;
    if (g_method == 1 && pb.k <= ksg::kMaxK) return kOk;
    const size_t lds_cap = 160 * 1024;
    const size_t rec_bytes = ksgw::walk_buffer_bytes(pb.d, pb.nmax);
    const bool fits = pb.k > ksg::kMaxK ? ksgw::bigk_lds_bytes(pb.nmax) <= lds_cap
                                        : rec_bytes + 256 <= lds_cap;
    const bool eligible = pb.nmax <= ksgw::kMaxFrames && fits &&
                          (pb.R == 1 || counts_out == nullptr);
    if (!eligible) {
        NSX_REQUIRE(pb.k <= ksg::kMaxK,
                    "ksg: k > 16 needs windows of at most 8192 frames and, with several "
                    "windows, no per-frame counts");
        NSX_REQUIRE(g_method != 2,
                    "ksg: the list path needs windows of at most 8192 frames whose two rows "
                    "fit in LDS (2*d*n floats <= 160 KiB)");
        return kOk;
    }
    WalkShape ws;
    int rc = walk_shape(pb, ws);
    if (rc) return rc;
    for (int t = 0; t < 4; ++t) g_last_phase_ms[t] = 0;
    PhaseMarks pm;
    pm.on = true;
    hipEvent_t ev_start = g_phase_events.get(0);
    if (ev_start) (void)hipEventRecord(ev_start, stream);

    ListLayout L;
    bool ok = false;
    rc = list_prepare(pb, L, stream, &ok, ListMem{});
    if (rc || !ok) return rc;
    hipEvent_t ev_prep = pm.ev();
    if (ev_prep) (void)hipEventRecord(ev_prep, stream);
    pm.add(0, ev_start, ev_prep);  // the node map and pair order count as table set-up
    rc = list_enqueue(pb, L, ws, d_psi, mi_out, counts_out, stream, &pm);
    if (rc) return rc;
    unsigned long long h_stats[2] = {0, 0};
    NSX_HIP(hipMemcpyAsync(h_stats, L.flags + 8, sizeof(h_stats), hipMemcpyDeviceToHost, stream),
            "walk stats readback");
    NSX_HIP(hipStreamSynchronize(stream), "walk sync");
    for (const PhaseMarks::Mark &m : pm.marks) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, m.a, m.b) == hipSuccess) g_last_phase_ms[m.phase] += ms;
    }
    *ran = true;
    g_last_method = 2;
    g_last_entries = (long long)h_stats[1];
    g_last_ties = (long long)h_stats[0];
    g_last_blocks = L.nb;
    g_last_slots = L.S;
    g_last_node_builds = L.node_builds * pb.R;
    return kOk;
}

// ---- chunked pair loop (checkpointing callers) ----------------------------------

// All pairs in chunks [ends[c-1], ends[c]): on the list path the tables are
// built once and every chunk is a range of the same walk layout; otherwise
// (all-frames method, blocked tables) each chunk is an eager call. After
// chunk c its MI values are copied to h_mi_out on a second stream and
// on_chunk(user, ends[c]) runs while the device works on chunks c + 1 and
// c + 2.
int run_eager(const Problem &pb, float *mi_out, int *counts_out, hipStream_t stream);

int run_chunked(const Problem &pb, float *mi_out, float *h_mi_out, const long long *ends,
                int nends, int (*on_chunk)(void *, long long), void *user, hipStream_t stream) {
    int rc = ensure_device();
    if (rc) return rc;
    std::vector<float> psi(pb.nmax + 1);
    psi_table(psi.data(), pb.nmax);
    float *d_psi = (float *)workspace(kWsPsi, psi.size() * sizeof(float));
    if (!d_psi) return kHipError;
    NSX_HIP(hipMemcpyAsync(d_psi, psi.data(), psi.size() * sizeof(float), hipMemcpyHostToDevice,
                           stream),
            "psi upload");
    bool lists = false;
    ListLayout L;
    WalkShape ws;
    const size_t rec_bytes = ksgw::walk_buffer_bytes(pb.d, pb.nmax);
    const bool fits = pb.k > ksg::kMaxK ? ksgw::bigk_lds_bytes(pb.nmax) <= 160 * 1024
                                        : rec_bytes + 256 <= 160 * 1024;
    if ((g_method != 1 || pb.k > ksg::kMaxK) && pb.nmax <= ksgw::kMaxFrames && fits) {
        rc = walk_shape(pb, ws);
        if (rc) return rc;
        bool ok = false;
        rc = list_prepare(pb, L, stream, &ok, ListMem{});
        if (rc) return rc;
        lists = ok && L.nb == 1;
    }
    hipStream_t cs = nullptr;
    NSX_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "copy stream");
    struct Cleanup {
        hipStream_t s;
        std::vector<hipEvent_t> ev;
        ~Cleanup() {
            if (s) {
                (void)hipStreamSynchronize(s);
                (void)hipStreamDestroy(s);
            }
            for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        }
    } clean{cs, {}};
    // Device batches: consecutive chunks grouped until a batch holds at
    // least kBatchPairsPerCU pairs per CU (the last may hold fewer). A walk
    // launch ends with its slowest workgroup and the psi sums of a few
    // thousand pairs cannot fill the device, so each launch costs about
    // half a millisecond beyond its work; the chunks' callbacks then run in
    // order once their batch is done.
    constexpr long long kBatchPairsPerCU = 128;
    const long long min_batch = kBatchPairsPerCU * cu_count();
    std::vector<int> last;  // last chunk of each batch
    for (int c = 0; c < nends; ++c) {
        const long long b0 = last.empty() ? 0 : ends[last.back()];
        if (ends[c] - b0 >= min_batch || c == nends - 1) last.push_back(c);
    }
    const int nb = (int)last.size();
    clean.ev.assign(nb, nullptr);
    for (int t = 0; t < nb; ++t)
        NSX_HIP(hipEventCreateWithFlags(&clean.ev[t], hipEventDisableTiming), "chunk event");
    // Hands batch t's chunks to the host one by one (each copied just
    // before its callback, so a callback sees no later values) and calls
    // back.
    auto finish = [&](int t) -> int {
        const int c0 = t ? last[t - 1] + 1 : 0;
        NSX_HIP(hipStreamWaitEvent(cs, clean.ev[t], 0), "chunk wait");
        for (int c = c0; c <= last[t]; ++c) {
            const long long b = c ? ends[c - 1] : 0, e = ends[c];
            NSX_HIP(hipMemcpyAsync(h_mi_out + b, mi_out + b, (size_t)(e - b) * sizeof(float),
                                   hipMemcpyDeviceToHost, cs),
                    "chunk copy");
            NSX_HIP(hipStreamSynchronize(cs), "chunk copy sync");
            if (on_chunk && on_chunk(user, ends[c]) != 0) {
                (void)hipStreamSynchronize(stream);
                set_error("ksg: stopped by the chunk callback");
                return kInternal;
            }
        }
        return kOk;
    };
    long long ties = 0;
    for (int c = 0; c < nb; ++c) {
        const long long b = c ? ends[last[c - 1]] : 0, e = ends[last[c]];
        if (lists) {
            rc = list_enqueue(pb, L, ws, d_psi, mi_out, nullptr, stream, nullptr, b, e, c == 0);
        } else {
            Problem sub = pb;
            sub.ab = pb.ab + 2 * b;
            sub.P = e - b;
            rc = run_eager(sub, mi_out + b, nullptr, stream);
            ties += g_last_ties;
        }
        if (rc) return rc;
        NSX_HIP(hipEventRecord(clean.ev[c], stream), "chunk record");
        // Two chunks stay queued behind the one being handed over, so the
        // device keeps working while the callback runs.
        if (c >= 2) {
            rc = finish(c - 2);
            if (rc) return rc;
        }
    }
    for (int c = std::max(0, nb - 2); c < nb; ++c) {
        rc = finish(c);
        if (rc) return rc;
    }
    if (lists) {
        unsigned long long h_stats[2] = {0, 0};
        NSX_HIP(hipMemcpyAsync(h_stats, L.flags + 8, sizeof(h_stats), hipMemcpyDeviceToHost,
                               stream),
                "walk stats readback");
        NSX_HIP(hipStreamSynchronize(stream), "walk sync");
        g_last_method = 2;
        g_last_entries = (long long)h_stats[1];
        g_last_ties = (long long)h_stats[0];
    } else {
        g_last_ties = ties;
    }
    return kOk;
}

// ---- eager path -------------------------------------------------------------

int run_eager(const Problem &pb, float *mi_out, int *counts_out, hipStream_t stream) {
    NSX_REQUIRE(mi_out, "ksg: null output pointer");
    int rc = ensure_device();
    if (rc) return rc;
    g_last_ties = 0;
    g_last_method = 0;
    g_last_entries = 0;
    if (pb.P == 0) return kOk;
    if (g_method != 1 || pb.k > ksg::kMaxK) {
        std::vector<float> psi(pb.nmax + 1);
        psi_table(psi.data(), pb.nmax);
        float *d_psi = (float *)workspace(kWsPsi, psi.size() * sizeof(float));
        if (!d_psi) return kHipError;
        NSX_HIP(hipMemcpyAsync(d_psi, psi.data(), psi.size() * sizeof(float),
                               hipMemcpyHostToDevice, stream), "psi upload");
        bool ran = false;
        rc = run_lists(pb, d_psi, mi_out, counts_out, stream, &ran);
        if (rc || ran) return rc;
        NSX_REQUIRE(pb.k <= ksg::kMaxK, "ksg: internal: no path for k > 16");
    }
    g_last_method = 1;
    const Config c = config_for(pb);
    Shape s;
    rc = shape_for(pb, c, (long long)pb.R * pb.P, s);
    if (rc) return rc;
    std::vector<float> psi(pb.nmax + 1);
    psi_table(psi.data(), pb.nmax);
    Buffers b;
    b.psi = (float *)workspace(kWsPsi, psi.size() * sizeof(float));
    b.stats = (unsigned long long *)workspace(kWsCounter, 64);
    b.scratch = (KeyIdx *)workspace(
        kWsTieScratch,
        (size_t)s.blocks * c.wpb * ksg::scratch_records(pb.nmax, c.sym != 0) * sizeof(KeyIdx));
    b.win = (int *)workspace(kWsWindows, pb.host_win.size() * sizeof(int));
    if (!b.psi || !b.stats || !b.scratch || !b.win) return kHipError;
    NSX_HIP(hipMemcpyAsync(b.psi, psi.data(), psi.size() * sizeof(float),
                           hipMemcpyHostToDevice, stream), "psi upload");
    NSX_HIP(hipMemcpyAsync(b.win, pb.host_win.data(), pb.host_win.size() * sizeof(int),
                           hipMemcpyHostToDevice, stream), "window upload");
    rc = fill_u32(b.stats, 0u, 2, stream);
    if (rc) return rc;
    rc = launch(s, c, make_args(pb, c, b, mi_out, counts_out, pb.P), stream);
    if (rc) return rc;
    unsigned long long ties = 0;
    NSX_HIP(hipMemcpyAsync(&ties, b.stats, sizeof(ties), hipMemcpyDeviceToHost, stream),
            "tie count readback");
    NSX_HIP(hipStreamSynchronize(stream), "ksg sync");
    g_last_ties = (long long)ties;
    return kOk;
}

// ---- allocation-free variant of nsx_ksg_mi -------------------------------------
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc): the all-frames
// kernel with every buffer it needs besides X, ab and the output in one
// static device array, `noalloc_scratch`, of kNoAllocBytes (each GPU holds
// its own copy for the life of the process). Layout: [0, 64) the tie
// counter (two u64), [64, 76) the window record (offset 0, n, replicate 0),
// [256, ...) the psi table, then per wave one n-record tie row. The launch
// is capped to the waves whose rows fit. Every call on a GPU shares that
// GPU's array: calls that overlap on one GPU race.

constexpr size_t kNoAllocBytes = size_t(64) << 20;
constexpr int kNoAllocMaxFrames = 1 << 18;  // psi (1 MiB) and 4 waves' rows (8 MiB each) fit

__device__ __align__(256) unsigned char noalloc_scratch[kNoAllocBytes];

// Device address of noalloc_scratch on each device (the array's name is a
// host-side stand-in; kernels take this address). Filled once per device,
// with that device current.
std::mutex g_noalloc_mutex;
void *g_noalloc_ptr[64] = {};

int noalloc_base(void **base) {
    int dev = 0;
    NSX_HIP(hipGetDevice(&dev), "hipGetDevice");
    NSX_REQUIRE(dev >= 0 && dev < 64, "nsx_ksg_mi_noalloc: device index above 63");
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        NSX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch)),
                "hipGetSymbolAddress(noalloc_scratch)");
        g_noalloc_ptr[dev] = p;
    }
    *base = g_noalloc_ptr[dev];
    return kOk;
}

// psi.cpp:6-19 on the device: the terms (float)(1.0 / (double)(float)i) in
// parallel, then one lane adds them in order from psi[1] =
// (float)-0.57721566490153 (the same float sums as psi_table on the host).
__global__ __launch_bounds__(256) void psi_table_kernel(float *psi, int n) {
    for (int i = (int)threadIdx.x + 1; i < n; i += (int)blockDim.x)
        psi[i + 1] = (float)(1.0 / (double)(float)i);
    __syncthreads();
    if (threadIdx.x == 0) {
        psi[0] = 0.0f;
        if (n >= 1) {
            float acc = (float)-0.57721566490153;
            psi[1] = acc;
            for (int i = 1; i < n; ++i) {
                acc = acc + psi[i + 1];
                psi[i + 1] = acc;
            }
        }
    }
}

int run_noalloc(const Problem &pb, float *mi_out, hipStream_t stream) {
    void *vbase = nullptr;
    int rc = noalloc_base(&vbase);
    if (rc) return rc;
    char *base = (char *)vbase;
    const int n = pb.nmax;
    const size_t psi_off = 256;
    const size_t rows_off = (psi_off + (size_t)(n + 1) * sizeof(float) + 255) & ~(size_t)255;
    Config c = config_for(pb);
    c.sym = 0;  // plain pass 2: one n-record row per wave
    Shape s;
    rc = shape_for(pb, c, pb.P, s);
    if (rc) return rc;
    const long long waves_fit =
        (long long)((kNoAllocBytes - rows_off) / ((size_t)n * sizeof(KeyIdx)));
    NSX_REQUIRE(waves_fit >= c.wpb, "nsx_ksg_mi_noalloc: the static scratch holds too few tie rows");
    s.blocks = std::max<long long>(1, std::min<long long>(s.blocks, waves_fit / c.wpb));
    Buffers b;
    b.stats = (unsigned long long *)base;
    b.win = (int *)(base + 64);
    b.psi = (float *)(base + psi_off);
    b.scratch = (KeyIdx *)(base + rows_off);
    rc = fill_u32(b.stats, 0u, 4, stream);
    if (!rc) rc = fill_u32(b.win, 0u, 1, stream);
    if (!rc) rc = fill_u32(b.win + 1, (uint32_t)n, 1, stream);
    if (!rc) rc = fill_u32(b.win + 2, 0u, 1, stream);
    if (rc) return rc;
    hipLaunchKernelGGL(psi_table_kernel, dim3(1), dim3(256), 0, stream, b.psi, n);
    NSX_HIP(hipGetLastError(), "psi table launch");
    return launch(s, c, make_args(pb, c, b, mi_out, nullptr, pb.P), stream);
}

}  // namespace
}  // namespace nsx

// ---- plans: hipGraph-captured batched launches --------------------------------

struct nsx_ksg_plan_s {
    nsx::Problem pb;
    nsx::Config cfg;
    nsx::Shape shape;
    nsx::Buffers buf;
    float *mi_out = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double tune_ms = 0;
    int method = 1;              // 1 all-frames graph, 2 list-path graph
    nsx::ListLayout L;           // list-path graph: layout over `owned`
    std::vector<void *> owned;   // list-path graph buffers
};

using namespace nsx;

namespace {

void destroy_plan(nsx_ksg_plan_s *p) {
    if (!p) return;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    (void)hipFree(p->buf.psi);
    (void)hipFree(p->buf.win);
    (void)hipFree(p->buf.scratch);
    (void)hipFree(p->buf.stats);
    for (void *q : p->owned) (void)hipFree(q);
    delete p;
}

}  // namespace

extern "C" {

int nsx_ksg_mi(const float *d_X, long long ld, long long dstride, int d, int n,
               const int *d_ab, long long P, int k, float *d_mi_out, int *d_counts_out,
               void *stream) {
    TraceRange trace("nsx_ksg_mi");
    clear_error();
    const auto t0 = std::chrono::steady_clock::now();
    const int off = 0;
    Problem pb;
    int rc = make_problem(pb, d_X, ld, dstride, d, &off, &n, 1, d_ab, P, k);
    if (!rc) rc = run_eager(pb, d_mi_out, d_counts_out, (hipStream_t)stream);
    g_last_call_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}

int nsx_ksg_mi_windows(const float *d_X, long long ld, long long dstride, int d,
                       const int *h_win_off, const int *h_win_n, int R, const int *d_ab,
                       long long P, int k, float *d_mi_out, void *stream) {
    TraceRange trace("nsx_ksg_mi_windows");
    clear_error();
    NSX_REQUIRE(h_win_off && h_win_n && R >= 1, "ksg windows: bad window arrays");
    const auto t0 = std::chrono::steady_clock::now();
    Problem pb;
    int rc = make_problem(pb, d_X, ld, dstride, d, h_win_off, h_win_n, R, d_ab, P, k);
    if (!rc) rc = run_eager(pb, d_mi_out, nullptr, (hipStream_t)stream);
    g_last_call_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}

int nsx_ksg_mi_noalloc(const float *d_X, long long ld, long long dstride, int d, int n,
                       const int *d_ab, long long P, int k, float *d_mi_out, void *stream) {
    TraceRange trace("nsx_ksg_mi_noalloc");
    clear_error();
    NSX_REQUIRE(d_mi_out, "nsx_ksg_mi_noalloc: null output pointer");
    NSX_REQUIRE(k <= ksg::kMaxK, "nsx_ksg_mi_noalloc: k must be at most 16");
    NSX_REQUIRE(n >= 1 && n <= kNoAllocMaxFrames,
                "nsx_ksg_mi_noalloc: n must be at most 262144 frames (the psi table and the "
                "per-wave tie rows live in a 64 MiB static scratch)");
    const int off = 0;
    Problem pb;
    int rc = make_problem(pb, d_X, ld, dstride, d, &off, &n, 1, d_ab, P, k);
    if (rc) return rc;
    rc = ensure_device();
    if (rc) return rc;
    if (P == 0) return kOk;
    return run_noalloc(pb, d_mi_out, (hipStream_t)stream);
}

int nsx_ksg_mi_chunked(const float *d_X, long long ld, long long dstride, int d, int n,
                       const int *d_ab, long long P, int k, float *d_mi_out, float *h_mi_out,
                       const long long *h_ends, int nends, nsx_chunk_fn on_chunk, void *user,
                       void *stream) {
    TraceRange trace("nsx_ksg_mi_chunked");
    clear_error();
    NSX_REQUIRE(d_mi_out && h_mi_out && h_ends && nends >= 1,
                "nsx_ksg_mi_chunked: invalid arguments");
    for (int c = 0; c < nends; ++c)
        NSX_REQUIRE(h_ends[c] > (c ? h_ends[c - 1] : 0) && h_ends[c] <= P,
                    "nsx_ksg_mi_chunked: chunk ends must increase strictly up to P");
    NSX_REQUIRE(h_ends[nends - 1] == P, "nsx_ksg_mi_chunked: the last chunk must end at P");
    const int off = 0;
    Problem pb;
    int rc = make_problem(pb, d_X, ld, dstride, d, &off, &n, 1, d_ab, P, k);
    if (rc) return rc;
    g_last_method = 0;
    return run_chunked(pb, d_mi_out, h_mi_out, h_ends, nends, on_chunk, user,
                       (hipStream_t)stream);
}

int nsx_ksg_last_tie_frames(long long *count) {
    if (count) *count = g_last_ties;
    return kOk;
}

int nsx_ksg_set_list_cap(long long entries) {
    clear_error();
    NSX_REQUIRE(entries >= 0 && entries <= ksgw::kMaxFrames,
                "nsx_ksg_set_list_cap: entries must be 0..8192");
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    g_list_cap = entries;
    return kOk;
}

int nsx_ksg_set_method(int method, int walk_waves, int walk_batch) {
    clear_error();
    NSX_REQUIRE(method >= 0 && method <= 2, "nsx_ksg_set_method: method must be 0, 1 or 2");
    NSX_REQUIRE(walk_waves >= 0 && walk_waves <= ksgw::kMaxWalkWaves,
                "nsx_ksg_set_method: walk_waves must be 0..16");
    NSX_REQUIRE(walk_batch >= 0 && walk_batch <= 64, "nsx_ksg_set_method: walk_batch must be 0..64");
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    g_method = method;
    g_walk_wpb = walk_waves ? walk_waves : 16;
    g_walk_batch = walk_batch ? walk_batch : 16;
    return kOk;
}

int nsx_ksg_last_method(int *method, long long *list_entries) {
    if (method) *method = g_last_method;
    if (list_entries) *list_entries = g_last_entries;
    return kOk;
}

int nsx_ksg_last_phase_ms(double *phase_ms) {
    if (phase_ms)
        for (int t = 0; t < 4; ++t) phase_ms[t] = g_last_phase_ms[t];
    return kOk;
}

int nsx_ksg_last_call_ms(double *wall_ms) {
    if (wall_ms) *wall_ms = g_last_call_ms;
    return kOk;
}

int nsx_ksg_last_list_info(int *blocks, long long *slots, long long *node_builds) {
    if (blocks) *blocks = g_last_method == 2 ? g_last_blocks : 0;
    if (slots) *slots = g_last_method == 2 ? g_last_slots : 0;
    if (node_builds) *node_builds = g_last_method == 2 ? g_last_node_builds : 0;
    return kOk;
}

int nsx_ksg_set_config(int qpl, int chunk, int wpb, int blocks_per_cu, int symmetric) {
    clear_error();
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    if (qpl == 0) {
        g_has_override = false;
        return kOk;
    }
    NSX_REQUIRE((qpl == 1 || qpl == 2) && (chunk == 64 || chunk == 128 || chunk == 256) &&
                    wpb >= 1 && wpb <= ksg::kMaxWavesPerBlock && blocks_per_cu >= 0 &&
                    (symmetric == 0 || symmetric == 1),
                "nsx_ksg_set_config: bad configuration");
    g_override = Config{qpl, chunk, wpb, blocks_per_cu, symmetric};
    g_has_override = true;
    return kOk;
}

int nsx_ksg_plan_create(nsx_ksg_plan *plan, const float *d_X, long long ld, long long dstride,
                        int d, const int *h_win_off, const int *h_win_n, int R,
                        const int *d_ab, long long P, int k, float *d_mi_out, int autotune) {
    TraceRange trace("nsx_ksg_plan_create");
    clear_error();
    NSX_REQUIRE(plan && d_mi_out && h_win_off && h_win_n && R >= 1 && P >= 1,
                "nsx_ksg_plan_create: invalid arguments");
    NSX_REQUIRE(k <= ksg::kMaxK, "nsx_ksg_plan_create: batched plans support 1 <= k <= 16");
    *plan = nullptr;
    int rc = ensure_device();
    if (rc) return rc;
    auto *p = new nsx_ksg_plan_s();
    rc = make_problem(p->pb, d_X, ld, dstride, d, h_win_off, h_win_n, R, d_ab, P, k);
    if (rc) {
        delete p;
        return rc;
    }
    const Problem &pb = p->pb;
    p->mi_out = d_mi_out;
    // Plan-owned buffers (allocated once, reused by every replay).
    std::vector<float> psi(pb.nmax + 1);
    psi_table(psi.data(), pb.nmax);
    const size_t scratch_bytes =
        (size_t)max_waves() * ksg::scratch_records(pb.nmax, true) * sizeof(KeyIdx);
    hipError_t e = hipMalloc(&p->buf.psi, psi.size() * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&p->buf.win, pb.host_win.size() * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&p->buf.scratch, scratch_bytes);
    if (e == hipSuccess) e = hipMalloc(&p->buf.stats, 64);
    if (e == hipSuccess)
        e = hipMemcpy(p->buf.psi, psi.data(), psi.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(p->buf.win, pb.host_win.data(), pb.host_win.size() * sizeof(int),
                      hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        set_error(std::string("nsx_ksg_plan_create: ") + hipGetErrorString(e));
        destroy_plan(p);
        return kHipError;
    }
    hipStream_t cs = nullptr;
    e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e != hipSuccess) {
        set_error(std::string("capture stream: ") + hipGetErrorString(e));
        destroy_plan(p);
        return kHipError;
    }
    // The list path when it applies (windows of at most 8192 frames whose
    // pair records fit LDS; method 0 or 2): sized now, with plan-owned
    // buffers, from this plan's fixed pairs.
    if (g_method != 1) {
        const size_t rec = (size_t)2 * pb.d * pb.nmax * sizeof(float);
        const bool eligible = pb.nmax <= ksgw::kMaxFrames && rec + 256 <= 160 * 1024;
        if (!eligible && g_method == 2) {
            set_error("nsx_ksg_plan_create: the list path needs windows of at most 8192 frames "
                      "whose two rows fit in LDS");
            rc = kInvalidArgument;
        }
        if (!rc && eligible) {
            bool ok = false;
            rc = list_prepare(pb, p->L, cs, &ok, ListMem{&p->owned});
            if (!rc && ok) p->method = 2;
        }
        if (rc) {
            (void)hipStreamSynchronize(cs);
            (void)hipStreamDestroy(cs);
            destroy_plan(p);
            return rc;
        }
    }
    if (autotune && p->method == 1) {
        Config best;
        rc = nsx::autotune(pb, p->buf, cs, best, &p->tune_ms);
        if (rc) {
            (void)hipStreamDestroy(cs);
            destroy_plan(p);
            return rc;
        }
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        store_tune(tune_key(pb), best);
        p->cfg = best;
    } else {
        p->cfg = config_for(pb);
    }
    rc = shape_for(pb, p->cfg, (long long)pb.R * pb.P, p->shape);
    if (!rc && (size_t)p->shape.blocks * p->cfg.wpb *
                       ksg::scratch_records(pb.nmax, p->cfg.sym != 0) * sizeof(KeyIdx) >
                   scratch_bytes) {
        set_error("nsx_ksg_plan_create: tie scratch too small for the chosen shape");
        rc = kInternal;
    }
    WalkShape ws;
    if (!rc && p->method == 2) rc = walk_shape(pb, ws);
    // Capture. All-frames: reset the tie counter (a fill kernel, not a
    // memset node), then the persistent launch. List path: clear the flags,
    // the spans and finite check of every window, every window's builds,
    // walks, ties and psi sums, then the all-frames launch guarded by the
    // non-finite flag (it returns at once unless a coordinate is not finite).
    if (!rc) {
        e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            if (p->method == 2) {
                rc = fill_u32(p->L.flags, 0u, 32, cs);
                if (!rc) rc = enqueue_node_stats(pb, p->L, cs);
                if (!rc) rc = list_enqueue(pb, p->L, ws, p->buf.psi, d_mi_out, nullptr, cs, nullptr);
            }
            if (!rc) rc = fill_u32(p->buf.stats, 0u, 2, cs);
            Args fa = make_args(pb, p->cfg, p->buf, d_mi_out, nullptr, pb.P);
            if (p->method == 2) fa.run_if = p->L.flags;
            if (!rc) rc = launch(p->shape, p->cfg, fa, cs);
            hipError_t e2 = hipStreamEndCapture(cs, &p->graph);
            if (!rc && e2 != hipSuccess) {
                set_error(std::string("end capture: ") + hipGetErrorString(e2));
                rc = kHipError;
            }
        } else {
            set_error(std::string("begin capture: ") + hipGetErrorString(e));
            rc = kHipError;
        }
    }
    if (!rc) {
        e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            set_error(std::string("graph instantiate: ") + hipGetErrorString(e));
            rc = kHipError;
        }
    }
    (void)hipStreamSynchronize(cs);
    (void)hipStreamDestroy(cs);
    if (rc) {
        destroy_plan(p);
        return rc;
    }
    *plan = p;
    return kOk;
}

int nsx_ksg_plan_launch(nsx_ksg_plan plan, void *stream) {
    TraceRange trace("nsx_ksg_plan_launch");
    clear_error();
    NSX_REQUIRE(plan && plan->exec, "nsx_ksg_plan_launch: no plan");
    NSX_HIP(hipGraphLaunch(plan->exec, (hipStream_t)stream), "hipGraphLaunch");
    return kOk;
}

int nsx_ksg_plan_tie_frames(nsx_ksg_plan plan, long long *count) {
    clear_error();
    NSX_REQUIRE(plan && count, "nsx_ksg_plan_tie_frames: invalid arguments");
    unsigned long long v = 0, w = 0;
    NSX_HIP(hipDeviceSynchronize(), "plan sync");
    NSX_HIP(hipMemcpy(&v, plan->buf.stats, sizeof(v), hipMemcpyDeviceToHost), "tie readback");
    if (plan->method == 2)
        NSX_HIP(hipMemcpy(&w, plan->L.flags + 8, sizeof(w), hipMemcpyDeviceToHost),
                "tie readback");
    *count = (long long)(v + w);
    return kOk;
}

int nsx_ksg_plan_method(nsx_ksg_plan plan, int *method) {
    clear_error();
    NSX_REQUIRE(plan && method, "nsx_ksg_plan_method: invalid arguments");
    *method = plan->method;
    return kOk;
}

int nsx_ksg_plan_config(nsx_ksg_plan plan, int *qpl, int *chunk, int *wpb, long long *blocks,
                        double *tune_ms, int *symmetric) {
    clear_error();
    NSX_REQUIRE(plan, "nsx_ksg_plan_config: no plan");
    if (symmetric) *symmetric = plan->cfg.sym;
    if (qpl) *qpl = plan->cfg.qpl;
    if (chunk) *chunk = plan->cfg.chunk;
    if (wpb) *wpb = plan->cfg.wpb;
    if (blocks) *blocks = plan->shape.blocks;
    if (tune_ms) *tune_ms = plan->tune_ms;
    return kOk;
}

int nsx_ksg_plan_destroy(nsx_ksg_plan plan) {
    if (plan) (void)hipStreamSynchronize(nullptr);
    destroy_plan(plan);
    return kOk;
}

}  // extern "C"
