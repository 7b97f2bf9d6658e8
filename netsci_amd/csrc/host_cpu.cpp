// CPU side of the `platform` flag (netcalc/include/platform.h:9-10):
// KSG MI over pairs, the GC transform, Hedetniemi APSP and
// correlation-to-adjacency on host threads. Results equal the GPU path's,
// which equal the reference CPU path's (mutual_information_cpu.cpp:6-121).
//
// The per-frame k-nearest selection uses sort_select.h, which reproduces the
// first k+1 positions of libstdc++'s std::sort without sorting the whole row,
// so tie order matches the reference while the selection costs O(n).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/netsci_hip.h"
#include "sort_select.h"

namespace nsx {
void set_error(const std::string &msg);
void clear_error();
void psi_table(float *psi, int n);
}  // namespace nsx

namespace {

// Worker count: the caller's, else OMP_NUM_THREADS when set (the usual way
// a job states its CPU share), else every hardware thread.
int pick_threads(int threads) {
    if (threads > 0) return threads;
    if (const char *env = std::getenv("OMP_NUM_THREADS")) {
        const int t = std::atoi(env);
        if (t > 0) return t;
    }
    unsigned hc = std::thread::hardware_concurrency();
    return hc ? (int)hc : 1;
}

// Run body(t) for t in [0, count) on `threads` workers, dynamically scheduled.
template <class F>
void parallel_for(long long count, int threads, F body) {
    threads = (int)std::min<long long>(pick_threads(threads), std::max<long long>(count, 1));
    if (threads <= 1) {
        for (long long t = 0; t < count; ++t) body(t);
        return;
    }
    std::atomic<long long> next{0};
    std::vector<std::thread> pool;
    for (int w = 0; w < threads; ++w) {
        pool.emplace_back([&]() {
            while (true) {
                long long t = next.fetch_add(1);
                if (t >= count) break;
                body(t);
            }
        });
    }
    for (auto &th : pool) th.join();
}

inline float marg(const float *x, long long ds, int d, int i, int j) {
    float r = std::abs(x[i] - x[j]);
    for (int di = 1; di < d; ++di) r = std::max(r, std::abs(x[i + di * ds] - x[j + di * ds]));
    return r;
}

// One pair: mutual_information_cpu.cpp:22-118 with the O(n) tie-exact selection.
float ksg_pair_cpu(const float *xa, const float *xb, long long ds, int d, int n, int k,
                   const float *psi, float inv_k, std::vector<nsx::KeyIdx> &row) {
    row.resize(n);
    float acc = 0.0f;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) {
            row[j].key = std::max(marg(xa, ds, d, i, j), marg(xb, ds, d, i, j));
            row[j].idx = j;
        }
        nsx::sort_prefix(row.data(), n, k + 1);
        float ea = 0.0f, eb = 0.0f;
        for (int m = 1; m <= k; ++m) {
            ea = std::max(ea, marg(xa, ds, d, i, row[m].idx));
            eb = std::max(eb, marg(xb, ds, d, i, row[m].idx));
        }
        int ca = 0, cb = 0;
        for (int j = 0; j < n; ++j) {
            if (marg(xa, ds, d, i, j) <= ea) ++ca;
            if (marg(xb, ds, d, i, j) <= eb) ++cb;
        }
        acc += (psi[ca - 1] + psi[cb - 1]);
    }
    acc /= (float)n;
    return psi[n] + psi[k] - inv_k - acc;
}

}  // namespace

extern "C" {

int nsx_cpu_ksg_mi(const float *h_X, long long ld, long long dstride, int d, int n,
                   const int *h_ab, long long P, int k, float *h_mi_out, int threads) {
    nsx::clear_error();
    if (!h_X || !h_ab || !h_mi_out || d < 1 || n < 1 || k < 1 || k >= n || P < 0 ||
        ld <= 0 || dstride < n || (long long)(d - 1) * dstride + n > ld) {
        nsx::set_error("nsx_cpu_ksg_mi: invalid arguments (need 1 <= k < n, d >= 1, "
                       "d*n within a row)");
        return 1;
    }
    std::vector<float> psi(n + 1);
    nsx::psi_table(psi.data(), n);
    const float inv_k = (float)(1.0 / (float)k);
    parallel_for(P, threads, [&](long long p) {
        thread_local std::vector<nsx::KeyIdx> row;
        const float *xa = h_X + (long long)h_ab[2 * p] * ld;
        const float *xb = h_X + (long long)h_ab[2 * p + 1] * ld;
        h_mi_out[p] = ksg_pair_cpu(xa, xb, dstride, d, n, k, psi.data(), inv_k, row);
    });
    return 0;
}

int nsx_select_prefix_host(const float *h_key, int n, int m, int *h_idx, int *h_heap_ranges) {
    nsx::clear_error();
    if (!h_key || !h_idx || n < 0 || m < 0 || m > n) {
        nsx::set_error("nsx_select_prefix_host: invalid arguments");
        return 1;
    }
    std::vector<nsx::KeyIdx> row(n);
    for (int j = 0; j < n; ++j) row[j] = nsx::KeyIdx{h_key[j], j};
    int heaps = 0;
    nsx::sort_prefix(row.data(), n, m, &heaps);
    for (int j = 0; j < m; ++j) h_idx[j] = row[j].idx;
    if (h_heap_ranges) *h_heap_ranges = heaps;
    return 0;
}

// generalized_correlation_cpu.cpp:8-27 / generalized_correlation_gpu.cpp:8-33.
int nsx_gc_from_mi_host(const float *h_mi, long long P, int d, float *h_gc) {
    nsx::clear_error();
    if (!h_mi || !h_gc || d < 1 || P < 0) {
        nsx::set_error("nsx_gc_from_mi_host: invalid arguments");
        return 1;
    }
    // Chunks of 32768 values on up to 16 threads (each value on its own, so
    // the split does not change a bit); short inputs stay on this thread.
    constexpr long long kChunk = 32768;
    const long long chunks = (P + kChunk - 1) / kChunk;
    const int threads = P < 2 * kChunk ? 1 : std::min(pick_threads(0), 16);
    parallel_for(chunks, threads, [&](long long c) {
        const long long end = std::min(P, (c + 1) * kChunk);
        for (long long p = c * kChunk; p < end; ++p) {
            const float mi = h_mi[p];
            if (mi <= 0.0) h_gc[p] = 0.0f;
            else h_gc[p] = (float)std::sqrt(1.0 - (float)std::exp(-(2.0 / (float)d) * mi));
        }
    });
    return 0;
}

// hedetniemi_gpu.cu:8-26 on the host.
// Row mean / standard deviation in the summation order of
// math_utils_gpu.cu:23-156 (1024 strided partial sums folded as a tree), so
// the CPU platform gives the GPU platform's bits. The reference's CPU
// branch computes nothing (math_utils.cpp:6-35).
static float row_fold(const float *row, long long n, float shift, bool square) {
    float s[1024];
    for (int t = 0; t < 1024; ++t) {
        float acc = 0.0f;
        for (long long i = t; i < n; i += 1024) {
            if (square) {
                const float d = row[i] - shift;
                acc += d * d;
            } else {
                acc += row[i];
            }
        }
        s[t] = acc;
    }
    for (int o = 512; o >= 1; o >>= 1)
        for (int t = 0; t < o; ++t) s[t] = s[t] + s[t + o];
    return s[0];
}

int nsx_cpu_row_stats(const float *h_a, long long m, long long n, float *h_u, float *h_sigma) {
    nsx::clear_error();
    if (!h_a || !h_u || m < 0 || n < 1) {
        nsx::set_error("nsx_cpu_row_stats: invalid arguments");
        return 1;
    }
    parallel_for(m, 0, [&](long long r) {
        const float *row = h_a + r * n;
        const float mean = row_fold(row, n, 0.0f, false) / (float)n;
        h_u[r] = mean;
        if (h_sigma) h_sigma[r] = std::sqrt(row_fold(row, n, mean, true) / (float)n);
    });
    return 0;
}

int nsx_cpu_corr_to_adj(const float *h_C, int N, float *h_A) {
    nsx::clear_error();
    if (!h_C || !h_A || N < 0) {
        nsx::set_error("nsx_cpu_corr_to_adj: invalid arguments");
        return 1;
    }
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            const long long e = (long long)i * N + j;
            if (i == j) h_A[e] = 0.0f;
            else if (h_C[e] == 0.0f) h_A[e] = INFINITY;
            else h_A[e] = 1.0f / h_C[e];
        }
    return 0;
}

// hedetniemi_gpu.cu:47-107 in Jacobi form on host threads (one output row
// per task). Same recurrence, k order, comparisons and backtrack rule as the
// GPU path. h_paths (nullable) is N*L*N ints.
int nsx_cpu_hedetniemi(const float *h_A, int N, int L, float tol, float *h_H, int *h_paths,
                       int threads) {
    nsx::clear_error();
    if (!h_A || !h_H || N < 1 || L < 0) {
        nsx::set_error("nsx_cpu_hedetniemi: invalid arguments");
        return 1;
    }
    const long long NN = (long long)N * N;
    std::vector<float> cur(h_A, h_A + NN), nxt(NN);
    std::vector<float> hist;
    if (h_paths) hist.resize((size_t)L * NN);
    for (int p = 0; p < L; ++p) {
        if (h_paths) std::copy(cur.begin(), cur.end(), hist.begin() + (size_t)p * NN);
        parallel_for(N, threads, [&](long long i) {
            // H_{p+1}[i][j] = min_k fl(A[j][k] + H_p[k][i]).
            std::vector<float> col(N);
            for (int kk = 0; kk < N; ++kk) col[kk] = cur[(long long)kk * N + i];
            for (int j = 0; j < N; ++j) {
                const float *arow = h_A + (long long)j * N;
                float c = INFINITY;
                for (int kk = 0; kk < N; ++kk) {
                    const float s = arow[kk] + col[kk];
                    if (s < c) c = s;
                }
                nxt[i * N + j] = c;
            }
        });
        cur.swap(nxt);
    }
    std::copy(cur.begin(), cur.end(), h_H);
    if (!h_paths) return 0;
    parallel_for(N, threads, [&](long long i) {
        for (int j = 0; j < N; ++j) {
            int *pp = h_paths + i * N * (long long)L + (long long)j * L;
            if (i == j) {
                for (int p = 0; p < L; ++p) pp[p] = (int)i;
                continue;
            }
            float len = cur[i * N + j];
            int k = j;
            for (int p = L - 1; p >= 0; --p) {
                const float *hrow = hist.data() + (size_t)p * NN + i * N;
                for (int q = 0; q < N; ++q) {
                    if (q == j) continue;
                    const float aq = h_A[(long long)q * N + k];
                    if (hrow[q] + aq - len < tol) {
                        len -= aq;
                        k = q;
                        break;
                    }
                }
                pp[p] = k;
            }
        }
    });
    return 0;
}

}  // extern "C"
