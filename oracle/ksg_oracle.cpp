// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A plain CPU restatement of the reference's KSG mutual-information,
// generalized-correlation and Hedetniemi semantics. It is the checker that
// the HIP path is compared against in tests/, in __graft_entry__.smoke() and
// in bench.py's cpu_baseline leg. No product code links, loads or calls it.
//
// Every function follows the reference file:line named in its comment, in the
// same arithmetic order and types, so that results are bit-identical:
//   * float32 throughout the KSG path, double only where the reference
//     promotes (the psi reciprocal, the GC transform);
//   * std::sort of an iota index array by joint distance (libstdc++'s
//     unstable introsort decides ties exactly as it does in the reference);
//   * counts with `<=` and self excluded by subtracting one;
//   * a sequential float32 accumulation of psi terms in frame order.
//
// Pinning (what checks this restatement, beyond re-reading it):
//   * orc_mi reproduces the reference's golden MI values
//     (tests/netcalc/python/test_MutualInformation.py:37,64) bit-exactly;
//   * brute-force k-NN on tiny inputs, the Gaussian closed form, self-pair
//     closed form and symmetry (tests/test_oracle.py);
//   * orc_hedetniemi distances agree with a textbook APSP (scipy, float64)
//     at convergence and paths decode to valid shortest paths
//     (tests/test_oracle.py). The reference's own Part-1 kernel updates H in
//     place (a race); this restatement is the deterministic Jacobi form.
//
// Build: oracle/Makefile (g++ -O3 -ffp-contract=off), output oracle/liboracle.so.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

extern "C" {

// psi.cpp:6-19 — psi[0] = 0 (CuArray::init zero-fills, cuarray.cu:35-39),
// psi[1] = (float)-0.57721566490153, psi[i+1] = psi[i] + (float)(1.0/(float)i).
void orc_psi(float *psi, int n) {
    for (int i = 0; i <= n; ++i) psi[i] = 0.0f;
    if (n >= 1) psi[1] = (float)-0.57721566490153;
    for (int i = 0; i < n; i++) {
        if (i > 0) {
            float inv = (float)(1.0 / static_cast<float>(i));
            psi[i + 1] = psi[i] + inv;
        }
    }
}

// mutual_information_cpu.cpp:27-47 — marginal max-norm distance of frames i, j.
static inline float marg_dist(const float *x, int i, int j, int n, int d) {
    float dx = std::abs(x[i] - x[j]);
    for (int di = 1; di < d; di++) {
        dx = std::max(dx, std::abs(x[i + di * n] - x[j + di * n]));
    }
    return dx;
}

// mutual_information_cpu.cpp:22-109 — per-frame marginal counts (already
// minus self) and the epsilons, exactly as the reference derives them.
void orc_ksg_counts(const float *xa, const float *xb, int k, int n, int d,
                    int *nxa, int *nxb, float *eps_a, float *eps_b) {
    std::vector<float> dz(n);
    std::vector<int> argmin(n);
    for (int i = 0; i < n; i++) {
        int ca = 0, cb = 0;
        for (int j = 0; j < n; j++) {
            float dxa = marg_dist(xa, i, j, n, d);
            float dxb = marg_dist(xb, i, j, n, d);
            dz[j] = std::max(dxa, dxb);
        }
        std::iota(argmin.begin(), argmin.end(), 0);
        std::sort(argmin.begin(), argmin.end(),
                  [&](int p, int q) { return dz[p] < dz[q]; });
        float ea = 0.0f, eb = 0.0f;
        for (int ki = 1; ki < k + 1; ki++) {
            ea = std::max(ea, marg_dist(xa, i, argmin[ki], n, d));
            eb = std::max(eb, marg_dist(xb, i, argmin[ki], n, d));
        }
        for (int j = 0; j < n; j++) {
            if (marg_dist(xa, i, j, n, d) <= ea) ca++;
            if (marg_dist(xb, i, j, n, d) <= eb) cb++;
        }
        nxa[i] = ca - 1;
        nxb[i] = cb - 1;
        if (eps_a) eps_a[i] = ea;
        if (eps_b) eps_b[i] = eb;
    }
}

// mutual_information_cpu.cpp:110-118 — sequential float32 epilogue.
float orc_mi_from_counts(const int *nxa, const int *nxb, int k, int n) {
    std::vector<float> psi(n + 1);
    orc_psi(psi.data(), n);
    float avg = 0.0f;
    for (int i = 0; i < n; i++) avg += (psi[nxa[i]] + psi[nxb[i]]);
    avg /= (float)n;
    return psi[n] + psi[k] - (float)(1.0 / (float)k) - avg;
}

// mutual_information_cpu.cpp:6-121 — mutualInformationCpu.
float orc_mi(const float *xa, const float *xb, int k, int n, int d) {
    std::vector<int> a(n), b(n);
    orc_ksg_counts(xa, xb, k, n, d, a.data(), b.data(), nullptr, nullptr);
    return orc_mi_from_counts(a.data(), b.data(), k, n);
}

// generalized_correlation_cpu.cpp:8-27 — GC transform in host double.
float orc_gc_from_mi(float mi, int d) {
    if (mi <= 0.0) return 0.0;
    return (float)std::sqrt(1.0 - (float)std::exp(-(2.0 / (float)d) * mi));
}

// generalized_correlation.cpp:104-173 / mutual_information.cpp:103-172 —
// the pair loop: row views of X (stride ld floats), one value per ab row.
void orc_pairs(const float *X, long long ld, const int *ab, long long P, int k,
               int n, int d, int gc, float *out) {
    for (long long p = 0; p < P; p++) {
        const float *xa = X + (long long)ab[2 * p] * ld;
        const float *xb = X + (long long)ab[2 * p + 1] * ld;
        float mi = orc_mi(xa, xb, k, n, d);
        out[p] = gc ? orc_gc_from_mi(mi, d) : mi;
    }
}

// hedetniemi_gpu.cu:8-26 — correlationToAdjacencyKernel.
void orc_corr_to_adj(const float *C, int N, float *A) {
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            long long e = (long long)i * N + j;
            if (i == j) A[e] = 0.0f;
            else if (C[e] == 0.0f) A[e] = INFINITY;
            else A[e] = 1.0f / C[e];
        }
}

// math_utils_gpu.cu:23-68 (meanKernel) and :70-156 (standardDeviationKernel):
// per row, thread t of 1024 sums a[t], a[t + 1024], ... in order; the 1024
// partial sums are folded as a tree, s[t] += s[t + o] for o = 512, ..., 1
// (the warp-synchronous tail of the reference is the same tree for s[0]);
// u = s[0] / (float)n. The deviation pass sums (a - u)^2 in the same order
// and sigma = sqrt(s[0] / (float)n). The reference squares with powf(x, 2);
// this restatement squares as fl(x * x) (parity with the original binary
// unpinned for sigma; every other step is exact float arithmetic).
static float row_tree_sum(const float *row, long long n, float shift, bool square) {
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

void orc_row_mean(const float *a, long long m, long long n, float *u) {
    for (long long r = 0; r < m; ++r) u[r] = row_tree_sum(a + r * n, n, 0.0f, false) / (float)n;
}

void orc_row_std(const float *a, long long m, long long n, float *u, float *sigma) {
    for (long long r = 0; r < m; ++r) {
        u[r] = row_tree_sum(a + r * n, n, 0.0f, false) / (float)n;
        sigma[r] = std::sqrt(row_tree_sum(a + r * n, n, u[r], true) / (float)n);
    }
}

// hedetniemi_gpu.cu:47-68 (Part-1) in Jacobi form, and :70-107 (Part-2).
// H_{p+1}[i][j] = min_k fl(A[j][k] + H_p[k][i]) taken with a strict `<`
// from +inf in ascending k; Hi[p] = H_p (H_0 = A). Part-2 backtracks each
// (i, j != i) from p = L-1 down to 0, taking the first q != j with
// fl(fl(Hi[p][i][q] + A[q][k]) - len) < tol. The diagonal is written as
// paths[i][i][*] = i (the reference writes paths[j*n+i], a slot of another
// pair). Hi (L*N*N floats) and paths (N*L*N ints) may be null.
void orc_hedetniemi(const float *A, int N, int L, float tol, float *H_out,
                    float *Hi_out, int *paths) {
    const long long NN = (long long)N * N;
    std::vector<float> H(A, A + NN), Hn(NN);
    std::vector<float> Hi;
    bool want_paths = paths != nullptr;
    if (want_paths || Hi_out) Hi.resize((size_t)L * NN);
    // HT[i][k] = H[k][i]: the same candidates in the same ascending k order,
    // read from a row; rows i are independent, so threads change nothing.
    std::vector<float> HT(NN);
    for (int p = 0; p < L; p++) {
        if (!Hi.empty()) std::copy(H.begin(), H.end(), Hi.begin() + (size_t)p * NN);
        for (int k = 0; k < N; k++)
            for (int i = 0; i < N; i++) HT[(long long)i * N + k] = H[(long long)k * N + i];
#pragma omp parallel for schedule(dynamic, 4)
        for (int i = 0; i < N; i++) {
            const float *hti = HT.data() + (long long)i * N;
            for (int j = 0; j < N; j++) {
                const float *aj = A + (long long)j * N;
                float c = INFINITY;
                for (int k = 0; k < N; k++) {
                    float s = aj[k] + hti[k];
                    if (s < c) c = s;
                }
                Hn[(long long)i * N + j] = c;
            }
        }
        H.swap(Hn);
    }
    std::copy(H.begin(), H.end(), H_out);
    if (Hi_out) std::copy(Hi.begin(), Hi.end(), Hi_out);
    if (!want_paths) return;
    std::fill(paths, paths + (size_t)N * L * N, -1);
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            int *pp = paths + (long long)i * N * L + (long long)j * L;
            if (i == j) {
                for (int p = 0; p < L; p++) pp[p] = i;
                continue;
            }
            float len = H[(long long)i * N + j];
            int k = j;
            for (int p = L - 1; p >= 0; --p) {
                const float *hip = Hi.data() + (size_t)p * NN + (long long)i * N;
                for (int q = 0; q < N; ++q) {
                    if (q == j) continue;
                    if (hip[q] + A[(long long)q * N + k] - len < tol) {
                        len -= A[(long long)q * N + k];
                        k = q;
                        break;
                    }
                }
                pp[p] = k;
            }
        }
}

// hedetniemi.cpp:71-108 — recoverSingleShortestPath. Writes at most L+2
// nodes into out and returns the count.
int orc_recover_path(const int *paths, int N, int L, int i, int j, int *out) {
    std::vector<int> path;
    path.push_back(i);
    for (int k = 0; k < L; k++) {
        int node = paths[(long long)i * L * N + (long long)j * L + k];
        if (std::find(path.begin(), path.end(), node) == path.end())
            path.push_back(node);
    }
    if (std::find(path.begin(), path.end(), j) == path.end()) path.push_back(j);
    std::copy(path.begin(), path.end(), out);
    return (int)path.size();
}

// std::sort of iota(0..n-1) by key with `<` — the permutation the reference's
// mutualInformationCpu produces for one frame (mutual_information_cpu.cpp:48-53).
void orc_argsort_like_reference(const float *key, int n, int *out) {
    std::iota(out, out + n, 0);
    std::sort(out, out + n, [&](int p, int q) { return key[p] < key[q]; });
}

// std::partial_sort(first, last, last) on the same comparator: the heap-sort
// fallback libstdc++'s introsort takes when its depth limit runs out.
void orc_heapsort_like_libstdcxx(const float *key, int n, int *out) {
    std::iota(out, out + n, 0);
    std::partial_sort(out, out + n, out + n,
                      [&](int p, int q) { return key[p] < key[q]; });
}

// McIlroy's adversary ("A Killer Adversary for Quicksort", Software:
// Practice and Experience 29(4), 1999), run against libstdc++'s std::sort:
// values are fixed lazily as comparisons reveal them, which yields an input
// that drives the introsort to its depth limit and heap-sort fallback.
void orc_antiqsort(int n, float *out) {
    std::vector<int> val(n), ptr(n);
    const int gas = n - 1;
    int nsolid = 0, candidate = 0;
    for (int i = 0; i < n; ++i) {
        ptr[i] = i;
        val[i] = gas;
    }
    auto cmp = [&](int x, int y) {
        if (val[x] == gas && val[y] == gas) {
            if (x == candidate) val[x] = nsolid++;
            else val[y] = nsolid++;
        }
        if (val[x] == gas) candidate = x;
        else if (val[y] == gas) candidate = y;
        return val[x] < val[y];
    };
    std::sort(ptr.begin(), ptr.end(), cmp);
    for (int i = 0; i < n; ++i) out[i] = (float)val[i];
}

}  // extern "C"
