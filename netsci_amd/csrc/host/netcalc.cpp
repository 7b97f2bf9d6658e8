// netcalc:: on libnsx. See netcalc.h for the reference each function follows.
#include "netcalc.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace netcalc {
namespace {

void check(int rc, const char *what) {
    if (rc != 0) throw std::runtime_error(std::string(what) + ": " + nsx_last_error());
}

// A device buffer owned for the duration of one call.
struct DeviceBuffer {
    void *ptr = nullptr;
    explicit DeviceBuffer(size_t bytes) { check(nsx_alloc(&ptr, bytes), "device allocation"); }
    ~DeviceBuffer() { nsx_free(ptr); }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    template <class T>
    T *as() const { return (T *)ptr; }
};

// The reference's GPU driver has kernels only for xd == 2 (two variables)
// and leaves its kernel pointer uninitialised otherwise
// (mutual_information_gpu.cu:1015-1030); its CPU path ignores xd
// (mutual_information_cpu.cpp:6-121), and so does the CPU platform here.
void validate_xd(int xd, bool gpu) {
    if (gpu && xd != 2)
        throw std::invalid_argument("xd must be 2 on the GPU platform (MI of two variables)");
}

void validate_pairs(CuArray<float> *X, CuArray<int> *ab, int k, int n, int d) {
    if (!X || !ab) throw std::invalid_argument("X and ab must be CuArrays");
    if (d < 1) throw std::invalid_argument("d must be >= 1");
    if (k < 1 || k >= n) throw std::invalid_argument("need 1 <= k < n");
    if ((long long)d * n > X->n())
        throw std::invalid_argument("d * n exceeds the row length of X (X.n())");
    if (ab->m() > 0 && ab->n() != 2) throw std::invalid_argument("ab must have two columns");
    for (int p = 0; p < ab->m(); ++p) {
        const int a = ab->get(p, 0), b = ab->get(p, 1);
        if (a < 0 || b < 0 || a >= X->m() || b >= X->m())
            throw std::invalid_argument("ab holds a node index outside X's rows");
    }
}

// MI of pairs ab[first .. first+count) on the GPU (X uploaded once per call).
void gpu_mi(const float *hX, long long M, long long ld, long long dstride, const int *hab,
            long long count, int k, int n, int d, float *out) {
    if (count == 0) return;
    DeviceBuffer dX((size_t)(M * ld) * sizeof(float));
    DeviceBuffer dab((size_t)count * 2 * sizeof(int));
    DeviceBuffer dmi((size_t)count * sizeof(float));
    check(nsx_memcpy_h2d(dX.ptr, hX, (size_t)(M * ld) * sizeof(float)), "upload X");
    check(nsx_memcpy_h2d(dab.ptr, hab, (size_t)count * 2 * sizeof(int)), "upload ab");
    check(nsx_ksg_mi(dX.as<float>(), ld, dstride, d, n, dab.as<int>(), count, k,
                     dmi.as<float>(), nullptr, nullptr),
          "KSG on the GPU");
    check(nsx_memcpy_d2h(out, dmi.ptr, (size_t)count * sizeof(float)), "download MI");
}

void cpu_mi(const float *hX, long long ld, long long dstride, const int *hab, long long count,
            int k, int n, int d, float *out) {
    if (count == 0) return;
    check(nsx_cpu_ksg_mi(hX, ld, dstride, d, n, hab, count, k, out, 0), "KSG on the CPU");
}

void to_gc(float *v, long long count, int d) {
    check(nsx_gc_from_mi_host(v, count, d, v), "GC transform");
}

// The pair loop of mutual_information.cpp:103-172 / generalized_correlation.cpp:104-173.
int pair_loop(CuArray<float> *X, CuArray<float> *R, CuArray<int> *ab, int k, int n, int xd,
              int d, int platform, bool gc) {
    R->init(1, ab->m());
    if (ab->m() == 0) return platform;
    if (platform != GPU_PLATFORM && platform != CPU_PLATFORM)
        throw std::runtime_error("Invalid platform");
    validate_xd(xd, platform == GPU_PLATFORM);
    validate_pairs(X, ab, k, n, d);
    if (platform == GPU_PLATFORM)
        gpu_mi(X->host(), X->m(), X->n(), n, ab->host(), ab->m(), k, n, d, R->host());
    else
        cpu_mi(X->host(), X->n(), n, ab->host(), ab->m(), k, n, d, R->host());
    if (gc) to_gc(R->host(), ab->m(), d);
    return platform;
}

void gpu_ckpt(CuArray<float> *X, CuArray<float> *R, CuArray<int> *ab, int k, int n, int d,
              bool gc, int freq, const std::string &name);

// Checkpointed pair loop (generalized_correlation.cpp:8-101): the same files
// as the reference — <base>_ab.npy, and after pair i with i % freq == 0 the
// whole R (zeros after i) as <base>_<i>.npy, removing <base>_<i-freq>.npy.
// Pairs run in batches that end exactly at those checkpoints.
int pair_loop_ckpt(CuArray<float> *X, CuArray<float> *R, CuArray<int> *ab, int k, int n, int xd,
                   int d, int platform, bool gc, int freq, std::string name) {
    if (freq <= 0)
        throw std::invalid_argument("checkpointFrequency must be >= 1 (the reference's -1 "
                                    "writes a file per pair and 0 divides by zero)");
    if (name.size() > 4 && name.substr(name.size() - 4, 4) == ".npy")
        name = name.substr(0, name.size() - 4);
    ab->save(name + "_ab.npy");
    R->init(1, ab->m());
    if (ab->m() == 0) return platform;
    if (platform != GPU_PLATFORM && platform != CPU_PLATFORM)
        throw std::runtime_error("Invalid platform");
    validate_xd(xd, platform == GPU_PLATFORM);
    validate_pairs(X, ab, k, n, d);
    const long long P = ab->m();
    if (platform == GPU_PLATFORM) {
        gpu_ckpt(X, R, ab, k, n, d, gc, freq, name);
        return platform;
    }
    long long start = 0;
    while (start < P) {
        // Next checkpoint index c >= start with c % freq == 0.
        long long c = ((start + freq - 1) / freq) * freq;
        long long end = std::min(P, c + 1);
        const long long count = end - start;
        float *out = R->host() + start;
        const int *pab = ab->host() + 2 * start;
        if (platform == GPU_PLATFORM)
            gpu_mi(X->host(), X->m(), X->n(), n, pab, count, k, n, d, out);
        else
            cpu_mi(X->host(), X->n(), n, pab, count, k, n, d, out);
        if (gc) to_gc(out, count, d);
        if (c < P) {
            std::remove((name + "_" + std::to_string(c - freq) + ".npy").c_str());
            R->save(name + "_" + std::to_string(c) + ".npy");
        }
        start = end;
    }
    return platform;
}

// The GPU side of the checkpointed loop: X and ab go to the device once, the
// pairs run as one nsx_ksg_mi_chunked call whose chunks end after each
// checkpoint index (the list path's tables are built once), and each chunk's
// callback forms its GC values in R and writes the checkpoint files while
// the device computes the next chunk.
struct CkptState {
    CuArray<float> *R;
    std::vector<float> *mi;
    int d;
    bool gc;
    long long freq;
    std::string name;
    long long prev;
    std::string error;
};

int ckpt_chunk_done(void *user, long long end) {
    CkptState *st = static_cast<CkptState *>(user);
    try {
        const long long b = st->prev;
        std::copy(st->mi->begin() + b, st->mi->begin() + end, st->R->host() + b);
        if (st->gc) to_gc(st->R->host() + b, end - b, st->d);
        const long long c = end - 1;  // the chunk ends after pair c
        if (c % st->freq == 0) {
            std::remove((st->name + "_" + std::to_string(c - st->freq) + ".npy").c_str());
            st->R->save(st->name + "_" + std::to_string(c) + ".npy");
        }
        st->prev = end;
        return 0;
    } catch (const std::exception &e) {
        st->error = e.what();
        return 1;
    }
}

void gpu_ckpt(CuArray<float> *X, CuArray<float> *R, CuArray<int> *ab, int k, int n, int d,
              bool gc, int freq, const std::string &name) {
    const long long P = ab->m(), M = X->m(), ld = X->n();
    std::vector<long long> ends;
    for (long long start = 0; start < P;) {
        const long long c = ((start + freq - 1) / freq) * freq;
        const long long end = std::min(P, c + 1);
        ends.push_back(end);
        start = end;
    }
    DeviceBuffer dX((size_t)(M * ld) * sizeof(float));
    DeviceBuffer dab((size_t)P * 2 * sizeof(int));
    DeviceBuffer dmi((size_t)P * sizeof(float));
    check(nsx_memcpy_h2d(dX.ptr, X->host(), (size_t)(M * ld) * sizeof(float)), "upload X");
    check(nsx_memcpy_h2d(dab.ptr, ab->host(), (size_t)P * 2 * sizeof(int)), "upload ab");
    std::vector<float> mi((size_t)P);
    CkptState st{R, &mi, d, gc, freq, name, 0, std::string()};
    const int rc = nsx_ksg_mi_chunked(dX.as<float>(), ld, n, d, n, dab.as<int>(), P, k,
                                      dmi.as<float>(), mi.data(), ends.data(), (int)ends.size(),
                                      ckpt_chunk_done, &st, nullptr);
    if (!st.error.empty()) throw std::runtime_error("checkpoint: " + st.error);
    check(rc, "KSG on the GPU");
}

// One pair from two flat arrays (Xa, Xb hold d*n values at indices i + di*n).
float single_pair(CuArray<float> *Xa, CuArray<float> *Xb, int k, int n, int xd, int d,
                  bool gpu) {
    if (!Xa || !Xb) throw std::invalid_argument("Xa and Xb must be CuArrays");
    validate_xd(xd, gpu);
    if (d < 1) throw std::invalid_argument("d must be >= 1");
    if (k < 1 || k >= n) throw std::invalid_argument("need 1 <= k < n");
    const long long len = (long long)d * n;
    if (Xa->size() < len || Xb->size() < len)
        throw std::invalid_argument("Xa and Xb need d * n values");
    std::vector<float> X2((size_t)(2 * len));
    std::copy(Xa->host(), Xa->host() + len, X2.begin());
    std::copy(Xb->host(), Xb->host() + len, X2.begin() + len);
    const int ab[2] = {0, 1};
    float mi = 0.0f;
    if (gpu) gpu_mi(X2.data(), 2, len, n, ab, 1, k, n, d, &mi);
    else cpu_mi(X2.data(), len, n, ab, 1, k, n, d, &mi);
    return mi;
}

float gc_of(float mi, int d) {
    float g = 0.0f;
    check(nsx_gc_from_mi_host(&mi, 1, d, &g), "GC transform");
    return g;
}

}  // namespace

int mutualInformation(CuArray<float> *X, CuArray<float> *I, CuArray<int> *ab, int k, int n,
                      int xd, int d, int platform) {
    return pair_loop(X, I, ab, k, n, xd, d, platform, false);
}

int mutualInformation(CuArray<float> *X, CuArray<float> *I, CuArray<int> *ab, int k, int n,
                      int xd, int d, int platform, int checkpointFrequency,
                      std::string checkpointFileName) {
    return pair_loop_ckpt(X, I, ab, k, n, xd, d, platform, false, checkpointFrequency,
                          checkpointFileName);
}

float mutualInformationGpu(CuArray<float> *Xa, CuArray<float> *Xb, int k, int n, int xd, int d) {
    return single_pair(Xa, Xb, k, n, xd, d, true);
}

float mutualInformationCpu(CuArray<float> *Xa, CuArray<float> *Xb, int k, int n, int xd, int d) {
    return single_pair(Xa, Xb, k, n, xd, d, false);
}

int generalizedCorrelation(CuArray<float> *X, CuArray<float> *R, CuArray<int> *ab, int k, int n,
                           int xd, int d, int platform) {
    return pair_loop(X, R, ab, k, n, xd, d, platform, true);
}

int generalizedCorrelation(CuArray<float> *X, CuArray<float> *R, CuArray<int> *ab, int k, int n,
                           int xd, int d, int platform, int checkpointFrequency,
                           std::string checkpointFileName) {
    return pair_loop_ckpt(X, R, ab, k, n, xd, d, platform, true, checkpointFrequency,
                          checkpointFileName);
}

float generalizedCorrelationGpu(CuArray<float> *Xa, CuArray<float> *Xb, int k, int n, int xd,
                                int d) {
    return gc_of(mutualInformationGpu(Xa, Xb, k, n, xd, d), d);
}

float generalizedCorrelationCpu(CuArray<float> *Xa, CuArray<float> *Xb, int k, int n, int xd,
                                int d) {
    return gc_of(mutualInformationCpu(Xa, Xb, k, n, xd, d), d);
}

void generateRestartAbFromCheckpointFile(CuArray<int> *ab, CuArray<int> *restartAb,
                                         const std::string &checkpointFileName) {
    // mutual_information.cpp:179-200: the index after the last '_' of the
    // file name (std::stoi stops at ".npy"); rows after it are copied.
    std::string name = checkpointFileName;
    const size_t slash = name.find_last_of('/');
    if (slash != std::string::npos) name = name.substr(slash + 1);
    const size_t us = name.find_last_of('_');
    const int last = std::stoi(name.substr(us + 1));
    restartAb->fromCuArrayDeepCopy(ab, last + 1, ab->m() - 1, ab->m() - last - 1, 2);
}

int windowedCorrelation(CuArray<float> *X, CuArray<float> *R, CuArray<int> *ab,
                        CuArray<int> *windows, int k, int nTotal, int d, int gc) {
    if (!windows || windows->n() != 2 || windows->m() < 1)
        throw std::invalid_argument("windows must be an R x 2 array of (offset, length)");
    const int nw = windows->m();
    std::vector<int> off(nw), len(nw);
    for (int r = 0; r < nw; ++r) {
        off[r] = windows->get(r, 0);
        len[r] = windows->get(r, 1);
        if (off[r] < 0 || len[r] <= k || off[r] + len[r] > nTotal)
            throw std::invalid_argument("window outside the trajectory or not longer than k");
    }
    validate_pairs(X, ab, k, nTotal, d);
    R->init(nw, ab->m());
    const long long P = ab->m();
    if (P == 0) return GPU_PLATFORM;
    const long long M = X->m(), ld = X->n();
    DeviceBuffer dX((size_t)(M * ld) * sizeof(float));
    DeviceBuffer dab((size_t)P * 2 * sizeof(int));
    DeviceBuffer dmi((size_t)nw * P * sizeof(float));
    check(nsx_memcpy_h2d(dX.ptr, X->host(), (size_t)(M * ld) * sizeof(float)), "upload X");
    check(nsx_memcpy_h2d(dab.ptr, ab->host(), (size_t)P * 2 * sizeof(int)), "upload ab");
    check(nsx_ksg_mi_windows(dX.as<float>(), ld, nTotal, d, off.data(), len.data(), nw,
                             dab.as<int>(), P, k, dmi.as<float>(), nullptr),
          "windowed KSG on the GPU");
    check(nsx_memcpy_d2h(R->host(), dmi.ptr, (size_t)nw * P * sizeof(float)), "download");
    if (gc) to_gc(R->host(), (long long)nw * P, d);
    return GPU_PLATFORM;
}

// ---- Hedetniemi ------------------------------------------------------------

namespace {
void validate_square(CuArray<float> *A) {
    if (!A || A->m() != A->n() || A->m() < 1)
        throw std::invalid_argument("A must be a non-empty square matrix");
}
}  // namespace

void hedetniemiAllShortestPathsGpu(CuArray<float> *A, CuArray<float> *H, CuArray<int> *paths,
                                   float tolerance, int maxPathLength) {
    validate_square(A);
    if (maxPathLength < 1) throw std::invalid_argument("maxPathLength must be >= 1");
    const long long N = A->n(), L = maxPathLength;
    H->fromCuArrayDeepCopy(A, 0, A->m() - 1, A->m(), A->n());
    paths->init(1, (int)(N * L * N));
    DeviceBuffer dA((size_t)(N * N) * sizeof(float));
    DeviceBuffer dH((size_t)(N * N) * sizeof(float));
    DeviceBuffer dHi((size_t)(L * N * N) * sizeof(float));
    DeviceBuffer dP((size_t)(N * L * N) * sizeof(int));
    check(nsx_memcpy_h2d(dA.ptr, A->host(), (size_t)(N * N) * sizeof(float)), "upload A");
    check(nsx_hedetniemi(dA.as<float>(), (int)N, (int)L, tolerance, dH.as<float>(),
                         dHi.as<float>(), dP.as<int>(), nullptr),
          "Hedetniemi on the GPU");
    check(nsx_memcpy_d2h(H->host(), dH.ptr, (size_t)(N * N) * sizeof(float)), "download H");
    check(nsx_memcpy_d2h(paths->host(), dP.ptr, (size_t)(N * L * N) * sizeof(int)),
          "download paths");
}

void hedetniemiAllShortestPathsCpu(CuArray<float> *A, CuArray<float> *H, CuArray<int> *paths,
                                   float tolerance, int maxPathLength) {
    validate_square(A);
    if (maxPathLength < 1) throw std::invalid_argument("maxPathLength must be >= 1");
    const long long N = A->n(), L = maxPathLength;
    H->init(A->m(), A->n());
    paths->init(1, (int)(N * L * N));
    check(nsx_cpu_hedetniemi(A->host(), (int)N, (int)L, tolerance, H->host(), paths->host(), 0),
          "Hedetniemi on the CPU");
}

void hedetniemiAllShortestPaths(CuArray<float> *A, CuArray<float> *H, CuArray<int> *paths,
                                int maxPathLength, float tolerance, int platform) {
    H->init(A->m(), A->n());
    if (platform == GPU_PLATFORM)
        hedetniemiAllShortestPathsGpu(A, H, paths, tolerance, maxPathLength);
    else if (platform == CPU_PLATFORM)
        hedetniemiAllShortestPathsCpu(A, H, paths, tolerance, maxPathLength);
    else
        throw std::invalid_argument("Invalid platform");
}

void hedetniemiAllShortestPathLengthsGpu(CuArray<float> *A, CuArray<float> *H,
                                         int maxPathLength) {
    validate_square(A);
    if (maxPathLength < 0) throw std::invalid_argument("maxPathLength must be >= 0");
    const long long N = A->n();
    H->fromCuArrayDeepCopy(A, 0, A->m() - 1, A->m(), A->n());
    DeviceBuffer dA((size_t)(N * N) * sizeof(float));
    DeviceBuffer dH((size_t)(N * N) * sizeof(float));
    check(nsx_memcpy_h2d(dA.ptr, A->host(), (size_t)(N * N) * sizeof(float)), "upload A");
    check(nsx_hedetniemi_lengths(dA.as<float>(), (int)N, maxPathLength, dH.as<float>(), nullptr),
          "Hedetniemi lengths on the GPU");
    check(nsx_memcpy_d2h(H->host(), dH.ptr, (size_t)(N * N) * sizeof(float)), "download H");
}

void hedetniemiAllShortestPathLengths(CuArray<float> *A, CuArray<float> *H, int maxPathLength,
                                      int platform) {
    H->init(A->m(), A->n());
    if (platform == GPU_PLATFORM) {
        hedetniemiAllShortestPathLengthsGpu(A, H, maxPathLength);
    } else if (platform == CPU_PLATFORM) {
        validate_square(A);
        check(nsx_cpu_hedetniemi(A->host(), A->n(), maxPathLength, 0.0f, H->host(), nullptr, 0),
              "Hedetniemi lengths on the CPU");
    } else {
        throw std::invalid_argument("Invalid platform");
    }
}

void correlationToAdjacencyGpu(CuArray<float> *A, CuArray<float> *C, int n) {
    if (!C || (long long)C->size() < (long long)n * n)
        throw std::invalid_argument("C must hold n x n values");
    if (A->size() != n * n) A->init(n, n);
    const size_t bytes = (size_t)n * n * sizeof(float);
    DeviceBuffer dC(bytes), dA(bytes);
    check(nsx_memcpy_h2d(dC.ptr, C->host(), bytes), "upload C");
    check(nsx_corr_to_adj(dC.as<float>(), n, dA.as<float>(), nullptr), "adjacency on the GPU");
    check(nsx_memcpy_d2h(A->host(), dA.ptr, bytes), "download A");
}

void correlationToAdjacency(CuArray<float> *A, CuArray<float> *C, int n, int platform) {
    A->init(n, n);
    if (platform == GPU_PLATFORM) {
        correlationToAdjacencyGpu(A, C, n);
    } else if (platform == CPU_PLATFORM) {
        if (!C || (long long)C->size() < (long long)n * n)
            throw std::invalid_argument("C must hold n x n values");
        check(nsx_cpu_corr_to_adj(C->host(), n, A->host()), "adjacency on the CPU");
    } else {
        throw std::invalid_argument("Invalid platform");
    }
}

// ---- row statistics (math_utils.cpp:6-35, math_utils_gpu.cu:158-206) ----

static void check_rows(CuArray<float> *a, int m, int n) {
    if (n < 1) throw std::invalid_argument("n must be >= 1");
    if (m < 0 || (long long)m * n > (long long)a->size())
        throw std::invalid_argument("a holds fewer than m x n values");
}

void meanGpu(CuArray<float> *a, CuArray<float> *u, int m, int n) {
    check_rows(a, m, n);
    if ((int)u->size() < m) u->init(1, m);
    if (m == 0) return;
    const size_t in = (size_t)m * n * sizeof(float), out = (size_t)m * sizeof(float);
    DeviceBuffer da(in), du(out);
    check(nsx_memcpy_h2d(da.ptr, a->host(), in), "upload a");
    check(nsx_row_mean(da.as<float>(), m, n, du.as<float>(), nullptr), "row mean on the GPU");
    check(nsx_memcpy_d2h(u->host(), du.ptr, out), "download u");
}

void standardDeviationGpu(CuArray<float> *a, CuArray<float> *u, CuArray<float> *sigma, int m,
                          int n) {
    check_rows(a, m, n);
    if ((int)u->size() < m) u->init(1, m);
    if ((int)sigma->size() < m) sigma->init(1, m);
    if (m == 0) return;
    const size_t in = (size_t)m * n * sizeof(float), out = (size_t)m * sizeof(float);
    DeviceBuffer da(in), du(out), ds(out);
    check(nsx_memcpy_h2d(da.ptr, a->host(), in), "upload a");
    check(nsx_row_std(da.as<float>(), m, n, du.as<float>(), ds.as<float>(), nullptr),
          "row standard deviation on the GPU");
    check(nsx_memcpy_d2h(u->host(), du.ptr, out), "download u");
    check(nsx_memcpy_d2h(sigma->host(), ds.ptr, out), "download sigma");
}

void mean(CuArray<float> *a, CuArray<float> *u, int m, int n, int platform) {
    if (n < 1) throw std::invalid_argument("n must be >= 1");
    m = (int)(a->size() / n);
    u->init(1, m);
    if (platform == GPU_PLATFORM) {
        meanGpu(a, u, m, n);
    } else if (platform == CPU_PLATFORM) {
        check(nsx_cpu_row_stats(a->host(), m, n, u->host(), nullptr), "row mean on the CPU");
    } else {
        throw std::invalid_argument("Invalid platform");
    }
}

void standardDeviation(CuArray<float> *a, CuArray<float> *u, CuArray<float> *sigma, int m, int n,
                       int platform) {
    if (n < 1) throw std::invalid_argument("n must be >= 1");
    m = (int)(a->size() / n);
    u->init(1, m);
    sigma->init(1, m);
    if (platform == GPU_PLATFORM) {
        standardDeviationGpu(a, u, sigma, m, n);
    } else if (platform == CPU_PLATFORM) {
        check(nsx_cpu_row_stats(a->host(), m, n, u->host(), sigma->host()),
              "row standard deviation on the CPU");
    } else {
        throw std::invalid_argument("Invalid platform");
    }
}

void recoverSingleShortestPath(int **NUMPY_ARRAY, int **NUMPY_ARRAY_DIM1, CuArray<int> *paths,
                               int maxPathLength, int i, int j) {
    if (!NUMPY_ARRAY || !NUMPY_ARRAY_DIM1)
        throw std::invalid_argument("recoverSingleShortestPath: null output pointers");
    const std::vector<int> path = recoverSingleShortestPath(paths, maxPathLength, i, j);
    *NUMPY_ARRAY_DIM1 = new int[1];
    (*NUMPY_ARRAY_DIM1)[0] = (int)path.size();
    *NUMPY_ARRAY = new int[path.size()];
    std::copy(path.begin(), path.end(), *NUMPY_ARRAY);
}

std::vector<int> recoverSingleShortestPath(CuArray<int> *paths, int maxPathLength, int i, int j) {
    if (maxPathLength < 1) throw std::invalid_argument("maxPathLength must be >= 1");
    const int n = (int)std::sqrt((double)(paths->n() / maxPathLength));
    if (i < 0 || j < 0 || i >= n || j >= n) throw std::out_of_range("node index out of range");
    std::vector<int> path;
    path.push_back(i);
    for (int k = 0; k < maxPathLength; k++) {
        const int node = paths->get(0, i * maxPathLength * n + j * maxPathLength + k);
        if (std::find(path.begin(), path.end(), node) == path.end()) path.push_back(node);
    }
    if (std::find(path.begin(), path.end(), j) == path.end()) path.push_back(j);
    return path;
}

}  // namespace netcalc
