// A C++ caller written against the reference's netcalc / cuarray headers
// (netcalc/include/{mutual_information,generalized_correlation,hedetniemi,
// platform}.h, cuarray/include/cuarray.h): every hot-path call below uses
// the reference's argument list, so this file compiling against
// netsci_amd/csrc/host/netcalc.h is the drop-in check. It then runs the
// parts that need no GPU: recoverSingleShortestPath in both forms, the CPU
// platform's pair loop (its return value is the platform), and the xd check
// of the GPU platform, which throws before any device call.
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../netsci_amd/csrc/host/netcalc.h"

using namespace netcalc;

static int failures = 0;
#define CHECK(cond, msg)                                          \
    do {                                                          \
        if (!(cond)) {                                            \
            std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, msg); \
            ++failures;                                           \
        }                                                         \
    } while (0)

// Never called: a list of the reference's call shapes, type-checked by the
// compiler.
[[maybe_unused]] static void signatures(CuArray<float> *X, CuArray<float> *R, CuArray<int> *ab,
                                        CuArray<float> *A, CuArray<float> *H, CuArray<int> *paths) {
    int k = 6, n = 100, xd = 2, d = 3;
    int r = mutualInformation(X, R, ab, k, n, xd, d, GPU_PLATFORM);
    r += mutualInformation(X, R, ab, k, n, xd, d, CPU_PLATFORM, 10, std::string("ckpt"));
    r += generalizedCorrelation(X, R, ab, k, n, xd, d, GPU_PLATFORM);
    r += generalizedCorrelation(X, R, ab, k, n, xd, d, CPU_PLATFORM, 10, std::string("ckpt"));
    float f = mutualInformationGpu(X, X, k, n, xd, d) + mutualInformationCpu(X, X, k, n, xd, d) +
              generalizedCorrelationGpu(X, X, k, n, xd, d) +
              generalizedCorrelationCpu(X, X, k, n, xd, d);
    generateRestartAbFromCheckpointFile(ab, ab, std::string("ckpt_10.npy"));
    hedetniemiAllShortestPaths(A, H, paths, 4, 1e-4f, GPU_PLATFORM);
    hedetniemiAllShortestPathsGpu(A, H, paths, 1e-4f, 4);
    hedetniemiAllShortestPathLengths(A, H, 4, GPU_PLATFORM);
    correlationToAdjacency(A, H, 10, GPU_PLATFORM);
    int *path = nullptr, *len = nullptr;
    recoverSingleShortestPath(&path, &len, paths, 4, 0, 1);
    (void)r;
    (void)f;
}

int main() {
    // A 3-node graph's paths array (N * L * N ints, L = 3): paths[i][j][.].
    const int N = 3, L = 3;
    CuArray<int> paths;
    paths.init(1, N * L * N);
    for (int t = 0; t < N * L * N; ++t) paths.set(-1, 0, t);
    auto put = [&](int i, int j, std::vector<int> v) {
        for (int s = 0; s < L; ++s) paths.set(v[s], 0, i * L * N + j * L + s);
    };
    put(0, 2, {0, 1, 2});
    put(2, 0, {2, 2, 2});
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            int *out = nullptr, *dim = nullptr;
            recoverSingleShortestPath(&out, &dim, &paths, L, i, j);
            const std::vector<int> want = recoverSingleShortestPath(&paths, L, i, j);
            CHECK(dim[0] == (int)want.size(), "path length");
            for (int s = 0; s < dim[0]; ++s) CHECK(out[s] == want[s], "path entry");
            delete[] out;
            delete[] dim;
        }
    {
        int *out = nullptr, *dim = nullptr;
        recoverSingleShortestPath(&out, &dim, &paths, L, 0, 2);
        CHECK(dim[0] == 3 && out[0] == 0 && out[1] == 1 && out[2] == 2, "path 0 -> 2");
        delete[] out;
        delete[] dim;
    }

    // CPU platform pair loop: returns the platform, xd is ignored as the
    // reference's CPU path ignores it.
    const int n = 64, d = 1, k = 4;
    CuArray<float> X, R;
    CuArray<int> ab;
    X.init(2, n);
    for (int f = 0; f < n; ++f) {
        X.set(std::sin(0.3f * f), 0, f);
        X.set(std::sin(0.3f * f) + 0.01f * std::cos(1.7f * f), 1, f);
    }
    ab.init(1, 2);
    ab.set(0, 0, 0);
    ab.set(1, 0, 1);
    CHECK(mutualInformation(&X, &R, &ab, k, n, 2, d, CPU_PLATFORM) == CPU_PLATFORM, "return");
    const float mi2 = R.get(0, 0);
    CHECK(mutualInformation(&X, &R, &ab, k, n, 3, d, CPU_PLATFORM) == CPU_PLATFORM, "return");
    CHECK(R.get(0, 0) == mi2 && mi2 > 0.5f, "CPU platform ignores xd");

    // GPU platform: xd != 2 is rejected before any device work.
    bool threw = false;
    try {
        mutualInformation(&X, &R, &ab, k, n, 3, d, GPU_PLATFORM);
    } catch (const std::invalid_argument &e) {
        threw = std::string(e.what()).find("xd") != std::string::npos;
    }
    CHECK(threw, "xd != 2 on the GPU platform throws std::invalid_argument");
    threw = false;
    try {
        CuArray<float> xa, xb;
        xa.init(1, n);
        xb.init(1, n);
        generalizedCorrelationGpu(&xa, &xb, k, n, 1, d);
    } catch (const std::invalid_argument &) {
        threw = true;
    }
    CHECK(threw, "single-pair GPU call with xd != 2 throws");

    std::printf(failures ? "reference signatures: %d failures\n" : "reference signatures: ok\n",
                failures);
    return failures ? 1 : 0;
}
