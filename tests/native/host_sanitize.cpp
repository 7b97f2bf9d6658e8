// Host-side code under AddressSanitizer and UndefinedBehaviorSanitizer
// (built and run by tests/test_host_sanitizers.py with g++): the CPU
// platform's KSG (libstdc++-order selection), the prefix selection on
// adversarial inputs, CPU Hedetniemi and adjacency, row statistics, the
// host side of CuArray (views, deep copies, .npy save/load, sort/argsort),
// and netchem's PDB / DCD readers on the golden files and on truncated,
// byte-swapped and malformed copies (argv[1] = tests/golden/netchem).
// Results are also checked against direct recomputations where that is
// short. The runtime symbols the host code calls into (error string, psi
// table, device allocator) are defined here; the device paths are not
// exercised.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/netsci_hip.h"
#include "../../netsci_amd/csrc/host/cuarray.h"
#include "../../netsci_amd/csrc/host/netchem.h"

namespace nsx {
static std::string g_err;
void set_error(const std::string &m) { g_err = m; }
void clear_error() { g_err.clear(); }
void psi_table(float *psi, int n) {
    for (int i = 0; i <= n; ++i) psi[i] = 0.0f;
    if (n >= 1) psi[1] = (float)-0.57721566490153;
    for (int i = 1; i < n; ++i) psi[i + 1] = psi[i] + (float)(1.0 / (float)i);
}
}  // namespace nsx

extern "C" {
const char *nsx_last_error(void) { return nsx::g_err.c_str(); }
int nsx_alloc(void **p, size_t) { *p = nullptr; nsx::g_err = "no device"; return 1; }
int nsx_free(void *) { return 0; }
int nsx_memcpy_h2d(void *, const void *, size_t) { return 1; }
int nsx_memcpy_d2h(void *, const void *, size_t) { return 1; }
}

static int failures = 0;
#define CHECK(cond, ...)                                  \
    do {                                                  \
        if (!(cond)) {                                    \
            std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            std::fprintf(stderr, __VA_ARGS__);            \
            std::fprintf(stderr, "\n");                   \
            ++failures;                                   \
        }                                                 \
    } while (0)

// Direct KSG counts with std::sort (mutual_information_cpu.cpp:22-109).
static float marg(const float *x, int n, int d, int i, int j) {
    float r = std::abs(x[i] - x[j]);
    for (int di = 1; di < d; ++di) r = std::max(r, std::abs(x[i + di * n] - x[j + di * n]));
    return r;
}

static float direct_mi(const float *xa, const float *xb, int n, int d, int k) {
    std::vector<float> psi(n + 1);
    nsx::psi_table(psi.data(), n);
    std::vector<float> dz(n);
    std::vector<int> order(n);
    float acc = 0.0f;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) dz[j] = std::max(marg(xa, n, d, i, j), marg(xb, n, d, i, j));
        std::iota(order.begin(), order.end(), 0);
        std::sort(order.begin(), order.end(), [&](int p, int q) { return dz[p] < dz[q]; });
        float ea = 0.0f, eb = 0.0f;
        for (int m = 1; m <= k; ++m) {
            ea = std::max(ea, marg(xa, n, d, i, order[m]));
            eb = std::max(eb, marg(xb, n, d, i, order[m]));
        }
        int ca = 0, cb = 0;
        for (int j = 0; j < n; ++j) {
            if (marg(xa, n, d, i, j) <= ea) ++ca;
            if (marg(xb, n, d, i, j) <= eb) ++cb;
        }
        acc += psi[ca - 1] + psi[cb - 1];
    }
    acc /= (float)n;
    return (psi[n] + psi[k]) - (float)(1.0 / (float)k) - acc;
}

static void test_ksg(std::mt19937 &rng) {
    std::normal_distribution<float> g(0.0f, 1.0f);
    const int cases[][4] = {{1, 40, 4, 0}, {2, 67, 6, 1}, {3, 131, 6, 1}, {3, 50, 16, 0},
                            {3, 30, 1, 1}};
    for (auto &c : cases) {
        const int d = c[0], n = c[1], k = c[2], grid = c[3];
        const int M = 3;
        std::vector<float> X((size_t)M * d * n);
        for (auto &v : X) v = grid ? std::round(g(rng) * 3.0f) / 3.0f : g(rng);
        const int ab[] = {0, 1, 1, 2, 2, 0};
        std::vector<float> mi(3);
        CHECK(nsx_cpu_ksg_mi(X.data(), (long long)d * n, n, d, n, ab, 3, k, mi.data(), 2) == 0,
              "cpu ksg rc: %s", nsx_last_error());
        for (int p = 0; p < 3; ++p) {
            const float want = direct_mi(&X[(size_t)ab[2 * p] * d * n],
                                         &X[(size_t)ab[2 * p + 1] * d * n], n, d, k);
            CHECK(mi[p] == want, "ksg d=%d n=%d k=%d pair %d: %.9g vs %.9g", d, n, k, p, mi[p],
                  want);
        }
    }
}

static void test_select_prefix(std::mt19937 &rng) {
    const int n = 700;
    std::vector<std::vector<float>> inputs(5, std::vector<float>(n));
    for (int i = 0; i < n; ++i) {
        inputs[0][i] = (float)i;                         // sorted
        inputs[1][i] = (float)(n - i);                   // reversed
        inputs[2][i] = 1.0f;                             // all equal
        inputs[3][i] = (float)std::min(i, n - 1 - i);    // organ pipe
        inputs[4][i] = (float)(rng() % 7);               // heavy ties
    }
    std::vector<int> idx(n), ranges(256);
    for (auto &key : inputs) {
        for (int m : {1, 7, 17, n}) {
            CHECK(nsx_select_prefix_host(key.data(), n, m, idx.data(), ranges.data()) == 0,
                  "select rc");
            std::vector<int> ref(n);
            std::iota(ref.begin(), ref.end(), 0);
            std::sort(ref.begin(), ref.end(), [&](int p, int q) { return key[p] < key[q]; });
            for (int t = 0; t < m; ++t) CHECK(idx[t] == ref[t], "select m=%d pos %d", m, t);
        }
    }
}

static void test_apsp_and_stats(std::mt19937 &rng) {
    const int N = 45, L = 6;
    std::uniform_real_distribution<float> u(0.1f, 3.0f), p01(0.0f, 1.0f);
    std::vector<float> A((size_t)N * N), H((size_t)N * N), C((size_t)N * N), Aj((size_t)N * N);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            A[(size_t)i * N + j] = (i == j) ? 0.0f : (p01(rng) < 0.2f ? u(rng) : INFINITY);
            C[(size_t)i * N + j] = p01(rng) < 0.3f ? 0.0f : p01(rng);
        }
    std::vector<int> paths((size_t)N * N * L);
    CHECK(nsx_cpu_hedetniemi(A.data(), N, L, 1e-4f, H.data(), paths.data(), 2) == 0, "apsp rc");
    for (int i = 0; i < N; ++i) CHECK(H[(size_t)i * N + i] == 0.0f, "apsp diagonal");
    for (size_t e = 0; e < paths.size(); ++e) CHECK(paths[e] >= 0 && paths[e] < N, "path range");
    CHECK(nsx_cpu_corr_to_adj(C.data(), N, Aj.data()) == 0, "adjacency rc");
    std::vector<float> rows((size_t)7 * 1500), mu(7), sd(7);
    for (auto &v : rows) v = u(rng);
    CHECK(nsx_cpu_row_stats(rows.data(), 7, 1500, mu.data(), sd.data()) == 0, "row stats rc");
    for (int r = 0; r < 7; ++r) CHECK(mu[r] > 0.1f && mu[r] < 3.0f && sd[r] > 0.0f, "stats range");
    CHECK(nsx_cpu_row_stats(rows.data(), 7, 0, mu.data(), sd.data()) != 0, "n = 0 rejected");
}

static void test_cuarray() {
    CuArray<float> a;
    a.init(20, 100);
    for (int i = 0; i < 20; ++i)
        for (int j = 0; j < 100; ++j) a.set((float)(i * 100 + j) * 0.5f, i, j);
    CHECK(a.set(1.0f, 20, 0) == 1, "set bounds");
    CuArray<float> view, deep;
    view.fromCuArrayShallowCopy(&a, 18, 19, 2, 100);
    deep.fromCuArrayDeepCopy(&a, 18, 19, 2, 100);
    CHECK(view.owner() == 0 && deep.owner() == 1, "owners");
    CHECK(view.get(1, 99) == a.get(19, 99) && deep.get(0, 0) == a.get(18, 0), "copies");
    const std::string f = "host_sanitize_tmp.npy";
    a.save(f);
    CuArray<float> b;
    CHECK(b.load(f) == 0 && b.m() == 20 && b.n() == 100, "npy shape");
    for (int e = 0; e < 2000; ++e) CHECK(b[e] == a[e], "npy value %d", e);
    std::remove(f.c_str());
    CuArray<float> *s = a.sort(3);
    CuArray<int> *as = a.argsort(3);
    CHECK(s->get(0, 0) == a.get(3, 99) && as->get(0, 0) == 99, "descending sort");
    delete s;
    delete as;
    CHECK(a.toDevice() == 1 && a.allocatedDevice() == 0, "toDevice without device");
}

static std::vector<char> slurp(const std::string &f) {
    std::ifstream in(f, std::ios::binary);
    return std::vector<char>(std::istreambuf_iterator<char>(in), {});
}

static void spill(const std::string &f, const std::vector<char> &b) {
    std::ofstream(f, std::ios::binary).write(b.data(), (std::streamsize)b.size());
}

template <class F>
static bool raises(F f) {
    try {
        f();
    } catch (const std::exception &) {
        return true;
    }
    return false;
}

static void test_netchem(const std::string &gold) {
    const std::string pdb = gold + "/test.pdb", dcd = gold + "/test.dcd";
    netchem::Network net;
    net.init(dcd, pdb, 0, 9, 1);
    CHECK(net.numNodes() == 290 && net.numFrames() == 10, "netchem shape");
    CHECK(net.nodeCoordinates()->m() == 290 && net.nodeCoordinates()->n() == 30, "coords");
    netchem::Network s3;
    s3.init(dcd, pdb, 0, -1, 3);
    for (int i = 0; i < 290; ++i)
        for (int j = 0; j < 4; ++j)
            CHECK(s3.nodeCoordinates()->get(i, j) == net.nodeCoordinates()->get(i, 3 * j),
                  "stride 3 node %d frame %d", i, j);
    netchem::Network pg;
    pg.init(gold + "/protG_short.dcd", gold + "/protG_short.pdb", 5, 97, 7);
    CHECK(pg.numNodes() == 56 && pg.numFrames() == 14, "protein G shape");
    // Byte-swapped copy: every 4-byte word of the file reversed.
    std::vector<char> b = slurp(dcd);
    std::vector<char> sw = b;
    // Words are 4 bytes except the unit-cell doubles (skipped as a unit) and
    // the title text; reversing each 4-byte word is a valid big-endian file
    // for everything that is read.
    for (size_t i = 0; i + 4 <= sw.size(); i += 4) std::reverse(sw.begin() + i, sw.begin() + i + 4);
    std::copy(b.begin() + 4, b.begin() + 8, sw.begin() + 4);  // "CORD" stays text
    spill("swapped.dcd", sw);
    netchem::Network be;
    be.init("swapped.dcd", pdb, 0, 9, 1);
    for (int e = 0; e < 290 * 30; ++e)
        CHECK((*be.nodeCoordinates())[e] == (*net.nodeCoordinates())[e], "swapped %d", e);
    // Truncations at every region of the header and the first frames.
    for (size_t cut : {(size_t)0, (size_t)3, (size_t)50, (size_t)95, (size_t)200, (size_t)280,
                       (size_t)300, (size_t)9000, (size_t)40000, b.size() - 1}) {
        spill("cut.dcd", std::vector<char>(b.begin(), b.begin() + cut));
        CHECK(raises([&] {
                  netchem::Network t;
                  t.init("cut.dcd", pdb, 0, 9, 1);
              }),
              "truncated DCD at %zu", cut);
    }
    // Corrupt record length in frame 4.
    std::vector<char> bad = b;
    bad[276 + 4 * 27800 + 56] ^= 0x10;
    spill("bad.dcd", bad);
    CHECK(raises([&] {
              netchem::Network t;
              t.init("bad.dcd", pdb, 0, 9, 1);
          }),
          "corrupt record");
    std::ofstream("short.pdb") << "ATOM      1  CA  MET X   1      35.611  24.523\n";
    CHECK(raises([&] {
              netchem::Network t;
              t.init(dcd, "short.pdb", 0, 0, 1);
          }),
          "short PDB record");
    CHECK(raises([] { netchem::parse_atom_record(
                          "ATOM      1  CA  MET X   Z      35.611  24.523  44.414  0.00  0.00"
                          "           C",
                          0); }),
          "bad residue id");
    for (const char *f : {"swapped.dcd", "cut.dcd", "bad.dcd", "short.pdb"})
        std::remove(f);
}

int main(int argc, char **argv) {
    std::mt19937 rng(12345);
    test_ksg(rng);
    test_select_prefix(rng);
    test_apsp_and_stats(rng);
    test_cuarray();
    if (argc > 1) test_netchem(argv[1]);
    if (failures) {
        std::fprintf(stderr, "%d failures\n", failures);
        return 1;
    }
    std::printf("host sanitizer run: ok\n");
    return 0;
}
