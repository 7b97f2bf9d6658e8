// libnsx runtime: status, device set-up, the HIP allocator with a pinned
// staging ring (behind CuArray), events, and small utility kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/netsci_hip.h"
#include "nsx_internal.h"

namespace nsx {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

static thread_local int g_device = -1;

int ensure_device() {
    if (g_device >= 0) return kOk;
    return nsx_init(0);
}

static int g_cu_count = 0;

int cu_count() { return g_cu_count > 0 ? g_cu_count : 256; }

// Grow-only workspaces, per process (one device per process in this build).
static std::mutex g_ws_mutex;
static void *g_ws_ptr[kWsCount] = {};
static size_t g_ws_bytes[kWsCount] = {};

void *workspace(int slot, size_t bytes) {
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    if (slot < 0 || slot >= kWsCount) {
        set_error("workspace: bad slot");
        return nullptr;
    }
    if (bytes == 0) bytes = 256;
    if (g_ws_bytes[slot] >= bytes) return g_ws_ptr[slot];
    if (g_ws_ptr[slot]) {
        // The previous buffer may still be in use by queued work.
        (void)hipDeviceSynchronize();
        (void)hipFree(g_ws_ptr[slot]);
        g_ws_ptr[slot] = nullptr;
        g_ws_bytes[slot] = 0;
    }
    size_t want = std::max(bytes, (size_t)(g_ws_bytes[slot] * 3 / 2));
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        e = hipMalloc(&p, bytes);
        want = bytes;
    }
    if (e != hipSuccess) {
        set_error(std::string("workspace hipMalloc: ") + hipGetErrorString(e));
        return nullptr;
    }
    g_ws_ptr[slot] = p;
    g_ws_bytes[slot] = want;
    return p;
}

static void *g_tab_ptr = nullptr;
static size_t g_tab_bytes = 0;
static size_t g_tab_budget = 0;  // 0: the default budget

void set_table_budget(size_t bytes) {
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    g_tab_budget = bytes;
}

size_t table_budget() {
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    return g_tab_budget;
}

int table_space(size_t slot_bytes, long long want_slots, long long min_slots, void **ptr,
                long long *slots) {
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    *ptr = nullptr;
    *slots = 0;
    if (slot_bytes == 0 || want_slots <= 0) return kOk;
    long long have = (long long)(g_tab_bytes / slot_bytes);
    if (have < want_slots) {
        size_t free_b = 0, total_b = 0;
        NSX_HIP(hipMemGetInfo(&free_b, &total_b), "memory info");
        const double avail = (double)free_b + (double)g_tab_bytes;
        double budget = 0.85 * avail - (double)(1ull << 30);
        if (g_tab_budget) budget = std::min(budget, (double)g_tab_budget);
        long long target = budget > 0 ? (long long)(budget / (double)slot_bytes) : 0;
        target = std::min(target, want_slots);
        if (target > have && target >= min_slots) {
            // The current buffer may still be read by queued work.
            NSX_HIP(hipDeviceSynchronize(), "table buffer sync");
            if (g_tab_ptr) (void)hipFree(g_tab_ptr);
            g_tab_ptr = nullptr;
            g_tab_bytes = 0;
            while (target >= min_slots && target > 0) {
                void *p = nullptr;
                const size_t bytes = (size_t)target * slot_bytes + 256;
                if (hipMalloc(&p, bytes) == hipSuccess) {
                    g_tab_ptr = p;
                    g_tab_bytes = bytes;
                    break;
                }
                (void)hipGetLastError();
                target = target * 3 / 4;
            }
            have = (long long)(g_tab_bytes / slot_bytes);
        }
    }
    // A budget below the buffer held applies at once.
    if (g_tab_budget) have = std::min(have, (long long)(g_tab_budget / slot_bytes));
    *ptr = g_tab_ptr;
    *slots = std::min(have, want_slots);
    return kOk;
}

void release_workspaces() {
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    for (int s = 0; s < kWsCount; ++s) {
        if (g_ws_ptr[s]) (void)hipFree(g_ws_ptr[s]);
        g_ws_ptr[s] = nullptr;
        g_ws_bytes[s] = 0;
    }
    if (g_tab_ptr) (void)hipFree(g_tab_ptr);
    g_tab_ptr = nullptr;
    g_tab_bytes = 0;
}

__global__ void fill_u32_kernel(uint32_t *p, uint32_t v, size_t count) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) p[i] = v;
}

int fill_u32(void *d_ptr, uint32_t value, size_t count, hipStream_t stream) {
    if (count == 0) return kOk;
    size_t blocks = std::min<size_t>((count + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)blocks), dim3(256), 0,
                       stream, (uint32_t *)d_ptr, value, count);
    NSX_HIP(hipGetLastError(), "fill_u32 launch");
    return kOk;
}

void psi_table(float *psi, int n) {
    // psi.cpp:6-19: zero-initialised, psi[1] = (float)-0.57721566490153,
    // psi[i+1] = psi[i] + (float)(1.0 / (float)i) in float32.
    for (int i = 0; i <= n; ++i) psi[i] = 0.0f;
    if (n >= 1) psi[1] = (float)-0.57721566490153;
    for (int i = 1; i < n; ++i) {
        float inv = (float)(1.0 / static_cast<float>(i));
        psi[i + 1] = psi[i] + inv;
    }
}

// Pinned staging ring for synchronous host<->device copies.
static std::mutex g_stage_mutex;
static void *g_stage[2] = {nullptr, nullptr};
static hipEvent_t g_stage_ev[2];
static hipStream_t g_stage_stream = nullptr;
static const size_t kStageBytes = 8u << 20;

static int stage_init() {
    if (g_stage[0]) return kOk;
    NSX_HIP(hipStreamCreateWithFlags(&g_stage_stream, hipStreamNonBlocking),
            "staging stream");
    for (int b = 0; b < 2; ++b) {
        NSX_HIP(hipHostMalloc(&g_stage[b], kStageBytes, hipHostMallocDefault),
                "staging hipHostMalloc");
        NSX_HIP(hipEventCreateWithFlags(&g_stage_ev[b], hipEventDisableTiming),
                "staging event");
    }
    return kOk;
}

// Coherent pinned host words that kernels write and the host polls.
static std::mutex g_flag_mutex;
static unsigned *g_flag_host = nullptr;
static unsigned *g_flag_dev = nullptr;

int host_flags(unsigned **host, unsigned **dev) {
    std::lock_guard<std::mutex> lock(g_flag_mutex);
    if (!g_flag_host) {
        void *p = nullptr;
        NSX_HIP(hipHostMalloc(&p, 256, hipHostMallocMapped | hipHostMallocCoherent),
                "flag hipHostMalloc");
        void *d = nullptr;
        NSX_HIP(hipHostGetDevicePointer(&d, p, 0), "flag device pointer");
        std::memset(p, 0, 256);
        g_flag_host = (unsigned *)p;
        g_flag_dev = (unsigned *)d;
    }
    *host = g_flag_host;
    *dev = g_flag_dev;
    return kOk;
}

}  // namespace nsx

using namespace nsx;

extern "C" {

const char *nsx_last_error(void) { return g_last_error.c_str(); }

int nsx_device_count(int *count) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    (void)hipGetLastError();
    if (count) *count = c;
    return kOk;
}

int nsx_init(int device) {
    int c = 0;
    nsx_device_count(&c);
    if (c <= 0) {
        set_error("no HIP device is visible: the GPU platform needs an MI355X (gfx950)");
        return kNoDevice;
    }
    NSX_REQUIRE(device >= 0 && device < c, "nsx_init: device index out of range");
    NSX_HIP(hipSetDevice(device), "hipSetDevice");
    hipDeviceProp_t prop;
    NSX_HIP(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_error(std::string("libnsx is built for gfx950; device is ") + prop.gcnArchName);
        return kNoDevice;
    }
    g_cu_count = prop.multiProcessorCount;
    g_device = device;
    return kOk;
}

int nsx_device_info(char *name, int len, int *cu, int *clock_khz) {
    int rc = ensure_device();
    if (rc) return rc;
    hipDeviceProp_t prop;
    NSX_HIP(hipGetDeviceProperties(&prop, g_device), "hipGetDeviceProperties");
    if (name && len > 0) {
        std::snprintf(name, (size_t)len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (cu) *cu = prop.multiProcessorCount;
    if (clock_khz) *clock_khz = prop.clockRate;
    return kOk;
}

int nsx_alloc(void **d_ptr, size_t bytes) {
    NSX_REQUIRE(d_ptr, "nsx_alloc: null output pointer");
    int rc = ensure_device();
    if (rc) return rc;
    *d_ptr = nullptr;
    if (bytes == 0) bytes = 4;
    NSX_HIP(hipMalloc(d_ptr, bytes), "hipMalloc");
    return kOk;
}

int nsx_free(void *d_ptr) {
    if (!d_ptr) return kOk;
    NSX_HIP(hipFree(d_ptr), "hipFree");
    return kOk;
}

int nsx_host_alloc(void **h_ptr, size_t bytes) {
    NSX_REQUIRE(h_ptr, "nsx_host_alloc: null output pointer");
    int rc = ensure_device();
    if (rc) return rc;
    NSX_HIP(hipHostMalloc(h_ptr, bytes ? bytes : 4, hipHostMallocDefault), "hipHostMalloc");
    return kOk;
}

int nsx_host_free(void *h_ptr) {
    if (!h_ptr) return kOk;
    NSX_HIP(hipHostFree(h_ptr), "hipHostFree");
    return kOk;
}

// Copies go through two pinned buffers: the host memcpy into one buffer
// overlaps the DMA of the other.
int nsx_memcpy_h2d(void *d_dst, const void *h_src, size_t bytes) {
    if (bytes == 0) return kOk;
    int rc = ensure_device();
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g_stage_mutex);
    rc = stage_init();
    if (rc) return rc;
    const char *src = (const char *)h_src;
    char *dst = (char *)d_dst;
    size_t off = 0;
    int b = 0;
    bool used[2] = {false, false};
    while (off < bytes) {
        size_t chunk = std::min(kStageBytes, bytes - off);
        if (used[b]) NSX_HIP(hipEventSynchronize(g_stage_ev[b]), "staging wait");
        std::memcpy(g_stage[b], src + off, chunk);
        NSX_HIP(hipMemcpyAsync(dst + off, g_stage[b], chunk, hipMemcpyHostToDevice,
                               g_stage_stream), "h2d copy");
        NSX_HIP(hipEventRecord(g_stage_ev[b], g_stage_stream), "staging record");
        used[b] = true;
        off += chunk;
        b ^= 1;
    }
    NSX_HIP(hipStreamSynchronize(g_stage_stream), "h2d sync");
    return kOk;
}

int nsx_memcpy_d2h(void *h_dst, const void *d_src, size_t bytes) {
    if (bytes == 0) return kOk;
    int rc = ensure_device();
    if (rc) return rc;
    // Work queued on the default stream must finish before reading results.
    NSX_HIP(hipDeviceSynchronize(), "d2h device sync");
    std::lock_guard<std::mutex> lock(g_stage_mutex);
    rc = stage_init();
    if (rc) return rc;
    char *dst = (char *)h_dst;
    const char *src = (const char *)d_src;
    size_t off = 0;
    // Two chunks in flight: copy out chunk c while chunk c+1 transfers.
    size_t pend_off[2] = {0, 0}, pend_len[2] = {0, 0};
    int b = 0;
    while (off < bytes || pend_len[0] || pend_len[1]) {
        if (pend_len[b]) {
            NSX_HIP(hipEventSynchronize(g_stage_ev[b]), "staging wait");
            std::memcpy(dst + pend_off[b], g_stage[b], pend_len[b]);
            pend_len[b] = 0;
        }
        if (off < bytes) {
            size_t chunk = std::min(kStageBytes, bytes - off);
            NSX_HIP(hipMemcpyAsync(g_stage[b], src + off, chunk, hipMemcpyDeviceToHost,
                                   g_stage_stream), "d2h copy");
            NSX_HIP(hipEventRecord(g_stage_ev[b], g_stage_stream), "staging record");
            pend_off[b] = off;
            pend_len[b] = chunk;
            off += chunk;
        }
        b ^= 1;
    }
    return kOk;
}

int nsx_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes, void *stream) {
    if (bytes == 0) return kOk;
    NSX_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice,
                           (hipStream_t)stream), "d2d copy");
    return kOk;
}

int nsx_stream_sync(void *stream) {
    NSX_HIP(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
    return kOk;
}

int nsx_device_sync(void) {
    NSX_HIP(hipDeviceSynchronize(), "hipDeviceSynchronize");
    return kOk;
}

int nsx_event_create(void **ev) {
    int rc = ensure_device();
    if (rc) return rc;
    hipEvent_t e;
    NSX_HIP(hipEventCreate(&e), "hipEventCreate");
    *ev = (void *)e;
    return kOk;
}

int nsx_event_destroy(void *ev) {
    NSX_HIP(hipEventDestroy((hipEvent_t)ev), "hipEventDestroy");
    return kOk;
}

int nsx_event_record(void *ev, void *stream) {
    NSX_HIP(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream), "hipEventRecord");
    return kOk;
}

int nsx_event_elapsed_ms(void *start, void *stop, float *ms) {
    NSX_HIP(hipEventSynchronize((hipEvent_t)stop), "hipEventSynchronize");
    NSX_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop),
            "hipEventElapsedTime");
    return kOk;
}

int nsx_release_workspace(void) {
    clear_error();
    if (g_device < 0) return kOk;
    NSX_HIP(hipDeviceSynchronize(), "release sync");
    release_workspaces();
    return kOk;
}

int nsx_ksg_set_table_budget(long long bytes) {
    clear_error();
    NSX_REQUIRE(bytes >= 0, "nsx_ksg_set_table_budget: bytes must be >= 0");
    set_table_budget((size_t)bytes);
    return kOk;
}

int nsx_psi_host(float *h_psi, int n) {
    NSX_REQUIRE(h_psi && n >= 0, "nsx_psi_host: bad arguments");
    psi_table(h_psi, n);
    return kOk;
}

}  // extern "C"
