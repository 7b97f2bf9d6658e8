// Internal helpers shared by the libnsx translation units.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

namespace nsx {

// A named roctx range for the lifetime of the object: each C-ABI entry
// point opens one, so `rocprofv3 --marker-trace` attributes device time to
// library calls. Without an attached profiler the calls return at once.
struct TraceRange {
    explicit TraceRange(const char *name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange &) = delete;
    TraceRange &operator=(const TraceRange &) = delete;
};

// Thread-local last-error message behind nsx_last_error().
void set_error(const std::string &msg);
void clear_error();

// Return codes.
enum Status : int {
    kOk = 0,
    kInvalidArgument = 1,
    kHipError = 2,
    kNoDevice = 3,
    kInternal = 4,
};

// Check a HIP call; on failure record "<what>: <hip message>" and return.
#define NSX_HIP(call, what)                                                    \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) {                                                \
            ::nsx::set_error(std::string(what) + ": " + hipGetErrorString(e_)); \
            return ::nsx::kHipError;                                           \
        }                                                                      \
    } while (0)

#define NSX_REQUIRE(cond, msg)                                                 \
    do {                                                                       \
        if (!(cond)) {                                                         \
            ::nsx::set_error(msg);                                             \
            return ::nsx::kInvalidArgument;                                    \
        }                                                                      \
    } while (0)

// Ensure a usable device is current (initialises once per thread).
int ensure_device();

// Number of CUs of the current device.
int cu_count();

// Grow-only device workspace owned by the library, one per slot name.
// Returns nullptr (and sets the error) on allocation failure.
void *workspace(int slot, size_t bytes);

// The sorted-list tables of the KSG list path: one grow-only device buffer
// counted in node slots of slot_bytes each. Returns in *ptr / *slots the
// buffer and the slots it holds, at most want_slots. The buffer grows (the
// only time device memory is queried or allocated) when it holds fewer than
// want_slots and the budget allows more; *slots < min_slots means no room.
// The budget is set_table_budget()'s, or 85% of the device memory free
// (counting the current buffer) less 1 GiB.
int table_space(size_t slot_bytes, long long want_slots, long long min_slots, void **ptr,
                long long *slots);
void set_table_budget(size_t bytes);
size_t table_budget();  // 0: no cap set
// Frees every workspace and the table buffer (the device must be idle).
void release_workspaces();

enum WorkspaceSlot : int {
    kWsPsi = 0,
    kWsCounter = 1,
    kWsTieScratch = 2,
    kWsStats = 3,
    kWsWindows = 4,
    kWsApsp0 = 5,
    kWsApsp1 = 6,
    kWsScreen0 = 7,
    kWsScreen1 = 8,
    kWsApspRows = 9,
    kWsApspCand = 10,
    kWsApspRowsA = 11,
    kWsApspCandA = 12,
    kWsApspG = 13,
    kWsListMap = 14,
    kWsListCnt = 15,
    kWsListStats = 16,
    kWsListPerm = 17,
    kWsListSpan = 18,
    kWsCount = 20,
};

// 64 coherent pinned host words (allocated once per process): `host` for the
// host, `dev` for kernels. Kernels write results there that the host polls
// instead of waiting for a stream synchronisation.
int host_flags(unsigned **host, unsigned **dev);

// Fill `count` 32-bit words at d_ptr with `value` on `stream` (a kernel, so
// it may be recorded into a HIP graph).
int fill_u32(void *d_ptr, uint32_t value, size_t count, hipStream_t stream);

// psi table of psi.cpp:6-19, host side.
void psi_table(float *psi, int n);

}  // namespace nsx
