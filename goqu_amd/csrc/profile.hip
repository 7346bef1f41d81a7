// profile.hip — kernel timing for benchmarks: when enabled, launches of the profiled kernels are
// bracketed by HIP events on their own stream; the events are resolved (and the device
// synchronised) only when the totals are read, so profiling adds no host synchronisation to the
// paths it measures.
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <vector>

#include "runtime.h"

namespace goqu {

namespace {

struct Span {
    hipEvent_t a, b;
    int bin;
    double bytes;
};

std::atomic<bool> g_on{false};
std::mutex g_mu;
std::vector<Span> g_pending;
double g_ms[PROF_BINS] = {}, g_bytes[PROF_BINS] = {};
int64_t g_launches[PROF_BINS] = {};

// folds the finished spans into the totals (caller holds g_mu)
int resolve_locked() {
    if (g_pending.empty()) return GOQU_OK;
    GOQU_HIP(hipDeviceSynchronize());
    for (Span &sp : g_pending) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, sp.a, sp.b) == hipSuccess) {
            g_ms[sp.bin] += ms;
            g_launches[sp.bin] += 1;
            g_bytes[sp.bin] += sp.bytes;
        }
        (void)hipEventDestroy(sp.a);
        (void)hipEventDestroy(sp.b);
    }
    g_pending.clear();
    return GOQU_OK;
}

}  // namespace

bool profiling() noexcept { return g_on.load(std::memory_order_relaxed); }

hipEvent_t prof_begin(hipStream_t s) {
    if (!profiling()) return nullptr;
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    if (hipEventRecord(e, s) != hipSuccess) {
        (void)hipEventDestroy(e);
        return nullptr;
    }
    return e;
}

void prof_end(hipEvent_t start, hipStream_t s, int bin, double bytes) {
    if (!start) return;
    hipEvent_t e = nullptr;
    if (bin < 0 || bin >= PROF_BINS || hipEventCreate(&e) != hipSuccess || hipEventRecord(e, s) != hipSuccess) {
        (void)hipEventDestroy(start);
        if (e) (void)hipEventDestroy(e);
        return;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    g_pending.push_back({start, e, bin, bytes});
}

}  // namespace goqu

using namespace goqu;

extern "C" {
int goqu_profile_enable(int32_t on) {
    GOQU_GUARD_BEGIN
    g_on.store(on != 0);
    return GOQU_OK;
    GOQU_GUARD_END
}

int goqu_profile_reset(void) {
    GOQU_GUARD_BEGIN
    std::lock_guard<std::mutex> lk(g_mu);
    int rc = resolve_locked();
    for (int b = 0; b < PROF_BINS; b++) g_ms[b] = g_bytes[b] = 0, g_launches[b] = 0;
    return rc;
    GOQU_GUARD_END
}

int goqu_profile_read_kernel(int32_t kernel, double *kernel_ms, int64_t *launches, double *bytes) {
    GOQU_GUARD_BEGIN
    if (kernel < 0 || kernel >= PROF_BINS) return fail(GOQU_ERR_INVALID_ARG, "unknown profiled kernel");
    std::lock_guard<std::mutex> lk(g_mu);
    int rc = resolve_locked();
    if (rc) return rc;
    if (kernel_ms) *kernel_ms = g_ms[kernel];
    if (launches) *launches = g_launches[kernel];
    if (bytes) *bytes = g_bytes[kernel];
    return GOQU_OK;
    GOQU_GUARD_END
}

int goqu_profile_read(double *kernel_ms, int64_t *launches, double *bytes) {
    GOQU_GUARD_BEGIN
    std::lock_guard<std::mutex> lk(g_mu);
    int rc = resolve_locked();
    if (rc) return rc;
    if (kernel_ms) *kernel_ms = g_ms[PROF_STAGE] + g_ms[PROF_STAGE_DM];
    if (launches) *launches = g_launches[PROF_STAGE] + g_launches[PROF_STAGE_DM];
    if (bytes) *bytes = g_bytes[PROF_STAGE] + g_bytes[PROF_STAGE_DM];
    return GOQU_OK;
    GOQU_GUARD_END
}
}
