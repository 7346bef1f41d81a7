// api.hip — the extern "C" entry points of libgoqu_hip (declared in include/goqu_hip.h).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <new>
#include <stdexcept>

#include <cstdio>

#include <mutex>

#include "runtime.h"

namespace goqu {

// The calling thread's last error message, in a fixed buffer: recording an error never allocates.
static thread_local char g_err[1024];

int fail_cstr(int code, const char *msg) noexcept {
    std::strncpy(g_err, msg ? msg : "", sizeof(g_err) - 1);
    g_err[sizeof(g_err) - 1] = 0;
    return code;
}
void set_error(const std::string &msg) { fail_cstr(0, msg.c_str()); }
int fail(int code, const std::string &msg) { return fail_cstr(code, msg.c_str()); }

int on_exception() noexcept {
    try {
        throw;
    } catch (const std::bad_alloc &) {
        return fail_cstr(GOQU_ERR_OUT_OF_MEMORY, "host memory exhausted");
    } catch (const std::length_error &) {
        return fail_cstr(GOQU_ERR_OUT_OF_MEMORY, "host container size limit exceeded");
    } catch (const std::exception &e) {
        return fail_cstr(GOQU_ERR_DEVICE, e.what());
    } catch (...) {
        return fail_cstr(GOQU_ERR_DEVICE, "unknown internal error");
    }
}

// ---- fault injection ------------------------------------------------------------------------------
static std::atomic<int> g_fault_kind{FAULT_NONE};
static std::atomic<int> g_fault_countdown{0};

static int arm_fault(int kind, int countdown) {
    if (kind < FAULT_NONE || kind > FAULT_HOST_ALLOC || countdown < 0) return GOQU_ERR_INVALID_ARG;
    g_fault_countdown.store(countdown);
    g_fault_kind.store(countdown > 0 ? kind : FAULT_NONE);
    return GOQU_OK;
}

// GOQU_FAULT_INJECT=device_alloc:<n> or host_alloc:<n> arms the hook when the library loads
[[maybe_unused]] static const bool g_fault_env = [] {
    if (const char *e = std::getenv("GOQU_FAULT_INJECT")) {
        const char *c = std::strchr(e, ':');
        const int n = c ? std::atoi(c + 1) : 1;
        if (!std::strncmp(e, "device_alloc", 12)) arm_fault(FAULT_DEVICE_ALLOC, n);
        if (!std::strncmp(e, "host_alloc", 10)) arm_fault(FAULT_HOST_ALLOC, n);
    }
    return true;
}();

bool fault_fires(int kind) noexcept {
    if (g_fault_kind.load() != kind) return false;
    if (g_fault_countdown.fetch_sub(1) != 1) return false;
    g_fault_kind.store(FAULT_NONE);
    return true;
}

void host_fault_point() {
    if (fault_fires(FAULT_HOST_ALLOC)) throw std::bad_alloc();
}

hipError_t dmalloc(void **p, size_t bytes) {
    if (fault_fires(FAULT_DEVICE_ALLOC)) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        cache_release_all();
        e = hipMalloc(p, bytes);
    }
    return e;
}

namespace {
constexpr int kCacheDevices = 64;
struct Kept {
    void *p = nullptr;
    size_t bytes = 0;
};
std::mutex g_cache_mu;
Kept g_kept[kCacheDevices][CACHE_KINDS];

int cache_device() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kCacheDevices) return -1;
    return dev;
}
}  // namespace

hipError_t cache_take(int kind, void **p, size_t bytes, size_t *got) {
    *p = nullptr;
    *got = 0;
    const int dev = cache_device();
    Kept old;
    if (dev >= 0 && kind >= 0 && kind < CACHE_KINDS) {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        Kept &k = g_kept[dev][kind];
        if (k.p && k.bytes >= bytes) {
            *p = k.p;
            *got = k.bytes;
            k = Kept();
            return hipSuccess;
        }
        old = k;
        k = Kept();
    }
    if (old.p) (void)hipFree(old.p);
    const hipError_t e = dmalloc(p, bytes);
    if (e == hipSuccess) *got = bytes;
    return e;
}

void cache_give(int kind, void *p, size_t bytes) noexcept {
    if (!p) return;
    const int dev = cache_device();
    void *drop = p;
    if (dev >= 0 && kind >= 0 && kind < CACHE_KINDS) {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        Kept &k = g_kept[dev][kind];
        if (!k.p || k.bytes < bytes) {  // keep the larger of the two
            drop = k.p;
            k.p = p;
            k.bytes = bytes;
        }
    }
    if (drop) (void)hipFree(drop);
}

void cache_release_all() noexcept {
    const int dev = cache_device();
    if (dev < 0) return;
    Kept all[CACHE_KINDS];
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        for (int k = 0; k < CACHE_KINDS; k++) {
            all[k] = g_kept[dev][k];
            g_kept[dev][k] = Kept();
        }
    }
    for (auto &k : all)
        if (k.p) (void)hipFree(k.p);
}

static int check_device() {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n < 1) return fail(GOQU_ERR_NO_DEVICE, "no HIP device visible");
    return GOQU_OK;
}

static int alloc_state(State &st, int nq, int prec) {
    st.nq = nq;
    st.prec = prec;
    st.elem = prec == GOQU_C128 ? 16 : 8;
    hipError_t e = dmalloc((void **)&st.amps, st.size() * st.elem);
    if (e != hipSuccess) {
        st.amps = nullptr;
        (void)hipGetLastError();
        return fail(GOQU_ERR_OUT_OF_MEMORY, "cannot allocate " + std::to_string(st.size() * st.elem) +
                                                " bytes of device memory: " + hipGetErrorString(e));
    }
    return GOQU_OK;
}

static void free_state(State &st) {
    if (st.amps) (void)hipFree(st.amps);
    st.amps = nullptr;
}

// i^n (sim/pauli/pauli.go:212-224)
static void ipow(int n, double &re, double &im) {
    switch (((n % 4) + 4) % 4) {
    case 1: re = 0, im = 1; break;
    case 2: re = -1, im = 0; break;
    case 3: re = 0, im = -1; break;
    default: re = 1, im = 0; break;
    }
}

// Expectations of Pauli terms: terms sharing an x mask share one sweep.
static int expect_terms(State &st, int nq_user, int nt, const uint64_t *x, const uint64_t *z, const int32_t *ny,
                        const double *coeffs, double *out, Scratch &scr, hipStream_t s) {
    const uint64_t valid = nq_user >= 64 ? ~0ull : ((1ull << nq_user) - 1);
    std::map<uint64_t, std::vector<int>> groups;
    for (int t = 0; t < nt; t++) {
        if ((x[t] & ~valid) || (z[t] & ~valid)) return fail(GOQU_ERR_INVALID_ARG, "Pauli mask beyond the qubit count");
        groups[x[t]].push_back(t);
    }
    for (auto &g : groups) {
        const std::vector<int> &idx = g.second;
        std::vector<uint64_t> zs(idx.size());
        for (size_t k = 0; k < idx.size(); k++) zs[k] = z[idx[k]];
        std::vector<double> sums(2 * idx.size());
        int rc = pauli_sums(st, g.first, zs.data(), (int)idx.size(), sums.data(), scr, s);
        if (rc) return rc;
        for (size_t k = 0; k < idx.size(); k++) {
            const int t = idx[k];
            double yr, yi;
            ipow(ny[t], yr, yi);
            // coeff * i^ny * S  (expect_sv.go:31-46)
            const double sr = sums[2 * k] * yr - sums[2 * k + 1] * yi;
            const double si = sums[2 * k] * yi + sums[2 * k + 1] * yr;
            const double cr = coeffs[2 * t], ci = coeffs[2 * t + 1];
            out[2 * t] = cr * sr - ci * si;
            out[2 * t + 1] = cr * si + ci * sr;
        }
    }
    return GOQU_OK;
}

template <typename R>
__global__ void k_to_f64(const cx<R> *__restrict__ src, cx<double> *__restrict__ dst, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = cx<double>{(double)src[i].x, (double)src[i].y};
}
__global__ void k_from_f64(const cx<double> *__restrict__ src, cx<float> *__restrict__ dst, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = cx<float>{(float)src[i].x, (float)src[i].y};
}

}  // namespace goqu

using namespace goqu;

extern "C" {

const char *goqu_last_error(void) { return g_err; }

int goqu_debug_inject_fault(int32_t kind, int32_t countdown) {
    GOQU_GUARD_BEGIN
    if (arm_fault(kind, countdown)) return fail_cstr(GOQU_ERR_INVALID_ARG, "bad fault kind or countdown");
    return GOQU_OK;
    GOQU_GUARD_END
}

int goqu_device_count(int32_t *out) {
    GOQU_GUARD_BEGIN
    if (!out) return fail(GOQU_ERR_INVALID_ARG, "null pointer");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *out = e == hipSuccess ? n : 0;
    return GOQU_OK;
    GOQU_GUARD_END
}

int goqu_device_name(char *buf, int32_t len) {
    GOQU_GUARD_BEGIN
    if (!buf || len < 1) return fail(GOQU_ERR_INVALID_ARG, "null or empty buffer");
    int dev = 0;
    hipDeviceProp_t p;
    GOQU_HIP(hipGetDevice(&dev));
    GOQU_HIP(hipGetDeviceProperties(&p, dev));
    std::snprintf(buf, (size_t)len, "%s", p.gcnArchName);
    return GOQU_OK;
    GOQU_GUARD_END
}

// ---- planning on the host (no device needed) -------------------------------------------------------
int goqu_plan_stats(int32_t num_qubits, int32_t density, const goqu_op_t *ops, int32_t num_ops, int32_t tile_bits,
                    int64_t *stages, int64_t *instrs) {
    GOQU_GUARD_BEGIN
    if (num_ops < 0 || (num_ops > 0 && !ops) || !stages || !instrs)
        return fail(GOQU_ERR_INVALID_ARG, "bad argument");
    const int maxq = density ? 16 : 34;
    if (num_qubits < 1 || num_qubits > maxq) return fail(GOQU_ERR_INVALID_ARG, "qubit count out of range");
    if (tile_bits != 0 && (tile_bits < 8 || tile_bits > kMaxTileBits))
        return fail(GOQU_ERR_INVALID_ARG, "tile_bits must be 0 (default) or 8..13");
    host_fault_point();
    std::vector<EOp> eops;
    std::string err;
    const int nqi = std::max(density ? 2 * num_qubits : num_qubits, kMinInternalQubits);
    int rc = density ? lower_dm_ops(ops, num_ops, num_qubits, eops, err)
                     : lower_sv_ops(ops, num_ops, num_qubits, eops, err);
    if (rc) return fail(rc, err);
    PlanConfig cfg = default_plan_config(GOQU_C128, nqi, false);
    if (tile_bits) cfg.tile_bits = tile_bits;
    const Program P = plan_program(eops, nqi, cfg);
    *stages = (int64_t)P.stages.size();
    *instrs = (int64_t)P.instrs.size();
    return GOQU_OK;
    GOQU_GUARD_END
}

// ---- statevector ----------------------------------------------------------------------------------
int goqu_sv_create(int32_t num_qubits, int32_t precision, goqu_sv_t **out) {
    GOQU_GUARD_BEGIN
    if (!out) return fail(GOQU_ERR_INVALID_ARG, "null output handle");
    *out = nullptr;
    if (num_qubits < 1 || num_qubits > 34)
        return fail(GOQU_ERR_INVALID_ARG, "statevector: numQubits " + std::to_string(num_qubits) + " out of range [1, 34]");
    if (precision != GOQU_C128 && precision != GOQU_C64) return fail(GOQU_ERR_INVALID_ARG, "unknown precision");
    int rc = check_device();
    if (rc) return rc;
    goqu_sv *h = new goqu_sv();
    h->nq_user = num_qubits;
    if (hipGetDevice(&h->device) != hipSuccess) {
        delete h;
        return fail(GOQU_ERR_DEVICE, "hipGetDevice failed");
    }
    rc = alloc_state(h->st, std::max(num_qubits, kMinInternalQubits), precision);
    if (rc) {
        delete h;
        return rc;
    }
    if (hipStreamCreate(&h->stream) != hipSuccess) {
        free_state(h->st);
        delete h;
        return fail(GOQU_ERR_DEVICE, "hipStreamCreate failed");
    }
    rc = fill_basis(h->st, 0, h->stream);
    if (!rc && hipStreamSynchronize(h->stream) != hipSuccess) rc = fail(GOQU_ERR_DEVICE, "init failed");
    if (rc) {
        goqu_sv_destroy(h);
        return rc;
    }
    *out = h;
    return GOQU_OK;
    GOQU_GUARD_END
}

int goqu_sv_destroy(goqu_sv_t *h) {
    GOQU_GUARD_BEGIN
    if (!h) return GOQU_OK;
    if (h->stream) {
        (void)hipStreamSynchronize(h->stream);
        (void)hipStreamDestroy(h->stream);
    }
    free_state(h->st);
    free_state(h->adj_lam);
    free_state(h->adj_work);
    delete h;
    return GOQU_OK;
    GOQU_GUARD_END
}

int goqu_sv_reset(goqu_sv_t *h) {
    GOQU_GUARD_BEGIN
    if (!h) return fail(GOQU_ERR_INVALID_ARG, "null handle");
    int rc = fill_basis(h->st, 0, h->stream);
    if (rc) return rc;
    GOQU_HIP(hipStreamSynchronize(h->stream));
    return GOQU_OK;
    GOQU_GUARD_END
}

// The content of a statevector op list as cache-key bytes: per op its kind, gate id, qubits and
// matrix entries.  False for records the key does not cover (the op list then runs uncached and
// the lowering reports any error).
static bool sv_ops_key(const goqu_op_t *ops, int nops, std::vector<unsigned char> &key) {
    key.clear();
    auto put = [&](const void *p, size_t n) {
        const unsigned char *c = static_cast<const unsigned char *>(p);
        key.insert(key.end(), c, c + n);
    };
    for (int i = 0; i < nops; i++) {
        const goqu_op_t &o = ops[i];
        if (o.num_qubits < 0 || o.num_qubits > GOQU_MAX_OP_QUBITS) return false;
        int mk = 0;  // the matrix acts on mk qubits
        if (o.kind == GOQU_OP_GATE) mk = o.num_qubits;
        else if (o.kind == GOQU_OP_CONTROLLED) mk = o.num_qubits - o.num_controls;
        else if (o.kind != GOQU_OP_RESET && o.kind != GOQU_OP_SKIP) return false;
        if ((o.kind == GOQU_OP_GATE || o.kind == GOQU_OP_CONTROLLED) && (mk < 1 || mk > kMaxWideTargets)) return false;
        put(&o.kind, 4);
        put(&o.gate_id, 4);
        put(&o.num_qubits, 4);
        put(&o.num_controls, 4);
        put(o.qubits, sizeof(int32_t) * (size_t)o.num_qubits);
        const unsigned char has = o.matrix != nullptr && mk > 0;
        put(&has, 1);
        if (has) put(o.matrix, sizeof(double) * (size_t(2) << (2 * mk)));
    }
    return true;
}

int goqu_sv_apply(goqu_sv_t *h, const goqu_op_t *ops, int32_t num_ops) {
    GOQU_GUARD_BEGIN
    if (!h || (num_ops > 0 && !ops)) return fail(GOQU_ERR_INVALID_ARG, "null argument");
    host_fault_point();
    ProgramCache &C = h->cache;
    std::vector<unsigned char> key;
    const bool keyed = num_ops > 0 && sv_ops_key(ops, num_ops, key);
    uint64_t hash = 1469598103934665603ull;  // FNV-1a
    if (keyed)
        for (unsigned char c : key) hash = (hash ^ c) * 1099511628211ull;
    auto launch = [&](const PreparedOps &p) -> int {
        int rc = launch_prepared(h->st, p, h->stream, nullptr);
        hipError_t e = hipStreamSynchronize(h->stream);
        if (rc) return rc;
        if (e != hipSuccess) return fail(GOQU_ERR_DEVICE, std::string("stage kernels: ") + hipGetErrorString(e));
        return GOQU_OK;
    };
    if (keyed)
        for (ProgramCache::Entry &E : C.entries)
            if (E.hash == hash && E.key == key) {
                E.used = ++C.clock;
                return launch(E.prep);
            }
    std::vector<EOp> eops;
    std::string err;
    int rc = lower_sv_ops(ops, num_ops, h->nq_user, eops, err);
    if (rc) return fail(rc, err);
    if (!keyed) return run_eops(h->st, eops, h->stream, h->plan_set ? &h->plan : nullptr);
    PreparedOps prep;
    const PlanConfig cfg = h->plan_set ? h->plan : default_plan_config(h->st.prec, h->st.nq, false);
    rc = prepare_eops(h->st, eops, cfg, h->stream, prep);
    if (rc) return rc;
    if (C.entries.size() >= ProgramCache::kMaxEntries) {  // the least recently used goes
        size_t lru = 0;
        for (size_t i = 1; i < C.entries.size(); i++)
            if (C.entries[i].used < C.entries[lru].used) lru = i;
        C.entries.erase(C.entries.begin() + (std::ptrdiff_t)lru);
    }
    C.entries.emplace_back();
    ProgramCache::Entry &E = C.entries.back();
    E.key = std::move(key);
    E.hash = hash;
    E.used = ++C.clock;
    E.prep = std::move(prep);
    return launch(E.prep);
    GOQU_GUARD_END
}

int goqu_sv_set_plan(goqu_sv_t *h, int32_t tile_bits, int32_t low_bits) {
    GOQU_GUARD_BEGIN
    if (!h) return fail_cstr(GOQU_ERR_INVALID_ARG, "null handle");
    h->cache.clear();  // cached programs follow the tile shape they were planned with
    if (tile_bits == 0 && low_bits == 0) {
        h->plan_set = false;
        return GOQU_OK;
    }
    if (tile_bits < 8 || tile_bits > kMaxTileBits || low_bits < 2 || low_bits > 6 || low_bits > tile_bits - kRegBits)
        return fail_cstr(GOQU_ERR_INVALID_ARG, "tile_bits must be 8..13 and low_bits 2..min(6, tile_bits - 4)");
    h->plan.tile_bits = tile_bits;
    h->plan.low_bits = low_bits;
    h->plan_set = true;
    return GOQU_OK;
    GOQU_GUARD_END
}

int goqu_sv_get_state(goqu_sv_t *h, double *amps, uint64_t offset, uint64_t count) {
    GOQU_GUARD_BEGIN
    if (!h || (count && !amps)) return fail(GOQU_ERR_INVALID_ARG, "null argument");
    const uint64_t N = 1ull << h->nq_user;
    if (offset > N || count > N - offset) return fail(GOQU_ERR_INVALID_ARG, "range beyond the state");
    if (!count) return GOQU_OK;
    if (h->st.prec == GOQU_C128) {
        GOQU_HIP(hipMemcpyAsync(amps, static_cast<char *>(h->st.amps) + offset * 16, count * 16, hipMemcpyDeviceToHost,
                                h->stream));
        GOQU_HIP(hipStreamSynchronize(h->stream));
        return GOQU_OK;
    }
    const uint64_t chunk = 1ull << 24;
    int rc = h->scr.ensure(chunk * 16);
    if (rc) return rc;
    for (uint64_t o = 0; o < count; o += chunk) {
        const uint64_t c = std::min(chunk, count - o);
        hipLaunchKernelGGL(k_to_f64<float>, dim3(1024), dim3(256), 0, h->stream,
                           static_cast<const cx<float> *>(h->st.amps) + offset + o,
                           static_cast<cx<double> *>(h->scr.ptr), c);
        GOQU_HIP(hipGetLastError());
        GOQU_HIP(hipMemcpyAsync(amps + 2 * o, h->scr.ptr, c * 16, hipMemcpyDeviceToHost, h->stream));
        GOQU_HIP(hipStreamSynchronize(h->stream));
    }
    return GOQU_OK;
    GOQU_GUARD_END
}

int goqu_sv_set_state(goqu_sv_t *h, const double *amps, uint64_t offset, uint64_t count) {
    GOQU_GUARD_BEGIN
    if (!h || (count && !amps)) return fail(GOQU_ERR_INVALID_ARG, "null argument");
    const uint64_t N = 1ull << h->nq_user;
    if (offset > N || count > N - offset) return fail(GOQU_ERR_INVALID_ARG, "range beyond the state");
    if (!count) return GOQU_OK;
    if (h->st.prec == GOQU_C128) {
        GOQU_HIP(hipMemcpyAsync(static_cast<char *>(h->st.amps) + offset * 16, amps, count * 16, hipMemcpyHostToDevice,
                                h->stream));
        GOQU_HIP(hipStreamSynchronize(h->stream));
        return GOQU_OK;
    }
    const uint64_t chunk = 1ull << 24;
    int rc = h->scr.ensure(chunk * 16);
    if (rc) return rc;
    for (uint64_t o = 0; o < count; o += chunk) {
        const uint64_t c = std::min(chunk, count - o);
        GOQU_HIP(hipMemcpyAsync(h->scr.ptr, amps + 2 * o, c * 16, hipMemcpyHostToDevice, h->stream));
        hipLaunchKernelGGL(k_from_f64, dim3(1024), dim3(256), 0, h->stream, static_cast<const cx<double> *>(h->scr.ptr),
                           static_cast<cx<float> *>(h->st.amps) + offset + o, c);
        GOQU_HIP(hipGetLastError());
        GOQU_HIP(hipStreamSynchronize(h->stream));
    }
    return GOQU_OK;
    GOQU_GUARD_END
}

int goqu_sv_sample(goqu_sv_t *h, const double *uniforms, int64_t shots, uint64_t *out_idx) {
    GOQU_GUARD_BEGIN
    if (!h || shots < 0 || (shots > 0 && (!uniforms || !out_idx))) return fail(GOQU_ERR_INVALID_ARG, "bad argument");
    return sample_state(h->st, 1ull << h->nq_user, uniforms, shots, out_idx, h->scr, h->stream);
    GOQU_GUARD_END
}

int goqu_sv_sample_noalloc(goqu_sv_t *h, const double *uniforms, int64_t shots, uint64_t *out_idx) {
    GOQU_GUARD_BEGIN
    if (!h || shots < 0 || (shots > 0 && (!uniforms || !out_idx))) return fail_cstr(GOQU_ERR_INVALID_ARG, "bad argument");
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return fail_cstr(GOQU_ERR_DEVICE, "hipGetDevice failed");
    if (dev != h->device)
        return fail_cstr(GOQU_ERR_INVALID_ARG, "sample without allocation: the handle's device is not the current device");
    return sample_state_static(h->st, dev, 1ull << h->nq_user, uniforms, shots, out_idx, h->stream);
    GOQU_GUARD_END
}

int goqu_sv_expect_paulis(goqu_sv_t *h, int32_t num_terms, const uint64_t *xmask, const uint64_t *zmask,
                          const int32_t *num_y, const double *coeffs, double *out) {
    GOQU_GUARD_BEGIN
    if (!h || num_terms < 0 || (num_terms > 0 && (!xmask || !zmask || !num_y || !coeffs || !out)))
        return fail(GOQU_ERR_INVALID_ARG, "bad argument");
    return expect_terms(h->st, h->nq_user, num_terms, xmask, zmask, num_y, coeffs, out, h->scr, h->stream);
    GOQU_GUARD_END
}

int goqu_sv_expectation_z(goqu_sv_t *h, uint64_t mask, double *out) {
    GOQU_GUARD_BEGIN
    if (!h || !out) return fail(GOQU_ERR_INVALID_ARG, "bad argument");
    const uint64_t x = 0;
    const int32_t ny = 0;
    const double c[2] = {1, 0};
    double r[2];
    int rc = expect_terms(h->st, h->nq_user, 1, &x, &mask, &ny, c, r, h->scr, h->stream);
    if (rc) return rc;
    *out = r[0];
    return GOQU_OK;
    GOQU_GUARD_END
}

int goqu_sv_measure_qubit(goqu_sv_t *h, int32_t qubit, double u, int32_t *outcome) {
    GOQU_GUARD_BEGIN
    if (!h || !outcome || qubit < 0 || qubit >= h->nq_user) return fail(GOQU_ERR_INVALID_ARG, "bad argument");
    double p0 = 0;
    int rc = prob_zero(h->st, qubit, &p0, h->scr, h->stream);
    if (rc) return rc;
    const int o = (u >= p0) ? 1 : 0;  // dynamic.go:195-198
    *outcome = o;
    const double norm = o == 0 ? std::sqrt(p0) : std::sqrt(1 - p0);
    if (norm < 1e-15) return GOQU_OK;
    rc = collapse(h->st, qubit, o, 1.0 / norm, h->stream);
    if (rc) return rc;
    GOQU_HIP(hipStreamSynchronize(h->stream));
    return GOQU_OK;
    GOQU_GUARD_END
}

int goqu_sv_norm2(goqu_sv_t *h, double *out) {
    GOQU_GUARD_BEGIN
    if (!h || !out) return fail(GOQU_ERR_INVALID_ARG, "bad argument");
    return prob_zero(h->st, -1, out, h->scr, h->stream);
    GOQU_GUARD_END
}

// ---- density matrix -------------------------------------------------------------------------------------
int goqu_dm_create(int32_t num_qubits, goqu_dm_t **out) {
    GOQU_GUARD_BEGIN
    if (!out) return fail(GOQU_ERR_INVALID_ARG, "null output handle");
    *out = nullptr;
    if (num_qubits < 1 || num_qubits > 16)
        return fail(GOQU_ERR_INVALID_ARG, "densitymatrix: numQubits " + std::to_string(num_qubits) + " out of range [1, 16]");
    int rc = check_device();
    if (rc) return rc;
    goqu_dm *h = new goqu_dm();
    h->n = num_qubits;
    rc = alloc_state(h->st, std::max(2 * num_qubits, kMinInternalQubits), GOQU_C128);
    if (rc) {
        delete h;
        return rc;
    }
    const size_t pb = (size_t(1) << num_qubits) * sizeof(double);
    if (hipStreamCreate(&h->stream) != hipSuccess || dmalloc((void **)&h->dprobs, pb) != hipSuccess ||
        dmalloc((void **)&h->dtmp, pb) != hipSuccess) {
        goqu_dm_destroy(h);
        return fail(GOQU_ERR_DEVICE, "density matrix set-up failed");
    }
    rc = fill_basis(h->st, 0, h->stream);
    if (!rc && hipStreamSynchronize(h->stream) != hipSuccess) rc = fail(GOQU_ERR_DEVICE, "init failed");
    if (rc) {
        goqu_dm_destroy(h);
        return rc;
    }
    *out = h;
    return GOQU_OK;
    GOQU_GUARD_END
}

int goqu_dm_destroy(goqu_dm_t *h) {
    GOQU_GUARD_BEGIN
    if (!h) return GOQU_OK;
    if (h->stream) {
        (void)hipStreamSynchronize(h->stream);
        (void)hipStreamDestroy(h->stream);
    }
    if (h->dprobs) (void)hipFree(h->dprobs);
    if (h->dtmp) (void)hipFree(h->dtmp);
    free_state(h->st);
    delete h;
    return GOQU_OK;
    GOQU_GUARD_END
}

int goqu_dm_reset(goqu_dm_t *h) {
    GOQU_GUARD_BEGIN
    if (!h) return fail(GOQU_ERR_INVALID_ARG, "null handle");
    int rc = fill_basis(h->st, 0, h->stream);
    if (rc) return rc;
    GOQU_HIP(hipStreamSynchronize(h->stream));
    return GOQU_OK;
    GOQU_GUARD_END
}

int goqu_dm_apply(goqu_dm_t *h, const goqu_op_t *ops, int32_t num_ops) {
    GOQU_GUARD_BEGIN
    if (!h || (num_ops > 0 && !ops)) return fail(GOQU_ERR_INVALID_ARG, "null argument");
    host_fault_point();
    std::vector<EOp> eops;
    std::string err;
    int rc = lower_dm_ops(ops, num_ops, h->n, eops, err);
    if (rc) return fail(rc, err);
    return run_eops(h->st, eops, h->stream);
    GOQU_GUARD_END
}

int goqu_dm_get(goqu_dm_t *h, double *rho, uint64_t offset, uint64_t count) {
    GOQU_GUARD_BEGIN
    if (!h || (count && !rho)) return fail(GOQU_ERR_INVALID_ARG, "null argument");
    const uint64_t N = 1ull << (2 * h->n);
    if (offset > N || count > N - offset) return fail(GOQU_ERR_INVALID_ARG, "range beyond the matrix");
    if (!count) return GOQU_OK;
    GOQU_HIP(hipMemcpyAsync(rho, static_cast<char *>(h->st.amps) + offset * 16, count * 16, hipMemcpyDeviceToHost,
                            h->stream));
    GOQU_HIP(hipStreamSynchronize(h->stream));
    return GOQU_OK;
    GOQU_GUARD_END
}

int goqu_dm_set(goqu_dm_t *h, const double *rho, uint64_t offset, uint64_t count) {
    GOQU_GUARD_BEGIN
    if (!h || (count && !rho)) return fail(GOQU_ERR_INVALID_ARG, "null argument");
    const uint64_t N = 1ull << (2 * h->n);
    if (offset > N || count > N - offset) return fail(GOQU_ERR_INVALID_ARG, "range beyond the matrix");
    if (!count) return GOQU_OK;
    GOQU_HIP(hipMemcpyAsync(static_cast<char *>(h->st.amps) + offset * 16, rho, count * 16, hipMemcpyHostToDevice,
                            h->stream));
    GOQU_HIP(hipStreamSynchronize(h->stream));
    return GOQU_OK;
    GOQU_GUARD_END
}

int goqu_dm_probs(goqu_dm_t *h, const double *p01, const double *p10, double *out) {
    GOQU_GUARD_BEGIN
    if (!h || !out) return fail(GOQU_ERR_INVALID_ARG, "null argument");
    int rc = dm_diag(h->st, h->n, h->dprobs, h->stream);
    if (rc) return rc;
    if (p01 && p10) {
        for (int q = 0; q < h->n; q++) {
            if (p01[q] < 0 || p10[q] < 0) continue;  // qubit without readout error
            rc = dm_readout(h->dprobs, h->dtmp, h->n, q, p01[q], p10[q], h->stream);
            if (rc) return rc;
        }
    }
    GOQU_HIP(hipMemcpyAsync(out, h->dprobs, (size_t(1) << h->n) * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    GOQU_HIP(hipStreamSynchronize(h->stream));
    return GOQU_OK;
    GOQU_GUARD_END
}

int goqu_dm_sample(goqu_dm_t *h, const double *probs, const double *uniforms, int64_t shots, uint64_t *out_idx) {
    GOQU_GUARD_BEGIN
    if (!h || !probs || shots < 0 || (shots > 0 && (!uniforms || !out_idx)))
        return fail(GOQU_ERR_INVALID_ARG, "bad argument");
    const uint64_t dim = 1ull << h->n;
    GOQU_HIP(hipMemcpyAsync(h->dtmp, probs, dim * sizeof(double), hipMemcpyHostToDevice, h->stream));
    return sample_probs(h->dtmp, dim, uniforms, shots, out_idx, h->scr, h->stream);
    GOQU_GUARD_END
}

int goqu_dm_expect_paulis(goqu_dm_t *h, int32_t num_terms, const uint64_t *xmask, const uint64_t *zmask,
                          const int32_t *num_y, const double *coeffs, double *out) {
    GOQU_GUARD_BEGIN
    if (!h || num_terms < 0 || (num_terms > 0 && (!xmask || !zmask || !num_y || !coeffs || !out)))
        return fail(GOQU_ERR_INVALID_ARG, "bad argument");
    const uint64_t valid = (1ull << h->n) - 1;
    for (int t = 0; t < num_terms; t++) {
        if ((xmask[t] & ~valid) || (zmask[t] & ~valid))
            return fail(GOQU_ERR_INVALID_ARG, "Pauli mask beyond the qubit count");
        double sres[2];
        int rc = dm_pauli(h->st, h->n, xmask[t], zmask[t], sres, h->scr, h->stream);
        if (rc) return rc;
        double yr, yi;
        ipow(num_y[t], yr, yi);
        const double sr = sres[0] * yr - sres[1] * yi, si = sres[0] * yi + sres[1] * yr;
        const double cr = coeffs[2 * t], ci = coeffs[2 * t + 1];
        out[2 * t] = cr * sr - ci * si;
        out[2 * t + 1] = cr * si + ci * sr;
    }
    return GOQU_OK;
    GOQU_GUARD_END
}

int goqu_dm_expectation_z(goqu_dm_t *h, uint64_t mask, double *out) {
    GOQU_GUARD_BEGIN
    if (!h || !out) return fail(GOQU_ERR_INVALID_ARG, "bad argument");
    if (mask & ~((1ull << h->n) - 1)) return fail(GOQU_ERR_INVALID_ARG, "mask beyond the qubit count");
    double r[2];
    int rc = dm_pauli(h->st, h->n, 0, mask, r, h->scr, h->stream);
    if (rc) return rc;
    *out = r[0];
    return GOQU_OK;
    GOQU_GUARD_END
}

int goqu_dm_purity(goqu_dm_t *h, double *out) {
    GOQU_GUARD_BEGIN
    if (!h || !out) return fail(GOQU_ERR_INVALID_ARG, "bad argument");
    return dm_purity(h->st, h->n, out, h->scr, h->stream);
    GOQU_GUARD_END
}

int goqu_dm_fidelity(goqu_dm_t *h, const double *pure, double *out) {
    GOQU_GUARD_BEGIN
    if (!h || !pure || !out) return fail(GOQU_ERR_INVALID_ARG, "bad argument");
    const uint64_t dim = 1ull << h->n;
    void *dpsi = nullptr;
    GOQU_HIP(dmalloc((void **)&dpsi, dim * 16));
    hipError_t e = hipMemcpyAsync(dpsi, pure, dim * 16, hipMemcpyHostToDevice, h->stream);
    int rc = e == hipSuccess ? dm_fidelity(h->st, h->n, dpsi, out, h->scr, h->stream)
                             : fail(GOQU_ERR_DEVICE, hipGetErrorString(e));
    (void)hipStreamSynchronize(h->stream);
    (void)hipFree(dpsi);
    return rc;
    GOQU_GUARD_END
}

}  // extern "C"
