// runtime.h — handles, error plumbing and kernel launchers shared by the C ABI files.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/goqu_hip.h"
#include "engine.h"

namespace goqu {

void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
// fail() without allocating: usable while std::bad_alloc is in flight
int fail_cstr(int code, const char *msg) noexcept;

// Entry-point bodies run inside GOQU_GUARD_BEGIN / GOQU_GUARD_END: a C++ exception (host memory
// exhausted in the planner's or lowering's containers, a length error) becomes a status code and
// never crosses the C ABI into a Go or Python host.
int on_exception() noexcept;
#define GOQU_GUARD_BEGIN try {
#define GOQU_GUARD_END \
    }                  \
    catch (...) { return ::goqu::on_exception(); }

// Fault injection for tests (goqu_debug_inject_fault, or GOQU_FAULT_INJECT=<kind>:<n> in the
// environment at load time): the n-th device allocation after arming reports a HIP failure, or the
// n-th lowering of an op list throws std::bad_alloc.
enum FaultKind { FAULT_NONE = 0, FAULT_DEVICE_ALLOC = 1, FAULT_HOST_ALLOC = 2 };
bool fault_fires(int kind) noexcept;
// hipMalloc through the fault-injection hook; when the device reports out-of-memory, the retained
// buffers below are handed back to it and the allocation is tried once more
hipError_t dmalloc(void **p, size_t bytes);

// Retained device buffers.  The batch entry points (QAOA instance pools, circuit batches,
// dynamic-circuit shot states) take their largest buffers from here and give them back when the
// call ends.  Handing tens of GiB back to the device and requesting them again makes the driver
// clear the recycled memory: on MI355X a 4,096-instance QAOA batch (a 42 GiB pool) paid 0.4-1.2 s
// in hipMalloc and about 1 s in hipFree per call once the device's fresh memory was used up, against
// 0.3 s for the batch itself.  So one given-back buffer per kind and device is kept for the next
// call; a later call needing more frees it and allocates anew.
enum CacheKind { CACHE_QAOA = 0, CACHE_BATCH_STATES, CACHE_BATCH_BLOBS, CACHE_DYN_STATES, CACHE_KINDS };
// a buffer of at least `bytes` (its true size in *got), from the cache or from dmalloc
hipError_t cache_take(int kind, void **p, size_t bytes, size_t *got);
// keeps `p` (of `bytes`) for the next cache_take of this kind; the device must be done with it
void cache_give(int kind, void *p, size_t bytes) noexcept;
// frees every retained buffer of the current device
void cache_release_all() noexcept;

// RAII holder: takes from the cache, gives back on destruction.  The owner calls idle() once the
// device has finished every use of the buffer (its stream synchronized); a holder destroyed without
// that (an error path) waits for the whole device first, so a kept buffer is never still in use.
struct CachedDevBuf {
    int kind;
    void *p = nullptr;
    size_t bytes = 0;
    bool is_idle = false;
    explicit CachedDevBuf(int k) : kind(k) {}
    CachedDevBuf(const CachedDevBuf &) = delete;
    CachedDevBuf &operator=(const CachedDevBuf &) = delete;
    hipError_t take(size_t need) { return cache_take(kind, &p, need, &bytes); }
    void idle() { is_idle = true; }
    ~CachedDevBuf() {
        if (!p) return;
        if (!is_idle) (void)hipDeviceSynchronize();
        cache_give(kind, p, bytes);
    }
};
// throws std::bad_alloc when a host-allocation fault is armed and due (before op-list lowering)
void host_fault_point();

#define GOQU_HIP(call)                                                                        \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return (void)hipGetLastError(), /* clear the sticky per-thread error */          \
                   ::goqu::fail(e_ == hipErrorOutOfMemory ? GOQU_ERR_OUT_OF_MEMORY : GOQU_ERR_DEVICE, \
                                std::string(#call) + ": " + hipGetErrorString(e_));          \
    } while (0)

template <typename R>
struct alignas(2 * sizeof(R)) cx {
    R x, y;
};

// Device-side copy of a planned program.
struct DeviceProgram {
    void *blob = nullptr;  // headers | instrs | pool (pool as R)
    size_t bytes = 0;
    std::vector<StageHeader> stages;  // host copy
    size_t hdr_off = 0, instr_off = 0, pool_off = 0;
};

// A state vector on the device: 2^nq internal amplitudes of type cx<R>.
struct State {
    int nq = 0;        // internal qubits (>= kMinInternalQubits)
    int prec = GOQU_C128;
    void *amps = nullptr;
    size_t elem = 16;  // bytes per amplitude
    uint64_t size() const { return 1ull << nq; }
};

struct Scratch {
    void *ptr = nullptr;
    size_t bytes = 0;
    int ensure(size_t need);
    ~Scratch();
};

// States batched along gridDim.y: state y starts y * stride amplitudes after the first; a state
// whose act byte is 0 is skipped by every kernel launched with the batch (act == nullptr: all
// run).  blob_stride != 0: state y runs its own program, a copy of the launched one's layout
// y * blob_stride bytes further on (headers, instructions and pool alike).
struct ShotBatch {
    int nshots = 1;
    uint64_t stride = 0;
    const uint8_t *act = nullptr;
    uint64_t blob_stride = 0;
};

// A dense op on 5..9 target qubits (EOP_WIDE) staged on the device for k_wide_mfma (wide.hip):
// dev holds the row bits, the free qubits and Re U | Im U | -Im U.
struct WideOp {
    void *dev = nullptr;
    int k = 0, nfree = 0;
    uint64_t cmask = 0, ncols = 0;
    size_t off_f = 0, off_m = 0;
    void release();
};
int prepare_wide(int nq, const EOp &e, hipStream_t s, WideOp &w);
int launch_wide(void *amps, int prec, const WideOp &w, hipStream_t s, const ShotBatch *b);
int run_wide(State &st, const EOp &e, hipStream_t s);  // prepare + launch + synchronise + release

// An op list planned and uploaded once, launchable any number of times: stage programs between
// wide ops.
struct PreparedOps {
    struct Piece {
        bool wide = false;
        DeviceProgram prog;
        WideOp w;
    };
    std::vector<Piece> pieces;
    int64_t stages = 0;  // stage-kernel launches per run
    void release();
    PreparedOps() = default;
    PreparedOps(const PreparedOps &) = delete;
    PreparedOps &operator=(const PreparedOps &) = delete;
    PreparedOps(PreparedOps &&o) noexcept : pieces(std::move(o.pieces)), stages(o.stages) { o.pieces.clear(); }
    PreparedOps &operator=(PreparedOps &&o) noexcept {
        if (this != &o) {
            release();
            pieces = std::move(o.pieces);
            stages = o.stages;
            o.pieces.clear();
            o.stages = 0;
        }
        return *this;
    }
    ~PreparedOps() { release(); }
};

// Planned, uploaded programs of a handle's recent Apply calls, keyed by the content of their op
// records (kinds, qubits, matrices): repeating an op list (Evolve of one circuit, sampling runs,
// repeated cost evaluations) launches the stored programs without lowering, planning or upload.
struct ProgramCache {
    struct Entry {
        std::vector<unsigned char> key;
        uint64_t hash = 0;
        uint64_t used = 0;
        PreparedOps prep;
    };
    static constexpr size_t kMaxEntries = 4;
    std::vector<Entry> entries;
    uint64_t clock = 0;
    void clear() { entries.clear(); }
};

// ---- kernel profiling (profile.hip; goqu_profile_*) -------------------------------------------
enum ProfBin {
    PROF_STAGE = 0,        // k_stage / k_stage_big (statevector stages, single or batched)
    PROF_STAGE_DM = 1,     // k_stage_dm
    PROF_STAGE_PAIR = 2,   // k_stage_pair (adjoint reverse sweeps)
    PROF_HAMILTONIAN = 3,  // k_pauli_apply_tile (lambda = H psi)
    PROF_WIDE = 4,         // k_wide_mfma (bytes: FP64 flops of the contraction)
    PROF_QAOA = 5,         // one batched-QAOA launch graph
    PROF_BINS = 6
};
bool profiling() noexcept;
// nullptr when profiling is off; otherwise an event recorded on s
hipEvent_t prof_begin(hipStream_t s);
// records the end on s: when the totals are read, the span's time, one launch and `bytes` go to bin
void prof_end(hipEvent_t start, hipStream_t s, int bin, double bytes);

// ---- launchers (stream-ordered) -------------------------------------------------------------
PlanConfig default_plan_config(int prec, int nq, bool paired);
int upload_program(const Program &P, int prec, hipStream_t s, DeviceProgram &dp);
// Device byte layout of a program: headers | instructions | pool (as the state's real type),
// each part 256-aligned.  Programs serialised with one common layout (sized for the largest) can
// sit at a fixed stride and run in one batched launch (ShotBatch::blob_stride).
struct ProgramLayout {
    size_t hdr_off = 0, instr_off = 0, pool_off = 0, bytes = 0;
};
ProgramLayout program_layout(size_t nh, size_t ni, size_t np, int prec);
// P serialised into out[0 .. L.bytes) (LDS offsets filled in); stages receives the final headers
int serialize_program(const Program &P, int prec, const ProgramLayout &L, unsigned char *out,
                      std::vector<StageHeader> *stages);
// a stage that changes nothing (no instructions, stores in load order): launches skip it
bool stage_is_identity(const StageHeader &H);
int run_program(State &st, const DeviceProgram &dp, hipStream_t s, const ShotBatch *b = nullptr);
// plan (with cfg, or the default configuration) + upload + run
int run_eops(State &st, const std::vector<EOp> &eops, hipStream_t s, const PlanConfig *cfg = nullptr);
// plan + upload only; st gives the qubit count and precision
int prepare_eops(const State &st, const std::vector<EOp> &eops, const PlanConfig &cfg, hipStream_t s,
                 PreparedOps &out);
// launches a prepared list (no synchronisation); b batches shots as above
int launch_prepared(State &st, const PreparedOps &p, hipStream_t s, const ShotBatch *b = nullptr);
// ops applied to psi and lam together; EOP_GRAD ops accumulate Re <lam| D |psi> per slot
int run_pair(State &psi, State &lam, const std::vector<EOp> &eops, int nslots, std::vector<double> &slot_vals,
             hipStream_t s);
// run_pair with the paired program already planned (plan_program(eops, psi.nq,
// default_plan_config(psi.prec, psi.nq, true)))
int run_pair_planned(State &psi, State &lam, const Program &P, int nslots, std::vector<double> &slot_vals,
                     hipStream_t s);

int fill_basis(State &st, uint64_t index, hipStream_t s);  // |index>
// nstates states of view.size() amplitudes from view.amps on, each |0...0>
int fill_basis_batch(const State &view, int nstates, hipStream_t s);
int copy_state(State &dst, const State &src, hipStream_t s);

// Pauli expectations: sums S_t = sum_i conj(psi[i^x]) (-1)^{|i&z_t|} psi[i] (no coefficient,
// no i^ny), all terms sharing one x mask.  out holds 2*nt doubles.
int pauli_sums(State &st, uint64_t x, const uint64_t *z, int nt, double *out, Scratch &scr, hipStream_t s);
// pauli_sums over nstates states at a stride (x-group terms z = dz[0..nt), a device array), with
// no synchronisation: state y's sum for term t goes to dout[y * out_stride + 2 * (out_off + t)]
// (re, im); the partials live in scr, which must stay untouched until the stream has run.
int pauli_sums_batch(const State &view, int nstates, uint64_t stride, uint64_t x, const uint64_t *dz, int nt,
                     double *dout, int out_stride, int out_off, Scratch &scr, hipStream_t s);
// Re/Im of <a|b> in FP64.
int inner_product(const State &a, const State &b, double out[2], Scratch &scr, hipStream_t s);
// sum over amplitudes with bit q == 0 of |a|^2 (q < 0: all)
int prob_zero(State &st, int q, double *out, Scratch &scr, hipStream_t s);
int collapse(State &st, int q, int outcome, double scale, hipStream_t s);
int sample_state(State &st, uint64_t nvalid, const double *u, int64_t shots, uint64_t *out, Scratch &scr,
                 hipStream_t s);
int sample_probs(const double *dprobs, uint64_t n, const double *u, int64_t shots, uint64_t *out, Scratch &scr,
                 hipStream_t s);
// sample_state through the device's static scratch array (no allocation); dev = the state's device
constexpr size_t kStaticScratchBytes = size_t(8) << 20;
constexpr int kMaxDevices = 64;
int sample_state_static(State &st, int dev, uint64_t nvalid, const double *u, int64_t shots, uint64_t *out,
                        hipStream_t s);
// lam += coeff * P psi over a group of terms sharing x: lam[i] += sum_t c_t s_t(i^x) psi[i^x]
int apply_pauli_group(const State &psi, State &lam, uint64_t x, const uint64_t *z, const double *c, int nt,
                      Scratch &scr, hipStream_t s);
// lam = sum_t c_t P_t psi (c_t with i^ny folded in) in passes over 2^12-amplitude tiles; done = false
// (lam untouched) when the state has fewer than 12 qubits or an x-group spans too many qubits
int apply_hamiltonian_tiled(const State &psi, State &lam, int nt, const uint64_t *x, const uint64_t *z,
                            const double *cfold, Scratch &scr, hipStream_t s, bool &done);
// <lam| (dM on targets, under cmask) |psi>, dM dense k-qubit (k<=4) in gate order
int sandwich(const State &lam, const State &psi, const EOp &dop, double out[2], Scratch &scr, hipStream_t s);
// the same for up to kMaxSandwich matrices on the same qubits in one pass: out[2j], out[2j+1]
constexpr int kMaxSandwich = 4;
int sandwich_multi(const State &lam, const State &psi, const std::vector<EOp> &ms, double *out, Scratch &scr,
                   hipStream_t s);
int scale_state(State &st, double re, double im, hipStream_t s);
int axpy_state(State &y, const State &x, double re, double im, hipStream_t s);  // y += c x

// density-matrix helpers (rho stored as a 2n-qubit vector, row bits high)
int dm_diag(const State &rho, int n, double *dprobs_out, hipStream_t s);
int dm_readout(double *dprobs, double *tmp, int n, int q, double p01, double p10, hipStream_t s);
int dm_pauli(const State &rho, int n, uint64_t x, uint64_t z, double out[2], Scratch &scr, hipStream_t s);
// batched over nstates density matrices at a stride, no synchronisation: (re, im) of state y at
// dout[y * out_stride + 2 * out_off]
int dm_pauli_batch(const State &view, int n, int nstates, uint64_t stride, uint64_t x, uint64_t z, double *dout,
                   int out_stride, int out_off, Scratch &scr, hipStream_t s);
// diagonals of nstates density matrices: dprobs[y * 2^n + i]
int dm_diag_batch(const State &view, int n, int nstates, uint64_t stride, double *dprobs, hipStream_t s);
int dm_purity(const State &rho, int n, double *out, Scratch &scr, hipStream_t s);
int dm_fidelity(const State &rho, int n, const void *dpsi, double *out, Scratch &scr, hipStream_t s);

}  // namespace goqu

// Opaque handle types of the C ABI.
struct goqu_sv {
    goqu::State st;
    int nq_user = 0;
    int device = 0;  // the device current when the handle was created
    bool plan_set = false;  // goqu_sv_set_plan: tile shape of this handle's Apply programs
    goqu::PlanConfig plan;
    goqu::State adj_lam, adj_work;  // adjoint work states, allocated on first use
    hipStream_t stream = nullptr;
    goqu::Scratch scr;
    goqu::ProgramCache cache;  // programs of recent Apply calls (cleared by goqu_sv_set_plan)
};

struct goqu_dm {
    goqu::State st;  // 2n-qubit vector, complex128
    int n = 0;
    hipStream_t stream = nullptr;
    goqu::Scratch scr;
    double *dprobs = nullptr, *dtmp = nullptr;
};
