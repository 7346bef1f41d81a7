// dynamic.hip — dynamic circuits (mid-circuit measurement, reset, feed-forward and structured
// control flow) with the shots batched on the device.
//
// The reference simulates shot after shot: reset the state, walk the op list, draw one uniform per
// measurement or reset, collapse, branch on the classical bits (sim/statevector/dynamic.go:14-230,
// sim/densitymatrix/dynamic.go:13-255).  Here B shots sit side by side in device memory (shot y at
// y * stride amplitudes) and walk one flattened program together: every kernel runs over all B
// shots along gridDim.y, and a shot that a branch does not take is masked out by one byte per shot
// and nesting level (act), which every kernel tests before touching that shot's state.  The host
// plans each gate run once, then issues launches only: measurement outcomes, clbits, branch masks
// and uniform draws stay on the device, and the host synchronises once per batch (and once per
// while-loop iteration, to learn whether any shot still loops).
//
// Shot s consumes its own slab of uniforms u[s * draws_per_shot + k] in order, as the
// reference's per-shot generator does (dynamic.go:26: a fresh PCG per shot).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

#include "runtime.h"

namespace goqu {

namespace {

constexpr int kDynThreads = 256;
constexpr int kMaxWhileIterations = 1000;  // ir.MaxControlFlowIterations (circuit/ir/circuit.go:129)

__device__ __forceinline__ uint64_t ins0(uint64_t v, int pos) {
    const uint64_t lo = v & ((1ull << pos) - 1);
    return ((v >> pos) << (pos + 1)) | lo;
}

__device__ __forceinline__ double block_total(double v, double *sh) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += sh[w];
    return t;
}

__global__ void k_dyn_fill(uint8_t *__restrict__ p, uint64_t n, uint8_t v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_dyn_zero_int(int *__restrict__ p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// Statevector P(0) of qubit q (measureQubit, dynamic.go:186-193): per shot and block, the sum of
// |a_i|^2 over the i with bit q clear, partials in part[shot * nb + block].
template <typename R>
__global__ void __launch_bounds__(kDynThreads) k_dyn_prob0(const cx<R> *__restrict__ a, uint64_t stride, uint64_t nhalf,
                                                           int q, const uint8_t *__restrict__ act,
                                                           double *__restrict__ part) {
    __shared__ double sh[kDynThreads / 64];
    const uint32_t shot = blockIdx.y;
    if (act && !act[shot]) return;
    a += shot * stride;
    double v = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nhalf;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const cx<R> p = a[ins0(i, q)];
        v += (double)p.x * p.x + (double)p.y * p.y;
    }
    const double t = block_total(v, sh);
    if (threadIdx.x == 0) part[(size_t)shot * gridDim.x + blockIdx.x] = t;
}

// Density-matrix P(0) (measureQubitDM, dynamic.go:211-219): the real diagonal rho[i][i] over the
// i with bit q clear; rho[r][c] sits at (r << n) | c.
__global__ void __launch_bounds__(kDynThreads) k_dyn_prob0_dm(const cx<double> *__restrict__ rho, uint64_t stride,
                                                              int n, int q, const uint8_t *__restrict__ act,
                                                              double *__restrict__ part) {
    __shared__ double sh[kDynThreads / 64];
    const uint32_t shot = blockIdx.y;
    if (act && !act[shot]) return;
    rho += shot * stride;
    const uint64_t nhalf = 1ull << (n - 1);
    double v = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nhalf;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t r = ins0(i, q);
        v += rho[(r << n) | r].x;
    }
    const double t = block_total(v, sh);
    if (threadIdx.x == 0) part[(size_t)shot * gridDim.x + blockIdx.x] = t;
}

// One thread per shot: P0 = the block partials added in order, outcome = (u >= P0)
// (dynamic.go:195-198), the collapse scale (statevector 1/sqrt(p), density matrix 1/p; 0 marks the
// reference's degenerate case, which leaves the state as it is), the clbit.
__global__ void k_dyn_decide(const double *__restrict__ part, int nb, const uint8_t *__restrict__ act, int nshots,
                             const double *__restrict__ u, int draws_per_shot, int *__restrict__ drawn,
                             uint8_t *__restrict__ clbits, int nclbits, int clbit, int *__restrict__ outcome,
                             double *__restrict__ scale, int *__restrict__ err, int density) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshots || (act && !act[s])) return;
    double p0 = 0;
    for (int b = 0; b < nb; b++) p0 += part[(size_t)s * nb + b];
    const int k = drawn[s];
    double r = 0;
    if (k < draws_per_shot) {
        r = u[(size_t)s * draws_per_shot + k];
        drawn[s] = k + 1;
    } else {
        err[0] = 1;  // the caller's slab is too short; the host reports it
    }
    const int o = r >= p0 ? 1 : 0;
    double sc;
    if (density) {
        const double p = o ? 1.0 - p0 : p0;  // dynamic.go (DM):229-235
        sc = p < 1e-15 ? 0.0 : 1.0 / p;
    } else {
        const double nrm = o ? sqrt(1.0 - p0) : sqrt(p0);  // dynamic.go:201-211
        sc = nrm < 1e-15 ? 0.0 : 1.0 / nrm;
    }
    outcome[s] = o;
    scale[s] = sc;
    if (clbit >= 0) clbits[(size_t)s * nclbits + clbit] = (uint8_t)o;
}

// Statevector collapse (dynamic.go:213-229); reset = measure then X on outcome 1 (dynamic.go:79-85),
// folded in: the kept half moves to |0>.
template <typename R>
__global__ void k_dyn_collapse(cx<R> *__restrict__ a, uint64_t stride, uint64_t npairs, int q,
                               const uint8_t *__restrict__ act, const int *__restrict__ outcome,
                               const double *__restrict__ scale, int reset) {
    const uint32_t shot = blockIdx.y;
    if (act && !act[shot]) return;
    a += shot * stride;
    const int o = outcome[shot];
    const double sc = scale[shot];
    if (sc == 0.0 && !(reset && o)) return;  // degenerate: untouched (and no X)
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < npairs;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i0 = ins0(i, q), i1 = i0 | (1ull << q);
        const cx<R> v0 = a[i0], v1 = a[i1];
        if (sc == 0.0) {  // degenerate outcome 1 of a reset: X alone
            a[i0] = v1;
            a[i1] = v0;
            continue;
        }
        const cx<R> keep = o ? v1 : v0;
        const cx<R> k = cx<R>{(R)(keep.x * sc), (R)(keep.y * sc)}, z = cx<R>{R(0), R(0)};
        if (reset || !o) {
            a[i0] = k;
            a[i1] = z;
        } else {
            a[i0] = z;
            a[i1] = k;
        }
    }
}

// Density-matrix collapse (dynamic.go (DM):237-252): rho[r][c] survives, scaled by 1/p, when bit q
// of both r and c equals the outcome; a reset then applies X on both sides, moving that block to
// r, c with bit q clear.  Groups of four elements: (row bit, column bit) of q.
__global__ void k_dyn_collapse_dm(cx<double> *__restrict__ rho, uint64_t stride, int n, int q,
                                  const uint8_t *__restrict__ act, const int *__restrict__ outcome,
                                  const double *__restrict__ scale, int reset) {
    const uint32_t shot = blockIdx.y;
    if (act && !act[shot]) return;
    rho += shot * stride;
    const int o = outcome[shot];
    const double sc = scale[shot];
    if (sc == 0.0 && !(reset && o)) return;
    const uint64_t ngroups = 1ull << (2 * n - 2);
    const uint64_t cb = 1ull << q, rb = 1ull << (n + q);
    const cx<double> z = {0, 0};
    for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < ngroups;
         g += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t b = ins0(ins0(g, q), n + q);
        const uint64_t e00 = b, e01 = b | cb, e10 = b | rb, e11 = b | rb | cb;
        if (sc == 0.0) {  // degenerate outcome 1 of a reset: X rho X
            const cx<double> a00 = rho[e00], a01 = rho[e01], a10 = rho[e10], a11 = rho[e11];
            rho[e00] = a11;
            rho[e11] = a00;
            rho[e01] = a10;
            rho[e10] = a01;
            continue;
        }
        const cx<double> keep = rho[o ? e11 : e00];
        const cx<double> k = {keep.x * sc, keep.y * sc};
        rho[e01] = z;
        rho[e10] = z;
        if (reset || !o) {
            rho[e00] = k;
            rho[e11] = z;
        } else {
            rho[e00] = z;
            rho[e11] = k;
        }
    }
}

__device__ __forceinline__ int read_value(const uint8_t *__restrict__ cl, const int *__restrict__ bits, int nbits) {
    int v = 0;  // readClassicalValue (dynamic.go:150-157): little-endian over the listed clbits
    for (int i = 0; i < nbits; i++) v |= (int)cl[bits[i]] << i;
    return v;
}

// A case of a branch group: act = parent & !taken & (value of the condition bits == value); the
// first matching case wins (executeControlFlow, dynamic.go:117-146).  mode 1 = the default body
// (no value test), mode 2 = opening the group (taken = 0, act = 0).
__global__ void k_dyn_case(const uint8_t *__restrict__ parent, uint8_t *__restrict__ act, uint8_t *__restrict__ taken,
                           int nshots, const uint8_t *__restrict__ clbits, int nclbits, const int *__restrict__ bits,
                           int nbits, int value, int mode) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshots) return;
    if (mode == 2) {
        taken[s] = 0;
        act[s] = 0;
        return;
    }
    const bool p = parent ? parent[s] != 0 : true;
    bool c = p && !taken[s];
    if (mode == 0) c = c && read_value(clbits + (size_t)s * nclbits, bits, nbits) == value;
    act[s] = c ? 1 : 0;
    if (c) taken[s] = 1;
}

// While-loop head: act &= (condition holds); any[0] = 1 when some shot still loops.  mode 1 opens
// the loop (act = parent first).
__global__ void k_dyn_while(const uint8_t *__restrict__ parent, uint8_t *__restrict__ act, int nshots,
                            const uint8_t *__restrict__ clbits, int nclbits, const int *__restrict__ bits, int nbits,
                            int value, int *__restrict__ any, int mode) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshots) return;
    bool a = mode == 1 ? (parent ? parent[s] != 0 : true) : act[s] != 0;
    a = a && read_value(clbits + (size_t)s * nclbits, bits, nbits) == value;
    act[s] = a ? 1 : 0;
    if (a) any[0] = 1;
}

unsigned blocks_for(uint64_t n, unsigned per) { return (unsigned)std::max<uint64_t>(1, (n + per - 1) / per); }

// The flattened program, checked, with the matching END of every BRANCH / WHILE / FOR.
struct DynProgram {
    std::vector<goqu_dyn_instr_t> ins;
    std::vector<int> end;
    int depth = 0;  // deepest nesting of BRANCH / WHILE levels
};

int check_program(const goqu_dyn_instr_t *prog, int ninstr, int num_ops, const int32_t *cond_bits, int ncond,
                  int nq_user, int nclbits, DynProgram &P) {
    P.ins.assign(prog, prog + ninstr);
    P.end.assign(ninstr, -1);
    struct Open {
        int at;
        bool seen_default;
    };
    std::vector<Open> stack;
    int levels = 0, depth = 0;  // open BRANCH / WHILE groups (FOR adds no level)
    auto bad = [&](int i, const char *what) {
        return fail(GOQU_ERR_INVALID_ARG, "dynamic program instruction " + std::to_string(i) + ": " + what);
    };
    auto cond_ok = [&](const goqu_dyn_instr_t &I) {
        if (I.a < 0 || I.b < 1 || I.b > 30 || I.a > ncond - I.b) return false;
        for (int j = 0; j < I.b; j++)
            if (cond_bits[I.a + j] < 0 || cond_bits[I.a + j] >= nclbits) return false;
        return true;
    };
    for (int i = 0; i < ninstr; i++) {
        const goqu_dyn_instr_t &I = P.ins[i];
        if (i > 0 && P.ins[i - 1].kind == GOQU_DYN_BRANCH && I.kind != GOQU_DYN_CASE &&
            I.kind != GOQU_DYN_DEFAULT && I.kind != GOQU_DYN_END)
            return bad(i, "a branch group opens with CASE or DEFAULT");
        switch (I.kind) {
        case GOQU_DYN_GATES:
            if (I.a < 0 || I.b < 0 || I.a > num_ops - I.b) return bad(i, "op range outside the op array");
            break;
        case GOQU_DYN_MEASURE:
            if (I.b < 0 || I.b >= nclbits) return bad(i, "clbit out of range");
            [[fallthrough]];
        case GOQU_DYN_RESET:
            if (I.a < 0 || I.a >= nq_user) return bad(i, "qubit out of range");
            break;
        case GOQU_DYN_BRANCH:
            stack.push_back({i, false});
            depth = std::max(depth, ++levels);
            break;
        case GOQU_DYN_CASE:
        case GOQU_DYN_DEFAULT:
            if (stack.empty() || P.ins[stack.back().at].kind != GOQU_DYN_BRANCH)
                return bad(i, "CASE / DEFAULT outside a branch group");
            if (stack.back().seen_default) return bad(i, "a body after DEFAULT");
            if (I.kind == GOQU_DYN_CASE && !cond_ok(I)) return bad(i, "condition bits out of range");
            if (I.kind == GOQU_DYN_DEFAULT) stack.back().seen_default = true;
            break;
        case GOQU_DYN_WHILE:
            if (!cond_ok(I)) return bad(i, "condition bits out of range");
            stack.push_back({i, false});
            depth = std::max(depth, ++levels);
            break;
        case GOQU_DYN_FOR:
            if (I.a < 0) return bad(i, "negative iteration count");
            stack.push_back({i, false});
            break;
        case GOQU_DYN_END:
            if (stack.empty()) return bad(i, "END without an open BRANCH, WHILE or FOR");
            P.end[stack.back().at] = i;
            if (P.ins[stack.back().at].kind != GOQU_DYN_FOR) levels--;
            stack.pop_back();
            break;
        default:
            return bad(i, "unknown instruction kind");
        }
    }
    if (!stack.empty()) return bad(stack.back().at, "not closed by END");
    P.depth = depth;
    return GOQU_OK;
}

// Device buffers of one batch of shots (sized for the largest batch of the call).
struct DynBuffers {
    CachedDevBuf amps_buf{CACHE_DYN_STATES};  // the shots' states: retained for the next call
    void *amps = nullptr;
    uint8_t *act = nullptr, *taken = nullptr, *clbits = nullptr;
    int *drawn = nullptr, *outcome = nullptr, *flags = nullptr, *bits = nullptr;
    double *u = nullptr, *part = nullptr, *scale = nullptr;
    ~DynBuffers() {
        for (void *p : {(void *)act, (void *)taken, (void *)clbits, (void *)drawn, (void *)outcome, (void *)flags,
                        (void *)bits, (void *)u, (void *)part, (void *)scale})
            if (p) (void)hipFree(p);
    }
};

struct DynRun {
    bool density = false;
    int n = 0;  // user qubits
    State view;  // shot 0's state; the batch follows at view.size() amplitude steps
    int nshots = 0;
    int nclbits = 0, draws = 0, nb = 1;
    hipStream_t s = nullptr;
    const DynProgram *P = nullptr;
    std::vector<PreparedOps> *prep = nullptr;
    DynBuffers *buf = nullptr;

    uint8_t *act_at(int level) const { return level == 0 ? nullptr : buf->act + (size_t)level * nshots; }
    uint8_t *taken_at(int level) const { return buf->taken + (size_t)level * nshots; }

    int measure(int q, int clbit, int reset, int level) {
        const uint8_t *act = act_at(level);
        const unsigned B = (unsigned)nshots;
        const uint64_t stride = view.size();
        if (density) {
            hipLaunchKernelGGL(k_dyn_prob0_dm, dim3(nb, B), dim3(kDynThreads), 0, s,
                               static_cast<const cx<double> *>(view.amps), stride, n, q, act, buf->part);
        } else if (view.prec == GOQU_C128) {
            hipLaunchKernelGGL(k_dyn_prob0<double>, dim3(nb, B), dim3(kDynThreads), 0, s,
                               static_cast<const cx<double> *>(view.amps), stride, stride / 2, q, act, buf->part);
        } else {
            hipLaunchKernelGGL(k_dyn_prob0<float>, dim3(nb, B), dim3(kDynThreads), 0, s,
                               static_cast<const cx<float> *>(view.amps), stride, stride / 2, q, act, buf->part);
        }
        GOQU_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_dyn_decide, dim3(blocks_for(B, 256)), dim3(256), 0, s, buf->part, nb, act, nshots, buf->u,
                           draws, buf->drawn, buf->clbits, nclbits, clbit, buf->outcome, buf->scale, buf->flags,
                           density ? 1 : 0);
        GOQU_HIP(hipGetLastError());
        const unsigned cgrid = blocks_for(density ? (1ull << (2 * n - 2)) : stride / 2, kDynThreads * 4);
        const unsigned cg = std::min(cgrid, 1024u);
        if (density)
            hipLaunchKernelGGL(k_dyn_collapse_dm, dim3(cg, B), dim3(kDynThreads), 0, s,
                               static_cast<cx<double> *>(view.amps), stride, n, q, act, buf->outcome, buf->scale,
                               reset);
        else if (view.prec == GOQU_C128)
            hipLaunchKernelGGL(k_dyn_collapse<double>, dim3(cg, B), dim3(kDynThreads), 0, s,
                               static_cast<cx<double> *>(view.amps), stride, stride / 2, q, act, buf->outcome,
                               buf->scale, reset);
        else
            hipLaunchKernelGGL(k_dyn_collapse<float>, dim3(cg, B), dim3(kDynThreads), 0, s,
                               static_cast<cx<float> *>(view.amps), stride, stride / 2, q, act, buf->outcome,
                               buf->scale, reset);
        GOQU_HIP(hipGetLastError());
        return GOQU_OK;
    }

    int exec(int i0, int i1, int level) {
        const unsigned sb = blocks_for((uint64_t)nshots, 256);
        for (int i = i0; i < i1; i++) {
            const goqu_dyn_instr_t &I = P->ins[i];
            switch (I.kind) {
            case GOQU_DYN_GATES: {
                ShotBatch b;
                b.nshots = nshots;
                b.stride = view.size();
                b.act = act_at(level);
                int rc = launch_prepared(view, (*prep)[i], s, &b);
                if (rc) return rc;
                break;
            }
            case GOQU_DYN_MEASURE: {
                int rc = measure(I.a, I.b, 0, level);
                if (rc) return rc;
                break;
            }
            case GOQU_DYN_RESET: {
                int rc = measure(I.a, -1, 1, level);
                if (rc) return rc;
                break;
            }
            case GOQU_DYN_BRANCH: {
                const int e = P->end[i], L = level + 1;
                hipLaunchKernelGGL(k_dyn_case, dim3(sb), dim3(256), 0, s, act_at(level), act_at(L), taken_at(L),
                                   nshots, buf->clbits, nclbits, buf->bits, 0, 0, 2);
                GOQU_HIP(hipGetLastError());
                int j = i + 1;
                while (j < e) {  // the bodies: CASE / DEFAULT up to the next CASE / DEFAULT at this depth
                    const goqu_dyn_instr_t &C = P->ins[j];
                    int k = j + 1;
                    while (k < e && P->ins[k].kind != GOQU_DYN_CASE && P->ins[k].kind != GOQU_DYN_DEFAULT)
                        k = (P->end[k] >= 0 ? P->end[k] : k) + 1;
                    hipLaunchKernelGGL(k_dyn_case, dim3(sb), dim3(256), 0, s, act_at(level), act_at(L), taken_at(L),
                                       nshots, buf->clbits, nclbits, buf->bits + (C.kind == GOQU_DYN_CASE ? C.a : 0),
                                       C.kind == GOQU_DYN_CASE ? C.b : 0, C.value,
                                       C.kind == GOQU_DYN_CASE ? 0 : 1);
                    GOQU_HIP(hipGetLastError());
                    int rc = exec(j + 1, k, L);
                    if (rc) return rc;
                    j = k;
                }
                i = e;
                break;
            }
            case GOQU_DYN_WHILE: {
                const int e = P->end[i], L = level + 1;
                int *any = buf->flags + 1;
                for (int it = 0; it < kMaxWhileIterations; it++) {
                    hipLaunchKernelGGL(k_dyn_zero_int, dim3(1), dim3(64), 0, s, any, 1);
                    hipLaunchKernelGGL(k_dyn_while, dim3(sb), dim3(256), 0, s, act_at(level), act_at(L), nshots,
                                       buf->clbits, nclbits, buf->bits + I.a, I.b, I.value, any, it == 0 ? 1 : 0);
                    GOQU_HIP(hipGetLastError());
                    int h_any = 0;
                    GOQU_HIP(hipMemcpyAsync(&h_any, any, sizeof(int), hipMemcpyDeviceToHost, s));
                    GOQU_HIP(hipStreamSynchronize(s));
                    if (!h_any) break;
                    int rc = exec(i + 1, e, L);
                    if (rc) return rc;
                }
                i = e;
                break;
            }
            case GOQU_DYN_FOR: {
                const int e = P->end[i];
                for (int it = 0; it < I.a; it++) {
                    int rc = exec(i + 1, e, level);
                    if (rc) return rc;
                }
                i = e;
                break;
            }
            default:
                break;  // CASE / DEFAULT / END are consumed by their group
            }
        }
        return GOQU_OK;
    }
};

// Shared driver of goqu_sv_run_dynamic and goqu_dm_run_dynamic.  `handle_state` receives the last
// shot's final state, as the reference's simulator keeps it after RunDynamic.
int run_dynamic(bool density, int n, State &handle_state, hipStream_t s, const PlanConfig &cfg,
                const goqu_op_t *ops, int num_ops, const goqu_dyn_instr_t *prog, int ninstr, const int32_t *cond_bits,
                int ncond, int nclbits, int64_t shots, const double *uniforms, int draws_per_shot,
                uint8_t *clbits_out) {
    if ((num_ops > 0 && !ops) || (ninstr > 0 && !prog) || (ncond > 0 && !cond_bits) || shots < 1 ||
        nclbits < 0 || draws_per_shot < 0 || (draws_per_shot > 0 && !uniforms) || (nclbits > 0 && !clbits_out))
        return fail(GOQU_ERR_INVALID_ARG, "bad argument");
    if (ninstr < 0 || num_ops < 0 || ncond < 0) return fail(GOQU_ERR_INVALID_ARG, "negative count");
    host_fault_point();
    DynProgram P;
    int rc = check_program(prog, ninstr, num_ops, cond_bits, ncond, n, nclbits, P);
    if (rc) return rc;

    // plan every gate run once
    std::vector<PreparedOps> prep(ninstr);
    for (int i = 0; i < ninstr; i++) {
        if (P.ins[i].kind != GOQU_DYN_GATES || P.ins[i].b == 0) continue;
        std::vector<EOp> eops;
        std::string err;
        rc = density ? lower_dm_ops(ops + P.ins[i].a, P.ins[i].b, n, eops, err)
                     : lower_sv_ops(ops + P.ins[i].a, P.ins[i].b, n, eops, err);
        if (rc) return fail(rc, "dynamic program instruction " + std::to_string(i) + ": " + err);
        rc = prepare_eops(handle_state, eops, cfg, s, prep[i]);
        if (rc) return rc;
    }

    // batch size: what a share of free device memory holds, within the grid's y range
    const uint64_t stride = handle_state.size();
    const size_t shot_bytes = stride * handle_state.elem + (size_t)draws_per_shot * 8 + nclbits + 64 +
                              2 * (size_t)(P.depth + 1) + 1024;
    size_t free_b = 0, total_b = 0;
    GOQU_HIP(hipMemGetInfo(&free_b, &total_b));
    int ymax = 65535;
    {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&v, hipDeviceAttributeMaxGridDimY, dev) ==
                                                     hipSuccess && v > 0)
            ymax = std::min(v, 65535);
    }
    const uint64_t fit = (uint64_t)(free_b / 2) / shot_bytes;
    if (fit < 1) return fail(GOQU_ERR_OUT_OF_MEMORY, "dynamic circuit: one shot's state does not fit");
    uint64_t cap = (uint64_t)ymax;
    if (const char *e = std::getenv("GOQU_DYN_MAX_BATCH")) {  // smaller batches (tests of the chunking)
        const long v = std::atol(e);
        if (v >= 1) cap = std::min<uint64_t>(cap, (uint64_t)v);
    }
    const int B = (int)std::min<uint64_t>({(uint64_t)shots, fit, cap});
    const int nb = (int)std::min<uint64_t>(64, std::max<uint64_t>(1, (density ? (1ull << (n - 1)) : stride / 2) /
                                                                         (kDynThreads * 8)));

    DynBuffers buf;
    const int levels = P.depth + 1;
    GOQU_HIP(buf.amps_buf.take((size_t)B * stride * handle_state.elem));
    buf.amps = buf.amps_buf.p;
    GOQU_HIP(dmalloc((void **)&buf.act, (size_t)levels * B));
    GOQU_HIP(dmalloc((void **)&buf.taken, (size_t)levels * B));
    GOQU_HIP(dmalloc((void **)&buf.clbits, (size_t)B * std::max(nclbits, 1)));
    GOQU_HIP(dmalloc((void **)&buf.drawn, (size_t)B * sizeof(int)));
    GOQU_HIP(dmalloc((void **)&buf.outcome, (size_t)B * sizeof(int)));
    GOQU_HIP(dmalloc((void **)&buf.flags, 4 * sizeof(int)));
    GOQU_HIP(dmalloc((void **)&buf.bits, (size_t)std::max(ncond, 1) * sizeof(int)));
    GOQU_HIP(dmalloc((void **)&buf.u, (size_t)B * std::max(draws_per_shot, 1) * sizeof(double)));
    GOQU_HIP(dmalloc((void **)&buf.part, (size_t)B * nb * sizeof(double)));
    GOQU_HIP(dmalloc((void **)&buf.scale, (size_t)B * sizeof(double)));
    if (ncond) GOQU_HIP(hipMemcpyAsync(buf.bits, cond_bits, (size_t)ncond * sizeof(int), hipMemcpyHostToDevice, s));

    DynRun R;
    R.density = density;
    R.n = n;
    R.view = handle_state;
    R.view.amps = buf.amps;
    R.nclbits = std::max(nclbits, 1);
    R.draws = draws_per_shot;
    R.nb = nb;
    R.s = s;
    R.P = &P;
    R.prep = &prep;
    R.buf = &buf;
    std::vector<uint8_t> host_cl;
    for (int64_t s0 = 0; s0 < shots; s0 += B) {
        const int nbatch = (int)std::min<int64_t>(B, shots - s0);
        R.nshots = nbatch;
        rc = fill_basis_batch(R.view, nbatch, s);
        if (rc) return rc;
        hipLaunchKernelGGL(k_dyn_fill, dim3(blocks_for((uint64_t)nbatch * R.nclbits, 256)), dim3(256), 0, s,
                           buf.clbits, (uint64_t)nbatch * R.nclbits, (uint8_t)0);
        hipLaunchKernelGGL(k_dyn_zero_int, dim3(blocks_for(nbatch, 256)), dim3(256), 0, s, buf.drawn, nbatch);
        hipLaunchKernelGGL(k_dyn_zero_int, dim3(1), dim3(64), 0, s, buf.flags, 4);
        GOQU_HIP(hipGetLastError());
        if (draws_per_shot)
            GOQU_HIP(hipMemcpyAsync(buf.u, uniforms + (size_t)s0 * draws_per_shot,
                                    (size_t)nbatch * draws_per_shot * sizeof(double), hipMemcpyHostToDevice, s));
        rc = R.exec(0, ninstr, 0);
        if (rc) {
            (void)hipStreamSynchronize(s);
            return rc;
        }
        int err = 0;
        GOQU_HIP(hipMemcpyAsync(&err, buf.flags, sizeof(int), hipMemcpyDeviceToHost, s));
        if (nclbits) {
            host_cl.resize((size_t)nbatch * nclbits);
            GOQU_HIP(hipMemcpyAsync(host_cl.data(), buf.clbits, host_cl.size(), hipMemcpyDeviceToHost, s));
        }
        GOQU_HIP(hipStreamSynchronize(s));
        if (err)
            return fail(GOQU_ERR_INVALID_ARG, "dynamic circuit: a shot needed more than draws_per_shot uniforms");
        if (nclbits) std::memcpy(clbits_out + (size_t)s0 * nclbits, host_cl.data(), host_cl.size());
        if (s0 + nbatch >= shots) {  // the simulator keeps the last shot's state (RunDynamic leaves it)
            GOQU_HIP(hipMemcpyAsync(handle_state.amps,
                                    static_cast<char *>(buf.amps) + (size_t)(nbatch - 1) * stride * handle_state.elem,
                                    stride * handle_state.elem, hipMemcpyDeviceToDevice, s));
            GOQU_HIP(hipStreamSynchronize(s));
        }
    }
    buf.amps_buf.idle();
    return GOQU_OK;
}

}  // namespace

}  // namespace goqu

using namespace goqu;

extern "C" {

int goqu_sv_run_dynamic(goqu_sv_t *h, const goqu_op_t *ops, int32_t num_ops, const goqu_dyn_instr_t *prog,
                        int32_t num_instr, const int32_t *cond_bits, int32_t num_cond_bits, int32_t num_clbits,
                        int64_t shots, const double *uniforms, int32_t draws_per_shot, uint8_t *clbits_out) {
    GOQU_GUARD_BEGIN
    if (!h) return fail(GOQU_ERR_INVALID_ARG, "null handle");
    const PlanConfig cfg = h->plan_set ? h->plan : default_plan_config(h->st.prec, h->st.nq, false);
    return run_dynamic(false, h->nq_user, h->st, h->stream, cfg, ops, num_ops, prog, num_instr, cond_bits,
                       num_cond_bits, num_clbits, shots, uniforms, draws_per_shot, clbits_out);
    GOQU_GUARD_END
}

int goqu_dm_run_dynamic(goqu_dm_t *h, const goqu_op_t *ops, int32_t num_ops, const goqu_dyn_instr_t *prog,
                        int32_t num_instr, const int32_t *cond_bits, int32_t num_cond_bits, int32_t num_clbits,
                        int64_t shots, const double *uniforms, int32_t draws_per_shot, uint8_t *clbits_out) {
    GOQU_GUARD_BEGIN
    if (!h) return fail(GOQU_ERR_INVALID_ARG, "null handle");
    const PlanConfig cfg = default_plan_config(GOQU_C128, h->st.nq, false);
    return run_dynamic(true, h->n, h->st, h->stream, cfg, ops, num_ops, prog, num_instr, cond_bits, num_cond_bits,
                       num_clbits, shots, uniforms, draws_per_shot, clbits_out);
    GOQU_GUARD_END
}

}  // extern "C"
