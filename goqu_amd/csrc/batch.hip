// batch.hip — many circuits on the same register size in one pass: the points of a parameter
// sweep (sweep.RunSim / RunDensitySim, sweep/run.go:27-171), the circuits a mitigation executor
// evaluates (algorithm/mitigation/executor.go:20-41), VQE parameter batches.
//
// The reference evaluates such a batch point by point, each on a fresh simulator (run.go:75-97),
// spread over GOMAXPROCS goroutines.  Here every circuit is lowered and planned on host threads,
// the circuits whose programs launch alike (the same number of stage sweeps, and per sweep the
// same tile size, kernel and skip decision) form a group, and a group runs as one launch per stage
// sweep over all its states along gridDim.y: each state's workgroups read their own program (its
// own instructions, coefficients and layouts) at a fixed byte stride.  Expectations and
// probabilities reduce all states of a launch together, and the host synchronises once per group
// chunk.  A circuit holding a wide op (5..9 targets) runs on its own.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "runtime.h"

namespace goqu {

namespace {

struct Planned {
    Program P;
    std::vector<EOp> eops;  // kept for circuits with wide ops
    bool wide = false;
    int rc = GOQU_OK;
    std::string err;
};

// what a launch of the group depends on: per stage sweep, the tile size, the kernel and whether
// the sweep is skipped
std::vector<int32_t> launch_signature(const Program &P) {
    std::vector<int32_t> sig;
    sig.reserve(P.stages.size());
    for (const StageHeader &H : P.stages)
        sig.push_back(H.m | ((H.flags & (SF_BIG | SF_DM)) << 8) | (stage_is_identity(H) ? 1 << 16 : 0));
    return sig;
}

void plan_range(bool density, int n, int nq_internal, const PlanConfig &cfg, const goqu_op_t *ops,
                const int64_t *off, int i0, int i1, std::vector<Planned> &out) {
    for (int i = i0; i < i1; i++) {
        Planned &p = out[i];
        try {
            std::vector<EOp> eops;
            const int nops = (int)(off[i + 1] - off[i]);
            p.rc = density ? lower_dm_ops(ops + off[i], nops, n, eops, p.err)
                           : lower_sv_ops(ops + off[i], nops, n, eops, p.err);
            if (p.rc) continue;
            for (const EOp &e : eops) p.wide = p.wide || e.kind == EOP_WIDE;
            if (p.wide)
                p.eops = std::move(eops);
            else
                p.P = plan_program(eops, nq_internal, cfg);
        } catch (const std::bad_alloc &) {
            p.rc = GOQU_ERR_OUT_OF_MEMORY;
            p.err = "host memory exhausted while planning";
        } catch (...) {
            p.rc = GOQU_ERR_DEVICE;
            p.err = "planning failed";
        }
    }
}

struct BatchJob {
    bool density = false;
    int n = 0;          // user qubits
    int prec = GOQU_C128;
    int nq = 0;         // internal qubits of one state
    PlanConfig cfg;
    // expectation terms (x-groups in ascending x, terms in input order inside a group)
    int nt = 0;
    const uint64_t *x = nullptr, *z = nullptr;
    const int32_t *ny = nullptr;
    const double *coeffs = nullptr;
    double *expect_out = nullptr;  // [instance][2 * nt]
    // sampling
    int64_t shots = 0;
    const double *uniforms = nullptr;  // [instance][shots]
    const double *p01 = nullptr, *p10 = nullptr;
    uint64_t *sample_out = nullptr;    // [instance][shots]
};

void ipow_host(int n, double &re, double &im) {
    switch (((n % 4) + 4) % 4) {
    case 0: re = 1, im = 0; break;
    case 1: re = 0, im = 1; break;
    case 2: re = -1, im = 0; break;
    default: re = 0, im = -1; break;
    }
}

class Batcher {
  public:
    Batcher(const BatchJob &job, hipStream_t s) : J(job), s(s) {}
    ~Batcher() {
        for (void *p : {(void *)dsums, (void *)dz, (void *)dprobs, (void *)dtmp})
            if (p) (void)hipFree(p);
    }

    // called once, before any launch: the states and the program blobs come from the retained
    // buffers of earlier batches when those are large enough
    int prepare(int max_states, size_t blob_bytes) {
        const size_t sb = (size_t(1) << J.nq) * (J.prec == GOQU_C128 ? 16 : 8);
        cap_states = max_states;
        cap_blob = blob_bytes * max_states;
        GOQU_HIP(cstates.take(sb * cap_states));
        GOQU_HIP(cblobs.take(cap_blob));
        states = cstates.p;
        blobs = cblobs.p;
        // reduction partials of the batched Pauli kernels, at their largest (no reallocation
        // while kernels are queued)
        int rc = scr.ensure((size_t)std::max(4096, chunk_cap) * 32 * 8 + 4096);
        if (rc) return rc;
        if (J.nt && !dsums) {
            GOQU_HIP(dmalloc((void **)&dsums, (size_t)chunk_cap * 2 * J.nt * sizeof(double)));
            GOQU_HIP(dmalloc((void **)&dz, (size_t)J.nt * sizeof(uint64_t)));
        }
        if (J.shots && J.density && !dprobs) {
            GOQU_HIP(dmalloc((void **)&dprobs, (size_t)chunk_cap << J.n << 3));
            GOQU_HIP(dmalloc((void **)&dtmp, (size_t(1) << J.n) * sizeof(double)));
        }
        return GOQU_OK;
    }

    // z masks in x-group order, uploaded once
    int upload_terms() {
        if (!J.nt) return GOQU_OK;
        for (int t = 0; t < J.nt; t++) groups[J.x[t]].push_back(t);
        std::vector<uint64_t> zs;
        for (auto &g : groups) {
            group_off.push_back((int)zs.size());
            for (int t : g.second) zs.push_back(J.z[t]);
        }
        GOQU_HIP(hipMemcpyAsync(dz, zs.data(), zs.size() * sizeof(uint64_t), hipMemcpyHostToDevice, s));
        GOQU_HIP(hipStreamSynchronize(s));
        return GOQU_OK;
    }

    // observables of `cnt` states at `view` (instances idx[0..cnt)), after their programs ran
    int measure(const State &view, const std::vector<int> &idx) {
        const int cnt = (int)idx.size();
        const uint64_t stride = view.size();
        if (J.nt) {
            int gi = 0;
            for (auto &g : groups) {
                const int off = group_off[gi++], k = (int)g.second.size();
                if (J.density) {
                    for (int j = 0; j < k; j++) {
                        int rc = dm_pauli_batch(view, J.n, cnt, stride, g.first, J.z[g.second[j]], dsums, 2 * J.nt,
                                                off + j, scr, s);
                        if (rc) return rc;
                    }
                } else {
                    int rc = pauli_sums_batch(view, cnt, stride, g.first, dz + off, k, dsums, 2 * J.nt, off, scr, s);
                    if (rc) return rc;
                }
            }
            std::vector<double> host((size_t)cnt * 2 * J.nt);
            GOQU_HIP(hipMemcpyAsync(host.data(), dsums, host.size() * sizeof(double), hipMemcpyDeviceToHost, s));
            GOQU_HIP(hipStreamSynchronize(s));
            for (int j = 0; j < cnt; j++) {
                const double *src = host.data() + (size_t)j * 2 * J.nt;
                double *dst = J.expect_out + (size_t)idx[j] * 2 * J.nt;
                int gi2 = 0;
                for (auto &g : groups) {
                    const int off = group_off[gi2++];
                    for (size_t k = 0; k < g.second.size(); k++) {
                        const int t = g.second[k];
                        double yr, yi;
                        ipow_host(J.ny[t], yr, yi);  // coeff * i^ny * S (expect_sv.go:31-46, expect_dm.go:9-30)
                        const double s0 = src[2 * (off + k)], s1 = src[2 * (off + k) + 1];
                        const double sr = s0 * yr - s1 * yi, si = s0 * yi + s1 * yr;
                        const double cr = J.coeffs[2 * t], ci = J.coeffs[2 * t + 1];
                        dst[2 * t] = cr * sr - ci * si;
                        dst[2 * t + 1] = cr * si + ci * sr;
                    }
                }
            }
        }
        if (J.shots) {
            const uint64_t dim = 1ull << J.n;
            if (J.density) {
                int rc = dm_diag_batch(view, J.n, cnt, stride, dprobs, s);
                if (rc) return rc;
            }
            for (int j = 0; j < cnt; j++) {
                const int i = idx[j];
                const double *u = J.uniforms + (size_t)i * J.shots;
                uint64_t *o = J.sample_out + (size_t)i * J.shots;
                int rc;
                if (J.density) {
                    double *p = dprobs + (size_t)j * dim;
                    if (J.p01 && J.p10)
                        for (int q = 0; q < J.n; q++) {
                            if (J.p01[q] < 0 || J.p10[q] < 0) continue;
                            rc = dm_readout(p, dtmp, J.n, q, J.p01[q], J.p10[q], s);
                            if (rc) return rc;
                        }
                    rc = sample_probs(p, dim, u, J.shots, o, scr, s);
                } else {
                    State v = view;
                    v.amps = static_cast<char *>(view.amps) + (size_t)j * stride * view.elem;
                    rc = sample_state(v, dim, u, J.shots, o, scr, s);
                }
                if (rc) return rc;
            }
        }
        return GOQU_OK;
    }

    const BatchJob &J;
    hipStream_t s;
    Scratch scr;
    CachedDevBuf cstates{CACHE_BATCH_STATES}, cblobs{CACHE_BATCH_BLOBS};
    void *states = nullptr, *blobs = nullptr;
    // every launch of the batch has completed (its stream synchronized)
    void idle() {
        cstates.idle();
        cblobs.idle();
    }
    int cap_states = 0;
    size_t cap_blob = 0;
    int chunk_cap = 0;
    double *dsums = nullptr, *dprobs = nullptr, *dtmp = nullptr;
    uint64_t *dz = nullptr;
    std::map<uint64_t, std::vector<int>> groups;
    std::vector<int> group_off;
};

int run_batch(const BatchJob &J, int ninst, const goqu_op_t *ops, const int64_t *op_offsets) {
    if (ninst < 0 || (ninst > 0 && (!ops || !op_offsets))) return fail(GOQU_ERR_INVALID_ARG, "bad argument");
    for (int i = 0; i < ninst; i++)
        if (op_offsets[i] < 0 || op_offsets[i + 1] < op_offsets[i] || op_offsets[i + 1] - op_offsets[i] > INT32_MAX)
            return fail(GOQU_ERR_INVALID_ARG, "op_offsets must be non-decreasing from 0");
    if (ninst == 0) return GOQU_OK;
    if (op_offsets[0] != 0) return fail(GOQU_ERR_INVALID_ARG, "op_offsets[0] must be 0");
    const uint64_t valid = (1ull << J.n) - 1;
    for (int t = 0; t < J.nt; t++)
        if ((J.x[t] & ~valid) || (J.z[t] & ~valid)) return fail(GOQU_ERR_INVALID_ARG, "Pauli mask beyond the qubit count");
    host_fault_point();

    // plan every circuit on host threads
    std::vector<Planned> plans(ninst);
    const int nthreads = std::max(1, std::min<int>({ninst, (int)std::thread::hardware_concurrency(), 16}));
    {
        std::vector<std::thread> pool;
        const int per = (ninst + nthreads - 1) / nthreads;
        for (int t = 0; t < nthreads; t++) {
            const int i0 = t * per, i1 = std::min(ninst, i0 + per);
            if (i0 >= i1) break;
            pool.emplace_back(plan_range, J.density, J.n, J.nq, std::cref(J.cfg), ops, op_offsets, i0, i1,
                              std::ref(plans));
        }
        for (std::thread &th : pool) th.join();
    }
    for (int i = 0; i < ninst; i++)
        if (plans[i].rc) return fail(plans[i].rc, "circuit " + std::to_string(i) + ": " + plans[i].err);

    // groups of circuits whose programs launch alike; wide circuits alone
    std::map<std::vector<int32_t>, std::vector<int>> groups;
    std::vector<int> singles;
    for (int i = 0; i < ninst; i++) {
        if (plans[i].wide)
            singles.push_back(i);
        else
            groups[launch_signature(plans[i].P)].push_back(i);
    }

    hipStream_t s = nullptr;
    GOQU_HIP(hipStreamCreate(&s));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    } sg{s};
    Batcher B(J, s);
    const size_t sb = (size_t(1) << J.nq) * (J.prec == GOQU_C128 ? 16 : 8);
    size_t free_b = 0, total_b = 0;
    GOQU_HIP(hipMemGetInfo(&free_b, &total_b));
    int ymax = 65535;
    if (const char *e = std::getenv("GOQU_BATCH_MAX_STATES")) {  // smaller chunks (tests of the chunking)
        const long v = std::atol(e);
        if (v >= 1) ymax = (int)std::min<long>(ymax, v);
    }
    // largest blob of any group, for the chunk size
    size_t max_blob = 0;
    for (auto &g : groups) {
        size_t nh = 0, ni = 0, np = 0;
        for (int i : g.second) {
            nh = std::max(nh, plans[i].P.stages.size());
            ni = std::max(ni, plans[i].P.instrs.size());
            np = std::max(np, plans[i].P.pool.size());
        }
        max_blob = std::max(max_blob, program_layout(nh, ni, np, J.prec).bytes);
    }
    const uint64_t per_state = sb + max_blob + 2 * (size_t)J.nt * 8 + (J.density && J.shots ? (8ull << J.n) : 0);
    const uint64_t fit = (uint64_t)(free_b / 2) / std::max<uint64_t>(per_state, 1);
    if (fit < 1) return fail(GOQU_ERR_OUT_OF_MEMORY, "batch: one circuit's state does not fit");
    B.chunk_cap = (int)std::min<uint64_t>({fit, (uint64_t)ymax, (uint64_t)ninst});
    int rc = B.prepare(B.chunk_cap, std::max<size_t>(max_blob, 256));
    if (rc) return rc;
    rc = B.upload_terms();
    if (rc) return rc;

    State view;
    view.nq = J.nq;
    view.prec = J.prec;
    view.elem = J.prec == GOQU_C128 ? 16 : 8;
    view.amps = B.states;
    std::vector<unsigned char> host;
    for (auto &g : groups) {
        const std::vector<int> &mem = g.second;
        size_t nh = 0, ni = 0, np = 0;
        for (int i : mem) {
            nh = std::max(nh, plans[i].P.stages.size());
            ni = std::max(ni, plans[i].P.instrs.size());
            np = std::max(np, plans[i].P.pool.size());
        }
        const ProgramLayout L = program_layout(nh, ni, np, J.prec);
        for (size_t c0 = 0; c0 < mem.size(); c0 += B.chunk_cap) {
            const int cnt = (int)std::min<size_t>(B.chunk_cap, mem.size() - c0);
            std::vector<int> idx(mem.begin() + c0, mem.begin() + c0 + cnt);
            host.assign((size_t)cnt * L.bytes, 0);
            DeviceProgram dp;
            for (int j = 0; j < cnt; j++) {
                rc = serialize_program(plans[idx[j]].P, J.prec, L, host.data() + (size_t)j * L.bytes,
                                       j == 0 ? &dp.stages : nullptr);
                if (rc) return rc;
            }
            GOQU_HIP(hipMemcpyAsync(B.blobs, host.data(), host.size(), hipMemcpyHostToDevice, s));
            rc = fill_basis_batch(view, cnt, s);
            if (rc) return rc;
            dp.blob = B.blobs;
            dp.bytes = L.bytes;
            dp.hdr_off = L.hdr_off;
            dp.instr_off = L.instr_off;
            dp.pool_off = L.pool_off;
            ShotBatch sbat;
            sbat.nshots = cnt;
            sbat.stride = view.size();
            sbat.blob_stride = L.bytes;
            rc = run_program(view, dp, s, &sbat);
            if (rc) return rc;
            rc = B.measure(view, idx);  // synchronises: the host blob buffer is free again after it
            if (rc) return rc;
            GOQU_HIP(hipStreamSynchronize(s));
        }
    }
    for (int i : singles) {
        PreparedOps prep;
        rc = prepare_eops(view, plans[i].eops, J.cfg, s, prep);
        if (rc) return rc;
        rc = fill_basis_batch(view, 1, s);
        if (rc) return rc;
        rc = launch_prepared(view, prep, s, nullptr);
        if (rc) return rc;
        rc = B.measure(view, std::vector<int>{i});
        if (rc) return rc;
        GOQU_HIP(hipStreamSynchronize(s));
    }
    GOQU_HIP(hipStreamSynchronize(s));
    B.idle();
    return GOQU_OK;
}

int check_terms(int nt, const uint64_t *x, const uint64_t *z, const int32_t *ny, const double *c, const double *out) {
    if (nt < 0 || (nt > 0 && (!x || !z || !ny || !c || !out))) return fail(GOQU_ERR_INVALID_ARG, "bad Pauli terms");
    return GOQU_OK;
}

int plan_of(int nq, int32_t tile_bits, int32_t low_bits, int prec, PlanConfig &cfg) {
    if (tile_bits == 0 && low_bits == 0) {
        cfg = default_plan_config(prec, nq, false);
        return GOQU_OK;
    }
    if (tile_bits < 8 || tile_bits > kMaxTileBits || low_bits < 2 || low_bits > 6 || low_bits > tile_bits - kRegBits)
        return fail(GOQU_ERR_INVALID_ARG, "tile_bits must be 8..13 and low_bits 2..min(6, tile_bits - 4)");
    cfg.tile_bits = tile_bits;
    cfg.low_bits = low_bits;
    return GOQU_OK;
}

}  // namespace

}  // namespace goqu

using namespace goqu;

extern "C" {

int goqu_sv_batch_expect(int32_t num_qubits, int32_t precision, int32_t tile_bits, int32_t low_bits,
                         int32_t num_instances, const goqu_op_t *ops, const int64_t *op_offsets, int32_t num_terms,
                         const uint64_t *xmask, const uint64_t *zmask, const int32_t *num_y, const double *coeffs,
                         double *out) {
    GOQU_GUARD_BEGIN
    if (num_qubits < 1 || num_qubits > 34 || (precision != GOQU_C128 && precision != GOQU_C64))
        return fail(GOQU_ERR_INVALID_ARG, "bad qubit count or precision");
    int rc = check_terms(num_terms, xmask, zmask, num_y, coeffs, out);
    if (rc) return rc;
    BatchJob J;
    J.n = num_qubits;
    J.prec = precision;
    J.nq = std::max(num_qubits, kMinInternalQubits);
    rc = plan_of(J.nq, tile_bits, low_bits, precision, J.cfg);
    if (rc) return rc;
    J.nt = num_terms;
    J.x = xmask;
    J.z = zmask;
    J.ny = num_y;
    J.coeffs = coeffs;
    J.expect_out = out;
    return run_batch(J, num_instances, ops, op_offsets);
    GOQU_GUARD_END
}

int goqu_sv_batch_sample(int32_t num_qubits, int32_t precision, int32_t tile_bits, int32_t low_bits,
                         int32_t num_instances, const goqu_op_t *ops, const int64_t *op_offsets, int64_t shots,
                         const double *uniforms, uint64_t *out_idx) {
    GOQU_GUARD_BEGIN
    if (num_qubits < 1 || num_qubits > 34 || (precision != GOQU_C128 && precision != GOQU_C64))
        return fail(GOQU_ERR_INVALID_ARG, "bad qubit count or precision");
    if (shots < 0 || (shots > 0 && (!uniforms || !out_idx))) return fail(GOQU_ERR_INVALID_ARG, "bad sampling buffers");
    BatchJob J;
    J.n = num_qubits;
    J.prec = precision;
    J.nq = std::max(num_qubits, kMinInternalQubits);
    int rc = plan_of(J.nq, tile_bits, low_bits, precision, J.cfg);
    if (rc) return rc;
    J.shots = shots;
    J.uniforms = uniforms;
    J.sample_out = out_idx;
    if (!shots) return GOQU_OK;
    return run_batch(J, num_instances, ops, op_offsets);
    GOQU_GUARD_END
}

int goqu_dm_batch_expect(int32_t num_qubits, int32_t num_instances, const goqu_op_t *ops, const int64_t *op_offsets,
                         int32_t num_terms, const uint64_t *xmask, const uint64_t *zmask, const int32_t *num_y,
                         const double *coeffs, double *out) {
    GOQU_GUARD_BEGIN
    if (num_qubits < 1 || num_qubits > 16) return fail(GOQU_ERR_INVALID_ARG, "densitymatrix: numQubits out of range");
    int rc = check_terms(num_terms, xmask, zmask, num_y, coeffs, out);
    if (rc) return rc;
    BatchJob J;
    J.density = true;
    J.n = num_qubits;
    J.nq = std::max(2 * num_qubits, kMinInternalQubits);
    J.cfg = default_plan_config(GOQU_C128, J.nq, false);
    J.nt = num_terms;
    J.x = xmask;
    J.z = zmask;
    J.ny = num_y;
    J.coeffs = coeffs;
    J.expect_out = out;
    return run_batch(J, num_instances, ops, op_offsets);
    GOQU_GUARD_END
}

int goqu_dm_batch_sample(int32_t num_qubits, int32_t num_instances, const goqu_op_t *ops, const int64_t *op_offsets,
                         const double *p01, const double *p10, int64_t shots, const double *uniforms,
                         uint64_t *out_idx) {
    GOQU_GUARD_BEGIN
    if (num_qubits < 1 || num_qubits > 16) return fail(GOQU_ERR_INVALID_ARG, "densitymatrix: numQubits out of range");
    if (shots < 0 || (shots > 0 && (!uniforms || !out_idx))) return fail(GOQU_ERR_INVALID_ARG, "bad sampling buffers");
    BatchJob J;
    J.density = true;
    J.n = num_qubits;
    J.nq = std::max(2 * num_qubits, kMinInternalQubits);
    J.cfg = default_plan_config(GOQU_C128, J.nq, false);
    J.shots = shots;
    J.uniforms = uniforms;
    J.p01 = p01;
    J.p10 = p10;
    J.sample_out = out_idx;
    if (!shots) return GOQU_OK;
    return run_batch(J, num_instances, ops, op_offsets);
    GOQU_GUARD_END
}

}  // extern "C"
