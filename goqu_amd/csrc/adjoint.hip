// adjoint.hip — adjoint differentiation (Jones & Gacon 2020, "Efficient calculation of
// gradients in classical simulations of variational quantum algorithms").
//
// Statevector: psi = U_N..U_1|0>, lam = H psi, E = Re<psi|lam>.  Going backwards,
// psi_{k-1} = U_k^dag psi_k and lam_{k-1} = U_k^dag lam_k; for every parameter entering U_k,
// dE/dtheta += scale * 2 Re <lam_k| dU_k |psi_{k-1}>.  The reverse sweep is one planned
// program run by the paired stage kernel, which moves psi and lam through the same tiles and
// reduces every derivative inside the stage where its op is uncomputed.  Derivatives on more
// than two target qubits fall back to uncomputing op by op with one reduction each.
#include <algorithm>
#include <cmath>
#include <complex>
#include <future>
#include <map>

#include "runtime.h"

namespace goqu {

using cd = std::complex<double>;

static void ipow(int n, double &re, double &im) {
    switch (((n % 4) + 4) % 4) {
    case 1: re = 0, im = 1; break;
    case 2: re = -1, im = 0; break;
    case 3: re = 0, im = -1; break;
    default: re = 1, im = 0; break;
    }
}

// lam = H psi
static int apply_hamiltonian(const State &psi, State &lam, int nt, const uint64_t *x, const uint64_t *z,
                             const int32_t *ny, const double *coeffs, Scratch &scr, hipStream_t s) {
    {  // tiled passes (psi and lam move once per pass of up to 12 tile qubits)
        std::vector<double> cf(2 * (size_t)nt);
        for (int t = 0; t < nt; t++) {
            double yr, yi;
            ipow(ny[t], yr, yi);
            const double cr = coeffs[2 * t], ci = coeffs[2 * t + 1];
            cf[2 * t] = cr * yr - ci * yi;
            cf[2 * t + 1] = cr * yi + ci * yr;
        }
        bool done = false;
        int rc = apply_hamiltonian_tiled(psi, lam, nt, x, z, cf.data(), scr, s, done);
        if (rc || done) return rc;
    }
    int rc0 = fill_basis(lam, ~0ull, s);  // all zeros
    if (rc0) return rc0;
    std::map<uint64_t, std::vector<int>> groups;
    for (int t = 0; t < nt; t++) groups[x[t]].push_back(t);
    for (auto &g : groups) {
        std::vector<uint64_t> zs;
        std::vector<double> cs;
        for (int t : g.second) {
            double yr, yi;
            ipow(ny[t], yr, yi);
            const double cr = coeffs[2 * t], ci = coeffs[2 * t + 1];
            zs.push_back(z[t]);
            cs.push_back(cr * yr - ci * yi);
            cs.push_back(cr * yi + ci * yr);
        }
        int rc = apply_pauli_group(psi, lam, g.first, zs.data(), cs.data(), (int)zs.size(), scr, s);
        if (rc) return rc;
    }
    return GOQU_OK;
}

// a work state shaped like ref, kept across calls (grow-only)
static int ensure_like(State &buf, const State &ref) {
    if (buf.amps && buf.nq == ref.nq && buf.prec == ref.prec) return GOQU_OK;
    if (buf.amps) (void)hipFree(buf.amps);
    buf.amps = nullptr;
    buf.nq = ref.nq;
    buf.prec = ref.prec;
    buf.elem = ref.elem;
    if (dmalloc((void **)&buf.amps, ref.size() * ref.elem) != hipSuccess) {
        (void)hipGetLastError();
        buf.amps = nullptr;
        return fail(GOQU_ERR_OUT_OF_MEMORY, "adjoint: cannot allocate a work state vector");
    }
    return GOQU_OK;
}

int adjoint_sv(State &st, int nq_user, const goqu_op_t *ops, int nops, const goqu_deriv_t *derivs, int nd, int np,
               int nt, const uint64_t *x, const uint64_t *z, const int32_t *ny, const double *coeffs, double *energy,
               double *grads, Scratch &scr, hipStream_t s, State &lam, State &work) {
    const uint64_t valid = (1ull << nq_user) - 1;
    for (int t = 0; t < nt; t++)
        if ((x[t] & ~valid) || (z[t] & ~valid)) return fail(GOQU_ERR_INVALID_ARG, "Pauli mask beyond the qubit count");
    std::vector<std::vector<int>> by_op(nops);
    for (int i = 0; i < nd; i++) {
        if (derivs[i].op < 0 || derivs[i].op >= nops || derivs[i].param < 0 || derivs[i].param >= np ||
            !derivs[i].dmatrix)
            return fail(GOQU_ERR_INVALID_ARG, "derivative entry " + std::to_string(i) + " out of range");
        by_op[derivs[i].op].push_back(i);
    }
    // lower every op on its own (needed for the op-by-op inverse)
    std::vector<std::vector<EOp>> low(nops);
    std::vector<EOp> all;
    std::string err;
    for (int k = 0; k < nops; k++) {
        int rc = lower_sv_ops(&ops[k], 1, nq_user, low[k], err);
        if (rc) return fail(rc, "op " + std::to_string(k) + ": " + err);
        all.insert(all.end(), low[k].begin(), low[k].end());
    }
    // The fused reverse program depends on the ops and derivatives alone: it is built and planned
    // on a second host thread while the device runs the forward pass and the Hamiltonian (planning
    // a 24-qubit EfficientSU2 reverse program costs milliseconds of host time).
    struct RevPlan {
        int rc = GOQU_OK;
        bool fused = false;
        std::string err;
        Program P;
    };
    std::future<RevPlan> rev_plan = std::async(std::launch::async, [&]() {
        RevPlan r;
        try {
            std::vector<EOp> rev;
            r.rc = build_sv_reverse(ops, nops, nq_user, derivs, nd, rev, r.fused, r.err);
            if (!r.rc && r.fused && !rev.empty())
                r.P = plan_program(rev, st.nq, default_plan_config(st.prec, st.nq, true));
        } catch (...) {
            r.rc = GOQU_ERR_OUT_OF_MEMORY;
            r.err = "host memory exhausted while planning the reverse pass";
        }
        return r;
    });
    int rc = fill_basis(st, 0, s);
    if (rc) return rc;
    rc = run_eops(st, all, s);
    if (rc) return rc;

    // lambda and the uncomputed copy of psi live in the handle between calls
    auto cleanup = [&]() {};
    rc = ensure_like(lam, st);
    if (!rc) rc = ensure_like(work, st);
    if (rc) return rc;
    rc = apply_hamiltonian(st, lam, nt, x, z, ny, coeffs, scr, s);
    double e2[2];
    if (!rc) rc = inner_product(st, lam, e2, scr, s);
    if (!rc) rc = copy_state(work, st, s);
    if (rc) {
        cleanup();
        return rc;
    }
    *energy = e2[0];
    for (int p = 0; p < np; p++) grads[p] = 0;

    // Fused reverse pass: one program, run on (psi, lam) together, in which every derivative is
    // an EOP_GRAD op placed right before the inverse of its op.
    RevPlan rp = rev_plan.get();
    const bool fused = rp.fused;
    rc = rp.rc;
    if (!rc && fused) {
        std::vector<double> vals;
        rc = run_pair_planned(work, lam, rp.P, nd, vals, s);
        if (!rc)
            for (int i = 0; i < nd; i++) grads[derivs[i].param] += derivs[i].scale * 2.0 * vals[i];
    } else if (rc) {
        rc = fail(rc, rp.err);
    }
    if (rc || fused) {
        cleanup();
        return rc;
    }

    // Op by op (derivatives on more than two targets): uncompute pending runs, then one
    // sandwich reduction per derivative.
    std::vector<EOp> pend_psi, pend_lam;
    for (int k = nops - 1; k >= 0 && !rc; k--) {
        std::vector<EOp> inv;
        for (int j = (int)low[k].size() - 1; j >= 0; j--) {
            EOp ie;
            rc = invert_eop(low[k][j], ie, err);
            if (rc) {
                cleanup();
                return fail(rc, err);
            }
            inv.push_back(ie);
        }
        pend_psi.insert(pend_psi.end(), inv.begin(), inv.end());
        if (!by_op[k].empty()) {
            rc = run_eops(work, pend_psi, s);
            if (!rc) rc = run_eops(lam, pend_lam, s);
            pend_psi.clear();
            pend_lam.clear();
            for (int i : by_op[k]) {
                if (rc) break;
                EOp dop;
                rc = deriv_eop(ops[k], derivs[i].dmatrix, dop, err);
                if (rc) {
                    rc = fail(rc, err);
                    break;
                }
                double sres[2];
                rc = sandwich(lam, work, dop, sres, scr, s);
                if (!rc) grads[derivs[i].param] += derivs[i].scale * 2.0 * sres[0];
            }
        }
        pend_lam.insert(pend_lam.end(), inv.begin(), inv.end());
    }
    cleanup();
    return rc;
}

// ---- density matrix ------------------------------------------------------------------------------
// E = Tr(H rho_N) = <h|rho_N> with h = vec(H) (h[r*dim + c] = H_rc, H Hermitian).  Going
// backwards h_{k-1} = S_k^dag h_k, where S_k is op k's superoperator (the conjugate transpose
// of every lowered op, in reverse order).  A gate U(theta) has S = U_row conj(U)_col and
// dE/dtheta = 2 Re <h_k| (dU_row conj(U)_col) |rho_{k-1}>.  rho_{k-1} is recomputed from the
// nearest stored checkpoint (every `every` ops), because channels cannot be uncomputed.

// h[(c ^ x) * dim + c] += coef * i^ny * (-1)^{|c & z|}
__global__ void k_add_pauli_matrix(cx<double> *__restrict__ h, uint64_t dim, uint64_t x, uint64_t z, double cr,
                                   double ci) {
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < dim; c += (uint64_t)gridDim.x * blockDim.x) {
        const bool neg = __popcll(c & z) & 1;
        cx<double> &e = h[(c ^ x) * dim + c];
        e.x += neg ? -cr : cr;
        e.y += neg ? -ci : ci;
    }
}

static int alloc_like(State &s, const State &ref) {
    s.nq = ref.nq;
    s.prec = ref.prec;
    s.elem = ref.elem;
    if (dmalloc((void **)&s.amps, ref.size() * ref.elem) != hipSuccess) {
        (void)hipGetLastError();
        s.amps = nullptr;
        return fail(GOQU_ERR_OUT_OF_MEMORY, "adjoint: cannot allocate a density-matrix buffer");
    }
    return GOQU_OK;
}

// h = vec(H): sum_t coeff_t i^{ny_t} P_t as a 2n-qubit vector (row bits high)
static int build_h(State &h, int n, int nt, const uint64_t *x, const uint64_t *z, const int32_t *ny,
                   const double *coeffs, hipStream_t s) {
    int rc = fill_basis(h, ~0ull, s);
    const uint64_t dim = 1ull << n;
    for (int t = 0; t < nt && !rc; t++) {
        double yr, yi;
        ipow(ny[t], yr, yi);
        const double cr = coeffs[2 * t] * yr - coeffs[2 * t + 1] * yi;
        const double ci = coeffs[2 * t] * yi + coeffs[2 * t + 1] * yr;
        const unsigned grid = (unsigned)std::min<uint64_t>((dim + 255) / 256, 4096);
        hipLaunchKernelGGL(k_add_pauli_matrix, dim3(grid), dim3(256), 0, s, static_cast<cx<double> *>(h.amps), dim,
                           x[t], z[t], cr, ci);
        if (hipGetLastError() != hipSuccess) rc = fail(GOQU_ERR_DEVICE, "pauli matrix kernel launch failed");
        // terms with the same x touch the same entries: keep launches ordered
    }
    return rc;
}

// The fused reverse sweep (when every op allows it, build_dm_reverse): the forward pass stores
// rho at the end of every segment; going backwards, rho is restored from the segment's end
// checkpoint (so uncomputing errors never cross a segment) and the segment's reverse program
// runs on (rho, h) by paired stage sweeps, with the derivative terms reduced inside the sweeps.
// fused = false: not applicable, nothing was run.
static int adjoint_dm_fused(State &rho, int n, const goqu_op_t *ops, int nops, const goqu_deriv_t *derivs, int nd,
                            int np, int nt, const uint64_t *x, const uint64_t *z, const int32_t *ny,
                            const double *coeffs, int every, const std::vector<std::vector<EOp>> &low,
                            double *energy, double *grads, Scratch &scr, hipStream_t s, bool &fused) {
    fused = false;
    (void)every;  // segments follow the error growth of uncomputing, not a fixed op count
    if (nops == 0) return GOQU_OK;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return fail(GOQU_ERR_DEVICE, "hipMemGetInfo failed");
    const size_t sbytes = rho.size() * rho.elem;
    const int64_t max_ck = fr > 4 * sbytes ? (int64_t)((fr - 4 * sbytes) / sbytes) : 0;
    // Segment boundaries: a segment ends before the bound on the error growth of uncomputing its
    // ops (product of the inverse channels' norms) would exceed kMaxGrowth; rho is restored
    // exactly at every boundary.
    constexpr double kMaxGrowth = 4.0;
    std::vector<int> bounds{0};
    double grow = 1.0;
    for (int k = 0; k < nops; k++) {
        const double g = dm_uncompute_growth(ops[k], n);
        if (!(g < kMaxGrowth)) return GOQU_OK;  // a reset, a singular or badly conditioned channel
        if (grow * g > kMaxGrowth && k > bounds.back()) {
            bounds.push_back(k);
            grow = 1.0;
        }
        grow *= g;
    }
    bounds.push_back(nops);
    const int nck = (int)bounds.size() - 1;
    if (nck - 1 > max_ck) return GOQU_OK;
    std::vector<std::vector<EOp>> rev(nck);
    std::string err;
    for (int c = 0; c < nck; c++) {
        bool ok = false;
        int rc = build_dm_reverse(ops, bounds[c], bounds[c + 1], n, derivs, nd, rev[c], ok, err);
        if (rc) return fail(rc, err);
        if (!ok) return GOQU_OK;
    }
    fused = true;
    std::vector<State> ck(nck - 1);
    State h;
    auto cleanup = [&]() {
        for (State &c : ck)
            if (c.amps) (void)hipFree(c.amps);
        if (h.amps) (void)hipFree(h.amps);
    };
    int rc = fill_basis(rho, 0, s);
    for (int c = 0; c < nck && !rc; c++) {
        std::vector<EOp> seg;
        for (int k = bounds[c]; k < bounds[c + 1]; k++) seg.insert(seg.end(), low[k].begin(), low[k].end());
        merge_channels(seg);  // the forward pass needs rho only at the segment ends
        rc = run_eops(rho, seg, s);
        if (!rc && c < nck - 1) rc = alloc_like(ck[c], rho);
        if (!rc && c < nck - 1) rc = copy_state(ck[c], rho, s);
    }
    if (!rc) rc = alloc_like(h, rho);
    if (!rc) rc = build_h(h, n, nt, x, z, ny, coeffs, s);
    double e2[2] = {0, 0};
    if (!rc) rc = inner_product(h, rho, e2, scr, s);
    if (rc) {
        cleanup();
        return rc;
    }
    *energy = e2[0];
    for (int p = 0; p < np; p++) grads[p] = 0;
    std::vector<double> vals;
    for (int c = nck - 1; c >= 0 && !rc; c--) {
        if (c < nck - 1) rc = copy_state(rho, ck[c], s);  // rho at the end of segment c, exactly
        if (!rc) rc = run_pair(rho, h, rev[c], nd, vals, s);
        for (int i = 0; i < nd && !rc; i++) grads[derivs[i].param] += derivs[i].scale * 2.0 * vals[i];
    }
    cleanup();
    return rc;
}

// Reverse sweep by segments of `every` ops: rho at the start of every segment is stored on the
// forward pass; going backwards, each segment that holds derivatives is replayed once from its
// checkpoint, keeping rho_{k-1} of each derivative op k in a snapshot, and the sandwiches of an op's
// derivatives (U3: three) share one pass over h and the snapshot.  Used when the fused sweep does
// not apply (resets, channels that are not parity-block or are badly conditioned).
int adjoint_dm(State &rho, int n, const goqu_op_t *ops, int nops, const goqu_deriv_t *derivs, int nd, int np,
               int nt, const uint64_t *x, const uint64_t *z, const int32_t *ny, const double *coeffs, int every,
               double *energy, double *grads, Scratch &scr, hipStream_t s) {
    const uint64_t valid = (1ull << n) - 1;
    for (int t = 0; t < nt; t++)
        if ((x[t] & ~valid) || (z[t] & ~valid)) return fail(GOQU_ERR_INVALID_ARG, "Pauli mask beyond the qubit count");
    std::vector<std::vector<int>> by_op(nops);
    for (int i = 0; i < nd; i++) {
        const goqu_deriv_t &d = derivs[i];
        if (d.op < 0 || d.op >= nops || d.param < 0 || d.param >= np || !d.dmatrix)
            return fail(GOQU_ERR_INVALID_ARG, "derivative entry " + std::to_string(i) + " out of range");
        const goqu_op_t &op = ops[d.op];
        if (op.kind != GOQU_OP_GATE || op.num_qubits > 2 || !op.matrix)
            return fail(GOQU_ERR_UNSUPPORTED, "density-matrix derivatives need 1- or 2-qubit gates with a matrix");
        by_op[d.op].push_back(i);
    }
    std::vector<std::vector<EOp>> low(nops);
    std::string err;
    for (int k = 0; k < nops; k++) {
        int rc = lower_dm_ops(&ops[k], 1, n, low[k], err);
        if (rc) return fail(rc, "op " + std::to_string(k) + ": " + err);
    }
    {
        bool fused = false;
        int rc = adjoint_dm_fused(rho, n, ops, nops, derivs, nd, np, nt, x, z, ny, coeffs, every, low, energy, grads,
                                  scr, s, fused);
        if (rc || fused) return rc;
    }
    if (every < 1) every = 1;
    // segment length: checkpoints plus one segment's snapshots must fit in free device memory
    // (h, the replay state and two more states stay free)
    auto snaps_of = [&](int ev) {
        int mx = 0;
        for (int c0 = 0; c0 < nops; c0 += ev) {
            int cnt = 0;
            for (int k = c0; k < std::min(nops, c0 + ev); k++) cnt += by_op[k].empty() ? 0 : 1;
            mx = std::max(mx, cnt);
        }
        return mx;
    };
    {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return fail(GOQU_ERR_DEVICE, "hipMemGetInfo failed");
        const size_t sbytes = rho.size() * rho.elem;
        const int64_t max_states = fr > 4 * sbytes ? (int64_t)((fr - 4 * sbytes) / sbytes) : 0;
        while (every < std::max(nops, 1) && (nops + every - 1) / every + snaps_of(every) > max_states) every++;
        if ((nops + every - 1) / every + snaps_of(every) > max_states)
            return fail(GOQU_ERR_OUT_OF_MEMORY, "adjoint: checkpoints do not fit in device memory");
    }
    const int nck = (nops + every - 1) / every;
    std::vector<State> ck(std::max(nck, 0)), snap(snaps_of(every));
    State h, work;
    auto cleanup = [&]() {
        for (State &c : ck)
            if (c.amps) (void)hipFree(c.amps);
        for (State &c : snap)
            if (c.amps) (void)hipFree(c.amps);
        if (h.amps) (void)hipFree(h.amps);
        if (work.amps) (void)hipFree(work.amps);
    };
    int rc = fill_basis(rho, 0, s);
    for (int c = 0; c < nck && !rc; c++) {
        rc = alloc_like(ck[c], rho);
        if (!rc) rc = copy_state(ck[c], rho, s);
        std::vector<EOp> seg;
        for (int k = c * every; k < std::min(nops, (c + 1) * every); k++) seg.insert(seg.end(), low[k].begin(), low[k].end());
        if (!rc) rc = run_eops(rho, seg, s);
    }
    for (size_t j = 0; j < snap.size() && !rc; j++) rc = alloc_like(snap[j], rho);
    if (!rc) rc = alloc_like(h, rho);
    if (!rc) rc = alloc_like(work, rho);
    if (!rc) rc = fill_basis(h, ~0ull, s);
    const uint64_t dim = 1ull << n;
    for (int t = 0; t < nt && !rc; t++) {
        double yr, yi;
        ipow(ny[t], yr, yi);
        const double cr = coeffs[2 * t] * yr - coeffs[2 * t + 1] * yi;
        const double ci = coeffs[2 * t] * yi + coeffs[2 * t + 1] * yr;
        const unsigned grid = (unsigned)std::min<uint64_t>((dim + 255) / 256, 4096);
        hipLaunchKernelGGL(k_add_pauli_matrix, dim3(grid), dim3(256), 0, s, static_cast<cx<double> *>(h.amps), dim,
                           x[t], z[t], cr, ci);
        if (hipGetLastError() != hipSuccess) rc = fail(GOQU_ERR_DEVICE, "pauli matrix kernel launch failed");
        // terms with the same x touch the same entries: keep launches ordered
    }
    double e2[2] = {0, 0};
    if (!rc) rc = inner_product(h, rho, e2, scr, s);
    if (rc) {
        cleanup();
        return rc;
    }
    *energy = e2[0];
    for (int p = 0; p < np; p++) grads[p] = 0;

    std::vector<EOp> pend;
    std::vector<double> sres(2 * kMaxSandwich);
    for (int c = nck - 1; c >= 0 && !rc; c--) {
        const int k0 = c * every, k1 = std::min(nops, (c + 1) * every);
        // replay the segment from its checkpoint up to its last derivative op, keeping rho_{k-1}
        // of every derivative op k
        std::vector<int> dks;
        for (int k = k0; k < k1; k++)
            if (!by_op[k].empty()) dks.push_back(k);
        if (!dks.empty()) {
            rc = copy_state(work, ck[c], s);
            int at = k0;
            for (size_t j = 0; j < dks.size() && !rc; j++) {
                std::vector<EOp> seg;
                for (int k = at; k < dks[j]; k++) seg.insert(seg.end(), low[k].begin(), low[k].end());
                rc = run_eops(work, seg, s);
                if (!rc) rc = copy_state(snap[j], work, s);
                at = dks[j];
            }
        }
        int sj = (int)dks.size() - 1;
        for (int k = k1 - 1; k >= k0 && !rc; k--) {
            if (!by_op[k].empty()) {
                rc = run_eops(h, pend, s);
                pend.clear();
                const goqu_op_t &op = ops[k];
                const int kg = op.num_qubits, d = 1 << kg, D = d * d;
                const std::vector<int> &ent = by_op[k];
                for (size_t e0 = 0; e0 < ent.size() && !rc; e0 += kMaxSandwich) {
                    std::vector<EOp> dops;
                    for (size_t e = e0; e < std::min(ent.size(), e0 + kMaxSandwich); e++) {
                        // dop = dU (x) conj(U) on (row qubits, column qubits)
                        EOp dop;
                        dop.kind = EOP_DENSE;
                        dop.k = 2 * kg;
                        for (int j = 0; j < kg; j++) {
                            dop.q[j] = op.qubits[j] + n;
                            dop.q[kg + j] = op.qubits[j];
                        }
                        dop.m.assign(2 * (size_t)D * D, 0.0);
                        const double *dU = derivs[ent[e]].dmatrix, *U = op.matrix;
                        for (int r1 = 0; r1 < d; r1++)
                            for (int r0 = 0; r0 < d; r0++)
                                for (int c1 = 0; c1 < d; c1++)
                                    for (int c0 = 0; c0 < d; c0++) {
                                        const double ar = dU[2 * (r1 * d + r0)], ai = dU[2 * (r1 * d + r0) + 1];
                                        const double br = U[2 * (c1 * d + c0)], bi = -U[2 * (c1 * d + c0) + 1];
                                        const size_t idx = (size_t)(r1 * d + c1) * D + (r0 * d + c0);
                                        dop.m[2 * idx] = ar * br - ai * bi;
                                        dop.m[2 * idx + 1] = ar * bi + ai * br;
                                    }
                        dops.push_back(std::move(dop));
                    }
                    rc = sandwich_multi(h, snap[sj], dops, sres.data(), scr, s);
                    for (size_t e = 0; e < dops.size() && !rc; e++) {
                        const goqu_deriv_t &dv = derivs[ent[e0 + e]];
                        grads[dv.param] += dv.scale * 2.0 * sres[2 * e];
                    }
                }
                sj--;
            }
            for (int j = (int)low[k].size() - 1; j >= 0 && !rc; j--) {
                EOp a;
                rc = invert_eop(low[k][j], a, err);  // conjugate transpose = adjoint superoperator
                if (rc) rc = fail(rc, err);
                pend.push_back(a);
            }
        }
    }
    cleanup();
    return rc;
}

}  // namespace goqu

using namespace goqu;

extern "C" {

int goqu_sv_adjoint(goqu_sv_t *h, const goqu_op_t *ops, int32_t num_ops, const goqu_deriv_t *derivs,
                    int32_t num_derivs, int32_t num_params, int32_t num_terms, const uint64_t *xmask,
                    const uint64_t *zmask, const int32_t *num_y, const double *coeffs, double *energy,
                    double *grads) {
    GOQU_GUARD_BEGIN
    if (!h || num_ops < 0 || num_derivs < 0 || num_params < 0 || num_terms < 0 || !energy ||
        (num_ops && !ops) || (num_derivs && !derivs) || (num_params && !grads) ||
        (num_terms && (!xmask || !zmask || !num_y || !coeffs)))
        return fail(GOQU_ERR_INVALID_ARG, "bad argument");
    return adjoint_sv(h->st, h->nq_user, ops, num_ops, derivs, num_derivs, num_params, num_terms, xmask, zmask, num_y,
                      coeffs, energy, grads, h->scr, h->stream, h->adj_lam, h->adj_work);
    GOQU_GUARD_END
}

int goqu_dm_adjoint(goqu_dm_t *h, const goqu_op_t *ops, int32_t num_ops, const goqu_deriv_t *derivs,
                    int32_t num_derivs, int32_t num_params, int32_t num_terms, const uint64_t *xmask,
                    const uint64_t *zmask, const int32_t *num_y, const double *coeffs, int32_t checkpoint_every,
                    double *energy, double *grads) {
    GOQU_GUARD_BEGIN
    if (!h || num_ops < 0 || num_derivs < 0 || num_params < 0 || num_terms < 0 || !energy ||
        (num_ops && !ops) || (num_derivs && !derivs) || (num_params && !grads) ||
        (num_terms && (!xmask || !zmask || !num_y || !coeffs)))
        return fail(GOQU_ERR_INVALID_ARG, "bad argument");
    return adjoint_dm(h->st, h->n, ops, num_ops, derivs, num_derivs, num_params, num_terms, xmask, zmask, num_y,
                      coeffs, checkpoint_every, energy, grads, h->scr, h->stream);
    GOQU_GUARD_END
}

}  // extern "C"
