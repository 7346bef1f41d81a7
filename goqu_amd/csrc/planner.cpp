// planner.cpp — lowering of goqu ops to elementary ops, and the stage planner.
//
// Lowering follows the reference's dispatch tiers (sim/statevector/sim.go:92-157,
// kernel2q.go:24-114, kernel3q.go:10-45, kernel_controlled.go:11-50): every tier becomes
// the elementary op with the same effect on the amplitudes (permutations move
// amplitudes, diagonal gates multiply, controlled gates act only where all controls are
// 1).  The planner then groups elementary ops into stages (one HBM sweep each) and, inside
// a stage, assigns register slots and the LDS re-distributions between them.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>

#include "../../include/goqu_hip.h"
#include "engine.h"

namespace goqu {

using cd = std::complex<double>;

static cd at(const double *m, int i) { return cd(m[2 * i], m[2 * i + 1]); }

static void push(std::vector<double> &v, cd c) {
    v.push_back(c.real());
    v.push_back(c.imag());
}

static bool is_diagonal(const double *m, int dim) {
    for (int r = 0; r < dim; r++)
        for (int c = 0; c < dim; c++)
            if (r != c && (m[2 * (r * dim + c)] != 0.0 || m[2 * (r * dim + c) + 1] != 0.0)) return false;
    return true;
}

static EOp make_diag(const int *q, int k, uint64_t cmask, const double *m, int dim) {
    EOp e;
    e.kind = EOP_DIAG;
    e.k = k;
    for (int i = 0; i < k; i++) e.q[i] = q[i];
    e.cmask = cmask;
    for (int i = 0; i < dim; i++) push(e.m, at(m, i * dim + i));
    return e;
}

static EOp make_wide(const int *q, int k, uint64_t cmask, const double *m) {
    EOp e;
    e.kind = EOP_WIDE;
    e.k = k;
    e.qw.assign(q, q + k);
    e.cmask = cmask;
    const size_t d = size_t(1) << k;
    e.m.assign(m, m + 2 * d * d);
    return e;
}

static EOp make_dense(const int *q, int k, uint64_t cmask, const double *m, int dim) {
    EOp e;
    e.kind = EOP_DENSE;
    e.k = k;
    for (int i = 0; i < k; i++) e.q[i] = q[i];
    e.cmask = cmask;
    e.m.assign(m, m + 2 * dim * dim);
    return e;
}

static EOp make_anti1(int t, uint64_t cmask, cd p0, cd p1) {
    EOp e;
    e.kind = EOP_ANTI1;
    e.k = 1;
    e.q[0] = t;
    e.cmask = cmask;
    push(e.m, p0);
    push(e.m, p1);
    e.pure_swap = (p0 == cd(1, 0) && p1 == cd(1, 0));
    return e;
}

static EOp make_swap2(int a, int b, uint64_t cmask, cd p) {
    EOp e;
    e.kind = EOP_SWAP2;
    e.k = 2;
    e.q[0] = a;
    e.q[1] = b;
    e.cmask = cmask;
    push(e.m, p);
    e.pure_swap = (p == cd(1, 0));
    return e;
}

// A matrix acting on k qubits (possibly under controls) -> the cheapest elementary op.
static void lower_matrix(const int *q, int k, uint64_t cmask, const double *m, std::vector<EOp> &out) {
    const int dim = 1 << k;
    if (is_diagonal(m, dim)) {
        out.push_back(make_diag(q, k, cmask, m, dim));
        return;
    }
    if (k == 1 && m[0] == 0 && m[1] == 0 && m[6] == 0 && m[7] == 0) {
        out.push_back(make_anti1(q[0], cmask, at(m, 1), at(m, 2)));
        return;
    }
    out.push_back(make_dense(q, k, cmask, m, dim));
}

int lower_sv_ops(const void *ops_v, int nops, int nq, std::vector<EOp> &out, std::string &err) {
    const goqu_op_t *ops = static_cast<const goqu_op_t *>(ops_v);
    out.clear();
    for (int i = 0; i < nops; i++) {
        const goqu_op_t &op = ops[i];
        if (op.kind == GOQU_OP_SKIP) continue;
        if (op.num_qubits < 1 || op.num_qubits > GOQU_MAX_OP_QUBITS) {
            err = "op " + std::to_string(i) + ": bad qubit count " + std::to_string(op.num_qubits);
            return GOQU_ERR_INVALID_ARG;
        }
        uint64_t seen = 0;
        for (int j = 0; j < op.num_qubits; j++) {
            int qq = op.qubits[j];
            if (qq < 0 || qq >= nq || (seen >> qq & 1)) {
                err = "op " + std::to_string(i) + ": qubit " + std::to_string(qq) + " invalid or repeated";
                return GOQU_ERR_INVALID_ARG;
            }
            seen |= 1ull << qq;
        }
        const int *q = op.qubits;
        const double *m = op.matrix;
        switch (op.kind) {
        case GOQU_OP_RESET: {
            EOp e;
            e.kind = EOP_RESET;
            e.k = 1;
            e.q[0] = q[0];
            out.push_back(e);
            continue;
        }
        case GOQU_OP_CONTROLLED: {  // dispatchControlled, kernel_controlled.go:11-50
            const int nc = op.num_controls, nt = op.num_qubits - nc;
            if (nc < 0 || nt < 1) {
                err = "op " + std::to_string(i) + ": bad control count";
                return GOQU_ERR_INVALID_ARG;
            }
            if (!m) continue;  // opaque inner gate: identity
            if (nt >= 3 && op.num_qubits > 10) {  // controlled.Matrix() cap, controlled.go:87-89
                err = "controlled gate with " + std::to_string(op.num_qubits) + " qubits exceeds 10";
                return GOQU_ERR_UNSUPPORTED;
            }
            uint64_t cm = 0;
            for (int j = 0; j < nc; j++) cm |= 1ull << q[j];
            if (nt > kRegBits) {  // kernel_controlled.go:249-292: full inner matrix on 5..9 targets
                if (is_diagonal(m, 1 << nt)) {
                    err = "controlled diagonal gate with more than 4 target qubits";
                    return GOQU_ERR_UNSUPPORTED;
                }
                out.push_back(make_wide(q + nc, nt, cm, m));
                continue;
            }
            lower_matrix(q + nc, nt, cm, m, out);
            continue;
        }
        case GOQU_OP_GATE:
            break;
        default:
            err = "op " + std::to_string(i) + ": kind " + std::to_string(op.kind) + " not valid for a statevector";
            return GOQU_ERR_INVALID_ARG;
        }
        const int k = op.num_qubits;
        if (k == 1) {
            if (!m) {
                err = "gate has no matrix representation";  // sim.go:139-142
                return GOQU_ERR_NO_MATRIX;
            }
            lower_matrix(q, 1, 0, m, out);
            continue;
        }
        if (k == 2) {  // dispatchGate2 tiers, kernel2q.go:24-114
            const uint64_t c0 = 1ull << q[0];
            switch (op.gate_id) {
            case GOQU_GATE_CNOT:
                out.push_back(make_anti1(q[1], c0, 1.0, 1.0));
                continue;
            case GOQU_GATE_CZ: {  // diag(1,-1) on q1 under control q0
                const double z[8] = {1, 0, 0, 0, 0, 0, -1, 0};
                out.push_back(make_diag(q + 1, 1, c0, z, 2));
                continue;
            }
            case GOQU_GATE_SWAP:
                out.push_back(make_swap2(q[0], q[1], 0, 1.0));
                continue;
            case GOQU_GATE_CY:  // |10> <- -i a3, |11> <- i a2
                out.push_back(make_anti1(q[1], c0, cd(0, -1), cd(0, 1)));
                continue;
            case GOQU_GATE_ISWAP:
                out.push_back(make_swap2(q[0], q[1], 0, cd(0, 1)));
                continue;
            case GOQU_GATE_SYCAMORE: {
                cd d11 = m ? at(m, 15) : std::exp(cd(0, -M_PI / 6));
                out.push_back(make_swap2(q[0], q[1], 0, cd(0, -1)));
                EOp e;
                e.kind = EOP_DIAG;
                e.k = 1;
                e.q[0] = q[1];
                e.cmask = c0;
                push(e.m, 1.0);
                push(e.m, d11);
                out.push_back(e);
                continue;
            }
            default:
                break;
            }
            if (!m) continue;  // opaque 2q gate: identity (kernel2q.go:106-108)
            if (op.gate_id == GOQU_GATE_CTRL_DIAG) {
                EOp e;
                e.kind = EOP_DIAG;
                e.k = 1;
                e.q[0] = q[1];
                e.cmask = c0;
                push(e.m, at(m, 10));
                push(e.m, at(m, 15));
                out.push_back(e);
                continue;
            }
            if (op.gate_id == GOQU_GATE_DIAG2) {
                EOp e;
                e.kind = EOP_DIAG;
                e.k = 2;
                e.q[0] = q[0];
                e.q[1] = q[1];
                push(e.m, at(m, 0));
                push(e.m, at(m, 5));
                push(e.m, at(m, 10));
                push(e.m, at(m, 15));
                out.push_back(e);
                continue;
            }
            if (op.gate_id == GOQU_GATE_CTRL_U) {
                const double sub[8] = {m[20], m[21], m[22], m[23], m[28], m[29], m[30], m[31]};
                lower_matrix(q + 1, 1, c0, sub, out);
                continue;
            }
            lower_matrix(q, 2, 0, m, out);
            continue;
        }
        if (k == 3) {  // dispatchGate3, kernel3q.go:10-45
            const uint64_t c01 = (1ull << q[0]) | (1ull << q[1]);
            switch (op.gate_id) {
            case GOQU_GATE_CCX:
                out.push_back(make_anti1(q[2], c01, 1.0, 1.0));
                continue;
            case GOQU_GATE_CSWAP:
                out.push_back(make_swap2(q[1], q[2], 1ull << q[0], 1.0));
                continue;
            case GOQU_GATE_CCZ: {
                EOp e;
                e.kind = EOP_DIAG;
                e.k = 1;
                e.q[0] = q[2];
                e.cmask = c01;
                push(e.m, 1.0);
                push(e.m, -1.0);
                out.push_back(e);
                continue;
            }
            default:
                break;
            }
            if (!m) continue;  // opaque 3q gate: identity (kernel3q.go:38-41)
            lower_matrix(q, 3, 0, m, out);
            continue;
        }
        err = "unsupported gate size: " + std::to_string(k) + " qubits";  // sim.go:152
        return GOQU_ERR_UNSUPPORTED;
    }
    return GOQU_OK;
}

// ---- density matrix: rho as a 2n-qubit vector (row bits n..2n-1, column bits 0..n-1) -------------
// U rho U^dagger = U on the row qubits, element-wise conj(U) on the column qubits
// (sim/densitymatrix/kernel1q.go:10-67, kernel2q.go:7-110).  A Kraus channel sum_e E rho E^dag
// is the superoperator sum_e E (x) conj(E) on (row qubits, column qubits)
// (sim/densitymatrix/sim.go:220-255); reset is the channel {|0><0|, |0><1|} (sim.go:195-217).
static std::vector<double> conj_of(const double *m, int n) {
    std::vector<double> r(m, m + 2 * n);
    for (int i = 0; i < n; i++) r[2 * i + 1] = -r[2 * i + 1];
    return r;
}

// A channel superoperator on (rows, cols) = q[0..k/2-1], q[k/2..k-1] (k = 2 or 4, q[0] = MSB of
// the index).  Channels whose superoperator keeps every qubit's row^column parity (Pauli and
// depolarizing channels, amplitude and phase damping, reset) become EOP_PARITY: block-diagonal
// over the parity classes, at 1/2 (one qubit) or 1/4 (two qubits) of the dense cost.
static bool parity_preserving(const double *m, int k) {
    const int D = 1 << k, h = k / 2;
    auto parity = [&](int x) { return (x >> h) ^ (x & ((1 << h) - 1)); };  // rows ^ cols, per qubit
    for (int o = 0; o < D; o++)
        for (int i = 0; i < D; i++)
            if (parity(o) != parity(i) && (m[2 * (o * D + i)] != 0.0 || m[2 * (o * D + i) + 1] != 0.0)) return false;
    return true;
}

static EOp make_superop(const int *q, int k, const double *m) {
    const int D = 1 << k;
    if (is_diagonal(m, D)) return make_diag(q, k, 0, m, D);
    EOp e = make_dense(q, k, 0, m, D);
    if (parity_preserving(m, k)) e.kind = EOP_PARITY;
    return e;
}

// Channel superoperators that meet with nothing in between merge: a parity-block channel e moves
// back past every op on other qubits (they commute) to the last op that touches its qubits; when
// that op is a parity-block channel on a superset of e's row and column qubits, e is folded into
// it (S <- E S, E embedded in the wider basis; the product of parity-keeping superoperators keeps
// parity).  A noise model that ends each layer with a one-qubit channel (amplitude damping on
// Delay) after a gate's own depolarizing channel then costs one parity-block op instead of two.
void merge_channels(std::vector<EOp> &ops) {
    std::vector<bool> dead(ops.size(), false);
    for (size_t i = 0; i < ops.size(); i++) {
        EOp &e = ops[i];
        if (e.kind != EOP_PARITY || e.cmask != 0) continue;
        const uint64_t et = e.touch();
        size_t j = i;
        bool found = false;
        while (j > 0) {
            j--;
            if (dead[j]) continue;
            if (ops[j].is_barrier()) break;
            if (ops[j].touch() & et) {
                found = true;
                break;
            }
        }
        if (!found) continue;
        EOp &t = ops[j];
        if (t.kind != EOP_PARITY || t.cmask != 0 || (t.touch() & et) != et) continue;
        // position in t's index (MSB first) of each of e's qubits
        const int K = t.k, D = 1 << K, ke = e.k, De = 1 << ke;
        int pos[4];
        for (int a = 0; a < ke; a++)
            for (int b = 0; b < K; b++)
                if (t.q[b] == e.q[a]) pos[a] = K - 1 - b;
        uint32_t emask = 0;
        for (int a = 0; a < ke; a++) emask |= 1u << pos[a];
        auto sub = [&](int x) {  // x's bits at e's qubits, in e's index order
            int v = 0;
            for (int a = 0; a < ke; a++) v |= ((x >> pos[a]) & 1) << (ke - 1 - a);
            return v;
        };
        std::vector<double> m(2 * (size_t)D * D, 0.0);
        for (int o = 0; o < D; o++)
            for (int c = 0; c < D; c++) {
                double re = 0, im = 0;
                for (int x = 0; x < D; x++) {
                    if ((o & ~emask) != (x & ~emask)) continue;  // E acts on e's qubits only
                    const double *ev = &e.m[2 * ((size_t)sub(o) * De + sub(x))];
                    const double *tv = &t.m[2 * ((size_t)x * D + c)];
                    re += ev[0] * tv[0] - ev[1] * tv[1];
                    im += ev[0] * tv[1] + ev[1] * tv[0];
                }
                m[2 * ((size_t)o * D + c)] = re;
                m[2 * ((size_t)o * D + c) + 1] = im;
            }
        t.m.swap(m);
        dead[i] = true;
    }
    size_t w = 0;
    for (size_t i = 0; i < ops.size(); i++)
        if (!dead[i]) {
            if (w != i) ops[w] = std::move(ops[i]);
            w++;
        }
    ops.resize(w);
}

int lower_dm_ops(const void *ops_v, int nops, int n, std::vector<EOp> &out, std::string &err) {
    const goqu_op_t *ops = static_cast<const goqu_op_t *>(ops_v);
    out.clear();
    for (int i = 0; i < nops; i++) {
        const goqu_op_t &op = ops[i];
        if (op.kind == GOQU_OP_SKIP) continue;
        const int k = op.num_qubits;
        if (k < 1 || k > GOQU_MAX_OP_QUBITS) {
            err = "op " + std::to_string(i) + ": bad qubit count";
            return GOQU_ERR_INVALID_ARG;
        }
        {
            uint64_t seen = 0;
            for (int j = 0; j < k; j++) {
                const int qq = op.qubits[j];
                if (qq < 0 || qq >= n || (seen >> qq & 1)) {
                    err = "op " + std::to_string(i) + ": qubit out of range or repeated";
                    return GOQU_ERR_INVALID_ARG;
                }
                seen |= 1ull << qq;
            }
        }
        // Gates on 3+ qubits (and controlled gates) act exactly: U on the row qubits and
        // conj(U) on the column qubits, controls as masks on the row / column bits.
        if ((op.kind == GOQU_OP_GATE && k > 2) || op.kind == GOQU_OP_CONTROLLED) {
            const int nc = op.kind == GOQU_OP_CONTROLLED ? op.num_controls : 0, nt = k - nc;
            if (nc < 0 || nt < 1) {
                err = "op " + std::to_string(i) + ": bad control count";
                return GOQU_ERR_INVALID_ARG;
            }
            if (nt > kMaxWideTargets) {
                err = "op " + std::to_string(i) + ": more than 9 target qubits";
                return GOQU_ERR_UNSUPPORTED;
            }
            if (!op.matrix) continue;
            uint64_t rc = 0, cc = 0;
            int rt[kMaxWideTargets], ct[kMaxWideTargets];
            for (int j = 0; j < nc; j++) {
                rc |= 1ull << (op.qubits[j] + n);
                cc |= 1ull << op.qubits[j];
            }
            for (int j = 0; j < nt; j++) {
                rt[j] = op.qubits[nc + j] + n;
                ct[j] = op.qubits[nc + j];
            }
            const int d = 1 << nt;
            std::vector<double> cm = conj_of(op.matrix, d * d);
            if (nt > kRegBits) {  // wide: U on the row qubits, conj(U) on the column qubits
                out.push_back(make_wide(rt, nt, rc, op.matrix));
                out.push_back(make_wide(ct, nt, cc, cm.data()));
                continue;
            }
            lower_matrix(rt, nt, rc, op.matrix, out);
            lower_matrix(ct, nt, cc, cm.data(), out);
            continue;
        }
        if (k > 2) {
            err = "op " + std::to_string(i) + ": channels act on 1 or 2 qubits";
            return GOQU_ERR_UNSUPPORTED;
        }
        const int rq[2] = {op.qubits[0] + n, k == 2 ? op.qubits[1] + n : 0};
        const int cq[2] = {op.qubits[0], k == 2 ? op.qubits[1] : 0};
        const int d = 1 << k;
        if (op.kind == GOQU_OP_RESET) {
            const double S[32] = {1, 0, 0, 0, 0, 0, 1, 0};  // rho00 <- rho00 + rho11
            const int q2[2] = {rq[0], cq[0]};
            out.push_back(make_superop(q2, 2, S));
            continue;
        }
        const bool channel = op.kind == GOQU_OP_CHANNEL && op.num_kraus > 1;
        if (op.kind != GOQU_OP_GATE && op.kind != GOQU_OP_CHANNEL) {
            err = "op " + std::to_string(i) + ": kind not valid for a density matrix";
            return GOQU_ERR_INVALID_ARG;
        }
        if (!op.matrix) continue;  // sim.go:147-150
        if (!channel) {  // unitary gate or single-Kraus channel: two passes
            lower_matrix(rq, k, 0, op.matrix, out);
            std::vector<double> cm = conj_of(op.matrix, d * d);
            lower_matrix(cq, k, 0, cm.data(), out);
            continue;
        }
        // superoperator S[(r',c'),(r,c)] = sum_e E[r',r] conj(E[c',c]); index = r*d + c
        const int D = d * d;
        std::vector<cd> S((size_t)D * D, cd(0, 0));
        for (int e = 0; e < op.num_kraus; e++) {
            const double *E = op.matrix + (size_t)2 * d * d * e;
            for (int r1 = 0; r1 < d; r1++)
                for (int c1 = 0; c1 < d; c1++)
                    for (int r0 = 0; r0 < d; r0++)
                        for (int c0 = 0; c0 < d; c0++)
                            S[(size_t)(r1 * d + c1) * D + (r0 * d + c0)] +=
                                at(E, r1 * d + r0) * std::conj(at(E, c1 * d + c0));
        }
        std::vector<double> Sm(2 * (size_t)D * D);
        for (size_t j = 0; j < S.size(); j++) {
            Sm[2 * j] = S[j].real();
            Sm[2 * j + 1] = S[j].imag();
        }
        int q4[4];
        if (k == 1) {
            q4[0] = rq[0];
            q4[1] = cq[0];
        } else {
            q4[0] = rq[0];
            q4[1] = rq[1];
            q4[2] = cq[0];
            q4[3] = cq[1];
        }
        out.push_back(make_superop(q4, 2 * k, Sm.data()));
    }
    merge_channels(out);
    return GOQU_OK;
}

// ---- peephole fusion of uncontrolled single-qubit ops on the same qubit ------------------
static void as_matrix2(const EOp &e, cd M[4]) {
    if (e.kind == EOP_DIAG) {
        M[0] = at(e.m.data(), 0);
        M[1] = 0;
        M[2] = 0;
        M[3] = at(e.m.data(), 1);
    } else if (e.kind == EOP_ANTI1) {
        M[0] = 0;
        M[1] = at(e.m.data(), 0);
        M[2] = at(e.m.data(), 1);
        M[3] = 0;
    } else {
        for (int i = 0; i < 4; i++) M[i] = at(e.m.data(), i);
    }
}

static bool fusable1(const EOp &e) {
    return e.cmask == 0 && e.k == 1 && (e.kind == EOP_DENSE || e.kind == EOP_DIAG || e.kind == EOP_ANTI1);
}

static std::vector<EOp> fuse_single_qubit(const std::vector<EOp> &in) {
    std::vector<EOp> out;
    out.reserve(in.size());
    int last[64];
    for (int i = 0; i < 64; i++) last[i] = -1;
    for (const EOp &e : in) {
        if (e.is_barrier()) {  // nothing fuses across a reset, on any qubit
            out.push_back(e);
            for (int i = 0; i < 64; i++) last[i] = (int)out.size() - 1;
            continue;
        }
        if (fusable1(e)) {
            int l = last[e.q[0]];
            if (l >= 0 && fusable1(out[l]) && out[l].q[0] == e.q[0]) {
                cd A[4], B[4], P[4];
                as_matrix2(out[l], A);
                as_matrix2(e, B);
                P[0] = B[0] * A[0] + B[1] * A[2];
                P[1] = B[0] * A[1] + B[1] * A[3];
                P[2] = B[2] * A[0] + B[3] * A[2];
                P[3] = B[2] * A[1] + B[3] * A[3];
                double pm[8];
                for (int i = 0; i < 4; i++) {
                    pm[2 * i] = P[i].real();
                    pm[2 * i + 1] = P[i].imag();
                }
                std::vector<EOp> tmp;
                lower_matrix(e.q, 1, 0, pm, tmp);
                out[l] = tmp[0];
                continue;
            }
        }
        out.push_back(e);
        const int idx = (int)out.size() - 1;
        uint64_t t = e.touch();
        while (t) {
            int b = __builtin_ctzll(t);
            t &= t - 1;
            last[b] = idx;
        }
    }
    return out;
}

// A unitary single-qubit matrix U = e^{ia} diag(1, e^{i phi}) R diag(1, e^{i lam}) with the real
// rotation R = [[r, -s], [s, r]], r = |u00|, s = |u10| (an angle in [0, pi/2]).  An uncontrolled
// one-qubit dense op becomes DIAG, real DENSE, DIAG: the emission runs the rotation as in-place
// shears and folds the diagonals into pending per-qubit phases applied in batches.
static std::vector<EOp> split_rotations(const std::vector<EOp> &in) {
    std::vector<EOp> out;
    out.reserve(in.size() * 3);
    for (const EOp &e : in) {
        if (!(e.kind == EOP_DENSE && e.k == 1 && e.cmask == 0)) {
            out.push_back(e);
            continue;
        }
        const cd u00(e.m[0], e.m[1]), u01(e.m[2], e.m[3]), u10(e.m[4], e.m[5]), u11(e.m[6], e.m[7]);
        const double r = std::abs(u00), s = std::abs(u10);
        // only unitary matrices with both a diagonal and an off-diagonal part
        const bool unitary = std::abs(std::norm(u00) + std::norm(u10) - 1.0) < 1e-12 &&
                             std::abs(std::norm(u01) + std::norm(u11) - 1.0) < 1e-12 &&
                             std::abs(std::conj(u00) * u01 + std::conj(u10) * u11) < 1e-12;
        if (!unitary || r < 1e-150 || s < 1e-150) {
            out.push_back(e);
            continue;
        }
        const cd ea = u00 / r, eaphi = u10 / s, ealam = -u01 / s;
        EOp b;  // applied first: e^{ia} diag(1, e^{i lam})
        b.kind = EOP_DIAG;
        b.k = 1;
        b.q[0] = e.q[0];
        push(b.m, ea);
        push(b.m, ealam);
        EOp rot;
        rot.kind = EOP_DENSE;
        rot.k = 1;
        rot.q[0] = e.q[0];
        rot.m = {r, 0.0, -s, 0.0, s, 0.0, r, 0.0};
        EOp a;  // applied last: diag(1, e^{i phi})
        a.kind = EOP_DIAG;
        a.k = 1;
        a.q[0] = e.q[0];
        push(a.m, cd(1, 0));
        push(a.m, eaphi / ea);
        // real orthogonal inputs (RY, H, ...) give diagonals of +-1; exact identities are dropped
        if (b.m != std::vector<double>{1.0, 0.0, 1.0, 0.0}) out.push_back(std::move(b));
        out.push_back(std::move(rot));
        if (a.m != std::vector<double>{1.0, 0.0, 1.0, 0.0}) out.push_back(std::move(a));
        (void)u11;
    }
    return out;
}

// ---- stage planning -----------------------------------------------------------------------
namespace {

struct StageOps {
    uint64_t T = 0;
    std::vector<int> ops;
    int ngrad = 0;
};

// Register allocation inside one stage.  Emits instructions into prog.
struct Layout {
    int m = 0;
    int lay[24];  // lay[0..3] register local bits; lay[4..m-1] thread local bits
};

// Thread bits in ascending order of the global qubits the local bits stand for (qmap): lanes
// 0..2 then sit on the lowest tile qubits, which keeps global loads/stores in contiguous runs.
// pref: up to two local bits for thread bits 4 and 5, the lane bits that trade places with a
// register slot by one permlane swap per word (bits already on lanes 0..2 stay there).
static Layout make_layout(int m, const int reg[4], const int *qmap, const int pref[2] = nullptr) {
    Layout L;
    L.m = m;
    for (int s = 0; s < kRegBits; s++) L.lay[s] = reg[s];
    int u = kRegBits;
    for (int b = 0; b < m; b++) {
        bool in = false;
        for (int s = 0; s < kRegBits; s++) in |= (reg[s] == b);
        if (!in) L.lay[u++] = b;
    }
    std::sort(L.lay + kRegBits, L.lay + m, [&](int x, int y) { return qmap[x] < qmap[y]; });
    if (pref && m - kRegBits > 5) {
        for (int k = 0; k < 2; k++) {
            if (pref[k] < 0) continue;
            int at = -1;
            for (int t = kRegBits; t < m; t++)
                if (L.lay[t] == pref[k]) at = t;
            if (at < kRegBits + 3) continue;  // a register slot or lanes 0..2
            const int want = kRegBits + 4 + k;
            std::swap(L.lay[at], L.lay[want]);
        }
    }
    return L;
}

}  // namespace

// Runs of consecutive I_SHEAR instructions on distinct register slots (from index `from` on)
// become one I_SHEARS: shears on different slots act on disjoint register pairs, so they
// commute and the merged op performs the same fused multiply-adds under one dispatch.
static void merge_shears(Program &P, int from) {
    std::vector<Instr> out;
    out.reserve(P.instrs.size() - (size_t)from);
    size_t i = (size_t)from;
    while (i < P.instrs.size()) {
        size_t j = i;
        uint32_t slots = 0;
        while (j < P.instrs.size() && P.instrs[j].op == I_SHEAR && !(slots >> P.instrs[j].s[0] & 1)) {
            slots |= 1u << P.instrs[j].s[0];
            j++;
        }
        if (j - i < 2) {
            out.push_back(P.instrs[i]);
            i++;
            continue;
        }
        Instr M;
        std::memset(&M, 0, sizeof(M));
        M.op = I_SHEARS;
        M.pat = 0xFFFFu;
        M.smask = (int32_t)slots;
        M.mat = (int32_t)(P.pool.size() / 2);
        double coef[12] = {0};
        for (size_t t = i; t < j; t++) {
            const Instr &S = P.instrs[t];
            M.k |= S.k << (2 * S.s[0]);
            for (int c = 0; c < 3; c++) coef[3 * S.s[0] + c] = shear_coef(S, c);
        }
        P.pool.insert(P.pool.end(), coef, coef + 12);
        out.push_back(M);
        i = j;
    }
    P.instrs.resize((size_t)from);
    P.instrs.insert(P.instrs.end(), out.begin(), out.end());
}

Program plan_program(const std::vector<EOp> &eops_in, int nq, const PlanConfig &cfg) {
    Program P;
    P.nq = nq;
    std::vector<EOp> eops = split_rotations(fuse_single_qubit(eops_in));
    P.num_eops = (int64_t)eops.size();
    const int m = std::min(cfg.tile_bits, nq);
    const int L = std::min(cfg.low_bits, m - kRegBits);
    const uint64_t all = (nq >= 64) ? ~0ull : ((1ull << nq) - 1);
    const uint64_t lowmask = (1ull << L) - 1;

    // 1. greedy stage formation
    std::vector<char> done(eops.size(), 0);
    std::vector<StageOps> stages;
    size_t first_open = 0;
    while (first_open < eops.size()) {
        StageOps S;
        S.T = lowmask;
        // qubits on which deferred ops act non-diagonally / only diagonally: a later op may pass
        // a deferred one when every qubit they share is diagonal for both (they commute)
        uint64_t b_nd = 0, b_dg = 0;
        for (size_t i = first_open; i < eops.size(); i++) {
            if (done[i]) continue;
            const EOp &e = eops[i];
            const uint64_t nd = e.is_barrier() ? all : e.nondiag(), dg = e.is_barrier() ? 0 : e.diag(),
                           tg = e.targets();
            const bool conflict = (nd & (b_nd | b_dg)) || (dg & b_nd);
            bool take = false;
            if (!conflict && !(e.is_grad() && S.ngrad >= kMaxGradPerStage)) {
                const uint64_t nt = S.T | tg;
                if (__builtin_popcountll(nt) <= m) {
                    S.T = nt;
                    S.ops.push_back((int)i);
                    S.ngrad += e.is_grad();
                    done[i] = 1;
                    take = true;
                }
            }
            if (!take) {
                b_nd |= nd;
                b_dg |= dg;
            }
            if ((b_nd & all) == all) break;
        }
        // fill the tile up to m qubits with the lowest unused qubits (longer contiguous runs)
        for (int b = 0; b < nq && __builtin_popcountll(S.T) < m; b++) S.T |= 1ull << b;
        stages.push_back(S);
        while (first_open < eops.size() && done[first_open]) first_open++;
    }

    // 2. per stage: local numbering, register allocation, instruction emission
    std::vector<cd> ph(nq > 0 ? nq : 1, cd(1, 0));  // pending phase diag(1, ph[q]) per qubit
    uint64_t pending = 0;
    cd gphase(1, 0);
    // Rotations may leave a real diagonal scale pending (two shears instead of three) unless the
    // program takes inner products on the way (gradient ops need the state at its true norm).
    bool defer_scale = true;
    for (const EOp &e : eops) defer_scale = defer_scale && e.kind != EOP_GRAD;
    size_t stage_index = 0;
    for (const StageOps &S : stages) {
        StageHeader H;
        std::memset(&H, 0, sizeof(H));
        H.m = m;
        H.instr_off = (int32_t)P.instrs.size();
        H.pool_off = (int32_t)(P.pool.size() / 2);
        H.grad_base = P.num_grad;
        H.ngrad = 0;
        // loc[q]: the local bit holding global qubit q; qmap[j]: the global qubit local bit j
        // holds.  Both start as the ascending tile order; an uncontrolled SWAP of two tile qubits
        // exchanges them instead of moving amplitudes, and the stage stores through qmap.
        int loc[64], qmap[16];
        for (int i = 0; i < 64; i++) loc[i] = -1;
        {
            int j = 0;
            for (int b = 0; b < nq; b++)
                if (S.T >> b & 1) {
                    H.T[j] = (uint8_t)b;
                    qmap[j] = b;
                    loc[b] = j++;
                }
        }
        // local bits holding global qubits 0..2 (the coalescing lanes of loads and stores)
        auto low_local = [&]() {
            uint32_t lw = 0;
            if (m - kRegBits >= 3)
                for (int q = 0; q < 3 && q < nq; q++) lw |= 1u << loc[q];
            return lw;
        };
        // targets (local bits) needed by each op, for register choice and look-ahead
        const size_t nops = S.ops.size();
        std::vector<uint32_t> need(nops, 0);
        for (size_t i = 0; i < nops; i++) {
            uint64_t t = eops[S.ops[i]].targets();
            uint32_t l = 0;
            while (t) {
                int b = __builtin_ctzll(t);
                t &= t - 1;
                l |= 1u << loc[b];
            }
            need[i] = l;
        }
        // List scheduling: an op is ready when it commutes with every earlier op of the stage
        // still waiting; ready ops whose targets are register slots run first, and a remap is
        // emitted only when no ready op can run.
        std::vector<char> fin(nops, 0);
        // Readiness per qubit: op i may run when, on every qubit it acts on non-diagonally, no
        // earlier unfinished op touches that qubit, and on every qubit it acts on diagonally, no
        // earlier unfinished op acts non-diagonally (ops commute when all shared qubits are
        // diagonal for both).  Per-qubit op lists with lazily advanced heads make a check cost
        // O(qubits of the op).
        // A reset (a barrier) counts as acting non-diagonally on every qubit: it waits for every
        // earlier op and every later op waits for it.
        auto sched_touch = [&](const EOp &x) { return x.is_barrier() ? all : x.touch(); };
        auto sched_nd = [&](const EOp &x) { return x.is_barrier() ? all : x.nondiag(); };
        std::vector<std::vector<uint32_t>> touch_list(64), nd_list(64);
        std::vector<size_t> touch_head(64, 0), nd_head(64, 0);
        for (size_t i = 0; i < nops; i++) {
            const EOp &x = eops[S.ops[i]];
            for (uint64_t t = sched_touch(x); t; t &= t - 1) touch_list[__builtin_ctzll(t)].push_back((uint32_t)i);
            for (uint64_t t = sched_nd(x); t; t &= t - 1) nd_list[__builtin_ctzll(t)].push_back((uint32_t)i);
        }
        auto first_open = [&](const std::vector<uint32_t> &L, size_t &h) -> size_t {
            while (h < L.size() && fin[L[h]]) h++;
            return h < L.size() ? L[h] : nops;
        };
        auto ready = [&](size_t i) {
            const EOp &x = eops[S.ops[i]];
            const uint64_t nd = sched_nd(x);
            for (uint64_t t = sched_touch(x); t; t &= t - 1) {
                const int q = __builtin_ctzll(t);
                const size_t f = (nd >> q & 1) ? first_open(touch_list[q], touch_head[q]) : first_open(nd_list[q], nd_head[q]);
                if (f < i) return false;
            }
            return true;
        };
        // new register set: targets of the first ready op that needs registers, then of the
        // other ready ops, then of the waiting ops in order; never the local bits of global
        // qubits 0..2 unless an op needs them
        auto choose = [&](int out[4], bool initial) {
            const uint32_t avoid = low_local();
            uint32_t set = 0;
            for (int pass = 0; pass < 2 && __builtin_popcount(set) < kRegBits; pass++)
                for (size_t i = 0; i < nops && __builtin_popcount(set) < kRegBits; i++) {
                    if (fin[i] || !need[i]) continue;
                    if (pass == 0 && !ready(i)) continue;
                    const uint32_t cand = need[i] & ~set & (initial ? ~avoid : ~0u);
                    if (initial && (need[i] & avoid)) continue;
                    if (__builtin_popcount(set | cand) <= kRegBits) set |= cand;
                }
            for (int b = m - 1; b >= 0 && __builtin_popcount(set) < kRegBits; b--)
                if (!(set >> b & 1) && !(avoid >> b & 1)) set |= 1u << b;
            for (int b = m - 1; b >= 0 && __builtin_popcount(set) < kRegBits; b--) set |= 1u << b;
            int s = 0;
            for (int b = 0; b < m; b++)
                if (set >> b & 1) out[s++] = b;
        };
        int reg[4];
        choose(reg, true);
        Layout cur = make_layout(m, reg, qmap);
        for (int i = 0; i < m; i++) {
            H.lay0[i] = (uint8_t)cur.lay[i];
            H.gq0[i] = H.T[cur.lay[i]];
        }

        auto slot_of = [&](int localbit) {
            for (int s = 0; s < kRegBits; s++)
                if (cur.lay[s] == localbit) return s;
            return -1;
        };
        // thread-id bit holding a local bit in the current layout (-1: a register slot)
        auto tbit_of = [&](int localbit) {
            for (int u = kRegBits; u < m; u++)
                if (cur.lay[u] == localbit) return u - kRegBits;
            return -1;
        };
        auto emit_remap = [&](const int nreg[4], const int *pref = nullptr) {
            cur = make_layout(m, nreg, qmap, pref);
            Instr I;
            std::memset(&I, 0, sizeof(I));
            I.op = I_REMAP;
            I.flags = F_OUTER;
            for (int i = 0; i < m; i++) I.lay[i] = (uint8_t)cur.lay[i];
            P.instrs.push_back(I);
        };
        // Register slot s trades places with thread-id bit u < 6 (a lane bit) by cross-lane moves.
        // Costs relative to a permlane swap (one instruction per word and register pair): a
        // lane bit below 4 takes four instructions per word (two DPP moves, two selects), an LDS
        // remap about ten (LDS bandwidth bound), whatever its size.
        auto xswap_cost = [&](int u) { return u >= 4 ? 1 : 4; };
        constexpr int kXswapBudget = 8;
        auto emit_xswap = [&](int sl, int u) {
            std::swap(cur.lay[sl], cur.lay[kRegBits + u]);
            Instr I;
            std::memset(&I, 0, sizeof(I));
            I.op = I_XSWAP;
            I.flags = F_OUTER;
            I.k = u;
            I.pat = 0xFFFFu;
            I.s[0] = (uint8_t)sl;
            for (int i = 0; i < m; i++) I.lay[i] = (uint8_t)cur.lay[i];
            P.instrs.push_back(I);
        };
        // reach the register set nreg by lane exchanges when they cost less than an LDS remap
        auto try_xswaps = [&](const int nreg[4]) {
            uint32_t rs = 0, ns = 0;
            for (int s2 = 0; s2 < kRegBits; s2++) {
                rs |= 1u << cur.lay[s2];
                ns |= 1u << nreg[s2];
            }
            const uint32_t in = ns & ~rs, out = rs & ~ns;
            int cost = 0;
            for (uint32_t t = in; t; t &= t - 1) {
                int u = -1;
                for (int k = kRegBits; k < m; k++)
                    if (cur.lay[k] == __builtin_ctz(t)) u = k - kRegBits;
                if (u < 0 || u > 5) return false;
                cost += xswap_cost(u);
            }
            if (cost > kXswapBudget) return false;
            uint32_t o = out;
            for (uint32_t t = in; t; t &= t - 1) {
                const int b = __builtin_ctz(t), ob = __builtin_ctz(o);
                o &= o - 1;
                int sl = -1, u = -1;
                for (int k = 0; k < kRegBits; k++)
                    if (cur.lay[k] == ob) sl = k;
                for (int k = kRegBits; k < m; k++)
                    if (cur.lay[k] == b) u = k - kRegBits;
                emit_xswap(sl, u);
            }
            return true;
        };
        // controls resolved against the current layout: register pattern, thread-id mask and
        // mask of qubits outside the tile
        auto set_controls = [&](Instr &I, uint64_t cm) {
            uint32_t pat = 0xFFFFu;
            while (cm) {
                const int q = __builtin_ctzll(cm);
                cm &= cm - 1;
                if (loc[q] < 0) {
                    I.omask |= 1ull << q;
                    continue;
                }
                const int sl = slot_of(loc[q]);
                if (sl >= 0) {
                    uint32_t pm = 0;
                    for (int r = 0; r < kRegAmps; r++)
                        if (r >> sl & 1) pm |= 1u << r;
                    pat &= pm;
                } else {
                    I.tmask |= 1u << tbit_of(loc[q]);
                }
            }
            I.pat = pat;
            if (pat != 0xFFFFu) I.flags |= F_PAT;
            if (I.tmask) I.flags |= F_TCTRL;
            if (I.omask) I.flags |= F_TILE;
        };

        // one elementary op -> one instruction, resolved against the current layout
        auto emit_eop = [&](const EOp &e) {
                Instr I;
                std::memset(&I, 0, sizeof(I));
                I.k = e.k;
                I.mat = (int32_t)(P.pool.size() / 2);
                I.pat = 0xFFFFu;
                switch (e.kind) {
                case EOP_DIAG: {
                    I.op = I_DIAG;
                    // register slots of the diagonal qubits and of the controls index the table; the
                    // other diagonal qubits become lane index bits, the other controls lane masks
                    int rm = 0, ctl_slots = 0, kl = 0, lane_j[4] = {0, 0, 0, 0};
                    for (int j = 0; j < e.k; j++) {
                        const int q = e.q[j];
                        const int sl = loc[q] >= 0 ? slot_of(loc[q]) : -1;
                        if (sl >= 0) {
                            rm |= 1 << sl;
                        } else {
                            I.ls[kl] = (uint8_t)(loc[q] < 0 ? 16 + q : tbit_of(loc[q]));
                            lane_j[kl++] = j;
                        }
                    }
                    uint64_t cm = e.cmask;
                    while (cm) {
                        const int q = __builtin_ctzll(cm);
                        cm &= cm - 1;
                        if (loc[q] < 0) {
                            I.omask |= 1ull << q;
                            continue;
                        }
                        const int sl = slot_of(loc[q]);
                        if (sl >= 0) {
                            rm |= 1 << sl;
                            ctl_slots |= 1 << sl;
                        } else {
                            I.tmask |= 1u << tbit_of(loc[q]);
                        }
                    }
                    if (I.tmask) I.flags |= F_TCTRL;
                    if (I.omask) I.flags |= F_TILE;
                    I.k = kl;
                    I.smask = rm;
                    // 16 entries per lane index, one per register index r
                    uint32_t nontriv = 0;
                    std::vector<double> tab((size_t)2 * 16 << kl);
                    for (int li = 0; li < (1 << kl); li++)
                        for (int r = 0; r < kRegAmps; r++) {
                            bool ok = true;
                            for (int sl = 0; sl < kRegBits; sl++)
                                if ((ctl_slots >> sl & 1) && !(r >> sl & 1)) ok = false;
                            int idx = 0;
                            for (int j = 0; j < e.k; j++) {
                                const int q = e.q[j];
                                const int sl = loc[q] >= 0 ? slot_of(loc[q]) : -1;
                                int bit = 0;
                                if (sl >= 0) {
                                    bit = (r >> sl) & 1;
                                } else {
                                    for (int p = 0; p < kl; p++)
                                        if (lane_j[p] == j) bit = (li >> (kl - 1 - p)) & 1;
                                }
                                idx = (idx << 1) | bit;
                            }
                            const double re = ok ? e.m[2 * idx] : 1.0, im = ok ? e.m[2 * idx + 1] : 0.0;
                            const int at = (li << kRegBits) | r;
                            tab[2 * at] = re;
                            tab[2 * at + 1] = im;
                            if (re != 1.0 || im != 0.0) nontriv |= 1u << r;
                        }
                    // Lane index bits that only gate the table (every entry with the bit clear is 1)
                    // become controls: thread bits join tmask, bits of the tile base join omask.
                    // Controlled phases between lane qubits (CZ, CP) then need no per-lane table.
                    for (int p = kl - 1; p >= 0; p--) {
                        const int bit = kl - 1 - p;  // ls[0] is the most significant index bit
                        bool gate = true;
                        for (int li = 0; li < (1 << kl) && gate; li++)
                            if (!(li >> bit & 1))
                                for (int r = 0; r < kRegAmps && gate; r++)
                                    gate = tab[2 * ((li << kRegBits) | r)] == 1.0 && tab[2 * ((li << kRegBits) | r) + 1] == 0.0;
                        if (!gate) continue;
                        std::vector<double> kept;
                        for (int li = 0; li < (1 << kl); li++)
                            if (li >> bit & 1)
                                kept.insert(kept.end(), tab.begin() + 2 * (li << kRegBits), tab.begin() + 2 * ((li + 1) << kRegBits));
                        tab.swap(kept);
                        const int src = I.ls[p];
                        if (src < 16) I.tmask |= 1u << src;
                        else I.omask |= 1ull << (src - 16);
                        for (int j = p; j + 1 < kl; j++) I.ls[j] = I.ls[j + 1];
                        I.ls[--kl] = 0;
                    }
                    if (I.tmask) I.flags |= F_TCTRL;
                    if (I.omask) I.flags |= F_TILE;
                    I.k = kl;
                    nontriv = 0;
                    for (int li = 0; li < (1 << kl); li++)
                        for (int r = 0; r < kRegAmps; r++)
                            if (tab[2 * ((li << kRegBits) | r)] != 1.0 || tab[2 * ((li << kRegBits) | r) + 1] != 0.0)
                                nontriv |= 1u << r;
                    I.pat = nontriv;
                    P.pool.insert(P.pool.end(), tab.begin(), tab.end());
                    break;
                }
                case EOP_ANTI1:
                case EOP_RESET: {
                    I.op = e.kind == EOP_ANTI1 ? I_ANTI1 : I_RESET;
                    I.s[0] = (uint8_t)slot_of(loc[e.q[0]]);
                    if (e.pure_swap) I.flags |= F_PURE;
                    if (e.kind == EOP_RESET) I.flags |= F_OUTER;
                    set_controls(I, e.cmask);
                    P.pool.insert(P.pool.end(), e.m.begin(), e.m.end());
                    break;
                }
                case EOP_SWAP2: {
                    I.op = I_SWAP2;
                    const int s0 = slot_of(loc[e.q[0]]), s1 = slot_of(loc[e.q[1]]);
                    I.s[0] = (uint8_t)std::min(s0, s1);
                    I.s[1] = (uint8_t)std::max(s0, s1);
                    if (e.pure_swap) I.flags |= F_PURE;
                    set_controls(I, e.cmask);
                    P.pool.insert(P.pool.end(), e.m.begin(), e.m.end());
                    break;
                }
                case EOP_PARITY: {
                    // blocks in the kernel's canonical form (see I_PARITY in engine.h): for each class
                    // and element, the slot bits, then the row/column bits they carry, then the
                    // superoperator index (q[0] = MSB)
                    I.op = I_PARITY;
                    int slot[4];
                    for (int j = 0; j < e.k; j++) slot[j] = slot_of(loc[e.q[j]]);
                    const int D = 1 << e.k;
                    auto entry = [&](int row_idx, int col_idx) { return cd(e.m[2 * (row_idx * D + col_idx)], e.m[2 * (row_idx * D + col_idx) + 1]); };
                    // superoperator index of the element whose register-slot bits are `bits`
                    auto sup_index = [&](int bits) {
                        int idx = 0;
                        for (int j = 0; j < e.k; j++) idx = (idx << 1) | ((bits >> slot[j]) & 1);
                        return idx;
                    };
                    std::vector<cd> blocks;
                    if (e.k == 2) {
                        const int lo = std::min(slot[0], slot[1]), hi = std::max(slot[0], slot[1]);
                        I.smask = (1 << lo) | (1 << hi);
                        for (int p = 0; p < 2; p++)
                            for (int i = 0; i < 2; i++)
                                for (int j = 0; j < 2; j++) {
                                    const int bi = (i << lo) | ((i ^ p) << hi), bj = (j << lo) | ((j ^ p) << hi);
                                    blocks.push_back(entry(sup_index(bi), sup_index(bj)));
                                }
                    } else {
                        // pairs (row a, col a) and (row b, col b); s0 = 0 and its partner, then the other two
                        const int pa_slots[2] = {slot[0], slot[2]}, pb_slots[2] = {slot[1], slot[3]};
                        const bool a_has0 = pa_slots[0] == 0 || pa_slots[1] == 0;
                        const int *P0 = a_has0 ? pa_slots : pb_slots, *P1 = a_has0 ? pb_slots : pa_slots;
                        const int s0 = 0, s1 = P0[0] == 0 ? P0[1] : P0[0];
                        const int s2 = std::min(P1[0], P1[1]), s3 = std::max(P1[0], P1[1]);
                        I.s[0] = (uint8_t)s0;
                        I.s[1] = (uint8_t)s1;
                        I.s[2] = (uint8_t)s2;
                        I.s[3] = (uint8_t)s3;
                        I.smask = 15;
                        for (int c = 0; c < 4; c++) {
                            const int pa = c >> 1, pb = c & 1;
                            auto bits_of = [&](int i) {
                                const int ia = i >> 1, ib = i & 1;
                                return (ia << s0) | ((ia ^ pa) << s1) | (ib << s2) | ((ib ^ pb) << s3);
                            };
                            for (int i = 0; i < 4; i++)
                                for (int j = 0; j < 4; j++) blocks.push_back(entry(sup_index(bits_of(i)), sup_index(bits_of(j))));
                        }
                    }
                    bool real = true;
                    for (const cd &x : blocks) {
                        real = real && x.imag() == 0.0;
                        P.pool.push_back(x.real());
                        P.pool.push_back(x.imag());
                    }
                    if (real) I.flags |= F_REAL;
                    if (e.side == 1) I.flags |= F_SIDE_A;
                    if (e.side == 2) I.flags |= F_SIDE_B;
                    break;
                }
                case EOP_GRAD:
                case EOP_DENSE: {
                    I.op = e.kind == EOP_GRAD ? I_GRAD : I_DENSE;
                    if (e.kind == EOP_GRAD) {
                        I.pad = H.ngrad++;
                        P.grad_slots.push_back(e.slot);
                        P.num_grad++;
                    }
                    const int k = e.k, dim = 1 << k;
                    int slot[4], smask = 0;
                    for (int j = 0; j < k; j++) {
                        slot[j] = slot_of(loc[e.q[j]]);
                        smask |= 1 << slot[j];
                    }
                    // rank of each gate qubit among the sorted slots: kernel index bit rank[j]
                    int rank[4];
                    for (int j = 0; j < k; j++) {
                        rank[j] = 0;
                        for (int t = 0; t < k; t++)
                            if (slot[t] < slot[j]) rank[j]++;
                    }
                    auto perm = [&](int g) {
                        int r = 0;
                        for (int j = 0; j < k; j++)
                            if (g >> (k - 1 - j) & 1) r |= 1 << rank[j];
                        return r;
                    };
                    std::vector<double> km(2 * dim * dim);
                    for (int gr = 0; gr < dim; gr++)
                        for (int gc = 0; gc < dim; gc++) {
                            const int kr = perm(gr), kc = perm(gc);
                            km[2 * (kr * dim + kc)] = e.m[2 * (gr * dim + gc)];
                            km[2 * (kr * dim + kc) + 1] = e.m[2 * (gr * dim + gc) + 1];
                        }
                    I.smask = smask;
                    set_controls(I, e.cmask);
                    bool real = true;
                    for (size_t t = 1; t < e.m.size(); t += 2) real = real && e.m[t] == 0.0;
                    if (real) I.flags |= F_REAL;
                    if (e.kind == EOP_DENSE && k > 2) {
                        I.flags |= F_OUTER;
                        H.flags |= SF_BIG;
                    }
                    P.pool.insert(P.pool.end(), km.begin(), km.end());
                    break;
                }

                }
                P.instrs.push_back(I);
            };
        // Pending phases: an uncontrolled one-qubit diagonal op is not emitted; its phase joins
        // ph[q] (diag(1, ph[q]), the rest goes to the global phase).  Phases are applied only
        // before an op that acts non-diagonally on their qubit, together with the pending phases
        // of the qubits in registers at that moment (one diagonal op over up to four qubits).
        auto emit_phase_group = [&](uint64_t group, cd scale) {
            EOp d;
            d.kind = EOP_DIAG;
            d.k = 0;
            for (uint64_t t = group; t; t &= t - 1) d.q[d.k++] = __builtin_ctzll(t);
            for (int idx = 0; idx < (1 << d.k); idx++) {
                cd v = scale;
                for (int j = 0; j < d.k; j++)
                    if (idx >> (d.k - 1 - j) & 1) v *= ph[d.q[j]];
                d.m.push_back(v.real());
                d.m.push_back(v.imag());
            }
            for (int j = 0; j < d.k; j++) ph[d.q[j]] = cd(1, 0);
            pending &= ~group;
            emit_eop(d);
        };
        auto next_group = [&](uint64_t must) {
            uint64_t group = 0;
            int cnt = 0;
            for (uint64_t t = must; t && cnt < 4; t &= t - 1, cnt++) group |= t & (~t + 1);
            for (int sl = 0; sl < kRegBits && cnt < 4; sl++) {
                const int qq = qmap[cur.lay[sl]];
                if ((pending >> qq & 1) && !(group >> qq & 1)) {
                    group |= 1ull << qq;
                    cnt++;
                }
            }
            return group;
        };
        // the global factor joins a phase group once its magnitude leaves [2^-30, 2^30], so the
        // stored amplitudes stay within a bounded factor of their true values
        auto flush = [&](uint64_t must) {
            must &= pending;
            while (must) {
                const uint64_t g = next_group(must);
                const double mag = std::abs(gphase);
                const bool fold = !(mag >= 0x1p-30 && mag <= 0x1p30);
                emit_phase_group(g, fold ? gphase : cd(1, 0));
                if (fold) gphase = cd(1, 0);
                must &= ~g;
            }
        };
        // an uncontrolled SWAP of two tile qubits: the two local bits trade global qubits
        auto relabel_swap = [&](const EOp &e) {
            const int qa = e.q[0], qb = e.q[1];
            std::swap(qmap[loc[qa]], qmap[loc[qb]]);
            std::swap(loc[qa], loc[qb]);
            std::swap(ph[qa], ph[qb]);
            const uint64_t ba = pending >> qa & 1, bb = pending >> qb & 1;
            pending = (pending & ~((1ull << qa) | (1ull << qb))) | (bb << qa) | (ba << qb);
            for (int q : {qa, qb})
                for (uint32_t i : touch_list[q]) {
                    if (fin[i]) continue;
                    uint32_t l = 0;
                    for (uint64_t t = eops[S.ops[i]].targets(); t; t &= t - 1) l |= 1u << loc[__builtin_ctzll(t)];
                    need[i] = l;
                }
        };
        auto is_relabel = [&](const EOp &e) { return e.kind == EOP_SWAP2 && e.cmask == 0 && e.pure_swap; };
        auto absorbable = [&](const EOp &e) {
            if (!(e.kind == EOP_DIAG && e.k == 1 && e.cmask == 0)) return false;
            const cd d0(e.m[0], e.m[1]), d1(e.m[2], e.m[3]);
            return std::abs(std::abs(d0) - 1.0) < 1e-13 && std::abs(std::abs(d1) - 1.0) < 1e-13;
        };
        auto absorb = [&](const EOp &e) {
            const cd d0(e.m[0], e.m[1]), d1(e.m[2], e.m[3]);
            const int q = e.q[0];
            gphase *= d0;
            ph[q] *= d1 / d0;
            if (ph[q] == cd(1, 0)) pending &= ~(1ull << q);
            else pending |= 1ull << q;
        };
        // uncontrolled real rotation [[r, -s], [s, r]] (r^2 + s^2 = 1)
        auto is_rotation = [&](const EOp &e) {
            if (!(e.kind == EOP_DENSE && e.k == 1 && e.cmask == 0)) return false;
            for (int t = 1; t < 8; t += 2)
                if (e.m[t] != 0.0) return false;
            return e.m[0] == e.m[6] && e.m[2] == -e.m[4] && std::abs(e.m[0] * e.m[0] + e.m[4] * e.m[4] - 1.0) < 1e-12;
        };
        // A rotation as in-place shears on its register slot.  With r >= 1/4 and a deferred
        // scale: R = diag(r, 1/r) L U with U = [[1, -s/r], [0, 1]], L = [[1, 0], [s r, 1]], and
        // diag(r, 1/r) = r diag(1, 1/r^2) joins the pending phases (errors grow at most by 1/r).
        // Otherwise the three-shear form with t = tan(angle/2) = s / (1 + r) <= 1:
        // R = [[1, -t], [0, 1]] [[1, 0], [s, 1]] [[1, -t], [0, 1]].  Pending phases on the qubit
        // have been applied before.
        auto emit_shear = [&](const EOp &e) {
            double r = e.m[0], s = e.m[4];
            if (r < 0) {  // R = -R' with R' an angle in (-pi/2, pi/2]
                r = -r;
                s = -s;
                gphase = -gphase;
            }
            Instr I;
            std::memset(&I, 0, sizeof(I));
            I.op = I_SHEAR;
            I.pat = 0xFFFFu;
            I.s[0] = (uint8_t)slot_of(loc[e.q[0]]);
            I.smask = 1 << I.s[0];
            if (defer_scale && r >= 0.25) {
                I.k = 2;
                set_shear_coefs(I, -s / r, s * r, 0.0);
                const int q = e.q[0];
                gphase *= r;
                ph[q] /= r * r;
                if (ph[q] == cd(1, 0)) pending &= ~(1ull << q);
                else pending |= 1ull << q;
            } else {
                const double t = s / (1.0 + r);
                I.k = 3;
                set_shear_coefs(I, -t, s, -t);
            }
            P.instrs.push_back(I);
        };
        size_t left = nops;
        size_t lo_open = 0;  // every op before it is finished
        while (left) {
            while (lo_open < nops && fin[lo_open]) lo_open++;
            // phases first (free, and they make the ops behind them ready together); the scan
            // looks a bounded window ahead
            bool any = false;
            for (size_t j = lo_open; j < nops && j < lo_open + 256; j++) {
                if (fin[j]) continue;
                const EOp &x = eops[S.ops[j]];
                const bool ab = absorbable(x), rl = !ab && is_relabel(x);
                if ((ab || rl) && ready(j)) {
                    fin[j] = 1;  // before relabel_swap, which refreshes the unfinished ops
                    if (ab) absorb(x);
                    else relabel_swap(x);
                    left--;
                    any = true;
                }
            }
            if (any) continue;
            // pick the first ready op that runs on the current registers
            uint32_t regset = 0;
            for (int s = 0; s < kRegBits; s++) regset |= 1u << cur.lay[s];
            size_t i = nops;
            for (size_t j = lo_open; j < nops; j++)
                if (!fin[j] && (need[j] & ~regset) == 0 && ready(j)) {
                    i = j;
                    break;
                }
            if (i == nops) {
                int nreg[4];
                choose(nreg, false);
                if (!try_xswaps(nreg)) {
                    // the next two local bits the waiting ops need go on lanes 4 and 5
                    uint32_t ns = 0;
                    for (int k = 0; k < kRegBits; k++) ns |= 1u << nreg[k];
                    const uint32_t lw = low_local();
                    int pref[2] = {-1, -1}, np = 0;
                    for (size_t j = lo_open; j < nops && np < 2; j++) {
                        if (fin[j]) continue;
                        for (uint32_t t = need[j] & ~ns & ~lw; t && np < 2; t &= t - 1) {
                            const int b = __builtin_ctz(t);
                            if (np == 0 || pref[0] != b) pref[np++] = b;
                        }
                    }
                    emit_remap(nreg, pref);
                }
                continue;
            }
            fin[i] = 1;
            left--;
            const EOp &e = eops[S.ops[i]];
            if (absorbable(e)) {
                absorb(e);
                continue;
            }
            if (is_relabel(e)) {
                relabel_swap(e);
                continue;
            }
            if (is_rotation(e)) {
                flush(e.nondiag());
                emit_shear(e);
                continue;
            }
            // Two uncontrolled one-qubit dense ops on different register slots run as one
            // two-slot op with the Kronecker product: the same 16 FMAs per amplitude, half the
            // register write-backs and one dispatch.
            if (e.kind == EOP_DENSE && e.k == 1 && e.cmask == 0) {
                size_t j = nops;
                for (size_t t = i + 1; t < nops; t++) {
                    if (fin[t]) continue;
                    const EOp &f = eops[S.ops[t]];
                    if (f.kind != EOP_DENSE || f.k != 1 || f.cmask != 0 || f.q[0] == e.q[0] || is_rotation(f)) continue;
                    if ((need[t] & ~regset) == 0 && ready(t)) {
                        j = t;
                        break;
                    }
                }
                if (j < nops) {
                    fin[j] = 1;
                    left--;
                    const EOp &f = eops[S.ops[j]];
                    flush(e.nondiag() | f.nondiag());
                    const int sa = slot_of(loc[e.q[0]]), sb = slot_of(loc[f.q[0]]);
                    const EOp &lo = sa < sb ? e : f, &hi = sa < sb ? f : e;
                    Instr I;
                    std::memset(&I, 0, sizeof(I));
                    I.op = I_DENSE;
                    I.k = 2;
                    I.mat = (int32_t)(P.pool.size() / 2);
                    I.pat = 0xFFFFu;
                    I.smask = (1 << sa) | (1 << sb);
                    bool real = true;
                    for (int t = 1; t < 8; t += 2) real = real && lo.m[t] == 0.0 && hi.m[t] == 0.0;
                    if (real) I.flags |= F_REAL;
                    for (int r = 0; r < 4; r++)
                        for (int c = 0; c < 4; c++) {  // bit 0 <-> lower slot, bit 1 <-> higher slot
                            const cd x = cd(lo.m[2 * ((r & 1) * 2 + (c & 1))], lo.m[2 * ((r & 1) * 2 + (c & 1)) + 1]) *
                                         cd(hi.m[2 * ((r >> 1) * 2 + (c >> 1))], hi.m[2 * ((r >> 1) * 2 + (c >> 1)) + 1]);
                            P.pool.push_back(x.real());
                            P.pool.push_back(x.imag());
                        }
                    P.instrs.push_back(I);
                    continue;
                }
            }
            // pending phases pass uncontrolled moves: a swap exchanges the two qubits' phases,
            // [a0, a1] <- [p0 a1, p1 a0] turns diag(1, f) into f diag(1, 1/f)
            if (e.is_barrier()) {
                // The reset keeps |(a0, a1)| of the state it sees: every pending phase and the
                // magnitude of the deferred global factor must be in the stored amplitudes first
                // (the reset's 1e-15 threshold compares true norms).  The global factor's phase
                // is discarded by the reset itself.
                const double mag = std::abs(gphase);
                bool scaled = false;
                for (uint64_t must = pending; must;) {
                    const uint64_t g = next_group(must);
                    emit_phase_group(g, scaled ? cd(1, 0) : cd(mag, 0));
                    scaled = true;
                    must &= ~g;
                }
                if (!scaled && mag != 1.0) {
                    ph[e.q[0]] = cd(1, 0);
                    emit_phase_group(1ull << e.q[0], cd(mag, 0));
                }
                gphase = cd(1, 0);
                emit_eop(e);
                continue;
            }
            if (e.kind == EOP_SWAP2 && e.cmask == 0) {
                const int qa = e.q[0], qb = e.q[1];
                std::swap(ph[qa], ph[qb]);
                const uint64_t ba = pending >> qa & 1, bb = pending >> qb & 1;
                pending = (pending & ~((1ull << qa) | (1ull << qb))) | (bb << qa) | (ba << qb);
            } else if (e.kind == EOP_ANTI1 && e.cmask == 0 && (pending >> e.q[0] & 1)) {
                gphase *= ph[e.q[0]];
                ph[e.q[0]] = 1.0 / ph[e.q[0]];
            } else {
                flush(e.nondiag());
            }
            emit_eop(e);
        }
        if (++stage_index == stages.size()) {  // the last stage applies every pending phase
            while (pending) {
                const uint64_t g = next_group(pending);
                const bool final_group = (pending & ~g) == 0;
                emit_phase_group(g, final_group ? gphase : cd(1, 0));
                if (final_group) gphase = cd(1, 0);
            }
            if (gphase != cd(1, 0)) {
                ph[qmap[cur.lay[0]]] = cd(1, 0);
                emit_phase_group(1ull << qmap[cur.lay[0]], gphase);
                gphase = cd(1, 0);
            }
        }
        // leave the tile with lanes 0..2 on global qubits 0..2 for coalesced stores: by lane
        // exchanges when those qubits sit in registers and cheap enough, else by an LDS remap in
        // which registers holding them trade places with the highest other local bits
        {
            const uint32_t lw = low_local();
            auto lanes_ok = [&]() {
                uint32_t l3 = 0;
                for (int k = 0; k < 3 && kRegBits + k < m; k++) l3 |= 1u << cur.lay[kRegBits + k];
                return l3 == lw;
            };
            if (lw && !lanes_ok()) {
                // plan exchanges: every register slot holding a low qubit with a lane 0..2 that does not
                int pairs[3][2], np = 0, cost = 0;
                uint32_t used = 0;
                for (int s2 = 0; s2 < kRegBits; s2++)
                    if (lw >> cur.lay[s2] & 1)
                        for (int u = 0; u < 3; u++)
                            if (!(used >> u & 1) && !(lw >> cur.lay[kRegBits + u] & 1)) {
                                used |= 1u << u;
                                pairs[np][0] = s2;
                                pairs[np][1] = u;
                                np++;
                                cost += xswap_cost(u);
                                break;
                            }
                Layout trial = cur;
                for (int k = 0; k < np; k++) std::swap(trial.lay[pairs[k][0]], trial.lay[kRegBits + pairs[k][1]]);
                uint32_t l3 = 0;
                for (int k = 0; k < 3; k++) l3 |= 1u << trial.lay[kRegBits + k];
                if (l3 == lw && cost <= kXswapBudget)
                    for (int k = 0; k < np; k++) emit_xswap(pairs[k][0], pairs[k][1]);
            }
            uint32_t rs = 0;
            for (int s = 0; s < kRegBits; s++) rs |= 1u << cur.lay[s];
            if (lw && !lanes_ok()) {
                uint32_t keep = rs & ~lw;
                for (int b = m - 1; b >= 0 && __builtin_popcount(keep) < kRegBits; b--)
                    if (!((rs | lw) >> b & 1)) keep |= 1u << b;
                int nreg[4], t = 0;
                for (int b = 0; b < m; b++)
                    if (keep >> b & 1) nreg[t++] = b;
                emit_remap(nreg);
            }
        }
        for (int i = 0; i < m; i++) {
            H.layN[i] = (uint8_t)cur.lay[i];
            H.gqN[i] = (uint8_t)qmap[cur.lay[i]];
        }
        merge_shears(P, H.instr_off);
        H.ninstr = (int32_t)P.instrs.size() - H.instr_off;
        {
            bool remap = false;
            for (int i = H.instr_off; i < (int)P.instrs.size(); i++) remap = remap || P.instrs[i].op == I_REMAP;
            if (!remap && std::memcmp(H.gq0, H.gqN, (size_t)m) != 0) H.flags |= SF_SYNC_STORE;
        }
        {
            Instr E;
            std::memset(&E, 0, sizeof(E));
            E.op = I_END;
            E.flags = F_OUTER | (K_OUTER << kCodeShift);
            P.instrs.push_back(E);
        }
        P.stages.push_back(H);
    }
    assign_codes(P);
    return P;
}

// The register slot s when a one-table phase op (16 entries) is a single uniform phase on the
// registers with bit s set (every other entry exactly 1), else -1.
int diag_single_slot(const Instr &I, const std::vector<double> &pool) {
    for (int s = 0; s < kRegBits; s++) {
        uint32_t want = 0;
        for (int r = 0; r < kRegAmps; r++)
            if (r >> s & 1) want |= 1u << r;
        if ((I.pat & 0xFFFFu) != want) continue;
        const double *t = pool.data() + 2 * (size_t)I.mat;
        const double pr = t[2 * (1 << s)], pi = t[2 * (1 << s) + 1];
        for (int r = 0; r < kRegAmps; r++) {
            const bool on = r >> s & 1;
            if (t[2 * r] != (on ? pr : 1.0) || t[2 * r + 1] != (on ? pi : 0.0)) return -1;
        }
        return s;
    }
    return -1;
}

void assign_codes(Program &P) {
    auto pair_index = [](int a, int b) {  // slot pairs 01 02 03 12 13 23
        static const int tab[4][4] = {{-1, 0, 1, 2}, {0, -1, 3, 4}, {1, 3, -1, 5}, {2, 4, 5, -1}};
        return tab[a][b];
    };
    static const int grad_masks[10] = {1, 2, 3, 4, 5, 6, 8, 9, 10, 12};
    for (StageHeader &H : P.stages) {
        for (int i = 0; i < H.ninstr; i++) {
            Instr &I = P.instrs[H.instr_off + i];
            int code = K_OUTER;
            const bool tc = I.flags & F_TCTRL, pure = I.flags & F_PURE, real = I.flags & F_REAL;
            if (!(I.flags & F_OUTER)) {
                switch (I.op) {
                case I_SHEAR: code = K_SHEAR + 2 * I.s[0] + (I.k == 3); break;
                case I_SHEARS: code = K_SHEARS; break;
                case I_DIAG: {
                    bool tile_bits = !tc;
                    for (int j = 0; j < I.k; j++) tile_bits = tile_bits && I.ls[j] >= 16;
                    const int slot = I.k == 0 ? diag_single_slot(I, P.pool) : -1;
                    if (I.k == 0 && !tc)
                        code = slot >= 0                                  ? K_DIAG_SLOT + slot
                               : __builtin_popcount(I.pat & 0xFFFEu) >= 11 ? K_DIAG_FULL
                                                                          : K_DIAG_MASK;
                    else if (I.k == 0 && slot >= 0)
                        code = K_DIAG_TSLOT + slot;
                    else
                        code = tile_bits ? K_DIAG_TILE : K_DIAG_LANE;
                    break;
                }
                case I_ANTI1: code = K_ANTI1 + 4 * I.s[0] + 2 * pure + tc; break;
                case I_SWAP2: code = K_SWAP2 + 4 * pair_index(I.s[0], I.s[1]) + 2 * pure + tc; break;
                case I_DENSE:
                    if (__builtin_popcount(I.smask) == 1) {
                        const int v = (I.flags & (F_TCTRL | F_PAT)) ? 0 : real ? 1 : 2;
                        code = K_DENSE1 + 3 * __builtin_ctz(I.smask) + v;
                    } else {
                        const int a = __builtin_ctz(I.smask), b = 31 - __builtin_clz(I.smask);
                        code = K_DENSE2 + 2 * pair_index(a, b) + ((I.flags & (F_TCTRL | F_PAT)) ? 0 : 1);
                    }
                    break;
                case I_PARITY:
                    if (I.k == 2) {
                        const int a = __builtin_ctz(I.smask), b = 31 - __builtin_clz(I.smask);
                        code = K_PARITY2 + 2 * pair_index(a, b) + !real;
                    } else {
                        code = K_PARITY4 + 2 * (I.s[1] - 1) + !real;
                    }
                    H.flags |= SF_DM;
                    break;
                case I_GRAD:
                    for (int g = 0; g < 10; g++)
                        if (grad_masks[g] == I.smask) code = K_GRAD + g;
                    break;
                default: break;
                }
            }
            I.flags = (I.flags & ((1 << kCodeShift) - 1)) | (code << kCodeShift);
        }
    }
}

void set_lds_offsets(Program &P) {
    auto offs = [&](const uint8_t *lay, uint16_t *out) {
        for (int r = 0; r < kRegAmps; r++) {
            uint32_t l = 0;
            for (int s = 0; s < kRegBits; s++)
                if (r >> s & 1) l |= 1u << lay[s];
            out[r] = (uint16_t)(lds_swizzle(l) * 8u);
        }
    };
    for (StageHeader &H : P.stages) {
        offs(H.lay0, H.k0);
        for (int i = 0; i < H.ninstr; i++) {
            Instr &I = P.instrs[H.instr_off + i];
            if (I.op == I_REMAP || I.op == I_XSWAP) offs(I.lay, I.kr);
        }
    }
}

}  // namespace goqu
