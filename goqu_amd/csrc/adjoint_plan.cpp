// adjoint_plan.cpp — host-side construction of adjoint-differentiation programs (Jones &
// Gacon 2020).  Plain C++ (no device code), so that the same programs can be executed by the
// CPU emulator in the tests.
//
// Statevector: psi_N = U_N..U_1|0>, lam_N = H psi_N.  Going backwards, for every derivative
// of op k,  dE/dtheta += scale * 2 Re <lam_k| dU_k |psi_{k-1}> = scale * 2 Re <lam_k| D |psi_k>
// with D = dU_k U_k^dag, and then psi_{k-1} = U_k^dag psi_k, lam_{k-1} = U_k^dag lam_k.  The
// whole reverse sweep is one program of EOP_GRAD ops and inverse ops applied to (psi, lam)
// together (reference: gradient.go's parameter-shift rule computes the same derivatives).
#include <algorithm>
#include <cmath>
#include <complex>
#include <string>
#include <vector>

#include "../../include/goqu_hip.h"
#include "engine.h"

namespace goqu {

using cd = std::complex<double>;

int invert_eop(const EOp &e, EOp &inv, std::string &err) {
    inv = e;
    switch (e.kind) {
    case EOP_DENSE:
    case EOP_WIDE:
    case EOP_PARITY: {  // conjugate transpose of the full matrix (keeps the parity blocks)
        const int d = 1 << e.k;
        for (int r = 0; r < d; r++)
            for (int c = 0; c < d; c++) {
                inv.m[2 * (r * d + c)] = e.m[2 * (c * d + r)];
                inv.m[2 * (r * d + c) + 1] = -e.m[2 * (c * d + r) + 1];
            }
        return GOQU_OK;
    }
    case EOP_DIAG:
        for (size_t i = 1; i < inv.m.size(); i += 2) inv.m[i] = -inv.m[i];
        return GOQU_OK;
    case EOP_ANTI1:  // n0 = p0 a1, n1 = p1 a0  =>  a0 = conj(p1) n1, a1 = conj(p0) n0
        inv.m = {e.m[2], -e.m[3], e.m[0], -e.m[1]};
        return GOQU_OK;
    case EOP_SWAP2:
        inv.m = {e.m[0], -e.m[1]};
        return GOQU_OK;
    default:
        err = "adjoint differentiation needs unitary ops (reset is not invertible)";
        return GOQU_ERR_UNSUPPORTED;
    }
}

int deriv_eop(const goqu_op_t &op, const double *dm, EOp &out, std::string &err) {
    out = EOp();
    out.kind = EOP_DENSE;
    const int nc = op.kind == GOQU_OP_CONTROLLED ? op.num_controls : 0;
    const int nt = op.num_qubits - nc;
    if (nt < 1 || nt > 4) {
        err = "derivative of an op with more than 4 target qubits";
        return GOQU_ERR_UNSUPPORTED;
    }
    out.k = nt;
    for (int j = 0; j < nt; j++) out.q[j] = op.qubits[nc + j];
    for (int j = 0; j < nc; j++) out.cmask |= 1ull << op.qubits[j];
    const int d = 1 << nt;
    out.m.assign(dm, dm + 2 * d * d);
    return GOQU_OK;
}

// D = dU U^dag on the op's targets; controls stay controls (for U = P0 (x) I + P1 (x) V,
// dU U^dag = P1 (x) dV V^dag).  False when the op has no matrix or more than two targets.
bool grad_eop(const goqu_op_t &op, const double *dm, int slot, EOp &out) {
    const int nc = op.kind == GOQU_OP_CONTROLLED ? op.num_controls : 0;
    const int nt = op.num_qubits - nc;
    if (!op.matrix || !dm || nt < 1 || nt > 2) return false;
    if (op.kind != GOQU_OP_GATE && op.kind != GOQU_OP_CONTROLLED) return false;
    const int d = 1 << nt;
    out = EOp();
    out.kind = EOP_GRAD;
    out.k = nt;
    out.slot = slot;
    for (int j = 0; j < nt; j++) out.q[j] = op.qubits[nc + j];
    for (int j = 0; j < nc; j++) out.cmask |= 1ull << op.qubits[j];
    out.m.assign(2 * d * d, 0.0);
    const double *U = op.matrix;
    for (int r = 0; r < d; r++)
        for (int c = 0; c < d; c++) {
            cd acc = 0;  // sum_j dU[r][j] conj(U[c][j])
            for (int j = 0; j < d; j++)
                acc += cd(dm[2 * (r * d + j)], dm[2 * (r * d + j) + 1]) *
                       std::conj(cd(U[2 * (c * d + j)], U[2 * (c * d + j) + 1]));
            out.m[2 * (r * d + c)] = acc.real();
            out.m[2 * (r * d + c) + 1] = acc.imag();
        }
    return true;
}

int build_sv_reverse(const goqu_op_t *ops, int nops, int nq_user, const goqu_deriv_t *derivs, int nd,
                     std::vector<EOp> &rev, bool &fused, std::string &err) {
    rev.clear();
    fused = true;
    std::vector<std::vector<int>> by_op(nops);
    for (int i = 0; i < nd; i++) by_op[derivs[i].op].push_back(i);
    for (int k = nops - 1; k >= 0; k--) {
        for (int i : by_op[k]) {
            EOp g;
            if (!grad_eop(ops[k], derivs[i].dmatrix, i, g)) {
                fused = false;
                rev.clear();
                return GOQU_OK;
            }
            rev.push_back(std::move(g));
        }
        std::vector<EOp> low;
        int rc = lower_sv_ops(&ops[k], 1, nq_user, low, err);
        if (rc) {
            err = "op " + std::to_string(k) + ": " + err;
            return rc;
        }
        for (int j = (int)low.size() - 1; j >= 0; j--) {
            if (low[j].kind == EOP_WIDE) {  // paired stage programs hold no wide ops
                fused = false;
                rev.clear();
                return GOQU_OK;
            }
            EOp ie;
            rc = invert_eop(low[j], ie, err);
            if (rc) return rc;
            rev.push_back(std::move(ie));
        }
    }
    return GOQU_OK;
}

bool inverse_superop(const EOp &e, EOp &inv, double bound) {
    if (e.kind != EOP_PARITY && e.kind != EOP_DENSE) return false;
    const int d = 1 << e.k;
    std::vector<cd> A((size_t)d * d), B((size_t)d * d, cd(0, 0));
    for (int i = 0; i < d * d; i++) A[i] = cd(e.m[2 * i], e.m[2 * i + 1]);
    for (int i = 0; i < d; i++) B[(size_t)i * d + i] = 1.0;
    double scale = 0;
    for (const cd &x : A) scale = std::max(scale, std::abs(x));
    if (!(scale > 0)) return false;
    for (int c = 0; c < d; c++) {  // Gauss-Jordan elimination with partial pivoting
        int p = c;
        for (int r = c + 1; r < d; r++)
            if (std::abs(A[(size_t)r * d + c]) > std::abs(A[(size_t)p * d + c])) p = r;
        if (std::abs(A[(size_t)p * d + c]) <= 1e-12 * scale) return false;
        if (p != c)
            for (int j = 0; j < d; j++) {
                std::swap(A[(size_t)p * d + j], A[(size_t)c * d + j]);
                std::swap(B[(size_t)p * d + j], B[(size_t)c * d + j]);
            }
        const cd piv = A[(size_t)c * d + c];
        for (int j = 0; j < d; j++) {
            A[(size_t)c * d + j] /= piv;
            B[(size_t)c * d + j] /= piv;
        }
        for (int r = 0; r < d; r++) {
            if (r == c) continue;
            const cd f = A[(size_t)r * d + c];
            if (f == cd(0, 0)) continue;
            for (int j = 0; j < d; j++) {
                A[(size_t)r * d + j] -= f * A[(size_t)c * d + j];
                B[(size_t)r * d + j] -= f * B[(size_t)c * d + j];
            }
        }
    }
    inv = e;
    for (int i = 0; i < d * d; i++) {
        if (!(std::abs(B[i]) <= bound)) return false;
        inv.m[2 * i] = B[i].real();
        inv.m[2 * i + 1] = B[i].imag();
    }
    return true;
}

double dm_uncompute_growth(const goqu_op_t &op, int n) {
    if (op.kind == GOQU_OP_RESET) return INFINITY;
    if (op.kind != GOQU_OP_CHANNEL || op.num_kraus < 2) return 1.0;  // unitary: exact inverse
    std::vector<EOp> low;
    std::string err;
    if (lower_dm_ops(&op, 1, n, low, err)) return INFINITY;
    double g = 1.0;
    for (EOp e : low) {
        if (e.kind == EOP_DIAG) {
            double mx = 0;
            for (size_t i = 0; i < e.m.size(); i += 2) {
                const double a = std::abs(cd(e.m[i], e.m[i + 1]));
                if (!(a > 0)) return INFINITY;
                mx = std::max(mx, 1.0 / a);
            }
            g *= mx;
            continue;
        }
        EOp inv;
        if (!inverse_superop(e, inv, 1e3)) return INFINITY;
        const int d = 1 << e.k;
        double mx = 0;  // infinity norm: the largest absolute row sum
        for (int r = 0; r < d; r++) {
            double rs = 0;
            for (int c = 0; c < d; c++) rs += std::abs(cd(inv.m[2 * (r * d + c)], inv.m[2 * (r * d + c) + 1]));
            mx = std::max(mx, rs);
        }
        g *= mx;
    }
    return g;
}

int build_dm_reverse(const goqu_op_t *ops, int k0, int k1, int n, const goqu_deriv_t *derivs, int nd,
                     std::vector<EOp> &rev, bool &fused, std::string &err) {
    rev.clear();
    fused = false;
    std::vector<std::vector<int>> by_op(k1 > k0 ? k1 - k0 : 0);
    for (int i = 0; i < nd; i++)
        if (derivs[i].op >= k0 && derivs[i].op < k1) by_op[derivs[i].op - k0].push_back(i);
    for (int k = k1 - 1; k >= k0; k--) {
        const goqu_op_t &op = ops[k];
        for (int i : by_op[k - k0]) {
            // D = dU U^dag acts on the row qubits (q + n) of rho's vector
            goqu_op_t rop = op;
            for (int j = 0; j < op.num_qubits; j++) rop.qubits[j] = op.qubits[j] + n;
            EOp g;
            if (op.kind != GOQU_OP_GATE || !grad_eop(rop, derivs[i].dmatrix, i, g)) {
                err = "op " + std::to_string(k) + ": derivative of a gate that is not a 1-2 qubit matrix";
                return GOQU_OK;
            }
            rev.push_back(std::move(g));
        }
        if (op.kind == GOQU_OP_RESET) {  // not invertible
            err = "op " + std::to_string(k) + ": reset";
            return GOQU_OK;
        }
        // a single Kraus operator runs as K (x) conj(K): its inverse is the conjugate transpose only
        // when K is unitary, which is not checked here
        if (op.kind == GOQU_OP_CHANNEL && op.num_kraus == 1) {
            err = "op " + std::to_string(k) + ": single-Kraus channel";
            return GOQU_OK;
        }
        std::vector<EOp> low;
        int rc = lower_dm_ops(&op, 1, n, low, err);
        if (rc) {
            err = "op " + std::to_string(k) + ": " + err;
            return rc;
        }
        const bool channel = op.kind == GOQU_OP_CHANNEL && op.num_kraus > 1;
        for (int j = (int)low.size() - 1; j >= 0; j--) {
            EOp e = low[j];
            if (channel && e.kind == EOP_DIAG && e.cmask == 0 && (e.k == 2 || e.k == 4)) {
                // a diagonal superoperator (phase damping ...) as a parity-block one, so that it can
                // act on one state of the pair
                const int d = 1 << e.k;
                std::vector<double> full(2 * (size_t)d * d, 0.0);
                for (int i = 0; i < d; i++) {
                    full[2 * ((size_t)i * d + i)] = e.m[2 * i];
                    full[2 * ((size_t)i * d + i) + 1] = e.m[2 * i + 1];
                }
                e.m = std::move(full);
                e.kind = EOP_PARITY;
            }
            if (e.kind == EOP_WIDE) {
                err = "op " + std::to_string(k) + ": gate on more than four targets";
                return GOQU_OK;
            }
            if (channel) {
                // rho moves through S^-1, h through S^dag; only parity-block channels (the paired
                // kernel applies channels one state at a time through its parity-block code)
                if (e.kind != EOP_PARITY) {
                    err = "op " + std::to_string(k) + ": channel superoperator is not parity-block";
                    return GOQU_OK;
                }
                EOp ri, hi;
                if (!inverse_superop(e, ri, 1e3)) {
                    err = "op " + std::to_string(k) + ": channel superoperator singular or badly conditioned";
                    return GOQU_OK;
                }
                rc = invert_eop(e, hi, err);
                if (rc) return rc;
                ri.side = 1;
                hi.side = 2;
                rev.push_back(std::move(ri));
                rev.push_back(std::move(hi));
                continue;
            }
            EOp ie;
            rc = invert_eop(e, ie, err);
            if (rc) return rc;
            rev.push_back(std::move(ie));
        }
    }
    fused = true;
    return GOQU_OK;
}

}  // namespace goqu
