// adjoint_plan.cpp — host-side construction of adjoint-differentiation programs (Jones &
// Gacon 2020).  Plain C++ (no device code), so that the same programs can be executed by the
// CPU emulator in the tests.
//
// Statevector: psi_N = U_N..U_1|0>, lam_N = H psi_N.  Going backwards, for every derivative
// of op k,  dE/dtheta += scale * 2 Re <lam_k| dU_k |psi_{k-1}> = scale * 2 Re <lam_k| D |psi_k>
// with D = dU_k U_k^dag, and then psi_{k-1} = U_k^dag psi_k, lam_{k-1} = U_k^dag lam_k.  The
// whole reverse sweep is one program of EOP_GRAD ops and inverse ops applied to (psi, lam)
// together (reference: gradient.go's parameter-shift rule computes the same derivatives).
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

}  // namespace goqu
