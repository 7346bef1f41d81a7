// stage_emulator.cpp — CPU lockstep emulation of the stage kernel's semantics, driven by the
// real planner (goqu_amd/csrc/planner.cpp).  Test infrastructure: it checks, without a GPU,
// that every planned program (tile qubits, register layouts, slot assignments, permuted
// matrices, remaps) computes the same state as the oracle, and that every index it touches
// stays inside the state.
#include <complex>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../goqu_amd/csrc/engine.h"
#include "../../include/goqu_hip.h"

using namespace goqu;
using cd = std::complex<double>;

namespace {

struct Lay {
    uint64_t tg;
    uint32_t tl;
    uint64_t rq[4];
    uint32_t rl[4];
};

// reads the global qubit of each entry from gq[] (as the kernel does) and checks it
// against the tile's qubit list
bool set_layout(Lay &L, const uint8_t *lay, const uint8_t *gq, const uint8_t *T, int m, uint32_t tid,
                uint64_t base) {
    bool ok = true;
    for (int i = 0; i < m; i++) ok &= gq[i] == T[lay[i]];
    for (int s = 0; s < 4; s++) {
        L.rl[s] = 1u << lay[s];
        L.rq[s] = 1ull << gq[s];
    }
    L.tg = base;
    L.tl = 0;
    for (int u = 0; u < m - 4; u++)
        if (tid >> u & 1) {
            L.tl |= 1u << lay[4 + u];
            L.tg |= 1ull << gq[4 + u];
        }
    return ok;
}

uint64_t rg(const Lay &L, int r) {
    uint64_t g = 0;
    for (int s = 0; s < 4; s++)
        if (r >> s & 1) g |= L.rq[s];
    return g;
}
uint32_t rlo(const Lay &L, int r) {
    uint32_t g = 0;
    for (int s = 0; s < 4; s++)
        if (r >> s & 1) g |= L.rl[s];
    return g;
}
int dep4(int v, int mask) {
    int r = 0, b = 0;
    for (int i = 0; i < 4; i++)
        if (mask >> i & 1) {
            if (v >> b & 1) r |= 1 << i;
            b++;
        }
    return r;
}

// st2 != nullptr: paired mode (every op on both states; I_GRAD adds Re <st2| M |st> of each
// tile to vals[grad_base + local slot])
int run(const Program &P, cd *st, int nq, std::string &err, cd *st2 = nullptr, std::vector<double> *vals = nullptr) {
    if (st2 && vals) vals->assign(P.num_grad, 0.0);
    const uint64_t N = 1ull << nq;
    for (const StageHeader &H : P.stages) {
        const int m = H.m;
        const int nthr = 1 << (m - 4);
        // tile qubits must be distinct, ascending, in range
        for (int j = 0; j < m; j++) {
            if (H.T[j] >= nq || (j && H.T[j] <= H.T[j - 1])) {
                err = "bad tile qubit list";
                return 1;
            }
        }
        const uint64_t ntiles = 1ull << (nq - m);
        std::vector<cd> lds(size_t(1) << m);
        std::vector<char> seen(size_t(1) << m);
        for (uint64_t tile = 0; tile < ntiles; tile++) {
            uint64_t base = tile;
            for (int j = 0; j < m; j++) {
                const int t = H.T[j];
                base = ((base >> t) << (t + 1)) | (base & ((1ull << t) - 1));
            }
            std::vector<Lay> L(nthr);
            std::vector<cd> a(size_t(nthr) * 16), b2(st2 ? size_t(nthr) * 16 : 0);
            for (int t = 0; t < nthr; t++) {
                if (!set_layout(L[t], H.lay0, H.gq0, H.T, m, t, base)) {
                    err = "layout global qubits disagree with the tile list";
                    return 1;
                }
                for (int r = 0; r < 16; r++) {
                    const uint64_t g = L[t].tg | rg(L[t], r);
                    if (g >= N) {
                        err = "load index out of range";
                        return 1;
                    }
                    a[t * 16 + r] = st[g];
                    if (st2) b2[t * 16 + r] = st2[g];
                }
            }
            for (int pc = 0; pc < H.ninstr; pc++) {
                const Instr &I = P.instrs[H.instr_off + pc];
                const double *pool = P.pool.data() + 2 * I.mat;
                if (I.op == I_REMAP) {
                    std::vector<cd> lds2(st2 ? lds.size() : 0);
                    std::fill(seen.begin(), seen.end(), 0);
                    for (int t = 0; t < nthr; t++)
                        for (int r = 0; r < 16; r++) {
                            uint32_t l = L[t].tl | rlo(L[t], r);
                            if (l >= lds.size() || seen[l]) {
                                err = "LDS write collision/out of range";
                                return 1;
                            }
                            seen[l] = 1;
                            lds[l] = a[t * 16 + r];
                            if (st2) lds2[l] = b2[t * 16 + r];
                        }
                    for (int t = 0; t < nthr; t++) {
                        if (!set_layout(L[t], I.lay, I.gq, H.T, m, t, base)) {
                            err = "remap global qubits disagree with the tile list";
                            return 1;
                        }
                        for (int r = 0; r < 16; r++) {
                            a[t * 16 + r] = lds[L[t].tl | rlo(L[t], r)];
                            if (st2) b2[t * 16 + r] = lds2[L[t].tl | rlo(L[t], r)];
                        }
                    }
                    continue;
                }
                for (int t = 0; t < nthr; t++)
                  for (int which = 0; which < (st2 ? 2 : 1); which++) {
                    cd *A = which ? &b2[t * 16] : &a[t * 16];
                    const Lay &LL = L[t];
                    auto ok = [&](int r) { return ((LL.tg | rg(LL, r)) & I.cmask) == I.cmask; };
                    if (I.op == I_RESET && which) break;  // the kernel resets only the first state
                    switch (I.op) {
                    case I_GRAD: {
                        if (!st2 || !vals) {
                            err = "I_GRAD outside a paired program";
                            return 1;
                        }
                        if (which) break;
                        const cd *B = &b2[t * 16];
                        const int MASK = I.smask, K = __builtin_popcount(MASK), D = 1 << K;
                        if (K < 1 || K > 2) {
                            err = "I_GRAD on more than two slots";
                            return 1;
                        }
                        if (H.grad_base + I.pad >= P.num_grad || I.pad >= H.ngrad) {
                            err = "gradient slot out of range";
                            return 1;
                        }
                        for (int o = 0; o < (16 >> K); o++) {
                            const int bb = dep4(o, 15 & ~MASK);
                            if (!ok(bb)) continue;
                            for (int i = 0; i < D; i++) {
                                cd y = 0;
                                for (int j = 0; j < D; j++)
                                    y += cd(pool[2 * (i * D + j)], pool[2 * (i * D + j) + 1]) * A[bb | dep4(j, MASK)];
                                (*vals)[H.grad_base + I.pad] += std::real(std::conj(B[bb | dep4(i, MASK)]) * y);
                            }
                        }
                        break;
                    }
                    case I_DENSE: {
                        const int MASK = I.smask, K = __builtin_popcount(MASK), D = 1 << K;
                        for (int o = 0; o < (16 >> K); o++) {
                            const int b = dep4(o, 15 & ~MASK);
                            if (!ok(b)) continue;
                            cd x[16], y[16];
                            for (int j = 0; j < D; j++) x[j] = A[b | dep4(j, MASK)];
                            for (int i = 0; i < D; i++) {
                                y[i] = 0;
                                for (int j = 0; j < D; j++)
                                    y[i] += cd(pool[2 * (i * D + j)], pool[2 * (i * D + j) + 1]) * x[j];
                            }
                            for (int i = 0; i < D; i++) A[b | dep4(i, MASK)] = y[i];
                        }
                        break;
                    }
                    case I_DIAG:
                        for (int r = 0; r < 16; r++) {
                            const uint64_t g = LL.tg | rg(LL, r);
                            if ((g & I.cmask) != I.cmask) continue;
                            int idx = 0;
                            for (int j = 0; j < I.k; j++) idx = (idx << 1) | int((g >> I.q[j]) & 1);
                            A[r] *= cd(pool[2 * idx], pool[2 * idx + 1]);
                        }
                        break;
                    case I_ANTI1: {
                        const int S = I.s[0];
                        for (int r = 0; r < 16; r++) {
                            if (r >> S & 1 || !ok(r)) continue;
                            const int r1 = r | (1 << S);
                            cd a0 = A[r], a1 = A[r1];
                            A[r] = cd(pool[0], pool[1]) * a1;
                            A[r1] = cd(pool[2], pool[3]) * a0;
                        }
                        break;
                    }
                    case I_SWAP2: {
                        const int S0 = I.s[0], S1 = I.s[1];
                        for (int r = 0; r < 16; r++) {
                            if ((r >> S0 & 1) || (r >> S1 & 1) || !ok(r)) continue;
                            const int i01 = r | (1 << S1), i10 = r | (1 << S0);
                            cd p(pool[0], pool[1]);
                            cd x01 = A[i01], x10 = A[i10];
                            A[i01] = p * x10;
                            A[i10] = p * x01;
                        }
                        break;
                    }
                    case I_RESET: {
                        const int S = I.s[0];
                        for (int r = 0; r < 16; r++) {
                            if (r >> S & 1) continue;
                            const int r1 = r | (1 << S);
                            double nrm = std::sqrt(std::norm(A[r]) + std::norm(A[r1]));
                            A[r] = nrm > 1e-15 ? cd(nrm, 0) : cd(0, 0);
                            A[r1] = 0;
                        }
                        break;
                    }
                    default:
                        err = "unknown instruction";
                        return 1;
                    }
                }
            }
            for (int t = 0; t < nthr; t++)
                for (int r = 0; r < 16; r++) {
                    st[L[t].tg | rg(L[t], r)] = a[t * 16 + r];
                    if (st2) st2[L[t].tg | rg(L[t], r)] = b2[t * 16 + r];
                }
        }
    }
    return 0;
}

thread_local std::string g_err;

}  // namespace

extern "C" {
// state: 2^nq_internal complex (nq_internal >= nq_user, padded qubits |0>)
int emu_apply(double *state, int nq_internal, int nq_user, const goqu_op_t *ops, int nops, int tile_bits,
              int low_bits, int dm, int64_t *num_stages) {
    std::vector<EOp> eops;
    std::string err;
    int rc = dm ? lower_dm_ops(ops, nops, nq_user, eops, err) : lower_sv_ops(ops, nops, nq_user, eops, err);
    if (rc) {
        g_err = err;
        return rc;
    }
    PlanConfig cfg;
    cfg.tile_bits = tile_bits;
    cfg.low_bits = low_bits;
    Program P = plan_program(eops, nq_internal, cfg);
    if (num_stages) *num_stages = (int64_t)P.stages.size();
    rc = run(P, reinterpret_cast<cd *>(state), nq_internal, err);
    if (rc) g_err = err;
    return rc ? 100 : 0;
}
const char *emu_error(void) { return g_err.c_str(); }

// The statevector adjoint's reverse sweep (build_sv_reverse + planner) on psi and lam in place;
// vals[i] = Re <lam_k| D_i |psi_k> of derivative i.  Returns 200 when the derivatives do not
// fit the fused program.
int emu_sv_reverse(double *psi, double *lam, int nq_internal, int nq_user, const goqu_op_t *ops, int nops,
                   const goqu_deriv_t *derivs, int nd, int tile_bits, int low_bits, double *vals,
                   int64_t *num_stages) {
    std::vector<EOp> rev;
    bool fused = false;
    std::string err;
    int rc = build_sv_reverse(ops, nops, nq_user, derivs, nd, rev, fused, err);
    if (rc) {
        g_err = err;
        return rc;
    }
    if (!fused) {
        g_err = "not fusable";
        return 200;
    }
    PlanConfig cfg;
    cfg.tile_bits = tile_bits;
    cfg.low_bits = low_bits;
    Program P = plan_program(rev, nq_internal, cfg);
    if (num_stages) *num_stages = (int64_t)P.stages.size();
    std::vector<double> pv;
    rc = run(P, reinterpret_cast<cd *>(psi), nq_internal, err, reinterpret_cast<cd *>(lam), &pv);
    if (rc) {
        g_err = err;
        return 100;
    }
    for (int i = 0; i < nd; i++) vals[i] = 0;
    if ((int)P.grad_slots.size() != P.num_grad) {
        g_err = "grad_slots size mismatch";
        return 100;
    }
    for (int j = 0; j < P.num_grad; j++) {
        const int slot = P.grad_slots[j];
        if (slot < 0 || slot >= nd) {
            g_err = "gradient slot outside the derivative list";
            return 100;
        }
        vals[slot] += pv[j];
    }
    return 0;
}
}

extern "C" int emu_plan_stats(const goqu_op_t *ops, int nops, int nq, int tile_bits, int low_bits, int64_t *stages,
                              int64_t *remaps, int64_t *instrs, int64_t *eops) {
    std::vector<EOp> e;
    std::string err;
    int rc = lower_sv_ops(ops, nops, nq, e, err);
    if (rc) return rc;
    PlanConfig cfg;
    cfg.tile_bits = tile_bits;
    cfg.low_bits = low_bits;
    Program P = plan_program(e, nq, cfg);
    *stages = (int64_t)P.stages.size();
    *instrs = (int64_t)P.instrs.size();
    *eops = P.num_eops;
    int64_t r = 0;
    for (const Instr &I : P.instrs) r += I.op == I_REMAP;
    *remaps = r;
    return 0;
}
