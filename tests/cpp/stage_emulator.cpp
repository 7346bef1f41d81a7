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

// Global addressing of a layout, as the kernel computes it from gq[] (checked against T[lay[]]).
struct GAddr {
    uint64_t tg;
    uint64_t rq[4];
};
// strict: gq must be the tile list in layout order (stage loads); otherwise any arrangement of the
// tile's qubits (stores: uncontrolled swaps inside a stage relabel local bits)
bool global_layout(GAddr &G, const uint8_t *lay, const uint8_t *gq, const uint8_t *T, int m, uint32_t tid,
                   uint64_t base, bool strict) {
    bool ok = true;
    if (strict) {
        for (int i = 0; i < m; i++) ok &= gq[i] == T[lay[i]];
    } else {
        uint64_t want = 0, got = 0;
        for (int i = 0; i < m; i++) {
            want |= 1ull << T[i];
            ok &= !(got >> gq[i] & 1);
            got |= 1ull << gq[i];
        }
        ok &= want == got;
    }
    for (int s = 0; s < 4; s++) G.rq[s] = 1ull << gq[s];
    G.tg = base;
    for (int u = 0; u < m - 4; u++)
        if (tid >> u & 1) G.tg |= 1ull << gq[4 + u];
    return ok;
}
uint64_t rg(const GAddr &G, int r) {
    uint64_t g = 0;
    for (int s = 0; s < 4; s++)
        if (r >> s & 1) g |= G.rq[s];
    return g;
}
// per-thread LDS byte offset of a layout (the kernel's lds_lane)
uint32_t lds_lane(const uint8_t *lay, int m, uint32_t tid) {
    uint32_t tl = 0;
    for (int u = 0; u < m - 4; u++)
        if (tid >> u & 1) tl |= 1u << lay[4 + u];
    return lds_swizzle(tl) * 8u;
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
int pext4(int v, int mask) {
    int r = 0, b = 0;
    for (int i = 0; i < 4; i++)
        if (mask >> i & 1) {
            if (v >> i & 1) r |= 1 << b;
            b++;
        }
    return r;
}

// instruction categories seen (kernel code paths), for coverage checks
enum Cov {
    C_DENSE, C_DENSE_TCTRL, C_DENSE_PAT, C_DENSE_TILE, C_DENSE_BIG, C_DIAG_UNIFORM, C_DIAG_KL1, C_DIAG_KL2,
    C_DIAG_TCTRL, C_DIAG_TILE, C_DIAG_REGS, C_ANTI_PURE, C_ANTI_PHASED, C_ANTI_TCTRL, C_ANTI_PAT, C_SWAP_PURE,
    C_SWAP_PHASED, C_SWAP_TCTRL, C_SWAP_PAT, C_REMAP, C_GRAD, C_GRAD_CTRL, C_RESET, C_PARITY2, C_PARITY4,
    C_PARITY_REAL, C_SHEAR2, C_SHEAR3, C_XSWAP_PERM, C_XSWAP_DPP, C_DIAG_FULL, C_DIAG_SLOT, C_DIAG_TSLOT, C_SHEARS,
    C_NCOV
};
int64_t g_cov[C_NCOV];
void count(const Instr &I) {
    const bool tc = I.flags & F_TCTRL, pt = I.flags & F_PAT, tl = I.flags & F_TILE;
    switch (I.op) {
    case I_DENSE:
        g_cov[__builtin_popcount(I.smask) > 2 ? C_DENSE_BIG : C_DENSE]++;
        if (tc) g_cov[C_DENSE_TCTRL]++;
        if (pt) g_cov[C_DENSE_PAT]++;
        if (tl) g_cov[C_DENSE_TILE]++;
        break;
    case I_DIAG: {
        const int code = instr_code(I.flags);
        if (code == K_DIAG_FULL) g_cov[C_DIAG_FULL]++;
        if (code >= K_DIAG_SLOT && code < K_DIAG_TSLOT) g_cov[C_DIAG_SLOT]++;
        if (code >= K_DIAG_TSLOT && code < K_ANTI1) g_cov[C_DIAG_TSLOT]++;
        g_cov[I.k == 0 ? (tc ? C_DIAG_TCTRL : C_DIAG_UNIFORM) : I.k == 1 ? C_DIAG_KL1 : C_DIAG_KL2]++;
        if (tc) g_cov[C_DIAG_TCTRL]++;
        if (tl) g_cov[C_DIAG_TILE]++;
        if (I.smask) g_cov[C_DIAG_REGS]++;
        break;
    }
    case I_ANTI1:
        g_cov[(I.flags & F_PURE) ? C_ANTI_PURE : C_ANTI_PHASED]++;
        if (tc) g_cov[C_ANTI_TCTRL]++;
        if (pt) g_cov[C_ANTI_PAT]++;
        break;
    case I_SWAP2:
        g_cov[(I.flags & F_PURE) ? C_SWAP_PURE : C_SWAP_PHASED]++;
        if (tc) g_cov[C_SWAP_TCTRL]++;
        if (pt) g_cov[C_SWAP_PAT]++;
        break;
    case I_REMAP: g_cov[C_REMAP]++; break;
    case I_GRAD: g_cov[(tc || pt || tl) ? C_GRAD_CTRL : C_GRAD]++; break;
    case I_RESET: g_cov[C_RESET]++; break;
    case I_SHEAR: g_cov[I.k == 2 ? C_SHEAR2 : C_SHEAR3]++; break;
    case I_SHEARS: g_cov[C_SHEARS]++; break;
    case I_XSWAP: g_cov[I.k >= 4 ? C_XSWAP_PERM : C_XSWAP_DPP]++; break;
    case I_PARITY:
        g_cov[I.k == 2 ? C_PARITY2 : C_PARITY4]++;
        if (I.flags & F_REAL) g_cov[C_PARITY_REAL]++;
        break;
    default: break;
    }
}

// The kernel dispatches on the code in flags bits 8..15 alone: decode it the way the kernel's
// switch does and check that the code path it selects is the one the instruction's fields describe.
bool check_code(const Instr &I, const std::vector<double> &pool, std::string &err) {
    const int code = instr_code(I.flags);
    static const int pairs[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    static const int grad_masks[10] = {1, 2, 3, 4, 5, 6, 8, 9, 10, 12};
    const bool tc = I.flags & F_TCTRL, pure = I.flags & F_PURE, real = I.flags & F_REAL;
    const bool ctrl = I.flags & (F_TCTRL | F_PAT);
    bool ok = false;
    if (I.flags & F_OUTER) ok = code == K_OUTER;
    else if (code == K_SHEARS) {
        ok = I.op == I_SHEARS && (I.k & ~0xFF) == 0 && I.smask > 0 && I.smask < 16;
        for (int sl = 0; sl < 4 && ok; sl++) {
            const int n = (I.k >> (2 * sl)) & 3;
            ok = (n == 1 ? false : (n != 0) == ((I.smask >> sl & 1) != 0));
        }
    } else if (code < K_SHEARS) ok = I.op == I_SHEAR && I.s[0] == code / 2 && I.k == 2 + code % 2;
    else if (code < K_ANTI1) {
        bool tile_bits = I.k > 0 && !tc;
        for (int j = 0; j < I.k; j++) tile_bits = tile_bits && I.ls[j] >= 16;
        // single-slot phases: one uniform entry on the registers of the slot, the rest exactly 1;
        // full tables: every register 1..15 is multiplied (trivial entries are exactly 1)
        const int slot = I.op == I_DIAG && I.k == 0 ? diag_single_slot(I, pool) : -1;
        if (code >= K_DIAG_TSLOT) ok = I.op == I_DIAG && I.k == 0 && tc && slot == code - K_DIAG_TSLOT;
        else if (code >= K_DIAG_SLOT) ok = I.op == I_DIAG && I.k == 0 && !tc && slot == code - K_DIAG_SLOT;
        else
            ok = I.op == I_DIAG && (code == K_DIAG_LANE   ? (I.k > 0 || tc) && !tile_bits
                                    : code == K_DIAG_TILE ? tile_bits
                                                          : (I.k == 0 && !tc)) &&
                 (code != K_DIAG_FULL || __builtin_popcount(I.pat & 0xFFFEu) >= 11);
    }
    else if (code < K_SWAP2) {
        const int c = code - K_ANTI1;
        ok = I.op == I_ANTI1 && I.s[0] == c / 4 && pure == ((c >> 1) & 1) && tc == (c & 1);
    } else if (code < K_DENSE1) {
        const int c = code - K_SWAP2;
        ok = I.op == I_SWAP2 && I.s[0] == pairs[c / 4][0] && I.s[1] == pairs[c / 4][1] && pure == ((c >> 1) & 1) &&
             tc == (c & 1);
    } else if (code < K_DENSE2) {
        const int c = code - K_DENSE1, v = ctrl ? 0 : real ? 1 : 2;
        ok = I.op == I_DENSE && I.smask == (1 << (c / 3)) && c % 3 == v;
    } else if (code < K_PARITY2) {
        const int c = code - K_DENSE2;
        ok = I.op == I_DENSE && I.smask == ((1 << pairs[c / 2][0]) | (1 << pairs[c / 2][1])) && (c % 2) == !ctrl;
    } else if (code < K_PARITY4) {
        const int c = code - K_PARITY2;
        ok = I.op == I_PARITY && I.k == 2 && I.smask == ((1 << pairs[c / 2][0]) | (1 << pairs[c / 2][1])) &&
             (c % 2) == !real;
    } else if (code < K_GRAD) {
        const int c = code - K_PARITY4;
        ok = I.op == I_PARITY && I.k == 4 && I.s[0] == 0 && I.s[1] == 1 + c / 2 && (c % 2) == !real;
    } else if (code < K_GRAD + 10) {
        ok = I.op == I_GRAD && I.smask == grad_masks[code - K_GRAD];
    }
    if ((I.flags & (F_SIDE_A | F_SIDE_B)) && I.op != I_PARITY) {
        ok = false;
        err = "a one-state flag on an op other than I_PARITY";
        return ok;
    }
    if (!ok) err = "instruction code " + std::to_string(code) + " does not select op " + std::to_string(I.op);
    return ok;
}

// Executes a program exactly as the stage kernel interprets its encoding (complex128: remaps move
// 8-byte LDS elements, the real parts in one pass and the imaginary parts in a second).
// st2 != nullptr: paired mode (every op on both states; I_GRAD adds Re <st2| M |st> of each
// tile to vals[grad_base + local slot]).
int run(Program P, cd *st, int nq, std::string &err, cd *st2 = nullptr, std::vector<double> *vals = nullptr) {
    const int elem = 8;
    set_lds_offsets(P);
    if (st2 && vals) vals->assign(P.num_grad, 0.0);
    const uint64_t N = 1ull << nq;
    std::vector<char> visited(N);
    // the thread that loaded each position: without a barrier between a tile's loads and its
    // stores (no LDS remap in the stage and no SF_SYNC_STORE), a thread may store only positions
    // it loaded itself, or it could overwrite data another wave has still to load
    std::vector<uint32_t> loader(N);
    for (const StageHeader &H : P.stages) {
        const int m = H.m;
        const int nthr = 1 << (m - 4);
        for (int pc = 0; pc < H.ninstr; pc++) {
            count(P.instrs[H.instr_off + pc]);
            if (!check_code(P.instrs[H.instr_off + pc], P.pool, err)) return 1;
            if (P.instrs[H.instr_off + pc].op == I_PARITY && !(H.flags & SF_DM)) {
                err = "a stage with channel ops is not flagged for the density-matrix kernel";
                return 1;
            }
        }
        for (int j = 0; j < m; j++) {
            if (H.T[j] >= nq || (j && H.T[j] <= H.T[j - 1])) {
                err = "bad tile qubit list";
                return 1;
            }
        }
        if ((size_t(1) << m) * elem > 65536) {
            err = "tile exceeds the LDS offset range";
            return 1;
        }
        // the final layout must be the one in force after the last instruction
        uint8_t lay[16];
        std::memcpy(lay, H.lay0, 16);
        for (int pc = 0; pc < H.ninstr; pc++) {
            const Instr &I = P.instrs[H.instr_off + pc];
            if (I.op == I_XSWAP) {  // the carried layout must be the exchange of slot s[0] and lane bit k
                if (I.s[0] > 3 || I.k < 0 || I.k > 5 || 4 + I.k >= m || !(I.flags & F_OUTER)) {
                    err = "malformed lane exchange";
                    return 1;
                }
                std::swap(lay[I.s[0]], lay[4 + I.k]);
                for (int i = 0; i < m; i++)
                    if (lay[i] != I.lay[i]) {
                        err = "lane exchange layout mismatch";
                        return 1;
                    }
            }
            if (I.op == I_REMAP) std::memcpy(lay, I.lay, 16);
        }
        for (int i = 0; i < m; i++)
            if (lay[i] != H.layN[i]) {
                err = "layN is not the layout after the last remap";
                return 1;
            }
        std::fill(visited.begin(), visited.end(), 0);
        bool barrier = (H.flags & SF_SYNC_STORE) != 0;
        for (int pc = 0; pc < H.ninstr; pc++) barrier = barrier || P.instrs[H.instr_off + pc].op == I_REMAP;
        const uint64_t ntiles = 1ull << (nq - m);
        std::vector<double> lds(size_t(1) << m);  // 8-byte elements
        std::vector<char> seen(size_t(1) << m);
        for (uint64_t tile = 0; tile < ntiles; tile++) {
            uint64_t base = tile;
            for (int j = 0; j < m; j++) {
                const int t = H.T[j];
                base = ((base >> t) << (t + 1)) | (base & ((1ull << t) - 1));
            }
            const int npass = st2 ? 2 : 1;
            std::vector<cd> arr[2];
            arr[0].assign(size_t(nthr) * 16, 0);
            arr[1].assign(size_t(nthr) * 16, 0);
            std::vector<uint32_t> lb(nthr);
            for (int t = 0; t < nthr; t++) {
                GAddr G;
                if (!global_layout(G, H.lay0, H.gq0, H.T, m, t, base, true)) {
                    err = "gq0 disagrees with the tile list";
                    return 1;
                }
                for (int r = 0; r < 16; r++) {
                    const uint64_t g = G.tg | rg(G, r);
                    if (g >= N || visited[g]) {
                        err = "load index out of range or loaded twice";
                        return 1;
                    }
                    visited[g] = 1;
                    loader[g] = (uint32_t)t;
                    arr[0][t * 16 + r] = st[g];
                    if (st2) arr[1][t * 16 + r] = st2[g];
                }
                lb[t] = lds_lane(H.lay0, m, t);
            }
            const uint16_t *K = H.k0;
            for (int pc = 0; pc < H.ninstr; pc++) {
                const Instr &I = P.instrs[H.instr_off + pc];
                const double *pool = P.pool.data() + 2 * I.mat;
                if (I.op == I_REMAP) {
                    for (int w = 0; w < npass; w++)
                        for (int comp = 0; comp < 2; comp++) {
                            std::fill(seen.begin(), seen.end(), 0);
                            for (int t = 0; t < nthr; t++)
                                for (int r = 0; r < 16; r++) {
                                    const uint32_t off = lb[t] ^ K[r];
                                    if (off % elem || off / elem >= lds.size() || seen[off / elem]) {
                                        err = "LDS write collision/out of range";
                                        return 1;
                                    }
                                    seen[off / elem] = 1;
                                    const cd v = arr[w][t * 16 + r];
                                    lds[off / elem] = comp ? v.imag() : v.real();
                                }
                            for (int t = 0; t < nthr; t++) {
                                const uint32_t lbn = lds_lane(I.lay, m, t);
                                for (int r = 0; r < 16; r++) {
                                    const uint32_t off = lbn ^ I.kr[r];
                                    if (off % elem || off / elem >= lds.size()) {
                                        err = "LDS read out of range";
                                        return 1;
                                    }
                                    cd &v = arr[w][t * 16 + r];
                                    v = comp ? cd(v.real(), lds[off / elem]) : cd(lds[off / elem], v.imag());
                                }
                            }
                        }
                    for (int t = 0; t < nthr; t++) lb[t] = lds_lane(I.lay, m, t);
                    K = I.kr;
                    continue;
                }
                if (I.op == I_XSWAP) {
                    // amplitude (thread t, register r) <- (t with lane bit u = bit S of r,
                    // r with bit S = lane bit u of t)
                    const int S = I.s[0], u = I.k;
                    for (int w = 0; w < npass; w++) {
                        std::vector<cd> old(arr[w]);
                        for (int t = 0; t < nthr; t++)
                            for (int r = 0; r < 16; r++) {
                                const int tb = t >> u & 1, rb = r >> S & 1;
                                const int ts = (t & ~(1 << u)) | (rb << u), rs = (r & ~(1 << S)) | (tb << S);
                                arr[w][t * 16 + r] = old[ts * 16 + rs];
                            }
                    }
                    for (int t = 0; t < nthr; t++) lb[t] = lds_lane(I.lay, m, t);
                    K = I.kr;
                    continue;
                }
                if ((I.flags & F_TILE) && (base & I.omask) != I.omask) continue;
                if (!(I.flags & F_TILE) && I.omask) {
                    err = "omask without F_TILE";
                    return 1;
                }
                for (int t = 0; t < nthr; t++) {
                    const bool lok = (uint32_t(t) & I.tmask) == I.tmask;
                    for (int w = 0; w < npass; w++) {
                        cd *A = &arr[w][t * 16];
                        switch (I.op) {
                        case I_GRAD: {
                            if (!st2 || !vals) {
                                err = "I_GRAD outside a paired program";
                                return 1;
                            }
                            if (w) break;
                            const cd *B = &arr[1][t * 16];
                            const int MASK = I.smask, Kq = __builtin_popcount(MASK), D = 1 << Kq;
                            if (Kq < 1 || Kq > 2) {
                                err = "I_GRAD on more than two slots";
                                return 1;
                            }
                            if (H.grad_base + I.pad >= P.num_grad || I.pad >= H.ngrad) {
                                err = "gradient slot out of range";
                                return 1;
                            }
                            if (!lok) break;
                            for (int o = 0; o < (16 >> Kq); o++) {
                                const int bb = dep4(o, 15 & ~MASK);
                                if (!(I.pat >> bb & 1)) continue;
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
                            const int MASK = I.smask, Kq = __builtin_popcount(MASK), D = 1 << Kq;
                            if (Kq < 1) {
                                err = "dense op without slots";
                                return 1;
                            }
                            // the lean stage kernel does not run 3-4-slot dense ops
                            if (Kq > 2 && (!(H.flags & SF_BIG) || !(I.flags & F_OUTER))) {
                                err = "3-4-slot dense op in a stage not flagged SF_BIG / F_OUTER";
                                return 1;
                            }
                            if (!lok) break;
                            for (int o = 0; o < (16 >> Kq); o++) {
                                const int b = dep4(o, 15 & ~MASK);
                                if (!(I.pat >> b & 1)) continue;
                                cd x[16], y[16];
                                for (int j = 0; j < D; j++) x[j] = A[b | dep4(j, MASK)];
                                // F_REAL (uncontrolled, at most two slots): the kernel reads real parts only
                                const bool real = (I.flags & F_REAL) && Kq <= 2 && !(I.flags & (F_TCTRL | F_PAT));
                                for (int i = 0; i < D; i++) {
                                    y[i] = 0;
                                    for (int j = 0; j < D; j++)
                                        y[i] += cd(pool[2 * (i * D + j)], real ? 0.0 : pool[2 * (i * D + j) + 1]) * x[j];
                                }
                                for (int i = 0; i < D; i++) A[b | dep4(i, MASK)] = y[i];
                            }
                            break;
                        }
                        case I_PARITY: {
                            // paired programs (density-matrix reverse sweeps): a channel on one
                            // state or the other, as the flags say
                            if ((I.flags & F_SIDE_A) && (I.flags & F_SIDE_B)) {
                                err = "I_PARITY on neither state";
                                return 1;
                            }
                            if ((w == 1 && (I.flags & F_SIDE_A)) || (w == 0 && (I.flags & F_SIDE_B))) break;
                            const bool real = I.flags & F_REAL;
                            auto B = [&](int off) { return real ? cd(pool[2 * off], 0) : cd(pool[2 * off], pool[2 * off + 1]); };
                            if (I.k == 2) {
                                const int lo = __builtin_ctz(I.smask), hi = 31 - __builtin_clz(I.smask);
                                if (__builtin_popcount(I.smask) != 2) {
                                    err = "I_PARITY k=2 needs two slots";
                                    return 1;
                                }
                                const int om = 15 & ~I.smask;
                                for (int g = 0; g < 4; g++) {
                                    const int base = dep4(g, om);
                                    for (int p = 0; p < 2; p++) {
                                        const int e0 = base | (p << hi), e1 = base | (1 << lo) | ((1 ^ p) << hi);
                                        const cd x0 = A[e0], x1 = A[e1];
                                        A[e0] = B(4 * p + 0) * x0 + B(4 * p + 1) * x1;
                                        A[e1] = B(4 * p + 2) * x0 + B(4 * p + 3) * x1;
                                    }
                                }
                            } else if (I.k == 4) {
                                const int s0 = I.s[0], s1 = I.s[1], s2 = I.s[2], s3 = I.s[3];
                                if (s0 != 0 || s2 >= s3 || ((1 << s0) | (1 << s1) | (1 << s2) | (1 << s3)) != 15) {
                                    err = "I_PARITY k=4 slots not canonical";
                                    return 1;
                                }
                                for (int c = 0; c < 4; c++) {
                                    const int pa = c >> 1, pb = c & 1;
                                    int e[4];
                                    cd x[4];
                                    for (int i = 0; i < 4; i++) {
                                        const int ia = i >> 1, ib = i & 1;
                                        e[i] = (ia << s0) | ((ia ^ pa) << s1) | (ib << s2) | ((ib ^ pb) << s3);
                                        x[i] = A[e[i]];
                                    }
                                    for (int i = 0; i < 4; i++) {
                                        cd y = 0;
                                        for (int j = 0; j < 4; j++) y += B(16 * c + 4 * i + j) * x[j];
                                        A[e[i]] = y;
                                    }
                                }
                            } else {
                                err = "I_PARITY with k other than 2 or 4";
                                return 1;
                            }
                            break;
                        }
                        case I_DIAG: {
                            const int RM = I.smask, KR = __builtin_popcount(RM), kl = I.k;
                            if (kl < 0 || kl > 4) {
                                err = "diag lane bits out of range";
                                return 1;
                            }
                            uint32_t li = 0;
                            for (int j = 0; j < kl; j++) {
                                const int src = I.ls[j];
                                const uint32_t bit = src < 16 ? ((uint32_t(t) >> src) & 1u)
                                                              : (uint32_t)((base >> (src - 16)) & 1u);
                                if (src < 16 && src >= m - 4) {
                                    err = "diag lane bit beyond the thread bits";
                                    return 1;
                                }
                                li = (li << 1) | bit;
                            }
                            (void)KR;
                            if (I.pat >> 16) {
                                err = "diag register mask beyond 16 registers";
                                return 1;
                            }
                            for (int r = 0; r < 16; r++) {
                                cd ph(pool[2 * ((li << 4) | r)], pool[2 * ((li << 4) | r) + 1]);
                                if (!(I.pat >> r & 1)) {
                                    if (ph != cd(1, 0)) {
                                        err = "diag register marked trivial holds an entry != 1";
                                        return 1;
                                    }
                                    continue;
                                }
                                if ((I.flags & F_TCTRL) && !lok) ph = 1;
                                A[r] *= ph;
                            }
                            break;
                        }
                        case I_SHEARS: {
                            // slot by slot, each the fused multiply-adds of one I_SHEAR
                            const double *Cf = P.pool.data() + 2 * (size_t)I.mat;
                            for (int S = 0; S < 4; S++) {
                                const int n = (I.k >> (2 * S)) & 3;
                                for (int p = 0; p < n; p++)
                                    for (int r = 0; r < 16; r++) {
                                        if (r >> S & 1) continue;
                                        cd &dst = (p & 1) ? A[r | (1 << S)] : A[r];
                                        const cd src = (p & 1) ? A[r] : A[r | (1 << S)];
                                        const double c = Cf[3 * S + p];
                                        dst = cd(std::fma(c, src.real(), dst.real()), std::fma(c, src.imag(), dst.imag()));
                                    }
                            }
                            break;
                        }
                        case I_SHEAR: {
                            const int S = I.s[0];
                            if (S > 3 || I.smask != (1 << S) || I.k < 2 || I.k > 3 || (I.flags & ((1 << kCodeShift) - 1)) != 0) {
                                err = "malformed shear";
                                return 1;
                            }
                            // in-place lifting steps, each rounded as one fused multiply-add
                            for (int p = 0; p < I.k; p++)
                                for (int r = 0; r < 16; r++) {
                                    if (r >> S & 1) continue;
                                    cd &dst = (p & 1) ? A[r | (1 << S)] : A[r];
                                    const cd src = (p & 1) ? A[r] : A[r | (1 << S)];
                                    const double c = shear_coef(I, p);
                                    dst = cd(std::fma(c, src.real(), dst.real()), std::fma(c, src.imag(), dst.imag()));
                                }
                            break;
                        }
                        case I_ANTI1: {
                            const int S = I.s[0];
                            if (!lok) break;
                            const bool pure = I.flags & F_PURE;
                            for (int r = 0; r < 16; r++) {
                                if (r >> S & 1 || !(I.pat >> r & 1)) continue;
                                const int r1 = r | (1 << S);
                                cd a0 = A[r], a1 = A[r1];
                                A[r] = pure ? a1 : cd(pool[0], pool[1]) * a1;
                                A[r1] = pure ? a0 : cd(pool[2], pool[3]) * a0;
                            }
                            break;
                        }
                        case I_SWAP2: {
                            const int S0 = I.s[0], S1 = I.s[1];
                            if (S0 >= S1) {
                                err = "swap slots not ascending";
                                return 1;
                            }
                            if (!lok) break;
                            const bool pure = I.flags & F_PURE;
                            for (int r = 0; r < 16; r++) {
                                if ((r >> S0 & 1) || (r >> S1 & 1) || !(I.pat >> r & 1)) continue;
                                const int i01 = r | (1 << S0), i10 = r | (1 << S1);
                                cd p = pure ? cd(1, 0) : cd(pool[0], pool[1]);
                                cd x01 = A[i01], x10 = A[i10];
                                A[i01] = p * x10;
                                A[i10] = p * x01;
                            }
                            break;
                        }
                        case I_RESET: {
                            if (w) break;  // the kernel resets only the first state
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
            }
            for (int t = 0; t < nthr; t++) {
                GAddr G;
                if (!global_layout(G, H.layN, H.gqN, H.T, m, t, base, false)) {
                    err = "gqN disagrees with the tile list";
                    return 1;
                }
                for (int r = 0; r < 16; r++) {
                    const uint64_t g = G.tg | rg(G, r);
                    if (g >= N || visited[g] != 1) {
                        err = "store index out of range or not loaded by this tile";
                        return 1;
                    }
                    if (!barrier && loader[g] != (uint32_t)t) {
                        err = "a thread stores a position another thread loaded, with no barrier in between";
                        return 1;
                    }
                    visited[g] = 2;
                    st[g] = arr[0][t * 16 + r];
                    if (st2) st2[g] = arr[1][t * 16 + r];
                }
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
    // stage programs between wide ops, as the device library runs them; a wide op is applied as
    // its dense matrix over every column whose controls hold
    cd *st = reinterpret_cast<cd *>(state);
    if (num_stages) *num_stages = 0;
    size_t i = 0;
    while (i < eops.size()) {
        size_t j = i;
        while (j < eops.size() && eops[j].kind != EOP_WIDE) j++;
        if (j > i) {
            Program P = plan_program(std::vector<EOp>(eops.begin() + i, eops.begin() + j), nq_internal, cfg);
            if (num_stages) *num_stages += (int64_t)P.stages.size();
            if (run(P, st, nq_internal, err)) {
                g_err = err;
                return 100;
            }
        }
        if (j < eops.size()) {
            const EOp &e = eops[j];
            const int k = e.k, D = 1 << k;
            uint64_t tmask = 0;
            for (int x : e.qw) tmask |= 1ull << x;
            std::vector<cd> v(D), w(D);
            for (uint64_t g = 0; g < (1ull << nq_internal); g++) {
                if ((g & tmask) || (g & e.cmask) != e.cmask) continue;  // one column per base index
                auto idx = [&](int r) {
                    uint64_t x = g;
                    for (int b = 0; b < k; b++)
                        if (r >> (k - 1 - b) & 1) x |= 1ull << e.qw[b];
                    return x;
                };
                for (int r = 0; r < D; r++) v[r] = st[idx(r)];
                for (int r = 0; r < D; r++) {
                    cd acc = 0;
                    for (int c = 0; c < D; c++) acc += cd(e.m[2 * (r * D + c)], e.m[2 * (r * D + c) + 1]) * v[c];
                    w[r] = acc;
                }
                for (int r = 0; r < D; r++) st[idx(r)] = w[r];
            }
            j++;
        }
        i = j;
    }
    return 0;
}
const char *emu_error(void) { return g_err.c_str(); }

// instruction-category counters since the last reset (see enum Cov)
int emu_coverage(int64_t *out, int n, int reset) {
    for (int i = 0; i < n && i < C_NCOV; i++) out[i] = g_cov[i];
    if (reset)
        for (int i = 0; i < C_NCOV; i++) g_cov[i] = 0;
    return C_NCOV;
}

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

// The density-matrix adjoint's fused reverse sweep over ops [k0, k1) on (rho, h) (2n-qubit
// vectors) in place; vals[i] receives Re <h_k| D_i rho_k> of derivative i.
extern "C" int emu_dm_reverse(double *rho, double *h, int n, const goqu_op_t *ops, int k0, int k1,
                              const goqu_deriv_t *derivs, int nd, int tile_bits, int low_bits, double *vals,
                              int64_t *num_stages) {
    std::vector<EOp> rev;
    bool fused = false;
    std::string err;
    int rc = build_dm_reverse(ops, k0, k1, n, derivs, nd, rev, fused, err);
    if (rc) {
        g_err = err;
        return rc;
    }
    if (!fused) {
        g_err = "not fusable: " + err;
        return 200;
    }
    PlanConfig cfg;
    cfg.tile_bits = tile_bits;
    cfg.low_bits = low_bits;
    const int nq = std::max(2 * n, kMinInternalQubits);
    Program P = plan_program(rev, nq, cfg);
    if (num_stages) *num_stages = (int64_t)P.stages.size();
    std::vector<double> pv;
    rc = run(P, reinterpret_cast<cd *>(rho), nq, err, reinterpret_cast<cd *>(h), &pv);
    if (rc) {
        g_err = err;
        return 100;
    }
    for (int i = 0; i < nd; i++) vals[i] = 0;
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
    for (const Instr &I : P.instrs) count(I);  // adds to the coverage counters
    return 0;
}
