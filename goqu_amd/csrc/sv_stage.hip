// sv_stage.hip — the tiled gate-application kernel for gfx950.
//
// One launch = one stage = one HBM sweep.  A workgroup of 2^(m-4) threads owns a tile of
// 2^m amplitudes: the tile's qubits are T[0..m-1] (the lowest L qubits plus the stage's
// chosen qubits); the remaining n-m qubit values are fixed per tile.  Each thread keeps 16
// amplitudes in registers: register slot s <-> tile-local bit lay[s], thread-id bit u <->
// tile-local bit lay[4+u].  Ops whose targets are register slots run entirely in
// registers; controls and diagonal gates may sit on any qubit (they only read index bits).
// I_REMAP re-distributes the tile through LDS so that other qubits become register slots.
// Global loads/stores use a layout whose lanes 0..2 walk the tile's lowest qubits, so a
// wave touches contiguous runs of >= 8 amplitudes (128 B at complex128).
#include <atomic>
#include <mutex>

#include "runtime.h"

namespace goqu {

__host__ __device__ constexpr int dep4(int v, int mask) {
    int r = 0, b = 0;
    for (int i = 0; i < 4; i++)
        if (mask >> i & 1) {
            if (v >> b & 1) r |= 1 << i;
            b++;
        }
    return r;
}
__host__ __device__ constexpr int popc4(int m) { return (m & 1) + (m >> 1 & 1) + (m >> 2 & 1) + (m >> 3 & 1); }

// XOR swizzle of LDS element indices: lanes that differ in any three tile bits whose
// positions are distinct mod 3 (mod 4 at complex64) hit distinct 16-byte bank groups.
template <typename R>
__device__ __forceinline__ uint32_t swz(uint32_t l) {
    if constexpr (sizeof(R) == 8) {
        uint32_t h = l >> 3;
        return l ^ ((h ^ (h >> 3) ^ (h >> 6) ^ (h >> 9)) & 7u);
    } else {
        uint32_t h = l >> 4;
        return l ^ ((h ^ (h >> 4) ^ (h >> 8)) & 15u);
    }
}

struct Lay {
    uint64_t tg;      // global index bits contributed by tile base + thread id
    uint32_t tl;      // tile-local bits contributed by thread id
    uint64_t rq[4];   // global bit of register slot s
    uint32_t rl[4];   // tile-local bit of register slot s
};

// byte j of four little-endian words (compile-time j after unrolling: scalar shifts)
__device__ __forceinline__ uint32_t byte_of(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3, int j) {
    const uint32_t w = j < 4 ? w0 : j < 8 ? w1 : j < 12 ? w2 : w3;
    return (w >> (8 * (j & 3))) & 255u;
}

// Layout from 16-byte arrays of tile-local bits and global qubits, read as four aligned words
// each (uniform addresses: scalar loads, no per-entry memory round trips).
__device__ __forceinline__ void set_layout(Lay &L, const uint8_t *__restrict__ lay, const uint8_t *__restrict__ gq,
                                           int m, uint32_t tid, uint64_t base) {
    const uint32_t *lw = reinterpret_cast<const uint32_t *>(lay);
    const uint32_t *gw = reinterpret_cast<const uint32_t *>(gq);
    const uint32_t l0 = lw[0], l1 = lw[1], l2 = lw[2], l3 = lw[3];
    const uint32_t g0 = gw[0], g1 = gw[1], g2 = gw[2], g3 = gw[3];
#pragma unroll
    for (int s = 0; s < 4; s++) {
        L.rl[s] = 1u << byte_of(l0, l1, l2, l3, s);
        L.rq[s] = 1ull << byte_of(g0, g1, g2, g3, s);
    }
    uint64_t tg = base;
    uint32_t tl = 0;
#pragma unroll
    for (int u = 0; u < 12; u++) {
        if (u < m - 4 && (tid >> u & 1)) {
            tl |= 1u << byte_of(l0, l1, l2, l3, 4 + u);
            tg |= 1ull << byte_of(g0, g1, g2, g3, 4 + u);
        }
    }
    L.tg = tg;
    L.tl = tl;
}

template <int r>
__device__ __forceinline__ uint64_t rgof(const Lay &L) {
    return ((r & 1) ? L.rq[0] : 0ull) | ((r & 2) ? L.rq[1] : 0ull) | ((r & 4) ? L.rq[2] : 0ull) |
           ((r & 8) ? L.rq[3] : 0ull);
}
template <int r>
__device__ __forceinline__ uint32_t rlof(const Lay &L) {
    return ((r & 1) ? L.rl[0] : 0u) | ((r & 2) ? L.rl[1] : 0u) | ((r & 4) ? L.rl[2] : 0u) | ((r & 8) ? L.rl[3] : 0u);
}

// compile-time loop helper
template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

template <typename R>
__device__ __forceinline__ cx<R> cmulp(cx<R> a, R br, R bi) {
    return {a.x * br - a.y * bi, a.x * bi + a.y * br};
}

// Register-slot pattern of a global qubit: bit r set iff register index r has that qubit = 1
// (uniform across the wavefront; zero when the qubit is not a register slot).
__device__ __forceinline__ uint32_t reg_pattern(const Lay &L, uint64_t qb) {
    return (L.rq[0] == qb ? 0xAAAAu : 0u) | (L.rq[1] == qb ? 0xCCCCu : 0u) | (L.rq[2] == qb ? 0xF0F0u : 0u) |
           (L.rq[3] == qb ? 0xFF00u : 0u);
}

// Controls split into a uniform pattern over register indices and one per-thread flag:
// the control holds for register index r iff tok && (pat >> r & 1).
struct Ctrl {
    uint32_t pat;
    bool tok;
};
__device__ __forceinline__ Ctrl ctrl_split(uint64_t cm, const Lay &L) {
    uint32_t pat = 0xFFFFu;
    if (L.rq[0] & cm) pat &= 0xAAAAu;
    if (L.rq[1] & cm) pat &= 0xCCCCu;
    if (L.rq[2] & cm) pat &= 0xF0F0u;
    if (L.rq[3] & cm) pat &= 0xFF00u;
    const uint64_t cmt = cm & ~(L.rq[0] | L.rq[1] | L.rq[2] | L.rq[3]);
    return {pat, (L.tg & cmt) == cmt};
}

template <int MASK, typename R>
__device__ __forceinline__ void op_dense(cx<R> (&a)[16], const R *__restrict__ M, uint64_t cm, const Lay &L) {
    constexpr int K = popc4(MASK);
    constexpr int D = 1 << K;
    constexpr int OUTER = 16 >> K;
    constexpr int OM = 15 & ~MASK;
    const Ctrl c = ctrl_split(cm, L);
    static_for<0, OUTER>([&](auto oc) __attribute__((always_inline)) {
        constexpr int base = dep4(decltype(oc)::value, OM);
        const bool ok = c.tok && ((c.pat >> base) & 1u);
        cx<R> x[D];
        static_for<0, D>([&](auto jc) __attribute__((always_inline)) { x[decltype(jc)::value] = a[base | dep4(decltype(jc)::value, MASK)]; });
        static_for<0, D>([&](auto ic) __attribute__((always_inline)) {
            constexpr int i = decltype(ic)::value;
            R re = 0, im = 0;
            static_for<0, D>([&](auto jc) __attribute__((always_inline)) {
                constexpr int j = decltype(jc)::value;
                const R mr = M[2 * (i * D + j)], mi = M[2 * (i * D + j) + 1];
                re = fma(mr, x[j].x, re);
                re = fma(-mi, x[j].y, re);
                im = fma(mr, x[j].y, im);
                im = fma(mi, x[j].x, im);
            });
            constexpr int dst = base | dep4(i, MASK);
            if (ok) a[dst] = cx<R>{re, im};
        });
    });
}

template <int S, typename R>
__device__ __forceinline__ void op_anti1(cx<R> (&a)[16], R p0r, R p0i, R p1r, R p1i, bool pure, uint64_t cm,
                                         const Lay &L) {
    const Ctrl c = ctrl_split(cm, L);
    static_for<0, 16>([&](auto rc) __attribute__((always_inline)) {
        constexpr int r = decltype(rc)::value;
        if constexpr (!(r >> S & 1)) {
            constexpr int r1 = r | (1 << S);
            const bool ok = c.tok && ((c.pat >> r) & 1u);
            cx<R> a0 = a[r], a1 = a[r1];
            cx<R> n0 = pure ? a1 : cmulp(a1, p0r, p0i);
            cx<R> n1 = pure ? a0 : cmulp(a0, p1r, p1i);
            if (ok) {
                a[r] = n0;
                a[r1] = n1;
            }
        }
    });
}

template <int S0, int S1, typename R>
__device__ __forceinline__ void op_swap2(cx<R> (&a)[16], R pr, R pi, bool pure, uint64_t cm, const Lay &L) {
    const Ctrl c = ctrl_split(cm, L);
    static_for<0, 16>([&](auto rc) __attribute__((always_inline)) {
        constexpr int r = decltype(rc)::value;
        if constexpr (!(r >> S0 & 1) && !(r >> S1 & 1)) {
            constexpr int i01 = r | (1 << S1), i10 = r | (1 << S0);
            const bool ok = c.tok && ((c.pat >> r) & 1u);
            cx<R> a01 = a[i01], a10 = a[i10];
            cx<R> n01 = pure ? a10 : cmulp(a10, pr, pi);
            cx<R> n10 = pure ? a01 : cmulp(a01, pr, pi);
            if (ok) {
                a[i01] = n01;
                a[i10] = n10;
            }
        }
    });
}

template <int S, typename R>
__device__ __forceinline__ void op_reset(cx<R> (&a)[16]) {
    static_for<0, 16>([&](auto rc) __attribute__((always_inline)) {
        constexpr int r = decltype(rc)::value;
        if constexpr (!(r >> S & 1)) {
            constexpr int r1 = r | (1 << S);
            cx<R> a0 = a[r], a1 = a[r1];
            R nrm = sqrt(a0.x * a0.x + a0.y * a0.y + a1.x * a1.x + a1.y * a1.y);
            a[r] = nrm > R(1e-15) ? cx<R>{nrm, R(0)} : cx<R>{R(0), R(0)};
            a[r1] = cx<R>{R(0), R(0)};
        }
    });
}

template <typename R>
__device__ __forceinline__ void op_diag(cx<R> (&a)[16], const Instr &I, const R *__restrict__ D, const Lay &L) {
    const int k = I.k;
    const uint64_t cm = I.cmask;
    const uint32_t qw = *reinterpret_cast<const uint32_t *>(I.q);
    const int q0 = qw & 255u, q1 = (qw >> 8) & 255u, q2 = (qw >> 16) & 255u, q3 = qw >> 24;
    if (k <= 2) {
        // index bits split like the controls: a uniform register pattern per diag qubit and a
        // per-thread part folded into a per-thread table e[ri] = D[tidx | ri]
        const uint32_t m0 = reg_pattern(L, 1ull << q0);
        const uint32_t m1 = k == 2 ? reg_pattern(L, 1ull << q1) : 0u;
        const bool t0 = (L.tg >> q0) & 1, t1 = k == 2 ? ((L.tg >> q1) & 1) : false;
        const Ctrl c = ctrl_split(cm, L);
        const R d0r = D[0], d0i = D[1], d1r = D[2], d1i = D[3];
        const R d2r = k == 2 ? D[4] : R(0), d2i = k == 2 ? D[5] : R(0);
        const R d3r = k == 2 ? D[6] : R(0), d3i = k == 2 ? D[7] : R(0);
        R e0r, e0i, e1r, e1i, e2r = 0, e2i = 0, e3r = 0, e3i = 0;
        if (k == 2) {  // bit 1 <- q0, bit 0 <- q1
            e0r = t0 ? (t1 ? d3r : d2r) : (t1 ? d1r : d0r);
            e0i = t0 ? (t1 ? d3i : d2i) : (t1 ? d1i : d0i);
            e1r = t0 ? d3r : d1r;
            e1i = t0 ? d3i : d1i;
            e2r = t1 ? d3r : d2r;
            e2i = t1 ? d3i : d2i;
            e3r = d3r;
            e3i = d3i;
        } else {
            e0r = t0 ? d1r : d0r;
            e0i = t0 ? d1i : d0i;
            e1r = d1r;
            e1i = d1i;
        }
        if (c.tok) {  // per-lane predicate; everything below branches on uniform values
            if ((m0 | m1) == 0) {  // no diag qubit in registers: one factor per thread
                static_for<0, 16>([&](auto rc) __attribute__((always_inline)) {
                    constexpr int r = decltype(rc)::value;
                    if ((c.pat >> r) & 1u) a[r] = cmulp(a[r], e0r, e0i);
                });
            } else {
                static_for<0, 16>([&](auto rc) __attribute__((always_inline)) {
                    constexpr int r = decltype(rc)::value;
                    if ((c.pat >> r) & 1u) {
                        const uint32_t ri = (k == 2) ? ((((m0 >> r) & 1u) << 1) | ((m1 >> r) & 1u)) : ((m0 >> r) & 1u);
                        if (ri == 0) a[r] = cmulp(a[r], e0r, e0i);
                        else if (ri == 1) a[r] = cmulp(a[r], e1r, e1i);
                        else if (ri == 2) a[r] = cmulp(a[r], e2r, e2i);
                        else a[r] = cmulp(a[r], e3r, e3i);
                    }
                });
            }
        }
    } else {
        static_for<0, 16>([&](auto rc) __attribute__((always_inline)) {
            constexpr int r = decltype(rc)::value;
            const uint64_t g = L.tg | rgof<r>(L);
            const bool ok = (g & cm) == cm;
            int idx = (int)((g >> q0) & 1);
            idx = (idx << 1) | (int)((g >> q1) & 1);
            idx = (idx << 1) | (int)((g >> q2) & 1);
            if (k == 4) idx = (idx << 1) | (int)((g >> q3) & 1);
            const R dr = D[2 * idx], di = D[2 * idx + 1];
            cx<R> v = cmulp(a[r], dr, di);
            if (ok) a[r] = v;
        });
    }
}

// Re sum over the op's groups of conj(b) (M a): the derivative term of the adjoint method.
template <int MASK, typename R>
__device__ __forceinline__ double op_grad(const cx<R> (&a)[16], const cx<R> (&b)[16], const R *__restrict__ M,
                                          uint64_t cm, const Lay &L) {
    constexpr int K = popc4(MASK);
    constexpr int D = 1 << K;
    constexpr int OUTER = 16 >> K;
    constexpr int OM = 15 & ~MASK;
    const Ctrl c = ctrl_split(cm, L);
    double acc = 0;
    static_for<0, OUTER>([&](auto oc) __attribute__((always_inline)) {
        constexpr int base = dep4(decltype(oc)::value, OM);
        if (c.tok && ((c.pat >> base) & 1u)) {
            static_for<0, D>([&](auto ic) __attribute__((always_inline)) {
                constexpr int i = decltype(ic)::value;
                R re = 0, im = 0;
                static_for<0, D>([&](auto jc) __attribute__((always_inline)) {
                    constexpr int j = decltype(jc)::value;
                    const cx<R> x = a[base | dep4(j, MASK)];
                    const R mr = M[2 * (i * D + j)], mi = M[2 * (i * D + j) + 1];
                    re = fma(mr, x.x, re);
                    re = fma(-mi, x.y, re);
                    im = fma(mr, x.y, im);
                    im = fma(mi, x.x, im);
                });
                const cx<R> bb = b[base | dep4(i, MASK)];
                acc += (double)bb.x * (double)re + (double)bb.y * (double)im;
            });
        }
    });
    return acc;
}

// One stage over one state (PAIR = false) or over psi and lambda together (PAIR = true: every op
// acts on both, and I_GRAD accumulates Re <lambda| M |psi> per gradient slot).
template <typename R, bool PAIR>
__device__ __forceinline__ void stage_body(cx<R> *__restrict__ st, cx<R> *__restrict__ st2,
                                           const StageHeader *__restrict__ hdr, const Instr *__restrict__ prog,
                                           const R *__restrict__ pool, uint64_t ntiles, double *__restrict__ gpart,
                                           double *gacc) {
    extern __shared__ unsigned char smem_raw[];
    cx<R> *lds = reinterpret_cast<cx<R> *>(smem_raw);
    const int m = hdr->m;
    const int ninstr = hdr->ninstr;
    const uint32_t tid = threadIdx.x;
    const Instr *__restrict__ P = prog + hdr->instr_off;
    const uint32_t *Tw = reinterpret_cast<const uint32_t *>(hdr->T);
    const uint32_t T0 = Tw[0], T1 = Tw[1], T2 = Tw[2], T3 = Tw[3];
    const int wave = tid >> 6, lane = tid & 63;
    int ngrad = 0;
    if constexpr (PAIR) {
        ngrad = hdr->ngrad;
        for (int i = tid; i < ngrad * 8; i += blockDim.x) gacc[i] = 0;
        __syncthreads();
    }

    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        // deposit the tile index into the qubits outside the tile (T ascending)
        uint64_t base = tile;
#pragma unroll
        for (int j = 0; j < 16; j++) {
            if (j < m) {
                const uint32_t t = byte_of(T0, T1, T2, T3, j);
                const uint64_t lo = base & ((1ull << t) - 1);
                base = ((base >> t) << (t + 1)) | lo;
            }
        }
        Lay L;
        set_layout(L, hdr->lay0, hdr->gq0, m, tid, base);
        cx<R> a[16], b[PAIR ? 16 : 1];
        static_for<0, 16>([&](auto rc) __attribute__((always_inline)) {
            constexpr int r = decltype(rc)::value;
            a[r] = st[L.tg | rgof<r>(L)];
            if constexpr (PAIR) b[r] = st2[L.tg | rgof<r>(L)];
        });

        for (int pc = 0; pc < ninstr; pc++) {
            const Instr &I = P[pc];
            const int op = I.op;
            // slots and diag qubits as one word each
            const uint32_t sw = *reinterpret_cast<const uint32_t *>(I.s);
            if (op == I_REMAP) {
                __syncthreads();
                static_for<0, 16>([&](auto rc) __attribute__((always_inline)) {
                    constexpr int r = decltype(rc)::value;
                    lds[swz<R>(L.tl | rlof<r>(L))] = a[r];
                });
                __syncthreads();
                Lay Ln;
                set_layout(Ln, I.lay, I.gq, m, tid, base);
                static_for<0, 16>([&](auto rc) __attribute__((always_inline)) {
                    constexpr int r = decltype(rc)::value;
                    a[r] = lds[swz<R>(Ln.tl | rlof<r>(Ln))];
                });
                if constexpr (PAIR) {
                    __syncthreads();
                    static_for<0, 16>([&](auto rc) __attribute__((always_inline)) {
                        constexpr int r = decltype(rc)::value;
                        lds[swz<R>(L.tl | rlof<r>(L))] = b[r];
                    });
                    __syncthreads();
                    static_for<0, 16>([&](auto rc) __attribute__((always_inline)) {
                        constexpr int r = decltype(rc)::value;
                        b[r] = lds[swz<R>(Ln.tl | rlof<r>(Ln))];
                    });
                }
                L = Ln;
            } else if (op == I_DENSE) {
                const R *__restrict__ M = pool + 2 * I.mat;
#define DENSE_CASE(MK)                                     \
    case MK:                                               \
        op_dense<MK>(a, M, I.cmask, L);                    \
        if constexpr (PAIR) op_dense<MK>(b, M, I.cmask, L); \
        break;
                switch (I.smask) {
                    DENSE_CASE(1)
                    DENSE_CASE(2)
                    DENSE_CASE(3)
                    DENSE_CASE(4)
                    DENSE_CASE(5)
                    DENSE_CASE(6)
                    DENSE_CASE(7)
                    DENSE_CASE(8)
                    DENSE_CASE(9)
                    DENSE_CASE(10)
                    DENSE_CASE(11)
                    DENSE_CASE(12)
                    DENSE_CASE(13)
                    DENSE_CASE(14)
                    DENSE_CASE(15)
                default:
                    break;
                }
#undef DENSE_CASE
            } else if (op == I_DIAG) {
                op_diag<R>(a, I, pool + 2 * I.mat, L);
                if constexpr (PAIR) op_diag<R>(b, I, pool + 2 * I.mat, L);
            } else if (op == I_ANTI1) {
                const R *__restrict__ p = pool + 2 * I.mat;
                const bool pure = I.flags & 1;
                const R p0r = p[0], p0i = p[1], p1r = p[2], p1i = p[3];
#define ANTI_CASE(S)                                                                      \
    case S:                                                                               \
        op_anti1<S>(a, p0r, p0i, p1r, p1i, pure, I.cmask, L);                             \
        if constexpr (PAIR) op_anti1<S>(b, p0r, p0i, p1r, p1i, pure, I.cmask, L);         \
        break;
                switch (sw & 255u) {
                    ANTI_CASE(0)
                    ANTI_CASE(1)
                    ANTI_CASE(2)
                    ANTI_CASE(3)
                default:
                    break;
                }
#undef ANTI_CASE
            } else if (op == I_SWAP2) {
                const R *__restrict__ p = pool + 2 * I.mat;
                const bool pure = I.flags & 1;
                const R pr = p[0], pi = p[1];
                const int sa = sw & 255u, sb = (sw >> 8) & 255u;
                const int s0 = sa < sb ? sa : sb, s1 = sa < sb ? sb : sa;
#define SWAP_CASE(A, B)                                                    \
    case A * 4 + B:                                                        \
        op_swap2<A, B>(a, pr, pi, pure, I.cmask, L);                       \
        if constexpr (PAIR) op_swap2<A, B>(b, pr, pi, pure, I.cmask, L);   \
        break;
                switch (s0 * 4 + s1) {
                    SWAP_CASE(0, 1)
                    SWAP_CASE(0, 2)
                    SWAP_CASE(0, 3)
                    SWAP_CASE(1, 2)
                    SWAP_CASE(1, 3)
                    SWAP_CASE(2, 3)
                default:
                    break;
                }
#undef SWAP_CASE
            } else if (op == I_RESET) {
                switch (sw & 255u) {
                case 0: op_reset<0>(a); break;
                case 1: op_reset<1>(a); break;
                case 2: op_reset<2>(a); break;
                default: op_reset<3>(a); break;
                }
            } else if (op == I_GRAD) {
                if constexpr (PAIR) {
                    const R *__restrict__ M = pool + 2 * I.mat;
                    double v = 0;
#define GRAD_CASE(MK)                                \
    case MK:                                         \
        v = op_grad<MK>(a, b, M, I.cmask, L);        \
        break;
                    switch (I.smask) {
                        GRAD_CASE(1)
                        GRAD_CASE(2)
                        GRAD_CASE(3)
                        GRAD_CASE(4)
                        GRAD_CASE(5)
                        GRAD_CASE(6)
                        GRAD_CASE(8)
                        GRAD_CASE(9)
                        GRAD_CASE(10)
                        GRAD_CASE(12)
                    default:
                        break;
                    }
#undef GRAD_CASE
                    // blocks of fewer than 64 threads only read active lanes
#pragma unroll
                    for (int off = 32; off >= 1; off >>= 1)
                        if (off < (int)blockDim.x) v += __shfl_xor(v, off, 64);
                    if (lane == 0) gacc[I.pad * 8 + wave] += v;  // one writer per (slot, wave)
                }
            }
        }
        static_for<0, 16>([&](auto rc) __attribute__((always_inline)) {
            constexpr int r = decltype(rc)::value;
            st[L.tg | rgof<r>(L)] = a[r];
            if constexpr (PAIR) st2[L.tg | rgof<r>(L)] = b[r];
        });
        if (tile + gridDim.x < ntiles) __syncthreads();
    }
    if constexpr (PAIR) {
        __syncthreads();
        const int nw = blockDim.x >> 6;
        for (int g = tid; g < ngrad; g += blockDim.x) {
            double s = 0;
            for (int w = 0; w < nw; w++) s += gacc[g * 8 + w];
            gpart[(size_t)g * gridDim.x + blockIdx.x] = s;
        }
    }
}

template <typename R>
__global__ void __launch_bounds__(512) k_stage(cx<R> *__restrict__ st, const StageHeader *__restrict__ hdr,
                                               const Instr *__restrict__ prog, const R *__restrict__ pool,
                                               uint64_t ntiles) {
    stage_body<R, false>(st, nullptr, hdr, prog, pool, ntiles, nullptr, nullptr);
}

template <typename R>
__global__ void __launch_bounds__(256) k_stage_pair(cx<R> *__restrict__ psi, cx<R> *__restrict__ lam,
                                                    const StageHeader *__restrict__ hdr,
                                                    const Instr *__restrict__ prog, const R *__restrict__ pool,
                                                    uint64_t ntiles, double *__restrict__ gpart) {
    __shared__ double gacc[kMaxGradPerStage * 8];
    stage_body<R, true>(psi, lam, hdr, prog, pool, ntiles, gpart, gacc);
}

// ---- profiling ---------------------------------------------------------------------------------
static std::atomic<bool> g_profile{false};
static std::mutex g_prof_mu;
static double g_prof_ms = 0, g_prof_bytes = 0;
static int64_t g_prof_launches = 0;

static void profile_add(double ms, int64_t launches, double bytes) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof_ms += ms;
    g_prof_launches += launches;
    g_prof_bytes += bytes;
}

// ---------------------------------------------------------------------------------------------
PlanConfig default_plan_config(int prec, int nq) {
    PlanConfig c;
    c.tile_bits = 12;
    c.low_bits = 3;
    (void)prec;
    (void)nq;
    return c;
}

int upload_program(const Program &P, int prec, hipStream_t s, DeviceProgram &dp) {
    const size_t rsz = prec == GOQU_C128 ? sizeof(double) : sizeof(float);
    const size_t nh = P.stages.size(), ni = P.instrs.size(), np = P.pool.size();
    dp.hdr_off = 0;
    dp.instr_off = (nh * sizeof(StageHeader) + 255) & ~size_t(255);
    dp.pool_off = (dp.instr_off + ni * sizeof(Instr) + 255) & ~size_t(255);
    const size_t bytes = dp.pool_off + np * rsz + 64;
    std::vector<unsigned char> host(bytes, 0);
    std::memcpy(host.data(), P.stages.data(), nh * sizeof(StageHeader));
    std::memcpy(host.data() + dp.instr_off, P.instrs.data(), ni * sizeof(Instr));
    if (prec == GOQU_C128) {
        std::memcpy(host.data() + dp.pool_off, P.pool.data(), np * sizeof(double));
    } else {
        float *f = reinterpret_cast<float *>(host.data() + dp.pool_off);
        for (size_t i = 0; i < np; i++) f[i] = (float)P.pool[i];
    }
    if (dp.bytes < bytes) {
        if (dp.blob) GOQU_HIP(hipFree(dp.blob));
        dp.blob = nullptr;
        GOQU_HIP(hipMalloc(&dp.blob, bytes));
        dp.bytes = bytes;
    }
    GOQU_HIP(hipMemcpyAsync(dp.blob, host.data(), bytes, hipMemcpyHostToDevice, s));
    // the host buffer must outlive the async copy
    GOQU_HIP(hipStreamSynchronize(s));
    dp.stages = P.stages;
    return GOQU_OK;
}

template <typename R>
static int run_program_t(State &st, const DeviceProgram &dp, hipStream_t s) {
    const unsigned char *blob = static_cast<const unsigned char *>(dp.blob);
    const StageHeader *hdrs = reinterpret_cast<const StageHeader *>(blob + dp.hdr_off);
    const Instr *instrs = reinterpret_cast<const Instr *>(blob + dp.instr_off);
    const R *pool = reinterpret_cast<const R *>(blob + dp.pool_off);
    const bool prof = g_profile.load();
    std::vector<hipEvent_t> ev;
    for (size_t i = 0; i < dp.stages.size(); i++) {
        const int m = dp.stages[i].m;
        if (dp.stages[i].ninstr == 0) continue;
        const uint64_t ntiles = 1ull << (st.nq - m);
        const unsigned threads = 1u << (m - kRegBits);
        const size_t lds = (size_t(1) << m) * sizeof(cx<R>);
        const uint64_t cap = 1ull << 20;
        const unsigned grid = (unsigned)(ntiles < cap ? ntiles : cap);
        if (prof) {
            ev.resize(ev.size() + 2);
            GOQU_HIP(hipEventCreate(&ev[ev.size() - 2]));
            GOQU_HIP(hipEventCreate(&ev[ev.size() - 1]));
            GOQU_HIP(hipEventRecord(ev[ev.size() - 2], s));
        }
        hipLaunchKernelGGL(k_stage<R>, dim3(grid), dim3(threads), lds, s, static_cast<cx<R> *>(st.amps),
                           hdrs + i, instrs, pool, ntiles);
        GOQU_HIP(hipGetLastError());
        if (prof) GOQU_HIP(hipEventRecord(ev.back(), s));
    }
    if (prof && !ev.empty()) {
        GOQU_HIP(hipStreamSynchronize(s));
        double ms_total = 0;
        for (size_t i = 0; i < ev.size(); i += 2) {
            float ms = 0;
            GOQU_HIP(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
            ms_total += ms;
            (void)hipEventDestroy(ev[i]);
            (void)hipEventDestroy(ev[i + 1]);
        }
        profile_add(ms_total, (int64_t)(ev.size() / 2), (double)(ev.size() / 2) * 2.0 * st.size() * sizeof(cx<R>));
    }
    return GOQU_OK;
}

int run_program(State &st, const DeviceProgram &dp, hipStream_t s) {
    return st.prec == GOQU_C128 ? run_program_t<double>(st, dp, s) : run_program_t<float>(st, dp, s);
}

// sums[slot] = sum over b of part[slot * nb + b], in a fixed order
__global__ void __launch_bounds__(256) k_sum_rows(const double *__restrict__ part, int64_t nb,
                                                  double *__restrict__ sums) {
    __shared__ double sh[256];
    const double *row = part + (size_t)blockIdx.x * nb;
    double v = 0;
    for (int64_t b = threadIdx.x; b < nb; b += blockDim.x) v += row[b];
    sh[threadIdx.x] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int i = 0; i < (int)blockDim.x; i++) t += sh[i];
        sums[blockIdx.x] = t;
    }
}

template <typename R>
static int run_pair_t(State &psi, State &lam, const Program &P, hipStream_t s, std::vector<double> &slot_vals) {
    static thread_local DeviceProgram dp;
    int rc = upload_program(P, psi.prec, s, dp);
    if (rc) return rc;
    const unsigned char *blob = static_cast<const unsigned char *>(dp.blob);
    const StageHeader *hdrs = reinterpret_cast<const StageHeader *>(blob + dp.hdr_off);
    const Instr *instrs = reinterpret_cast<const Instr *>(blob + dp.instr_off);
    const R *pool = reinterpret_cast<const R *>(blob + dp.pool_off);
    const int m = P.stages.empty() ? 0 : P.stages[0].m;
    if (m > 12) return fail(GOQU_ERR_UNSUPPORTED, "paired stages need tiles of at most 2^12 amplitudes");
    const uint64_t ntiles = 1ull << (psi.nq - m);
    const unsigned grid = (unsigned)std::min<uint64_t>(ntiles, 1ull << 20);
    const int ng = P.num_grad;
    // per-stage partials (slot-major, one per workgroup), reduced after every stage into gsum
    double *gpart = nullptr, *gsum = nullptr;
    if (ng) {
        GOQU_HIP(hipMalloc(&gpart, ((size_t)kMaxGradPerStage * grid + ng) * sizeof(double)));
        gsum = gpart + (size_t)kMaxGradPerStage * grid;
    }
    for (size_t i = 0; i < P.stages.size() && !rc; i++) {
        const StageHeader &H = P.stages[i];
        if (H.ninstr == 0) continue;
        const unsigned threads = 1u << (m - kRegBits);
        const size_t lds = (size_t(1) << m) * sizeof(cx<R>);
        hipLaunchKernelGGL(k_stage_pair<R>, dim3(grid), dim3(threads), lds, s, static_cast<cx<R> *>(psi.amps),
                           static_cast<cx<R> *>(lam.amps), hdrs + i, instrs, pool, ntiles, gpart);
        if (hipGetLastError() != hipSuccess) rc = fail(GOQU_ERR_DEVICE, "paired stage launch failed");
        if (!rc && H.ngrad) {
            hipLaunchKernelGGL(k_sum_rows, dim3(H.ngrad), dim3(256), 0, s, gpart, (int64_t)grid, gsum + H.grad_base);
            if (hipGetLastError() != hipSuccess) rc = fail(GOQU_ERR_DEVICE, "gradient reduction launch failed");
        }
    }
    slot_vals.assign(ng, 0.0);
    hipError_t e = hipSuccess;
    if (!rc && ng) e = hipMemcpyAsync(slot_vals.data(), gsum, ng * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (gpart) (void)hipFree(gpart);
    if (!rc && e != hipSuccess) rc = fail(GOQU_ERR_DEVICE, std::string("paired stages: ") + hipGetErrorString(e));
    return rc;
}

// Plans `eops` (which may hold EOP_GRAD ops) and runs it on psi and lam together; slot_vals
// receives Re <lam| D |psi> per gradient slot, indexed like the EOps' `slot` (slots absent from
// the program stay 0).
int run_pair(State &psi, State &lam, const std::vector<EOp> &eops, int nslots, std::vector<double> &slot_vals,
             hipStream_t s) {
    slot_vals.assign(nslots, 0.0);
    if (eops.empty()) return GOQU_OK;
    Program P = plan_program(eops, psi.nq, default_plan_config(psi.prec, psi.nq));
    std::vector<double> prog_vals;
    int rc = psi.prec == GOQU_C128 ? run_pair_t<double>(psi, lam, P, s, prog_vals)
                                   : run_pair_t<float>(psi, lam, P, s, prog_vals);
    if (rc) return rc;
    for (size_t j = 0; j < prog_vals.size(); j++) {
        const int slot = P.grad_slots[j];
        if (slot >= 0 && slot < nslots) slot_vals[slot] += prog_vals[j];
    }
    return GOQU_OK;
}

int run_eops(State &st, const std::vector<EOp> &eops, hipStream_t s) {
    if (eops.empty()) return GOQU_OK;
    Program P = plan_program(eops, st.nq, default_plan_config(st.prec, st.nq));
    // grow-only program buffer per host thread: every call ends synchronised, so the next call
    // on this thread may overwrite it
    static thread_local DeviceProgram dp;
    int rc = upload_program(P, st.prec, s, dp);
    if (rc) return rc;
    rc = run_program(st, dp, s);
    hipError_t e = hipStreamSynchronize(s);
    if (rc) return rc;
    if (e != hipSuccess) return fail(GOQU_ERR_DEVICE, std::string("stage kernels: ") + hipGetErrorString(e));
    return GOQU_OK;
}

}  // namespace goqu

extern "C" {
int goqu_profile_enable(int32_t on) {
    goqu::g_profile.store(on != 0);
    return GOQU_OK;
}
int goqu_profile_reset(void) {
    std::lock_guard<std::mutex> lk(goqu::g_prof_mu);
    goqu::g_prof_ms = goqu::g_prof_bytes = 0;
    goqu::g_prof_launches = 0;
    return GOQU_OK;
}
int goqu_profile_read(double *kernel_ms, int64_t *launches, double *bytes) {
    std::lock_guard<std::mutex> lk(goqu::g_prof_mu);
    if (kernel_ms) *kernel_ms = goqu::g_prof_ms;
    if (launches) *launches = goqu::g_prof_launches;
    if (bytes) *bytes = goqu::g_prof_bytes;
    return GOQU_OK;
}
}
