// sv_stage.hip — the tiled gate-application kernel for gfx950.
//
// One launch = one stage = one HBM sweep.  A workgroup of 2^(m-4) threads owns a tile of
// 2^m amplitudes: the tile's qubits are T[0..m-1] (the lowest L qubits plus the stage's
// chosen qubits); the remaining n-m qubit values are fixed per tile.  Each thread keeps 16
// amplitudes in registers: register slot s <-> tile-local bit lay[s], thread-id bit u <->
// tile-local bit lay[4+u].  Every instruction arrives resolved against the layout in force
// (register slots as compile-time register indices, controls as a register pattern, a
// thread-id mask and a tile-uniform mask), so the per-op work is the arithmetic itself:
// uncontrolled ops are straight-line FMAs on registers, thread-bit controls select the matrix
// (or the identity) once per op, register controls skip whole register groups.  I_REMAP
// re-distributes the tile through LDS; its offsets split into a per-thread part and a
// host-computed per-register part (the swizzle is XOR-linear).  Global loads/stores use layouts
// whose lanes 0..2 walk the tile's lowest qubits (runs of >= 8 amplitudes, 128 B at complex128).
#include <atomic>
#include <cstdlib>
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
// the bits of v at the positions of mask, packed (inverse of dep4)
__host__ __device__ constexpr int pext4(int v, int mask) {
    int r = 0, b = 0;
    for (int i = 0; i < 4; i++)
        if (mask >> i & 1) {
            if (v >> i & 1) r |= 1 << b;
            b++;
        }
    return r;
}
__host__ __device__ constexpr int popc4(int m) { return (m & 1) + (m >> 1 & 1) + (m >> 2 & 1) + (m >> 3 & 1); }

// the LDS swizzle of lds_swizzle() for 8-byte elements.  Remaps move 8-byte LDS elements: a
// complex64 amplitude, or one component of a complex128 amplitude (real parts in one pass,
// imaginary parts in a second), so a 2^12 tile needs 32 KiB of LDS at either precision.
__device__ __forceinline__ uint32_t swz8(uint32_t l) {
    const uint32_t h = l >> 4;
    return l ^ ((h ^ (h >> 4) ^ (h >> 8)) & 15u);
}

// byte j of four little-endian words (compile-time j after unrolling: scalar shifts)
__device__ __forceinline__ uint32_t byte_of(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3, int j) {
    const uint32_t w = j < 4 ? w0 : j < 8 ? w1 : j < 12 ? w2 : w3;
    return (w >> (8 * (j & 3))) & 255u;
}

// Global addressing of a layout (tile loads and stores): amplitude index of register index r
// is tg | rq(r), with tg the tile base plus the thread bits' qubits.
struct GAddr {
    uint64_t tg;
    uint64_t rq[4];
};
__device__ __forceinline__ void global_layout(GAddr &G, const uint8_t *__restrict__ gq, int m, uint32_t tid,
                                              uint64_t base) {
    const uint32_t *gw = reinterpret_cast<const uint32_t *>(gq);
    const uint32_t g0 = gw[0], g1 = gw[1], g2 = gw[2], g3 = gw[3];
#pragma unroll
    for (int s = 0; s < 4; s++) G.rq[s] = 1ull << byte_of(g0, g1, g2, g3, s);
    // thread-id bits deposited arithmetically (a per-bit select would keep one lane mask per bit
    // in SGPRs for the whole kernel)
    uint64_t tg = base;
#pragma unroll
    for (int u = 0; u < 12; u++)
        if (u < m - 4) tg |= (uint64_t)((tid >> u) & 1u) << byte_of(g0, g1, g2, g3, 4 + u);
    G.tg = tg;
}
template <int r>
__device__ __forceinline__ uint64_t rgof(const GAddr &G) {
    return ((r & 1) ? G.rq[0] : 0ull) | ((r & 2) ? G.rq[1] : 0ull) | ((r & 4) ? G.rq[2] : 0ull) |
           ((r & 8) ? G.rq[3] : 0ull);
}

// Tile loads and stores stream: each amplitude is read and written once per sweep, so they
// carry the nontemporal hint.
template <typename R>
__device__ __forceinline__ cx<R> ld_stream(const cx<R> *p) {
    typedef R v2 __attribute__((ext_vector_type(2)));
    const v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2 *>(p));
    return {v.x, v.y};
}
template <typename R>
__device__ __forceinline__ void st_stream(cx<R> *p, cx<R> a) {
    typedef R v2 __attribute__((ext_vector_type(2)));
    v2 v;
    v.x = a.x;
    v.y = a.y;
    __builtin_nontemporal_store(v, reinterpret_cast<v2 *>(p));
}

// per-thread part of the LDS byte offsets of a layout
template <typename R>
__device__ __forceinline__ uint32_t lds_lane(const uint8_t *__restrict__ lay, int m, uint32_t tid) {
    const uint32_t *lw = reinterpret_cast<const uint32_t *>(lay);
    const uint32_t l0 = lw[0], l1 = lw[1], l2 = lw[2], l3 = lw[3];
    uint32_t tl = 0;
#pragma unroll
    for (int u = 0; u < 12; u++)
        if (u < m - 4) tl |= ((tid >> u) & 1u) << byte_of(l0, l1, l2, l3, 4 + u);
    return swz8(tl) * 8u;
}

// per-register LDS byte offsets, two 16-bit values per word
struct KOff {
    uint32_t w[8];
};
__device__ __forceinline__ void load_koff(KOff &K, const uint16_t *__restrict__ k) {
    const uint32_t *kw = reinterpret_cast<const uint32_t *>(k);
#pragma unroll
    for (int i = 0; i < 8; i++) K.w[i] = kw[i];
}
template <int r>
__device__ __forceinline__ uint32_t koff(const KOff &K) {
    return (K.w[r >> 1] >> (16 * (r & 1))) & 0xFFFFu;
}

// The first 64 bytes of an Instr (everything the inner op loop reads), fetched as one batch of
// scalar loads.  I_SHEAR keeps its three coefficients (doubles) at bytes 40..63.
struct IHead {
    int32_t op, k, mat, flags;
    uint32_t pat, tmask;
    uint64_t omask;
    uint32_t sw;     // s[0..3]
    int32_t smask;
    uint32_t w10;    // pad (I_GRAD slot) / low word of shear coefficient 0
    uint32_t lsw;    // ls[0..3] / high word of shear coefficient 0
    uint32_t w12, w13, w14, w15;
};
static_assert(offsetof(Instr, lay) == 48 && offsetof(Instr, ls) == 44, "IHead mirrors Instr");
__device__ __forceinline__ IHead load_head(const Instr *__restrict__ P, int pc) {
    typedef uint32_t u16v __attribute__((ext_vector_type(16)));
    const u16v w = *reinterpret_cast<const u16v *>(P + pc);
    IHead h;
    h.op = (int32_t)w[0];
    h.k = (int32_t)w[1];
    h.mat = (int32_t)w[2];
    h.flags = (int32_t)w[3];
    h.pat = w[4];
    h.tmask = w[5];
    h.omask = (uint64_t)w[6] | ((uint64_t)w[7] << 32);
    h.sw = w[8];
    h.smask = (int32_t)w[9];
    h.w10 = w[10];
    h.lsw = w[11];
    h.w12 = w[12];
    h.w13 = w[13];
    h.w14 = w[14];
    h.w15 = w[15];
    return h;
}
__device__ __forceinline__ double shear_coef(const IHead &h, int i) {
    const uint32_t lo = i == 0 ? h.w10 : i == 1 ? h.w12 : h.w14;
    const uint32_t hi = i == 0 ? h.lsw : i == 1 ? h.w13 : h.w15;
    return __builtin_bit_cast(double, (uint64_t)lo | ((uint64_t)hi << 32));
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

// Register discipline of the instruction loop.  Every op leaves each amplitude in the register
// it arrived in: results are written back through tied ("+v") operands.  Without this the
// compiler keeps a second copy of the 16 amplitudes and moves all of them at the loop's
// control-flow joins, which doubles the VALU work of a gate.
__device__ __forceinline__ void commit(double &dst, double val) { asm("v_mov_b64 %0, %1" : "+v"(dst) : "v"(val)); }
__device__ __forceinline__ void commit(float &dst, float val) { asm("v_mov_b32 %0, %1" : "+v"(dst) : "v"(val)); }
template <typename R>
__device__ __forceinline__ void commit(cx<R> &dst, cx<R> val) {
    commit(dst.x, val.x);
    commit(dst.y, val.y);
}
// a <- a * (pr + i pi) in place; P in SGPRs (uniform phase) or VGPRs (per-lane phase).  The two
// products of the old imaginary part are plain code (the compiler schedules them early); the two
// FMAs that overwrite the amplitude are single tied instructions.
template <bool UNIFORM, typename R>
__device__ __forceinline__ void cmul_inplace(R &re, R &im, R pr, R pi) {
    const R t = im * pi, u = im * pr;
    if constexpr (sizeof(R) == 8) {
        if constexpr (UNIFORM) {
            asm("v_fma_f64 %0, %1, %2, %3" : "+v"(im) : "v"(re), "s"(pi), "v"(u));
            asm("v_fma_f64 %0, %0, %1, -%2" : "+v"(re) : "s"(pr), "v"(t));
        } else {
            asm("v_fma_f64 %0, %1, %2, %3" : "+v"(im) : "v"(re), "v"(pi), "v"(u));
            asm("v_fma_f64 %0, %0, %1, -%2" : "+v"(re) : "v"(pr), "v"(t));
        }
    } else {
        if constexpr (UNIFORM) {
            asm("v_fma_f32 %0, %1, %2, %3" : "+v"(im) : "v"(re), "s"(pi), "v"(u));
            asm("v_fma_f32 %0, %0, %1, -%2" : "+v"(re) : "s"(pr), "v"(t));
        } else {
            asm("v_fma_f32 %0, %1, %2, %3" : "+v"(im) : "v"(re), "v"(pi), "v"(u));
            asm("v_fma_f32 %0, %0, %1, -%2" : "+v"(re) : "v"(pr), "v"(t));
        }
    }
}
// acc <- acc + x * c in place, c uniform (SGPRs)
__device__ __forceinline__ void fma_inplace(double &acc, double x, double c) {
    asm("v_fma_f64 %0, %1, %2, %0" : "+v"(acc) : "v"(x), "s"(c));
}
__device__ __forceinline__ void fma_inplace(float &acc, float x, float c) {
    asm("v_fma_f32 %0, %1, %2, %0" : "+v"(acc) : "v"(x), "s"(c));
}
// a double swap as two v_swap_b32 on its halves (the halves are the double's own registers)
__device__ __forceinline__ void swap_inplace(double &x, double &y) {
    typedef uint32_t u2 __attribute__((ext_vector_type(2)));
    u2 a = __builtin_bit_cast(u2, x), b = __builtin_bit_cast(u2, y);
    uint32_t a0 = a[0], a1 = a[1], b0 = b[0], b1 = b[1];
    asm("v_swap_b32 %0, %1" : "+v"(a0), "+v"(b0));
    asm("v_swap_b32 %0, %1" : "+v"(a1), "+v"(b1));
    a[0] = a0;
    a[1] = a1;
    b[0] = b0;
    b[1] = b1;
    x = __builtin_bit_cast(double, a);
    y = __builtin_bit_cast(double, b);
}
__device__ __forceinline__ void swap_inplace(float &x, float &y) { asm("v_swap_b32 %0, %1" : "+v"(x), "+v"(y)); }
template <typename R>
__device__ __forceinline__ void swap_inplace(cx<R> &x, cx<R> &y) {
    swap_inplace(x.x, y.x);
    swap_inplace(x.y, y.y);
}

// Dense op on the register slots of MASK (at most two).  REAL: the matrix is real (rotations
// split off single-qubit gates), 2 FMAs per entry instead of 4.  CTRL: register groups outside `pat` are
// skipped; lanes whose thread-bit controls fail keep their amplitudes.  For one slot those lanes
// get the identity matrix (selected once per op); for two slots the outputs are selected, which
// keeps the 16-entry matrix in scalar registers (a per-lane copy would cost 64 VGPRs and lower
// the occupancy of the whole kernel).
template <int MASK, bool CTRL, bool PAIR, typename R, bool REAL = false>
__device__ __forceinline__ void op_dense(cx<R> (&a)[16], cx<R> (&b)[16], const R *__restrict__ M, uint32_t pat,
                                         bool lok) {
    constexpr int K = popc4(MASK);
    constexpr int D = 1 << K;
    constexpr int OUTER = 16 >> K;
    constexpr int OM = 15 & ~MASK;
    constexpr bool SEL_MATRIX = CTRL && K == 1;
    constexpr bool SEL_OUTPUT = CTRL && K == 2;
    R mm[2 * D * D];
    static_for<0, 2 * D * D>([&](auto ic) __attribute__((always_inline)) { mm[decltype(ic)::value] = M[decltype(ic)::value]; });
    if constexpr (SEL_MATRIX) {
        static_for<0, D * D>([&](auto ic) __attribute__((always_inline)) {
            constexpr int e = decltype(ic)::value;
            constexpr bool dg = (e / D) == (e % D);
            mm[2 * e] = lok ? mm[2 * e] : R(dg ? 1 : 0);
            mm[2 * e + 1] = lok ? mm[2 * e + 1] : R(0);
        });
    }
    static_for<0, OUTER>([&](auto oc) __attribute__((always_inline)) {
        constexpr int base = dep4(decltype(oc)::value, OM);
        if (!CTRL || ((pat >> base) & 1u)) {
            auto apply = [&](cx<R>(&v)[16]) __attribute__((always_inline)) {
                cx<R> x[D], y[D];
                static_for<0, D>([&](auto jc) __attribute__((always_inline)) {
                    x[decltype(jc)::value] = v[base | dep4(decltype(jc)::value, MASK)];
                });
                static_for<0, D>([&](auto ic) __attribute__((always_inline)) {
                    constexpr int i = decltype(ic)::value;
                    R re = 0, im = 0;
                    static_for<0, D>([&](auto jc) __attribute__((always_inline)) {
                        constexpr int j = decltype(jc)::value;
                        const R mr = mm[2 * (i * D + j)];
                        re = fma(mr, x[j].x, re);
                        im = fma(mr, x[j].y, im);
                        if constexpr (!REAL) {
                            const R mi = mm[2 * (i * D + j) + 1];
                            re = fma(-mi, x[j].y, re);
                            im = fma(mi, x[j].x, im);
                        }
                    });
                    if constexpr (SEL_OUTPUT) y[i] = cx<R>{lok ? re : x[i].x, lok ? im : x[i].y};
                    else y[i] = cx<R>{re, im};
                });
                static_for<0, D>([&](auto ic) __attribute__((always_inline)) {
                    commit(v[base | dep4(decltype(ic)::value, MASK)], y[decltype(ic)::value]);
                });
            };
            apply(a);
            if constexpr (PAIR) apply(b);
        }
    });
}

// A real rotation on slot S as NSH in-place shears (lifting steps): a0 += c0 a1, a1 += c1 a0,
// then (NSH = 3) a0 += c2 a1.  One FMA per amplitude component and shear, no temporaries and no
// write-backs; the planner picks coefficients whose products stay bounded.
// shears P0 .. P1-1 of a rotation on slot S (even: a0 += c a1, odd: a1 += c a0)
template <int S, int P0, int P1, bool PAIR, typename R>
__device__ __forceinline__ void shear_steps(cx<R> (&a)[16], cx<R> (&b)[16], const R *c) {
    auto apply = [&](cx<R>(&v)[16]) __attribute__((always_inline)) {
        static_for<P0, P1>([&](auto pc) __attribute__((always_inline)) {
            constexpr int p = decltype(pc)::value;
            static_for<0, 16>([&](auto rc) __attribute__((always_inline)) {
                constexpr int r = decltype(rc)::value;
                if constexpr (!(r >> S & 1)) {
                    constexpr int r1 = r | (1 << S);
                    if constexpr (p & 1) {
                        fma_inplace(v[r1].x, v[r].x, c[p]);
                        fma_inplace(v[r1].y, v[r].y, c[p]);
                    } else {
                        fma_inplace(v[r].x, v[r1].x, c[p]);
                        fma_inplace(v[r].y, v[r1].y, c[p]);
                    }
                }
            });
        });
    };
    apply(a);
    if constexpr (PAIR) apply(b);
}

template <int S, int NSH, bool PAIR, typename R>
__device__ __forceinline__ void op_shear(cx<R> (&a)[16], cx<R> (&b)[16], const IHead &h) {
    const R c[3] = {R(shear_coef(h, 0)), R(shear_coef(h, 1)), R(shear_coef(h, 2))};
    auto apply = [&](cx<R>(&v)[16]) __attribute__((always_inline)) {
        static_for<0, NSH>([&](auto pc) __attribute__((always_inline)) {
            constexpr int p = decltype(pc)::value;
            static_for<0, 16>([&](auto rc) __attribute__((always_inline)) {
                constexpr int r = decltype(rc)::value;
                if constexpr (!(r >> S & 1)) {
                    constexpr int r1 = r | (1 << S);
                    if constexpr (p & 1) {
                        fma_inplace(v[r1].x, v[r].x, c[p]);
                        fma_inplace(v[r1].y, v[r].y, c[p]);
                    } else {
                        fma_inplace(v[r].x, v[r1].x, c[p]);
                        fma_inplace(v[r].y, v[r1].y, c[p]);
                    }
                }
            });
        });
    };
    apply(a);
    if constexpr (PAIR) apply(b);
}

// Shears on several register slots (I_SHEARS): slot s takes (cnt >> 2s) & 3 shears (0, 2 or 3)
// with coefficients C[3s..3s+2], the same fused multiply-adds as one I_SHEAR per slot.
template <bool PAIR, typename R>
__device__ __forceinline__ void op_shears(cx<R> (&a)[16], cx<R> (&b)[16], const R *__restrict__ C, uint32_t cnt) {
    R c[12];
    if constexpr (sizeof(R) == 8) {  // 96 bytes: two batches of scalar loads
        typedef uint32_t v16 __attribute__((ext_vector_type(16)));
        typedef uint32_t v8 __attribute__((ext_vector_type(8)));
        const v16 w0 = *reinterpret_cast<const v16 *>(C);
        const v8 w1 = *reinterpret_cast<const v8 *>(C + 8);
        static_for<0, 8>([&](auto ic) __attribute__((always_inline)) {
            constexpr int i = decltype(ic)::value;
            c[i] = __builtin_bit_cast(R, (uint64_t)w0[2 * i] | ((uint64_t)w0[2 * i + 1] << 32));
        });
        static_for<0, 4>([&](auto ic) __attribute__((always_inline)) {
            constexpr int i = decltype(ic)::value;
            c[8 + i] = __builtin_bit_cast(R, (uint64_t)w1[2 * i] | ((uint64_t)w1[2 * i + 1] << 32));
        });
    } else {
        typedef uint32_t v8 __attribute__((ext_vector_type(8)));
        typedef uint32_t v4 __attribute__((ext_vector_type(4)));
        const v8 w0 = *reinterpret_cast<const v8 *>(C);
        const v4 w1 = *reinterpret_cast<const v4 *>(C + 8);
        static_for<0, 8>([&](auto ic) __attribute__((always_inline)) {
            const uint32_t x = w0[decltype(ic)::value];
            c[decltype(ic)::value] = __builtin_bit_cast(R, x);
        });
        static_for<0, 4>([&](auto ic) __attribute__((always_inline)) {
            const uint32_t x = w1[decltype(ic)::value];
            c[8 + decltype(ic)::value] = __builtin_bit_cast(R, x);
        });
    }
    static_for<0, 4>([&](auto sc) __attribute__((always_inline)) {
        constexpr int S = decltype(sc)::value;
        const uint32_t n = (cnt >> (2 * S)) & 3u;
        if (n) {
            shear_steps<S, 0, 2, PAIR>(a, b, c + 3 * S);
            if (n == 3) shear_steps<S, 2, 3, PAIR>(a, b, c + 3 * S);
        }
    });
}

// Dense op on three or four register slots: computed for every lane, kept where the
// thread-bit controls hold (a per-amplitude select; these ops are rare).
template <int MASK, bool PAIR, typename R>
__device__ __forceinline__ void op_dense_big(cx<R> (&a)[16], cx<R> (&b)[16], const R *__restrict__ M, uint32_t pat,
                                             bool lok) {
    constexpr int K = popc4(MASK);
    constexpr int D = 1 << K;
    constexpr int OUTER = 16 >> K;
    constexpr int OM = 15 & ~MASK;
    static_for<0, OUTER>([&](auto oc) __attribute__((always_inline)) {
        constexpr int base = dep4(decltype(oc)::value, OM);
        if ((pat >> base) & 1u) {
            auto apply = [&](cx<R>(&v)[16]) __attribute__((always_inline)) {
                cx<R> x[D], y[D];
                static_for<0, D>([&](auto jc) __attribute__((always_inline)) {
                    x[decltype(jc)::value] = v[base | dep4(decltype(jc)::value, MASK)];
                });
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
                    y[i] = cx<R>{lok ? re : x[i].x, lok ? im : x[i].y};
                });
                static_for<0, D>([&](auto ic) __attribute__((always_inline)) {
                    commit(v[base | dep4(decltype(ic)::value, MASK)], y[decltype(ic)::value]);
                });
            };
            apply(a);
            if constexpr (PAIR) apply(b);
        }
    });
}

// [a0, a1] <- [p0 a1, p1 a0] on slot S (PURE: p0 = p1 = 1).  TC: thread-bit controls.
template <int S, bool PURE, bool TC, bool PAIR, typename R>
__device__ __forceinline__ void op_anti1(cx<R> (&a)[16], cx<R> (&b)[16], const R *__restrict__ p, uint32_t pat,
                                         bool lok) {
    R p0r = 1, p0i = 0, p1r = 1, p1i = 0;
    if constexpr (!PURE) {
        p0r = p[0];
        p0i = p[1];
        p1r = p[2];
        p1i = p[3];
    }
    static_for<0, 16>([&](auto rc) __attribute__((always_inline)) {
        constexpr int r = decltype(rc)::value;
        if constexpr (!(r >> S & 1)) {
            constexpr int r1 = r | (1 << S);
            if ((pat >> r) & 1u) {
                auto apply = [&](cx<R>(&v)[16]) __attribute__((always_inline)) {
                    if constexpr (PURE && !TC) {
                        swap_inplace(v[r], v[r1]);
                    } else {
                        const cx<R> a0 = v[r], a1 = v[r1];
                        cx<R> n0 = a1, n1 = a0;
                        if constexpr (!PURE) {
                            n0 = cmulp(a1, p0r, p0i);
                            n1 = cmulp(a0, p1r, p1i);
                        }
                        if constexpr (TC) {
                            n0.x = lok ? n0.x : a0.x;
                            n0.y = lok ? n0.y : a0.y;
                            n1.x = lok ? n1.x : a1.x;
                            n1.y = lok ? n1.y : a1.y;
                        }
                        commit(v[r], n0);
                        commit(v[r1], n1);
                    }
                };
                apply(a);
                if constexpr (PAIR) apply(b);
            }
        }
    });
}

// |01> <- p |10>, |10> <- p |01> on slots S0 < S1.
template <int S0, int S1, bool PURE, bool TC, bool PAIR, typename R>
__device__ __forceinline__ void op_swap2(cx<R> (&a)[16], cx<R> (&b)[16], const R *__restrict__ p, uint32_t pat,
                                         bool lok) {
    R pr = 1, pi = 0;
    if constexpr (!PURE) {
        pr = p[0];
        pi = p[1];
    }
    static_for<0, 16>([&](auto rc) __attribute__((always_inline)) {
        constexpr int r = decltype(rc)::value;
        if constexpr (!(r >> S0 & 1) && !(r >> S1 & 1)) {
            constexpr int i01 = r | (1 << S0), i10 = r | (1 << S1);
            if ((pat >> r) & 1u) {
                auto apply = [&](cx<R>(&v)[16]) __attribute__((always_inline)) {
                    if constexpr (PURE && !TC) {
                        swap_inplace(v[i01], v[i10]);
                    } else {
                        const cx<R> x01 = v[i01], x10 = v[i10];
                        cx<R> n01 = x10, n10 = x01;
                        if constexpr (!PURE) {
                            n01 = cmulp(x10, pr, pi);
                            n10 = cmulp(x01, pr, pi);
                        }
                        if constexpr (TC) {
                            n01.x = lok ? n01.x : x01.x;
                            n01.y = lok ? n01.y : x01.y;
                            n10.x = lok ? n10.x : x10.x;
                            n10.y = lok ? n10.y : x10.y;
                        }
                        commit(v[i01], n01);
                        commit(v[i10], n10);
                    }
                };
                apply(a);
                if constexpr (PAIR) apply(b);
            }
        }
    });
}

// Register slot S trades places with lane bit U (I_XSWAP): for each register pair (x: bit S = 0,
// y: bit S = 1), lanes with bit U set take y from lane - 2^U into x, lanes with it clear take x
// from lane + 2^U into y.  Lane bits 5 and 4 are exactly v_permlane32_swap / v_permlane16_swap
// (one instruction per 32-bit word and pair); lane bits 0..3 stay inside a 16-lane row, where
// row_shr / row_shl DPP moves fetch the partner and a select keeps the other half.
__device__ __forceinline__ void xswap_word(uint32_t &x, uint32_t &y, int U, bool hi) {
    if (U == 5) {
        const auto p = __builtin_amdgcn_permlane32_swap(x, y, false, false);
        x = p[0];
        y = p[1];
    } else if (U == 4) {
        const auto p = __builtin_amdgcn_permlane16_swap(x, y, false, false);
        x = p[0];
        y = p[1];
    } else {
        // bound_ctrl: lanes without a source in the row read 0; the select discards them
        uint32_t nx, ny;
        switch (U) {
        case 0:
            nx = __builtin_amdgcn_mov_dpp(y, 0x111, 0xF, 0xF, true);  // row_shr:1
            ny = __builtin_amdgcn_mov_dpp(x, 0x101, 0xF, 0xF, true);  // row_shl:1
            break;
        case 1:
            nx = __builtin_amdgcn_mov_dpp(y, 0x112, 0xF, 0xF, true);
            ny = __builtin_amdgcn_mov_dpp(x, 0x102, 0xF, 0xF, true);
            break;
        case 2:
            nx = __builtin_amdgcn_mov_dpp(y, 0x114, 0xF, 0xF, true);
            ny = __builtin_amdgcn_mov_dpp(x, 0x104, 0xF, 0xF, true);
            break;
        default:
            nx = __builtin_amdgcn_mov_dpp(y, 0x118, 0xF, 0xF, true);
            ny = __builtin_amdgcn_mov_dpp(x, 0x108, 0xF, 0xF, true);
            break;
        }
        x = hi ? nx : x;
        y = hi ? y : ny;
    }
}
template <int S, int U, typename R>
__device__ __forceinline__ void op_xswap(cx<R> (&v)[16], uint32_t lane) {
    const bool hi = (lane >> U) & 1u;
    static_for<0, 16>([&](auto rc) __attribute__((always_inline)) {
        constexpr int r = decltype(rc)::value;
        if constexpr (!(r >> S & 1)) {
            constexpr int r1 = r | (1 << S);
            constexpr int W = (int)sizeof(cx<R>) / 4;
            typedef uint32_t words __attribute__((ext_vector_type(W)));
            words x = __builtin_bit_cast(words, v[r]), y = __builtin_bit_cast(words, v[r1]);
            static_for<0, W>([&](auto wc) __attribute__((always_inline)) {
                constexpr int w = decltype(wc)::value;
                uint32_t xw = x[w], yw = y[w];
                xswap_word(xw, yw, U, hi);
                x[w] = xw;
                y[w] = yw;
            });
            v[r] = __builtin_bit_cast(cx<R>, x);
            v[r1] = __builtin_bit_cast(cx<R>, y);
        }
    });
}
template <typename R>
__device__ __forceinline__ void xswap_dispatch(cx<R> (&v)[16], int S, int U, uint32_t lane) {
#define XS_CASE(SS, UU) \
    case SS * 8 + UU: op_xswap<SS, UU>(v, lane); break;
#define XS_ROW(SS) XS_CASE(SS, 0) XS_CASE(SS, 1) XS_CASE(SS, 2) XS_CASE(SS, 3) XS_CASE(SS, 4) XS_CASE(SS, 5)
    switch (S * 8 + U) {
        XS_ROW(0)
        XS_ROW(1)
        XS_ROW(2)
        XS_ROW(3)
    default:
        break;
    }
#undef XS_ROW
#undef XS_CASE
}

template <int S, typename R>
__device__ __forceinline__ void op_reset(cx<R> (&a)[16]) {
    static_for<0, 16>([&](auto rc) __attribute__((always_inline)) {
        constexpr int r = decltype(rc)::value;
        if constexpr (!(r >> S & 1)) {
            constexpr int r1 = r | (1 << S);
            cx<R> a0 = a[r], a1 = a[r1];
            R nrm = sqrt(a0.x * a0.x + a0.y * a0.y + a1.x * a1.x + a1.y * a1.y);
            commit(a[r], nrm > R(1e-15) ? cx<R>{nrm, R(0)} : cx<R>{R(0), R(0)});
            commit(a[r1], cx<R>{R(0), R(0)});
        }
    });
}

// N consecutive complex table entries at a uniform address, fetched as one scalar load
template <typename R, int N>
__device__ __forceinline__ void load_cx_batch(const cx<R> *__restrict__ p, cx<R> (&out)[N]) {
    constexpr int W = N * (int)sizeof(cx<R>) / 4;
    typedef uint32_t vec __attribute__((ext_vector_type(W)));
    const vec w = *reinterpret_cast<const vec *>(p);
    static_for<0, N>([&](auto jc) __attribute__((always_inline)) {
        constexpr int j = decltype(jc)::value;
        if constexpr (sizeof(R) == 8) {
            out[j].x = __builtin_bit_cast(double, (uint64_t)w[4 * j] | ((uint64_t)w[4 * j + 1] << 32));
            out[j].y = __builtin_bit_cast(double, (uint64_t)w[4 * j + 2] | ((uint64_t)w[4 * j + 3] << 32));
        } else {
            // copy the elements to scalars first: a bit_cast applied to the vector element
            // expressions was compiled to read element 2j for both parts
            const uint32_t xr = w[2 * j], xi = w[2 * j + 1];
            out[j].x = __builtin_bit_cast(float, xr);
            out[j].y = __builtin_bit_cast(float, xi);
        }
    });
}

// Diagonal op: register index r is multiplied by table[(lane_index << 4) | r] where bit r of
// `pat` is set (K_DIAG_*): uniform tables go four entries per scalar load; K_DIAG_LANE selects
// entries by lane bits (thread-id bits or bits of the tile base) and keeps lanes whose thread-bit
// controls fail unchanged.
template <bool PAIR, typename R>
__device__ __forceinline__ void op_diag(cx<R> (&a)[16], cx<R> (&b)[16], const IHead &I, const R *__restrict__ pool,
                                        uint32_t tid, uint64_t base, bool lok, int code) {
    const cx<R> *__restrict__ tab = reinterpret_cast<const cx<R> *>(pool) + I.mat;
    const uint32_t nontriv = I.pat;
    if (code == K_DIAG_TILE) {  // the table of this tile: index bits of the (uniform) tile base
        const uint32_t lsw = I.lsw;
        uint32_t li = 0;
        for (int j = 0; j < I.k; j++) li = (li << 1) | (uint32_t)((base >> (((lsw >> (8 * j)) & 255u) - 16)) & 1u);
        tab += __builtin_amdgcn_readfirstlane(li) << 4;
    }
    auto mul = [&](auto rc, auto uc, cx<R> ph) __attribute__((always_inline)) {
        constexpr int r = decltype(rc)::value;
        constexpr bool U = decltype(uc)::value;
        cmul_inplace<U>(a[r].x, a[r].y, ph.x, ph.y);
        if constexpr (PAIR) cmul_inplace<U>(b[r].x, b[r].y, ph.x, ph.y);
    };
    if (code >= K_DIAG_SLOT) {  // one phase on the eight registers of one slot
        cx<R> ph[1];
        load_cx_batch<R, 1>(tab + (1 << (code & 3)), ph);
        const bool tc = code >= K_DIAG_TSLOT;  // lanes whose thread-bit controls fail keep 1
        cx<R> pl;
        pl.x = (!tc || lok) ? ph[0].x : R(1);
        pl.y = (!tc || lok) ? ph[0].y : R(0);
        auto slot = [&](auto sc, auto tcc) __attribute__((always_inline)) {
            constexpr int S = decltype(sc)::value;
            static_for<0, 16>([&](auto rc) __attribute__((always_inline)) {
                constexpr int r = decltype(rc)::value;
                if constexpr (r >> S & 1) {
                    if constexpr (decltype(tcc)::value) mul(rc, std::false_type{}, pl);
                    else mul(rc, std::true_type{}, ph[0]);
                }
            });
        };
        using U = std::false_type;  // uniform phase
        using L = std::true_type;   // per-lane phase (thread-bit controls)
        switch (code - K_DIAG_SLOT) {
        case 0: slot(std::integral_constant<int, 0>{}, U{}); break;
        case 1: slot(std::integral_constant<int, 1>{}, U{}); break;
        case 2: slot(std::integral_constant<int, 2>{}, U{}); break;
        case 3: slot(std::integral_constant<int, 3>{}, U{}); break;
        case 4: slot(std::integral_constant<int, 0>{}, L{}); break;
        case 5: slot(std::integral_constant<int, 1>{}, L{}); break;
        case 6: slot(std::integral_constant<int, 2>{}, L{}); break;
        default: slot(std::integral_constant<int, 3>{}, L{}); break;
        }
        return;
    }
    if (code == K_DIAG_FULL) {  // registers 1..15 without tests, four entries per scalar load
        static_for<0, 4>([&](auto cc) __attribute__((always_inline)) {
            constexpr int c0 = 4 * decltype(cc)::value;
            cx<R> ph[4];
            load_cx_batch<R, 4>(tab + c0, ph);
            static_for<0, 4>([&](auto jc) __attribute__((always_inline)) {
                constexpr int r = c0 + decltype(jc)::value;
                if (r > 0 || (nontriv & 1u)) mul(std::integral_constant<int, r>{}, std::true_type{}, ph[decltype(jc)::value]);
            });
        });
        return;
    }
    if (code != K_DIAG_LANE) {  // a uniform table on the registers of `pat`: four entries per scalar load
        static_for<0, 4>([&](auto cc) __attribute__((always_inline)) {
            constexpr int c0 = 4 * decltype(cc)::value;
            if ((nontriv >> c0) & 15u) {
                cx<R> ph[4];
                load_cx_batch<R, 4>(tab + c0, ph);
                static_for<0, 4>([&](auto jc) __attribute__((always_inline)) {
                    constexpr int r = c0 + decltype(jc)::value;
                    if ((nontriv >> r) & 1u)
                        mul(std::integral_constant<int, r>{}, std::true_type{}, ph[decltype(jc)::value]);
                });
            }
        });
        return;
    }
    const int kl = I.k;
    const bool tc = I.flags & F_TCTRL;
    const uint32_t lsw = I.lsw;
    uint32_t li = 0;
    for (int j = 0; j < kl; j++) {
        const uint32_t src = (lsw >> (8 * j)) & 255u;
        const uint32_t bit = src < 16 ? ((tid >> src) & 1u) : (uint32_t)((base >> (src - 16)) & 1u);
        li = (li << 1) | bit;
    }
    static_for<0, 16>([&](auto rc) __attribute__((always_inline)) {
        constexpr int r = decltype(rc)::value;
        if ((nontriv >> r) & 1u) {
            cx<R> t;
            if (kl == 0) {
                t = tab[r];
            } else if (kl == 1) {  // two uniform candidates
                const cx<R> t0 = tab[r], t1 = tab[16 + r];
                t.x = li ? t1.x : t0.x;
                t.y = li ? t1.y : t0.y;
            } else {
                t = tab[(li << 4) | r];  // per-lane load from a small, cached table
            }
            if (tc) {
                t.x = lok ? t.x : R(1);
                t.y = lok ? t.y : R(0);
            }
            mul(rc, std::false_type{}, t);
        }
    });
}

// Parity-block channel superoperators (I_PARITY, canonical form in engine.h).  y = B x within each
// parity class; REAL reads only the real parts of B (2 FMAs per entry instead of 4).
template <int N, bool REAL, typename R>
__device__ __forceinline__ void block_apply(const R *__restrict__ B, const cx<R> (&x)[N], cx<R> (&y)[N]) {
    static_for<0, N>([&](auto ic) __attribute__((always_inline)) {
        constexpr int i = decltype(ic)::value;
        R re = 0, im = 0;
        static_for<0, N>([&](auto jc) __attribute__((always_inline)) {
            constexpr int j = decltype(jc)::value;
            const R br = B[2 * (i * N + j)];
            re = fma(br, x[j].x, re);
            im = fma(br, x[j].y, im);
            if constexpr (!REAL) {
                const R bi = B[2 * (i * N + j) + 1];
                re = fma(-bi, x[j].y, re);
                im = fma(bi, x[j].x, im);
            }
        });
        y[i] = cx<R>{re, im};
    });
}

// one-qubit channel: slots LO < HI, classes p = bit LO ^ bit HI, 2x2 blocks
template <int LO, int HI, bool REAL, typename R>
__device__ __forceinline__ void op_parity2(cx<R> (&a)[16], const R *__restrict__ M) {
    constexpr int OM = 15 & ~((1 << LO) | (1 << HI));
    static_for<0, 4>([&](auto gc) __attribute__((always_inline)) {
        constexpr int base = dep4(decltype(gc)::value, OM);
        static_for<0, 2>([&](auto pc) __attribute__((always_inline)) {
            constexpr int p = decltype(pc)::value;
            constexpr int e0 = base | (p << HI), e1 = base | (1 << LO) | ((1 ^ p) << HI);
            cx<R> x[2] = {a[e0], a[e1]}, y[2];
            block_apply<2, REAL>(M + 8 * p, x, y);
            commit(a[e0], y[0]);
            commit(a[e1], y[1]);
        });
    });
}

// two-qubit channel: slot pairs (0, S1) and (S2, S3), classes c = 2 pa + pb, 4x4 blocks
template <int S1, int S2, int S3, bool REAL, typename R>
__device__ __forceinline__ void op_parity4(cx<R> (&a)[16], const R *__restrict__ M) {
    static_for<0, 4>([&](auto cc) __attribute__((always_inline)) {
        constexpr int c = decltype(cc)::value;
        constexpr int pa = c >> 1, pb = c & 1;
        auto bits = [](int i, int pa_, int pb_) constexpr {
            const int ia = i >> 1, ib = i & 1;
            return ia | ((ia ^ pa_) << S1) | (ib << S2) | ((ib ^ pb_) << S3);
        };
        constexpr int e0 = bits(0, pa, pb), e1 = bits(1, pa, pb), e2 = bits(2, pa, pb), e3 = bits(3, pa, pb);
        cx<R> x[4] = {a[e0], a[e1], a[e2], a[e3]}, y[4];
        block_apply<4, REAL>(M + 32 * c, x, y);
        commit(a[e0], y[0]);
        commit(a[e1], y[1]);
        commit(a[e2], y[2]);
        commit(a[e3], y[3]);
    });
}

// Re sum over the op's groups of conj(b) (M a): the derivative term of the adjoint method.
// Lanes whose thread-bit controls fail and groups outside `pat` contribute nothing.
template <int MASK, typename R>
__device__ __forceinline__ double op_grad(const cx<R> (&a)[16], const cx<R> (&b)[16], const R *__restrict__ M,
                                          uint32_t pat, bool lok) {
    constexpr int K = popc4(MASK);
    constexpr int D = 1 << K;
    constexpr int OUTER = 16 >> K;
    constexpr int OM = 15 & ~MASK;
    double acc = 0;
    static_for<0, OUTER>([&](auto oc) __attribute__((always_inline)) {
        constexpr int base = dep4(decltype(oc)::value, OM);
        if ((pat >> base) & 1u) {
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
    return lok ? acc : 0.0;
}

// Workgroup barrier for LDS traffic only: waits for this wave's LDS accesses, then for every
// wave.  Unlike __syncthreads it leaves outstanding global memory operations in flight.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// LDS byte address lb ^ k of a remap element.  asm volatile keeps each pass's addresses next to
// the pass (recomputed, not held in 32 VGPRs across both passes of a complex128 remap); the LDS
// pointer is formed from the integer directly (the stage kernels' dynamic LDS starts at 0).
__device__ __forceinline__ uint32_t lds_xaddr(uint32_t lb, uint32_t k) {
    uint32_t a;
    asm volatile("v_xor_b32 %0, %1, %2" : "=v"(a) : "s"(k), "v"(lb));
    return a;
}
template <typename T>
__device__ __forceinline__ __attribute__((address_space(3))) T *lds_ptr(uint32_t a) {
    return reinterpret_cast<__attribute__((address_space(3))) T *>(static_cast<size_t>(a));
}

// One remap through LDS: the 16 amplitudes leave by the old offsets (lb ^ K[r]) and arrive by
// the new ones.  complex64: one pass of 8-byte amplitudes.  complex128: the real parts, then the
// imaginary parts (each pass writes only registers the other pass does not read).
template <typename R>
__device__ __forceinline__ void lds_remap(unsigned char *smem, cx<R> (&v)[16], uint32_t lb, const KOff &K,
                                          uint32_t lbn, const KOff &Kn) {
    (void)smem;
    if constexpr (sizeof(R) == 4) {
        lds_barrier();
        typedef R v2 __attribute__((ext_vector_type(2)));
        static_for<0, 16>([&](auto rc) __attribute__((always_inline)) {
            constexpr int r = decltype(rc)::value;
            v2 x;
            x.x = v[r].x;
            x.y = v[r].y;
            *lds_ptr<v2>(lds_xaddr(lb, koff<r>(K))) = x;
        });
        lds_barrier();
        static_for<0, 16>([&](auto rc) __attribute__((always_inline)) {
            constexpr int r = decltype(rc)::value;
            const v2 x = *lds_ptr<v2>(lds_xaddr(lbn, koff<r>(Kn)));
            v[r].x = x.x;
            v[r].y = x.y;
        });
    } else {
        auto pass = [&](auto comp) __attribute__((always_inline)) {
            constexpr int c = decltype(comp)::value;
            lds_barrier();
            static_for<0, 16>([&](auto rc) __attribute__((always_inline)) {
                constexpr int r = decltype(rc)::value;
                *lds_ptr<R>(lds_xaddr(lb, koff<r>(K))) = c ? v[r].y : v[r].x;
            });
            lds_barrier();
            static_for<0, 16>([&](auto rc) __attribute__((always_inline)) {
                constexpr int r = decltype(rc)::value;
                (c ? v[r].y : v[r].x) = *lds_ptr<R>(lds_xaddr(lbn, koff<r>(Kn)));
            });
        };
        pass(std::integral_constant<int, 0>{});
        pass(std::integral_constant<int, 1>{});
    }
}

// One stage over one state (PAIR = false) or over psi and lambda together (PAIR = true: every op
// acts on both, and I_GRAD accumulates Re <lambda| M |psi> per gradient slot).
template <typename R, bool PAIR, bool BIG, bool DM>
__device__ __forceinline__ void stage_body(cx<R> *__restrict__ st, cx<R> *__restrict__ st2,
                                           const StageHeader *__restrict__ hdr, const Instr *__restrict__ prog,
                                           const R *__restrict__ pool, uint64_t ntiles, double *__restrict__ gpart,
                                           double *gacc) {
    extern __shared__ unsigned char smem_raw[];
    const int m = hdr->m;
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

    // base index of a tile: the tile number deposited into the qubits outside the tile (T ascending)
    auto tile_base = [&](uint64_t tile) __attribute__((always_inline)) {
        uint64_t base = tile;
#pragma unroll
        for (int j = 0; j < 16; j++) {
            if (j < m) {
                const uint32_t t = byte_of(T0, T1, T2, T3, j);
                const uint64_t lo = base & ((1ull << t) - 1);
                base = ((base >> t) << (t + 1)) | lo;
            }
        }
        return base;
    };
    // the tile's loads issue ahead of other waves' arithmetic (wave priority)
    auto load_tile = [&](cx<R>(&v)[16], cx<R>(&w)[16], uint64_t base) __attribute__((always_inline)) {
        __builtin_amdgcn_s_setprio(2);
        GAddr G;
        global_layout(G, hdr->gq0, m, tid, base);
        static_for<0, 16>([&](auto rc) __attribute__((always_inline)) {
            constexpr int r = decltype(rc)::value;
            v[r] = ld_stream(st + (G.tg | rgof<r>(G)));
            if constexpr (PAIR) w[r] = ld_stream(st2 + (G.tg | rgof<r>(G)));
        });
        __builtin_amdgcn_s_setprio(0);
    };

    // One tile per workgroup (the grid covers every tile), with no loop, so that the compiler does
    // not hoist tile-invariant values into SGPRs it then spills (one tile is all a workgroup ever
    // runs).  Fetching the next tile while the ops run on the current one would need a second
    // register set: two waves per SIMD instead of four, measured 45 % slower.
    auto run_tile = [&](const uint64_t tile) __attribute__((always_inline)) {
        const uint64_t base = tile_base(tile);
        cx<R> a[16], b[16];  // b: lambda in paired mode, unused otherwise
        load_tile(a, b, base);
        uint32_t lb = lds_lane<R>(hdr->lay0, m, tid);
        KOff K;
        load_koff(K, hdr->k0);

        // The head of the current instruction and its code path.  Every stage's list ends with
        // an I_END (code K_OUTER), so the inner loop below has its only exit at its head and no
        // count test.  The instruction's byte offset goes through readfirstlane so that it stays
        // provably uniform (instruction fields then come through scalar loads).
        uint32_t ioff = 0;
        auto at = [&]() __attribute__((always_inline)) {
            return reinterpret_cast<const Instr *>(reinterpret_cast<const unsigned char *>(P) + ioff);
        };
        IHead cur = load_head(at(), 0);
        int code = (cur.flags >> kCodeShift) & 255;
        auto next = [&]() __attribute__((always_inline)) {
            ioff = __builtin_amdgcn_readfirstlane(ioff + (uint32_t)sizeof(Instr));
            cur = load_head(at(), 0);
            code = (cur.flags >> kCodeShift) & 255;
        };
        while (true) {
            // The ops up to the next remap.  The remap lives outside this loop: with it inside,
            // the compiler splits the amplitudes into two register sets and copies all of them
            // on every op.
            for (; code != K_OUTER; next()) {
                // controls on qubits off the tile (omask = 0 unless F_TILE): uniform per workgroup
                if ((base & cur.omask) != cur.omask) continue;
                const bool lok = (tid & cur.tmask) == cur.tmask;
                const uint32_t pat = cur.pat;
                const R *__restrict__ M = pool + 2 * cur.mat;
                // one switch per op range; the hot ranges (shears, phase tables) come first
                if (code < K_DIAG_FULL) {
                    switch (code) {
                    case 0: op_shear<0, 2, PAIR>(a, b, cur); break;
                    case 1: op_shear<0, 3, PAIR>(a, b, cur); break;
                    case 2: op_shear<1, 2, PAIR>(a, b, cur); break;
                    case 3: op_shear<1, 3, PAIR>(a, b, cur); break;
                    case 4: op_shear<2, 2, PAIR>(a, b, cur); break;
                    case 5: op_shear<2, 3, PAIR>(a, b, cur); break;
                    case 6: op_shear<3, 2, PAIR>(a, b, cur); break;
                    case 7: op_shear<3, 3, PAIR>(a, b, cur); break;
                    default: op_shears<PAIR>(a, b, M, (uint32_t)cur.k); break;  // K_SHEARS
                    }
                } else if (code < K_ANTI1) {
                    op_diag<PAIR>(a, b, cur, pool, tid, base, lok, code);
                } else if (code < K_DENSE1) {
#define ANTI_CASE(S)                                                                          \
    case K_ANTI1 + S * 4 + 0: op_anti1<S, false, false, PAIR>(a, b, M, pat, lok); break;        \
    case K_ANTI1 + S * 4 + 1: op_anti1<S, false, true, PAIR>(a, b, M, pat, lok); break;         \
    case K_ANTI1 + S * 4 + 2: op_anti1<S, true, false, PAIR>(a, b, M, pat, lok); break;         \
    case K_ANTI1 + S * 4 + 3: op_anti1<S, true, true, PAIR>(a, b, M, pat, lok); break;
#define SWAP_CASE(P, A, B)                                                                    \
    case K_SWAP2 + P * 4 + 0: op_swap2<A, B, false, false, PAIR>(a, b, M, pat, lok); break;     \
    case K_SWAP2 + P * 4 + 1: op_swap2<A, B, false, true, PAIR>(a, b, M, pat, lok); break;      \
    case K_SWAP2 + P * 4 + 2: op_swap2<A, B, true, false, PAIR>(a, b, M, pat, lok); break;      \
    case K_SWAP2 + P * 4 + 3: op_swap2<A, B, true, true, PAIR>(a, b, M, pat, lok); break;
                    switch (code) {
                        ANTI_CASE(0)
                        ANTI_CASE(1)
                        ANTI_CASE(2)
                        ANTI_CASE(3)
                        SWAP_CASE(0, 0, 1)
                        SWAP_CASE(1, 0, 2)
                        SWAP_CASE(2, 0, 3)
                        SWAP_CASE(3, 1, 2)
                        SWAP_CASE(4, 1, 3)
                        SWAP_CASE(5, 2, 3)
                    default:
                        break;
                    }
#undef ANTI_CASE
#undef SWAP_CASE
                } else if (code < K_PARITY2) {
#define DENSE1_CASE(S)                                                                                   \
    case K_DENSE1 + 3 * S + 0: op_dense<1 << S, true, PAIR>(a, b, M, pat, lok); break;                     \
    case K_DENSE1 + 3 * S + 1: op_dense<1 << S, false, PAIR, R, true>(a, b, M, pat, lok); break;           \
    case K_DENSE1 + 3 * S + 2: op_dense<1 << S, false, PAIR>(a, b, M, pat, lok); break;
#define DENSE2_CASE(P, MK)                                                                               \
    case K_DENSE2 + 2 * P + 0: op_dense<MK, true, PAIR>(a, b, M, pat, lok); break;                         \
    case K_DENSE2 + 2 * P + 1: op_dense<MK, false, PAIR>(a, b, M, pat, lok); break;
                    switch (code) {
                        DENSE1_CASE(0)
                        DENSE1_CASE(1)
                        DENSE1_CASE(2)
                        DENSE1_CASE(3)
                        DENSE2_CASE(0, 3)
                        DENSE2_CASE(1, 5)
                        DENSE2_CASE(2, 9)
                        DENSE2_CASE(3, 6)
                        DENSE2_CASE(4, 10)
                        DENSE2_CASE(5, 12)
                    default:
                        break;
                    }
#undef DENSE1_CASE
#undef DENSE2_CASE
                } else if (code < K_GRAD) {
                    if constexpr (DM) {  // density-matrix channel superoperators
                        // paired programs: F_SIDE_A / F_SIDE_B restrict a channel to one state
                        auto par = [&](auto fn) __attribute__((always_inline)) {
                            if (!(cur.flags & F_SIDE_B)) fn(a);
                            if constexpr (PAIR)
                                if (!(cur.flags & F_SIDE_A)) fn(b);
                        };
#define PAR2_CASE(P, LO, HI)                                                                             \
    case K_PARITY2 + 2 * P + 0: par([&](cx<R>(&v)[16]) __attribute__((always_inline)) { op_parity2<LO, HI, true>(v, M); }); break; \
    case K_PARITY2 + 2 * P + 1: par([&](cx<R>(&v)[16]) __attribute__((always_inline)) { op_parity2<LO, HI, false>(v, M); }); break;
#define PAR4_CASE(C, S1, S2, S3)                                                                         \
    case K_PARITY4 + 2 * C + 0: par([&](cx<R>(&v)[16]) __attribute__((always_inline)) { op_parity4<S1, S2, S3, true>(v, M); }); break; \
    case K_PARITY4 + 2 * C + 1: par([&](cx<R>(&v)[16]) __attribute__((always_inline)) { op_parity4<S1, S2, S3, false>(v, M); }); break;
                        switch (code) {
                            PAR2_CASE(0, 0, 1)
                            PAR2_CASE(1, 0, 2)
                            PAR2_CASE(2, 0, 3)
                            PAR2_CASE(3, 1, 2)
                            PAR2_CASE(4, 1, 3)
                            PAR2_CASE(5, 2, 3)
                            PAR4_CASE(0, 1, 2, 3)
                            PAR4_CASE(1, 2, 1, 3)
                            PAR4_CASE(2, 3, 1, 2)
                        default: break;
                        }
#undef PAR4_CASE
#undef PAR2_CASE
                    }
                } else {
                    if constexpr (PAIR) {
                        double v = 0;
                        switch (code - K_GRAD) {
                        case 0: v = op_grad<1>(a, b, M, pat, lok); break;
                        case 1: v = op_grad<2>(a, b, M, pat, lok); break;
                        case 2: v = op_grad<3>(a, b, M, pat, lok); break;
                        case 3: v = op_grad<4>(a, b, M, pat, lok); break;
                        case 4: v = op_grad<5>(a, b, M, pat, lok); break;
                        case 5: v = op_grad<6>(a, b, M, pat, lok); break;
                        case 6: v = op_grad<8>(a, b, M, pat, lok); break;
                        case 7: v = op_grad<9>(a, b, M, pat, lok); break;
                        case 8: v = op_grad<10>(a, b, M, pat, lok); break;
                        default: v = op_grad<12>(a, b, M, pat, lok); break;
                        }
                        // blocks of fewer than 64 threads only read active lanes
    #pragma unroll
                        for (int off = 32; off >= 1; off >>= 1)
                            if (off < (int)blockDim.x) v += __shfl_xor(v, off, 64);
                        if (lane == 0) gacc[cur.w10 * 8 + wave] += v;  // one writer per (slot, wave)
                    }
                }
            }
            if (cur.op == I_END) break;
            // an outer instruction: its head is in `h`, the rest (layout, LDS offsets) at *I
            const Instr *__restrict__ I = at();
            const IHead h = cur;
            next();
            const int s0 = (int)(h.sw & 255u);
            if (h.op == I_RESET) {
                switch (s0) {
                case 0: op_reset<0>(a); break;
                case 1: op_reset<1>(a); break;
                case 2: op_reset<2>(a); break;
                default: op_reset<3>(a); break;
                }
                continue;
            }
            if (h.op == I_DENSE) {  // three or four slots (stages flagged SF_BIG only)
                if constexpr (!BIG) continue;
                if ((base & h.omask) != h.omask) continue;
                const bool lok = (tid & h.tmask) == h.tmask;
                const R *__restrict__ M = pool + 2 * h.mat;
                switch (h.smask) {
                case 7: op_dense_big<7, PAIR>(a, b, M, h.pat, lok); break;
                case 11: op_dense_big<11, PAIR>(a, b, M, h.pat, lok); break;
                case 13: op_dense_big<13, PAIR>(a, b, M, h.pat, lok); break;
                case 14: op_dense_big<14, PAIR>(a, b, M, h.pat, lok); break;
                case 15: op_dense_big<15, PAIR>(a, b, M, h.pat, lok); break;
                default: break;
                }
                continue;
            }
            if (h.op == I_XSWAP) {  // cross-lane exchange; the LDS offsets follow the new layout
                xswap_dispatch(a, s0, h.k, (uint32_t)lane);
                if constexpr (PAIR) xswap_dispatch(b, s0, h.k, (uint32_t)lane);
                lb = lds_lane<R>(I->lay, m, tid);
                load_koff(K, I->kr);
                continue;
            }
            // remap: the tile goes through LDS into the next layout
            const uint32_t lbn = lds_lane<R>(I->lay, m, tid);
            KOff Kn;
            load_koff(Kn, I->kr);
            lds_remap<R>(smem_raw, a, lb, K, lbn, Kn);
            if constexpr (PAIR) lds_remap<R>(smem_raw, b, lb, K, lbn, Kn);
            lb = lbn;
            K = Kn;
        }
        // stores through a permuted mapping may hit positions other waves still have to load
        if (hdr->flags & SF_SYNC_STORE) __syncthreads();
        {
            __builtin_amdgcn_s_setprio(2);
            GAddr G;
            global_layout(G, hdr->gqN, m, tid, base);
            static_for<0, 16>([&](auto rc) __attribute__((always_inline)) {
                constexpr int r = decltype(rc)::value;
                st_stream(st + (G.tg | rgof<r>(G)), a[r]);
                if constexpr (PAIR) st_stream(st2 + (G.tg | rgof<r>(G)), b[r]);
            });
            __builtin_amdgcn_s_setprio(0);
        }
    };
    if (blockIdx.x < ntiles) run_tile(blockIdx.x);
    if constexpr (PAIR) {
        __syncthreads();
        const int nw = (blockDim.x + 63) >> 6;  // a block of fewer than 64 threads is one wave
        for (int g = tid; g < ngrad; g += blockDim.x) {
            double s = 0;
            for (int w = 0; w < nw; w++) s += gacc[g * 8 + w];
            gpart[(size_t)g * gridDim.x + blockIdx.x] = s;
        }
    }
}

template <typename T>
__device__ __forceinline__ const T *blob_at(const T *p, uint64_t byte_off) {
    return reinterpret_cast<const T *>(reinterpret_cast<const char *>(p) + byte_off);
}

// Single-state instantiations: k_stage runs statevector stages without 3-4-slot dense ops (few
// enough VGPRs for four waves per SIMD, and code small enough for short branches; a fifth wave,
// forced by spilling, measured slower), k_stage_big the stages that hold them; k_stage_dm /
// k_stage_dm_big the density-matrix stages without / with them (channel superoperators as well).
template <typename R>
__global__ void __launch_bounds__(512) k_stage(cx<R> *__restrict__ st, const StageHeader *__restrict__ hdr,
                                               const Instr *__restrict__ prog, const R *__restrict__ pool,
                                               uint64_t ntiles, const uint8_t *__restrict__ act, uint64_t shot_stride,
                                               uint64_t blob_stride) {
    if (act && !act[blockIdx.y]) return;  // a shot that this program does not reach
    const uint64_t bo = blockIdx.y * blob_stride;  // state y's own program (0: one shared program)
    stage_body<R, false, false, false>(st + blockIdx.y * shot_stride, nullptr, blob_at(hdr, bo), blob_at(prog, bo),
                                       blob_at(pool, bo), ntiles, nullptr, nullptr);
}
template <typename R>
__global__ void __launch_bounds__(512) k_stage_big(cx<R> *__restrict__ st, const StageHeader *__restrict__ hdr,
                                                   const Instr *__restrict__ prog, const R *__restrict__ pool,
                                                   uint64_t ntiles, const uint8_t *__restrict__ act, uint64_t shot_stride,
                                                   uint64_t blob_stride) {
    if (act && !act[blockIdx.y]) return;  // a shot that this program does not reach
    const uint64_t bo = blockIdx.y * blob_stride;  // state y's own program (0: one shared program)
    stage_body<R, false, true, false>(st + blockIdx.y * shot_stride, nullptr, blob_at(hdr, bo), blob_at(prog, bo),
                                      blob_at(pool, bo), ntiles, nullptr, nullptr);
}
template <typename R>
__global__ void __launch_bounds__(512) k_stage_dm(cx<R> *__restrict__ st, const StageHeader *__restrict__ hdr,
                                                  const Instr *__restrict__ prog, const R *__restrict__ pool,
                                                  uint64_t ntiles, const uint8_t *__restrict__ act, uint64_t shot_stride,
                                                  uint64_t blob_stride) {
    if (act && !act[blockIdx.y]) return;  // a shot that this program does not reach
    const uint64_t bo = blockIdx.y * blob_stride;  // state y's own program (0: one shared program)
    stage_body<R, false, false, true>(st + blockIdx.y * shot_stride, nullptr, blob_at(hdr, bo), blob_at(prog, bo),
                                     blob_at(pool, bo), ntiles, nullptr, nullptr);
}
template <typename R>
__global__ void __launch_bounds__(512) k_stage_dm_big(cx<R> *__restrict__ st, const StageHeader *__restrict__ hdr,
                                                  const Instr *__restrict__ prog, const R *__restrict__ pool,
                                                  uint64_t ntiles, const uint8_t *__restrict__ act, uint64_t shot_stride,
                                                  uint64_t blob_stride) {
    if (act && !act[blockIdx.y]) return;  // a shot that this program does not reach
    const uint64_t bo = blockIdx.y * blob_stride;  // state y's own program (0: one shared program)
    stage_body<R, false, true, true>(st + blockIdx.y * shot_stride, nullptr, blob_at(hdr, bo), blob_at(prog, bo),
                                     blob_at(pool, bo), ntiles, nullptr, nullptr);
}

// Paired stages (adjoint reverse sweeps: psi and lambda, or rho and h, through the same tiles).
// The lean kernels run stages without 3-4-slot dense ops in at most 168 VGPRs (three waves per
// SIMD; the two states' amplitudes alone take 128); the _big kernels run the stages that hold
// them (SF_BIG).  The gradient partials follow the tile's remap area in dynamic LDS (the remaps
// address LDS from 0, so the kernels hold no static LDS).
#define GOQU_PAIR_KERNEL(NAME, ATTR, BIG, DM)                                                          \
    template <typename R>                                                                             \
    __global__ void __launch_bounds__(256) ATTR NAME(cx<R> *__restrict__ psi, cx<R> *__restrict__ lam,  \
                                                     const StageHeader *__restrict__ hdr,             \
                                                     const Instr *__restrict__ prog,                  \
                                                     const R *__restrict__ pool, uint64_t ntiles,      \
                                                     double *__restrict__ gpart) {                    \
        extern __shared__ unsigned char smem_raw[];                                                   \
        double *gacc = reinterpret_cast<double *>(smem_raw + (size_t(8) << hdr->m));                 \
        stage_body<R, true, BIG, DM>(psi, lam, hdr, prog, pool, ntiles, gpart, gacc);                 \
    }
#define GOQU_THREE_WAVES __attribute__((amdgpu_waves_per_eu(3, 8)))
GOQU_PAIR_KERNEL(k_stage_pair, GOQU_THREE_WAVES, false, false)
GOQU_PAIR_KERNEL(k_stage_pair_big, , true, false)
// density-matrix reverse sweeps: rho and h = vec(H), channel superoperators on one state or the
// other (F_SIDE_A / F_SIDE_B)
GOQU_PAIR_KERNEL(k_stage_pair_dm, GOQU_THREE_WAVES, false, true)
GOQU_PAIR_KERNEL(k_stage_pair_dm_big, , true, true)
#undef GOQU_THREE_WAVES
#undef GOQU_PAIR_KERNEL

// The twelve stage kernels take minutes each to compile.  build.py compiles this file once per
// kernel with GOQU_STAGE_KERNEL = 1..12 (that kernel's code only), in parallel, and once without
// it: the host code below, which refers to the kernels as extern instantiations.
#define GOQU_STAGE_INST(EXT, K, R)                                                                \
    EXT template __global__ void K<R>(cx<R> *__restrict__, const StageHeader *__restrict__,          \
                                      const Instr *__restrict__, const R *__restrict__, uint64_t,      \
                                      const uint8_t *__restrict__, uint64_t, uint64_t);
#define GOQU_PAIR_INST(EXT, K, R)                                                                  \
    EXT template __global__ void K<R>(cx<R> *__restrict__, cx<R> *__restrict__,                     \
                                      const StageHeader *__restrict__, const Instr *__restrict__,     \
                                      const R *__restrict__, uint64_t, double *__restrict__);
#ifdef GOQU_STAGE_KERNEL
#if GOQU_STAGE_KERNEL == 1
GOQU_STAGE_INST(, k_stage, double)
#elif GOQU_STAGE_KERNEL == 2
GOQU_STAGE_INST(, k_stage, float)
#elif GOQU_STAGE_KERNEL == 3
GOQU_STAGE_INST(, k_stage_big, double)
#elif GOQU_STAGE_KERNEL == 4
GOQU_STAGE_INST(, k_stage_big, float)
#elif GOQU_STAGE_KERNEL == 5
GOQU_PAIR_INST(, k_stage_pair, double)
#elif GOQU_STAGE_KERNEL == 6
GOQU_PAIR_INST(, k_stage_pair, float)
#elif GOQU_STAGE_KERNEL == 7
GOQU_STAGE_INST(, k_stage_dm, double)
#elif GOQU_STAGE_KERNEL == 8
GOQU_PAIR_INST(, k_stage_pair_dm, double)
#elif GOQU_STAGE_KERNEL == 9
GOQU_PAIR_INST(, k_stage_pair_big, double)
#elif GOQU_STAGE_KERNEL == 10
GOQU_PAIR_INST(, k_stage_pair_big, float)
#elif GOQU_STAGE_KERNEL == 11
GOQU_PAIR_INST(, k_stage_pair_dm_big, double)
#elif GOQU_STAGE_KERNEL == 12
GOQU_STAGE_INST(, k_stage_dm_big, double)
#endif
}  // namespace goqu
#else  // the host-code object
GOQU_STAGE_INST(extern, k_stage, double)
GOQU_STAGE_INST(extern, k_stage, float)
GOQU_STAGE_INST(extern, k_stage_big, double)
GOQU_STAGE_INST(extern, k_stage_big, float)
GOQU_PAIR_INST(extern, k_stage_pair, double)
GOQU_PAIR_INST(extern, k_stage_pair, float)
GOQU_STAGE_INST(extern, k_stage_dm, double)
GOQU_PAIR_INST(extern, k_stage_pair_dm, double)
GOQU_PAIR_INST(extern, k_stage_pair_big, double)
GOQU_PAIR_INST(extern, k_stage_pair_big, float)
GOQU_PAIR_INST(extern, k_stage_pair_dm_big, double)
GOQU_STAGE_INST(extern, k_stage_dm_big, double)

// ---------------------------------------------------------------------------------------------
// Tiles of 2^12 amplitudes (256 threads, 32 KiB of LDS) holding the 4 lowest qubits.  On the
// 30-qubit depth-40 circuit: 61 sweeps of 6.89 ms; 2^13 tiles cut the sweeps to 53 but each takes
// 7.46 ms, which the autotuner weighs per circuit.
// GOQU_TILE_BITS (8..13) and GOQU_LOW_BITS (2..6) override the default, goqu_sv_set_plan sets a
// handle's shape (the autotuner's choice).  Paired programs (adjoint reverse sweeps, two states
// per tile) stay at most 2^12: their kernel holds twice the amplitudes.
PlanConfig default_plan_config(int prec, int nq, bool paired) {
    PlanConfig c;
    c.tile_bits = 12;
    c.low_bits = 4;
    if (const char *e = std::getenv("GOQU_TILE_BITS")) {
        const int t = std::atoi(e);
        if (t >= 8 && t <= (paired ? 12 : 13)) c.tile_bits = t;
    }
    if (const char *e = std::getenv("GOQU_LOW_BITS")) {
        const int l = std::atoi(e);
        if (l >= 2 && l <= 6) c.low_bits = l;
    }
    (void)prec;
    (void)nq;
    return c;
}

ProgramLayout program_layout(size_t nh, size_t ni, size_t np, int prec) {
    const size_t rsz = prec == GOQU_C128 ? sizeof(double) : sizeof(float);
    ProgramLayout L;
    L.hdr_off = 0;
    L.instr_off = (nh * sizeof(StageHeader) + 255) & ~size_t(255);
    L.pool_off = (L.instr_off + ni * sizeof(Instr) + 255) & ~size_t(255);
    // 64 spare bytes: the kernel's instruction-head fetch (64 B) stays inside the blob even for
    // a stage without instructions at the end of the array
    L.bytes = (L.pool_off + np * rsz + 64 + 255) & ~size_t(255);
    return L;
}

int serialize_program(const Program &P0, int prec, const ProgramLayout &L, unsigned char *out,
                      std::vector<StageHeader> *stages) {
    for (const StageHeader &H : P0.stages) {
        if ((size_t(1) << H.m) * 8 > 65536)
            return fail(GOQU_ERR_UNSUPPORTED, "stage tile larger than the LDS offset range");
        // the stage kernels run a stage's ops up to its closing I_END
        const size_t e = (size_t)H.instr_off + (size_t)H.ninstr;
        if (H.instr_off < 0 || H.ninstr < 0 || e >= P0.instrs.size() || P0.instrs[e].op != I_END)
            return fail(GOQU_ERR_INVALID_ARG, "stage instruction list not closed by I_END");
    }
    const size_t rsz = prec == GOQU_C128 ? sizeof(double) : sizeof(float);
    const size_t nh = P0.stages.size(), ni = P0.instrs.size(), np = P0.pool.size();
    if (L.instr_off < nh * sizeof(StageHeader) || L.pool_off < L.instr_off + ni * sizeof(Instr) ||
        L.bytes < L.pool_off + np * rsz + 64)
        return fail(GOQU_ERR_INVALID_ARG, "program larger than its layout");
    Program P;
    P.stages = P0.stages;
    P.instrs = P0.instrs;
    set_lds_offsets(P);
    std::memset(out, 0, L.bytes);
    std::memcpy(out + L.hdr_off, P.stages.data(), nh * sizeof(StageHeader));
    std::memcpy(out + L.instr_off, P.instrs.data(), ni * sizeof(Instr));
    if (prec == GOQU_C128) {
        std::memcpy(out + L.pool_off, P0.pool.data(), np * sizeof(double));
    } else {
        float *f = reinterpret_cast<float *>(out + L.pool_off);
        for (size_t i = 0; i < np; i++) f[i] = (float)P0.pool[i];
    }
    if (stages) *stages = P.stages;
    return GOQU_OK;
}

int upload_program(const Program &P0, int prec, hipStream_t s, DeviceProgram &dp) {
    const ProgramLayout L = program_layout(P0.stages.size(), P0.instrs.size(), P0.pool.size(), prec);
    std::vector<unsigned char> host(L.bytes);
    std::vector<StageHeader> stages;
    int rc = serialize_program(P0, prec, L, host.data(), &stages);
    if (rc) return rc;
    if (dp.bytes < L.bytes) {
        if (dp.blob) GOQU_HIP(hipFree(dp.blob));
        dp.blob = nullptr;
        dp.bytes = 0;
        GOQU_HIP(dmalloc((void **)&dp.blob, L.bytes));
        dp.bytes = L.bytes;
    }
    dp.hdr_off = L.hdr_off;
    dp.instr_off = L.instr_off;
    dp.pool_off = L.pool_off;
    GOQU_HIP(hipMemcpyAsync(dp.blob, host.data(), L.bytes, hipMemcpyHostToDevice, s));
    // the host buffer must outlive the async copy
    GOQU_HIP(hipStreamSynchronize(s));
    dp.stages = std::move(stages);
    return GOQU_OK;
}

// A stage without instructions that stores in its load order changes nothing (a stage of
// relabelled SWAPs alone still permutes its tile and runs).
bool stage_is_identity(const StageHeader &H) {
    return H.ninstr == 0 && std::memcmp(H.gq0, H.gqN, (size_t)H.m) == 0;
}

template <typename R>
static int run_program_t(State &st, const DeviceProgram &dp, hipStream_t s, const ShotBatch *b) {
    const unsigned char *blob = static_cast<const unsigned char *>(dp.blob);
    const StageHeader *hdrs = reinterpret_cast<const StageHeader *>(blob + dp.hdr_off);
    const Instr *instrs = reinterpret_cast<const Instr *>(blob + dp.instr_off);
    const R *pool = reinterpret_cast<const R *>(blob + dp.pool_off);
    const unsigned nshots = b ? (unsigned)b->nshots : 1u;
    const uint8_t *act = b ? b->act : nullptr;
    const uint64_t stride = b ? b->stride : 0, bstride = b ? b->blob_stride : 0;
    const double sweep_bytes = 2.0 * st.size() * sizeof(cx<R>) * nshots;  // read + write of every state
    for (size_t i = 0; i < dp.stages.size(); i++) {
        const int m = dp.stages[i].m;
        if (stage_is_identity(dp.stages[i])) continue;
        const uint64_t ntiles = 1ull << (st.nq - m);
        const unsigned threads = 1u << (m - kRegBits);
        const size_t lds = (size_t(1) << m) * 8;  // 8-byte LDS elements
        // one tile per workgroup, the grid covers every tile (the kernels run one tile each):
        // measured faster than persistent workgroups sized to the device, whose waves stay in
        // step and load in bursts
        if (ntiles > (1ull << 31) / threads) return fail(GOQU_ERR_UNSUPPORTED, "state too large for one launch");
        const unsigned grid = (unsigned)ntiles;
        hipEvent_t ev0 = prof_begin(s);
        const dim3 g(grid, nshots);
        if (dp.stages[i].flags & SF_DM) {
            if constexpr (sizeof(R) == 8)
                hipLaunchKernelGGL((dp.stages[i].flags & SF_BIG) ? k_stage_dm_big<double> : k_stage_dm<double>, g,
                                   dim3(threads), lds, s, static_cast<cx<double> *>(st.amps), hdrs + i, instrs, pool,
                                   ntiles, act, stride, bstride);
            else
                return fail(GOQU_ERR_UNSUPPORTED, "density-matrix stages run in complex128 only");
        } else if (dp.stages[i].flags & SF_BIG)
            hipLaunchKernelGGL(k_stage_big<R>, g, dim3(threads), lds, s, static_cast<cx<R> *>(st.amps), hdrs + i,
                               instrs, pool, ntiles, act, stride, bstride);
        else
            hipLaunchKernelGGL(k_stage<R>, g, dim3(threads), lds, s, static_cast<cx<R> *>(st.amps), hdrs + i, instrs,
                               pool, ntiles, act, stride, bstride);
        GOQU_HIP(hipGetLastError());
        prof_end(ev0, s, (dp.stages[i].flags & SF_DM) ? PROF_STAGE_DM : PROF_STAGE, sweep_bytes);
    }
    return GOQU_OK;
}

int run_program(State &st, const DeviceProgram &dp, hipStream_t s, const ShotBatch *b) {
    return st.prec == GOQU_C128 ? run_program_t<double>(st, dp, s, b) : run_program_t<float>(st, dp, s, b);
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
    // one tile per workgroup; the per-workgroup gradient partials (64 doubles each) bound the grid
    if (ntiles > (1ull << 21)) return fail(GOQU_ERR_UNSUPPORTED, "paired stages: state too large");
    const unsigned grid = (unsigned)ntiles;
    const int ng = P.num_grad;
    // per-stage partials (slot-major, one per workgroup), reduced after every stage into gsum;
    // grow-only per host thread (every call ends synchronised)
    static thread_local Scratch gbuf;
    double *gpart = nullptr, *gsum = nullptr;
    if (ng) {
        rc = gbuf.ensure(((size_t)kMaxGradPerStage * grid + ng) * sizeof(double));
        if (rc) return rc;
        gpart = static_cast<double *>(gbuf.ptr);
        gsum = gpart + (size_t)kMaxGradPerStage * grid;
    }
    for (size_t i = 0; i < P.stages.size() && !rc; i++) {
        const StageHeader &H = P.stages[i];
        if (stage_is_identity(H)) continue;
        const unsigned threads = 1u << (m - kRegBits);
        // 8-byte LDS elements, then the per-wave gradient partials
        const size_t lds = (size_t(1) << m) * 8 + kMaxGradPerStage * 8 * sizeof(double);
        hipEvent_t ev0 = prof_begin(s);
        const bool big = H.flags & SF_BIG;
        if (H.flags & SF_DM) {  // channel superoperators: the paired density-matrix kernels (complex128)
            if constexpr (sizeof(R) == 8)
                hipLaunchKernelGGL(big ? k_stage_pair_dm_big<double> : k_stage_pair_dm<double>, dim3(grid),
                                   dim3(threads), lds, s, static_cast<cx<double> *>(psi.amps),
                                   static_cast<cx<double> *>(lam.amps), hdrs + i, instrs, pool, ntiles, gpart);
            else
                rc = fail(GOQU_ERR_UNSUPPORTED, "paired density-matrix stages need complex128");
        } else {
            hipLaunchKernelGGL(big ? k_stage_pair_big<R> : k_stage_pair<R>, dim3(grid), dim3(threads), lds, s,
                               static_cast<cx<R> *>(psi.amps), static_cast<cx<R> *>(lam.amps), hdrs + i, instrs, pool,
                               ntiles, gpart);
        }
        if (hipGetLastError() != hipSuccess) rc = fail(GOQU_ERR_DEVICE, "paired stage launch failed");
        prof_end(ev0, s, PROF_STAGE_PAIR, 4.0 * psi.size() * sizeof(cx<R>));  // read + write of psi and lambda
        if (!rc && H.ngrad) {
            hipLaunchKernelGGL(k_sum_rows, dim3(H.ngrad), dim3(256), 0, s, gpart, (int64_t)grid, gsum + H.grad_base);
            if (hipGetLastError() != hipSuccess) rc = fail(GOQU_ERR_DEVICE, "gradient reduction launch failed");
        }
    }
    slot_vals.assign(ng, 0.0);
    hipError_t e = hipSuccess;
    if (!rc && ng) e = hipMemcpyAsync(slot_vals.data(), gsum, ng * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
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
    return run_pair_planned(psi, lam, plan_program(eops, psi.nq, default_plan_config(psi.prec, psi.nq, true)), nslots,
                            slot_vals, s);
}

int run_pair_planned(State &psi, State &lam, const Program &P, int nslots, std::vector<double> &slot_vals,
                     hipStream_t s) {
    slot_vals.assign(nslots, 0.0);
    if (P.stages.empty()) return GOQU_OK;
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

static int run_segment(State &st, const std::vector<EOp> &eops, hipStream_t s, const PlanConfig *cfg) {
    if (eops.empty()) return GOQU_OK;
    Program P = plan_program(eops, st.nq, cfg ? *cfg : default_plan_config(st.prec, st.nq, false));
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

// Stage programs between wide ops (dense on 5..9 targets), which run as MFMA contractions.
int run_eops(State &st, const std::vector<EOp> &eops, hipStream_t s, const PlanConfig *cfg) {
    bool wide = false;
    for (const EOp &e : eops) wide = wide || e.kind == EOP_WIDE;
    if (!wide) return run_segment(st, eops, s, cfg);
    size_t i = 0;
    while (i < eops.size()) {
        size_t j = i;
        while (j < eops.size() && eops[j].kind != EOP_WIDE) j++;
        if (j > i) {
            int rc = run_segment(st, std::vector<EOp>(eops.begin() + i, eops.begin() + j), s, cfg);
            if (rc) return rc;
        }
        if (j < eops.size()) {
            int rc = run_wide(st, eops[j], s);
            if (rc) return rc;
            j++;
        }
        i = j;
    }
    return GOQU_OK;
}

void PreparedOps::release() {
    for (Piece &p : pieces) {
        if (p.prog.blob) (void)hipFree(p.prog.blob);
        p.prog.blob = nullptr;
        p.w.release();
    }
    pieces.clear();
    stages = 0;
}

int prepare_eops(const State &st, const std::vector<EOp> &eops, const PlanConfig &cfg, hipStream_t s,
                 PreparedOps &out) {
    out.release();
    size_t i = 0;
    while (i < eops.size()) {
        size_t j = i;
        while (j < eops.size() && eops[j].kind != EOP_WIDE) j++;
        if (j > i) {
            Program P = plan_program(std::vector<EOp>(eops.begin() + i, eops.begin() + j), st.nq, cfg);
            out.pieces.emplace_back();
            int rc = upload_program(P, st.prec, s, out.pieces.back().prog);
            if (rc) return rc;
            for (const StageHeader &H : P.stages) out.stages += stage_is_identity(H) ? 0 : 1;
        }
        if (j < eops.size()) {
            out.pieces.emplace_back();
            out.pieces.back().wide = true;
            int rc = prepare_wide(st.nq, eops[j], s, out.pieces.back().w);
            if (rc) return rc;
            j++;
        }
        i = j;
    }
    return GOQU_OK;
}

int launch_prepared(State &st, const PreparedOps &p, hipStream_t s, const ShotBatch *b) {
    for (const PreparedOps::Piece &pc : p.pieces) {
        int rc = pc.wide ? launch_wide(st.amps, st.prec, pc.w, s, b) : run_program(st, pc.prog, s, b);
        if (rc) return rc;
    }
    return GOQU_OK;
}

}  // namespace goqu

#endif  // GOQU_STAGE_KERNEL
