// qaoa.hip — batched QAOA MaxCut: energy and adjoint gradient of thousands of instances.
//
// Circuit (algorithm/qaoa/qaoa.go:66-95 with maxcut.go:13-36): H on every qubit, then per
// layer k the cost layer RZZ(gamma_k) on every edge and the mixer RX(2 beta_k) on every qubit.
// The cost layer is one diagonal phase, exp(-i gamma/2 * S(z)) with S(z) = |E| - 2 cut(z), so
// it costs one table lookup per amplitude; cut(z) = sum_i popcount(U_i & (z_i ? ~z : z)) with
// U_i the neighbours j > i of vertex i.  H = sum_edges (-1/2 + Z_i Z_j / 2) is diagonal with
// value -cut(z).
//
// Gradient (adjoint): lambda = H psi; going backwards through layer k,
//   dE/dbeta_k  = 2 Im <lambda| sum_q X_q |psi>   (mixer = exp(-i beta_k sum_q X_q)),
//   dE/dgamma_k = Im <lambda| S |psi>            (cost  = exp(-i gamma_k S / 2)),
// then both states are uncomputed through the layer.  Both generators commute with their own
// layer, so each term may be taken anywhere inside its layer's uncompute.
//
// Instances with n <= 12 run entirely in one workgroup's LDS (psi and lambda together).  Larger
// instances are swept in passes over tiles of 4096 amplitudes (A tiles: qubits 0..11; B tiles: qubits
// 0..L-1 and 12..n-1), p + 1 passes forward and p + 1 reverse, each applying the second half of
// one layer's mixer, the next cost layer and the first half of the next mixer; the passes of all
// instances advance together, one launch per step, and the whole sequence is captured in a single
// hipGraph.  A tile lives in registers, 16 amplitudes per thread: the mixer runs on register bits,
// and LDS relayouts bring the other tile bits into registers.  cut(z) is computed once per
// instance (first cost layer) and kept as one byte per amplitude; the energy and lambda = H psi
// are formed when the reverse sweep first loads psi.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <type_traits>

#include "runtime.h"

namespace goqu {

constexpr int kQMaxQubits = 22;
constexpr int kQMaxLayers = 16;
constexpr int kQTileBits = 12;
constexpr int kQThreads = 256;
constexpr int kQMaxEdges = kQMaxQubits * (kQMaxQubits - 1) / 2;

struct QInst {
    int32_t n, p, nedges, slot_stride;  // slot_stride = 1 + 3p (large instances)
    uint32_t up[kQMaxQubits];           // neighbours j > i
    double gamma[kQMaxLayers], beta[kQMaxLayers];
    uint64_t psi, lam;                  // element offsets into the amplitude pool (large)
    uint64_t cut;                       // byte offset of cut(z), z = 0..2^n-1, in the cut pool (large)
    int64_t part;                       // offset into the partials (large)
    int64_t tab;                        // offset into the layer tables (k_qaoa_tables)
    int32_t ntiles, small;
};


// cut(z) from the upper adjacency masks staged in LDS
__device__ __forceinline__ int cut_of(const uint32_t *up, int n, uint32_t z) {
    int c = 0;
    for (int i = 0; i < n; i++) c += __popc(up[i] & ((z >> i & 1) ? ~z : z));
    return c;
}

__device__ __forceinline__ void stage_up(uint32_t *sup, const QInst &I) {
    for (int i = threadIdx.x; i < kQMaxQubits; i += blockDim.x) sup[i] = I.up[i];
    __syncthreads();
}

struct c2d {
    double x, y;
};
__device__ __forceinline__ c2d cmul(c2d a, c2d b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }

// RX(2*sb-angle) pass over local bit `b` of an LDS array of 2^m amplitudes: [a0,a1] <-
// [c a0 - i s a1, -i s a0 + c a1]; s < 0 uncomputes.
__device__ __forceinline__ void rx_pass(c2d *A, int m, int b, double c, double s) {
    const uint32_t half = 1u << (m - 1);
    for (uint32_t t = threadIdx.x; t < half; t += blockDim.x) {
        const uint32_t lo = t & ((1u << b) - 1);
        const uint32_t i0 = ((t >> b) << (b + 1)) | lo, i1 = i0 | (1u << b);
        const c2d a0 = A[i0], a1 = A[i1];
        A[i0] = {c * a0.x + s * a1.y, c * a0.y - s * a1.x};
        A[i1] = {c * a1.x + s * a0.y, c * a1.y - s * a0.x};
    }
    __syncthreads();
}

// sum over pairs of conj(L0) P1 + conj(L1) P0 (= <lambda| X_b |psi> restricted to the array)
__device__ __forceinline__ c2d x_overlap(const c2d *P, const c2d *L, int m, int b) {
    c2d w = {0, 0};
    const uint32_t half = 1u << (m - 1);
    for (uint32_t t = threadIdx.x; t < half; t += blockDim.x) {
        const uint32_t lo = t & ((1u << b) - 1);
        const uint32_t i0 = ((t >> b) << (b + 1)) | lo, i1 = i0 | (1u << b);
        const c2d p0 = P[i0], p1 = P[i1], l0 = L[i0], l1 = L[i1];
        w.x += l0.x * p1.x + l0.y * p1.y + l1.x * p0.x + l1.y * p0.y;
        w.y += l0.x * p1.y - l0.y * p1.x + l1.x * p0.y - l1.y * p0.x;
    }
    return w;
}

__device__ __forceinline__ double block_sum1(double v, double *sh) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += sh[w];
    __syncthreads();
    return t;
}

// The layer tables of every instance, computed once per batch: for layer k at I.tab + k (ne + 3),
// the cost phases exp(-i gamma_k / 2 (ne - 2c)) for c = 0..ne, then (cos beta_k, sin beta_k), then
// the shear form of the mixer rotation (tan(b/2), sin b) with b = beta_k folded by pi into
// cos b >= 0 (RX(2 beta) = -RX(2 (beta - pi)); the tile kernels drop that global sign, which
// multiplies psi and lambda alike and cancels in the energy and every gradient term).  The tile
// kernels copy a layer's phases into LDS instead of evaluating ne + 3 sines and cosines per
// workgroup.
constexpr int kQTabExtra = 3;
__global__ void k_qaoa_tables(const QInst *__restrict__ insts, c2d *__restrict__ tabs) {
    const QInst &I = insts[blockIdx.x];
    const int ne = I.nedges, T = ne + kQTabExtra;
    for (int e = threadIdx.x; e < I.p * T; e += blockDim.x) {
        const int k = e / T, c = e % T;
        double sn, cs;
        if (c <= ne) {
            sincos(-0.5 * I.gamma[k] * (double)(ne - 2 * c), &sn, &cs);
        } else {
            sincos(I.beta[k], &sn, &cs);
            if (c == ne + 2) {
                if (cs < 0) cs = -cs, sn = -sn;
                cs = sn / (1.0 + cs);  // tan(b/2), |.| <= 1
            }
        }
        tabs[I.tab + e] = {cs, sn};
    }
}
// layer k's cost phases into LDS
__device__ __forceinline__ void load_tab(c2d *tab, const QInst &I, const c2d *__restrict__ tabs, int k) {
    const c2d *src = tabs + I.tab + (int64_t)k * (I.nedges + kQTabExtra);
    for (int c = threadIdx.x; c <= I.nedges; c += blockDim.x) tab[c] = src[c];
    __syncthreads();
}
// (cos beta_k, sin beta_k) of layer k
__device__ __forceinline__ void beta_cs(const QInst &I, const c2d *__restrict__ tabs, int k, double &cb, double &sb) {
    const c2d v = tabs[I.tab + (int64_t)k * (I.nedges + kQTabExtra) + I.nedges + 1];
    cb = v.x;
    sb = v.y;
}
// (tan(b/2), sin b) of layer k's folded mixer angle
__device__ __forceinline__ void beta_shear(const QInst &I, const c2d *__restrict__ tabs, int k, double &tb, double &sb) {
    const c2d v = tabs[I.tab + (int64_t)k * (I.nedges + kQTabExtra) + I.nedges + 2];
    tb = v.x;
    sb = v.y;
}

// ---- small instances: everything in LDS ---------------------------------------------------------
__global__ void __launch_bounds__(kQThreads) k_qaoa_small(const QInst *__restrict__ insts,
                                                          const c2d *__restrict__ tabs,
                                                          const int32_t *__restrict__ ids,
                                                          double *__restrict__ energies,
                                                          const int64_t *__restrict__ goff,
                                                          double *__restrict__ grads) {
    extern __shared__ unsigned char smem[];
    const int id = ids[blockIdx.x];
    const QInst &I = insts[id];
    const int n = I.n, p = I.p, ne = I.nedges;
    const uint32_t N = 1u << n;
    c2d *P = reinterpret_cast<c2d *>(smem);
    c2d *L = P + N;
    c2d *tab = L + N;
    __shared__ double sh[16];
    __shared__ uint32_t sup[kQMaxQubits];
    stage_up(sup, I);
    // H^n |0>: every amplitude (1/sqrt2)^n, multiplied out as the H kernels would
    double a = 1.0;
    for (int q = 0; q < n; q++) a *= 0.70710678118654752440;
    for (uint32_t z = threadIdx.x; z < N; z += blockDim.x) P[z] = {a, 0.0};
    __syncthreads();
    for (int k = 0; k < p; k++) {
        load_tab(tab, I, tabs, k);
        for (uint32_t z = threadIdx.x; z < N; z += blockDim.x) P[z] = cmul(P[z], tab[cut_of(sup, n, z)]);
        __syncthreads();
        double sb, cb;
        beta_cs(I, tabs, k, cb, sb);
        for (int q = 0; q < n; q++) rx_pass(P, n, q, cb, sb);
    }
    // energy and lambda = H psi (H diagonal: -cut)
    double e = 0;
    for (uint32_t z = threadIdx.x; z < N; z += blockDim.x) {
        const double h = -(double)cut_of(sup, n, z);
        const c2d v = P[z];
        e += h * (v.x * v.x + v.y * v.y);
        L[z] = {h * v.x, h * v.y};
    }
    __syncthreads();
    e = block_sum1(e, sh);
    if (threadIdx.x == 0) energies[id] = e;
    double *g = grads + goff[id];
    for (int k = p - 1; k >= 0; k--) {
        double sb, cb;
        beta_cs(I, tabs, k, cb, sb);
        double wy = 0;
        for (int q = 0; q < n; q++) wy += x_overlap(P, L, n, q).y;
        wy = block_sum1(wy, sh);
        for (int q = 0; q < n; q++) {
            rx_pass(P, n, q, cb, -sb);
            rx_pass(L, n, q, cb, -sb);
        }
        load_tab(tab, I, tabs, k);
        double gy = 0;
        for (uint32_t z = threadIdx.x; z < N; z += blockDim.x) {
            const int c = cut_of(sup, n, z);
            const double S = (double)(ne - 2 * c);
            const c2d pz = P[z], lz = L[z];
            gy += S * (lz.x * pz.y - lz.y * pz.x);  // Im(conj(l) S p)
            const c2d t = {tab[c].x, -tab[c].y};
            P[z] = cmul(pz, t);
            L[z] = cmul(lz, t);
        }
        __syncthreads();
        gy = block_sum1(gy, sh);
        if (threadIdx.x == 0) {
            g[2 * k] = gy;
            g[2 * k + 1] = 2.0 * wy;
        }
    }
}

// ---- large instances: one tile of 4096 amplitudes per workgroup, one launch per pass -------------
// A pass sweeps an instance's state once, in tiles of one of two kinds: A tiles hold qubits 0..11,
// B tiles qubits 0..L-1 and 12..n-1 (L = 24 - n; the tile index supplies the rest).  Layer k's
// mixer is the product of its A half M_k^A (RX on qubits 0..11) and its B half M_k^B (qubits
// 12..n-1), which commute, and each cost layer C_k is diagonal, so the forward circuit regroups into
// p + 1 passes of alternating kinds,
//   [C_0 M_0^A] [M_0^B C_1 M_1^B] [M_1^A C_2 M_2^A] ... [M_{p-1}^X],
// each applying the second half of one layer's mixer, the next cost layer and the first half of
// the next mixer; the reverse sweep runs the same passes inverted, last to first.  With one pass per
// half layer instead, a batch moved (192 p - 16) bytes per amplitude; this way it moves 96 p + 32.
struct QWork {
    int32_t inst;
    int8_t half;               // 0: A tiles, 1: B tiles
    int8_t pre, cost, post;    // in forward order: mixer half of layer `pre`, cost layer `cost`, mixer
                               // half of layer `post` (-1: absent)
    int32_t flags;             // Q_*
    int32_t first_block;       // prefix over the launch
};
enum : int32_t {
    Q_INIT = 1,     // first forward pass: psi starts as H^n |0>, cut(z) is computed and kept
    Q_ENERGY = 2,   // first reverse pass: the energy, and lambda = H psi
    Q_NOSTORE = 4,  // last reverse pass: psi and lambda are not needed afterwards
};

// (an instance index: below 2^22, so 32 bits; loads and stores add it to the instance's uniform base)
__device__ __forceinline__ uint32_t tile_global(const QInst &I, int half, uint32_t tile, uint32_t l) {
    if (!half) return (tile << kQTileBits) | l;
    // B: local bits 0..L-1 -> qubits 0..L-1, local bits L..11 -> qubits 12..n-1; the tile index
    // supplies qubits L..11
    const int Lb = 2 * kQTileBits - I.n;
    const uint32_t lo = l & ((1u << Lb) - 1), hi = l >> Lb;
    return lo | (tile << Lb) | (hi << kQTileBits);
}

// Register tiles: thread t (lane, wave) holds 16 amplitudes of the tile; which 4 local bits its
// register index r covers depends on the layout:
//   0: r -> local bits 8-11, lane -> 0-5, wave -> 6-7   (loads and stores: lanes on the lowest bits)
//   1: r -> local bits 0-3,  lane bits 4-5 -> 4-5, wave -> 6-7, lane bits 0-3 -> 8-11
//   2: r -> local bits 4-7,  lane bits 4-5 -> 0-1, wave -> 2-3, lane bits 0-3 -> 8-11
// The mixer acts on the register bits of each layout in turn; layouts change through LDS.  A local
// index is the thread's part OR the register's part (disjoint bits), and the LDS swizzle is
// XOR-linear, so an LDS address is (thread part) XOR (compile-time register part).
template <int LAY>
__device__ __forceinline__ uint32_t qthread(uint32_t lane, uint32_t wave) {
    if constexpr (LAY == 0) return lane | (wave << 6);
    else if constexpr (LAY == 1) return ((lane >> 4) << 4) | (wave << 6) | ((lane & 15) << 8);
    else return (lane >> 4) | (wave << 2) | ((lane & 15) << 8);
}
template <int LAY>
__device__ __forceinline__ constexpr uint32_t qreg(uint32_t r) {
    return LAY == 0 ? (r << 8) : LAY == 1 ? r : (r << 4);
}
// LDS element of local index l: low bits XOR bits 8-11 (spreads the layouts' strided accesses)
__device__ __forceinline__ constexpr uint32_t qswz(uint32_t l) { return l ^ ((l >> 8) & 15u); }

// In-place register arithmetic.  Each amplitude keeps its own registers through inline asm with
// tied operands: written as plain expressions, the scheduler computed all 16 (or 32) results into
// fresh registers and copied them back at every branch merge, and the reverse kernel needed over
// 256 VGPRs (one wave per SIMD).
//
// RX: [a0, a1] <- [c a0 - i s a1, -i s a0 + c a1] (up to the global sign of the folded angle, with
// t = tan(b/2)), which is two real rotations, of (x0, y1) and of (y0, x1), each as three in-place
// shears (x0 += t y1, y1 -= s x0, x0 += t y1; y0 -= t x1, x1 += s y0, y0 -= t x1): six FMAs and no
// temporaries.  The inverse is the same with (-t, -s).
__device__ __forceinline__ void rx_pair(c2d &a0, c2d &a1, double t, double s) {
    asm volatile(
        "v_fma_f64 %[x0], %[t], %[y1], %[x0]\n\t"
        "v_fma_f64 %[y0], -%[t], %[x1], %[y0]\n\t"
        "v_fma_f64 %[y1], -%[s], %[x0], %[y1]\n\t"
        "v_fma_f64 %[x1], %[s], %[y0], %[x1]\n\t"
        "v_fma_f64 %[x0], %[t], %[y1], %[x0]\n\t"
        "v_fma_f64 %[y0], -%[t], %[x1], %[y0]"
        : [x0] "+v"(a0.x), [y0] "+v"(a0.y), [x1] "+v"(a1.x), [y1] "+v"(a1.y)
        : [t] "v"(t), [s] "v"(s));
}
// a <- a * (tc + i ts), or a * (tc - i ts) when CONJ
template <bool CONJ>
__device__ __forceinline__ void cmul_in(c2d &a, double tc, double ts) {
    double t0, t1;
    if constexpr (CONJ)
        asm volatile(
            "v_mul_f64 %[t0], %[ts], %[y]\n\t"
            "v_mul_f64 %[t1], %[ts], %[x]\n\t"
            "v_fma_f64 %[x], %[tc], %[x], %[t0]\n\t"
            "v_fma_f64 %[y], %[tc], %[y], -%[t1]"
            : [x] "+v"(a.x), [y] "+v"(a.y), [t0] "=&v"(t0), [t1] "=&v"(t1)
            : [tc] "v"(tc), [ts] "v"(ts));
    else
        asm volatile(
            "v_mul_f64 %[t0], %[ts], %[y]\n\t"
            "v_mul_f64 %[t1], %[ts], %[x]\n\t"
            "v_fma_f64 %[x], %[tc], %[x], -%[t0]\n\t"
            "v_fma_f64 %[y], %[tc], %[y], %[t1]"
            : [x] "+v"(a.x), [y] "+v"(a.y), [t0] "=&v"(t0), [t1] "=&v"(t1)
            : [tc] "v"(tc), [ts] "v"(ts));
}

// RX on register bit J of a 16-amplitude tile
template <int J>
__device__ __forceinline__ void rx_reg(c2d (&A)[16], double c, double s) {
#pragma unroll
    for (int r = 0; r < 16; r++) {
        if (r & (1 << J)) continue;
        rx_pair(A[r], A[r | (1 << J)], c, s);
    }
}
// Im <lambda| X_J |psi> over the register pairs of bit J
template <int J>
__device__ __forceinline__ double xov_reg(const c2d (&P)[16], const c2d (&L)[16]) {
    double w = 0;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        if (r & (1 << J)) continue;
        const c2d p0 = P[r], p1 = P[r | (1 << J)], l0 = L[r], l1 = L[r | (1 << J)];
        // one fused chain, ordered before the in-place rotations of the same registers (free to
        // move, it would need copies of the amplitudes the rotations overwrite)
        asm volatile(
            "v_fma_f64 %[w], %[l0x], %[p1y], %[w]\n\t"
            "v_fma_f64 %[w], -%[l0y], %[p1x], %[w]\n\t"
            "v_fma_f64 %[w], %[l1x], %[p0y], %[w]\n\t"
            "v_fma_f64 %[w], -%[l1y], %[p0x], %[w]"
            : [w] "+v"(w)
            : [l0x] "v"(l0.x), [l0y] "v"(l0.y), [l1x] "v"(l1.x), [l1y] "v"(l1.y), [p0x] "v"(p0.x),
              [p0y] "v"(p0.y), [p1x] "v"(p1.x), [p1y] "v"(p1.y));
    }
    return w;
}

// moves a register tile from layout FROM to layout TO through LDS (real parts, then imaginary)
template <int FROM, int TO>
__device__ __forceinline__ void relayout(c2d (&A)[16], double *buf, uint32_t lane, uint32_t wave) {
    // byte offsets into buf (8-byte elements)
    const uint32_t wt0 = qswz(qthread<FROM>(lane, wave)) << 3, rt0 = qswz(qthread<TO>(lane, wave)) << 3;
    unsigned char *const bb = reinterpret_cast<unsigned char *>(buf);
#pragma unroll
    for (int part = 0; part < 2; part++) {
        // opaque copies per pass: the 16 addresses of a pass are recomputed (one XOR each), not
        // held in registers across passes and states
        uint32_t wt, rt;
        asm volatile("v_mov_b32 %0, %1" : "=v"(wt) : "v"(wt0));
#pragma unroll
        for (int r = 0; r < 16; r++)
            *reinterpret_cast<double *>(bb + (wt ^ (qswz(qreg<FROM>(r)) << 3))) = part ? A[r].y : A[r].x;
        __syncthreads();
        asm volatile("v_mov_b32 %0, %1" : "=v"(rt) : "v"(rt0));
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const double v = *reinterpret_cast<const double *>(bb + (rt ^ (qswz(qreg<TO>(r)) << 3)));
            if (part)
                A[r].y = v;
            else
                A[r].x = v;
        }
        __syncthreads();
    }
}

// the mixer on the register bits of layout LAY that carry local bits >= b0 (forward), or the
// overlaps and the uncompute of both states (reverse); (cb, sb) = (tan(b/2), sin b) of the layer
template <bool REV, int LAY>
__device__ __forceinline__ void mix_group(c2d (&P)[16], c2d (&L)[16], int b0, double cb, double sb, double &wy) {
    constexpr int base = LAY == 0 ? 8 : LAY == 1 ? 0 : 4;
    auto bit = [&](auto jc) __attribute__((always_inline)) {
        constexpr int J = decltype(jc)::value;
        if (base + J < b0) return;
        if constexpr (REV) {
            wy += xov_reg<J>(P, L);
            rx_reg<J>(P, -cb, -sb);
            rx_reg<J>(L, -cb, -sb);
        } else {
            rx_reg<J>(P, cb, sb);
        }
    };
    bit(std::integral_constant<int, 0>{});
    bit(std::integral_constant<int, 1>{});
    bit(std::integral_constant<int, 2>{});
    bit(std::integral_constant<int, 3>{});
}

template <bool REV, int FROM, int TO>
__device__ __forceinline__ void relayout_states(c2d (&P)[16], c2d (&L)[16], double *buf, uint32_t lane,
                                                uint32_t wave) {
    relayout<FROM, TO>(P, buf, lane, wave);
    if constexpr (REV) relayout<FROM, TO>(L, buf, lane, wave);
}

// Sum of v[0..3] over the workgroup, in a fixed order; the totals are valid in thread 0
__device__ __forceinline__ void block_sum4(double (&v)[4], double *sh) {
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v[i] += __shfl_xor(v[i], off, 64);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int i = 0; i < 4; i++) sh[(threadIdx.x >> 6) * 4 + i] = v[i];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            double t = 0;
            for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += sh[w * 4 + i];
            v[i] = t;
        }
}

// gy += S Im(conj(l) p), ordered before the in-place updates of the same registers
__device__ __forceinline__ void cost_overlap(double &gy, double S, const c2d &l, const c2d &p) {
    double t;
    asm volatile(
        "v_mul_f64 %[t], %[S], %[lx]\n\t"
        "v_fma_f64 %[g], %[t], %[py], %[g]\n\t"
        "v_mul_f64 %[t], %[S], %[ly]\n\t"
        "v_fma_f64 %[g], -%[t], %[px], %[g]"
        : [g] "+v"(gy), [t] "=&v"(t)
        : [S] "v"(S), [lx] "v"(l.x), [ly] "v"(l.y), [px] "v"(p.x), [py] "v"(p.y));
}

// One pass of every work item of the launch (REV: the launch holds only reverse passes, over psi and
// lambda; otherwise only forward passes, over psi: the forward kernel keeps half the registers and
// runs more waves).
template <bool REV>
__global__ void __launch_bounds__(kQThreads) k_qaoa_pass(const QInst *__restrict__ insts,
                                                         const c2d *__restrict__ tabs,
                                                         const QWork *__restrict__ work, int nwork,
                                                         c2d *__restrict__ pool, uint8_t *__restrict__ cuts,
                                                         double *__restrict__ parts) {
    // 32 KiB of LDS in all, so five forward workgroups fit a CU: the relayout buffer also holds the
    // cost table, the adjacency masks and the reduction slots, which are never live during a relayout
    // (a cost layer ends with a barrier before the next relayout writes; a relayout ends with one
    // after its reads)
    __shared__ double buf[1 << kQTileBits];
    c2d *const tab = reinterpret_cast<c2d *>(buf);                                  // 3,712 B
    uint32_t *const sup = reinterpret_cast<uint32_t *>(buf) + 928;                 // 88 B
    double *const sh = buf + 480;                                                   // 128 B
    static_assert((kQMaxEdges + 1) * sizeof(c2d) <= 928 * 4 && 928 * 4 + kQMaxQubits * 4 <= 480 * 8 &&
                      (480 + 4 * (kQThreads / 64)) * 8 <= sizeof(double) << kQTileBits,
                  "LDS carve");
    // which work item: last with first_block <= blockIdx.x
    int lo = 0, hi = nwork;
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (work[mid].first_block <= (int)blockIdx.x) lo = mid;
        else hi = mid;
    }
    const QWork W = work[lo];
    const QInst &I = insts[W.inst];
    const int n = I.n, ne = I.nedges, half = W.half, kpre = W.pre, kcost = W.cost, kpost = W.post;
    // B tiles of 22-qubit instances read and write 64-byte runs (qubits 0-1), half a 128-byte cache
    // line; the other half belongs to the tile next in qubit 2.  Workgroups go to the eight XCDs
    // round-robin by index, so the pair is given to workgroups b and b + 8 (one XCD, one L2, running
    // together) instead of b and b + 1 (two XCDs).
    uint32_t tile = blockIdx.x - W.first_block;
    if (half && 2 * kQTileBits - n < 3 && I.ntiles >= 16)
        tile = (tile & ~15u) | ((tile & 7u) << 1) | ((tile >> 3) & 1u);
    const int flags = W.flags;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    c2d *psi = pool + I.psi, *lam = pool + I.lam;
    uint8_t *cut = cuts + I.cut;
    double *part = parts + I.part + (int64_t)tile * I.slot_stride;
    // the first forward pass computes cut(z); adj/deg: neighbours and degree of vertices 8..11
    uint32_t adj[4] = {0, 0, 0, 0};
    int deg[4] = {0, 0, 0, 0};
    if (!REV && (flags & Q_INIT)) {
        stage_up(sup, I);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int v = 8 + j;
            uint32_t a = v < n ? I.up[v] : 0u;
            for (int i = 0; i < v && i < n; i++) a |= ((I.up[i] >> v) & 1u) << i;
            adj[j] = a;
            deg[j] = __popc(a);
        }
    }
    const int b0 = half ? 2 * kQTileBits - n : 0;  // local bits b0..11 carry this pass's mixer qubits
    c2d P[16], L[16];

    // Instance index of register r in layout LAY: the tile mapping deposits bits, so it is the
    // thread's part OR the register index's part (four uniform bit deposits).  Each group of
    // loads or stores derives its 16 indices from its own opaque copy of the thread's part, so they
    // are recomputed (one OR each) instead of being held in registers across the kernel.
    struct Idx {
        uint32_t t, d0, d1, d2, d3;
    };
    auto make_idx = [&](auto layc) __attribute__((always_inline)) {
        constexpr int LAY = decltype(layc)::value;
        constexpr int base = LAY == 0 ? 8 : LAY == 1 ? 0 : 4;
        uint32_t t;
        asm volatile("v_mov_b32 %0, %1" : "=v"(t) : "v"(tile_global(I, half, tile, qthread<LAY>(lane, wave))));
        return Idx{t, tile_global(I, half, 0, 1u << base), tile_global(I, half, 0, 2u << base),
                   tile_global(I, half, 0, 4u << base), tile_global(I, half, 0, 8u << base)};
    };
    auto at = [](const Idx &x, int r) __attribute__((always_inline)) {
        return x.t | ((r & 1) ? x.d0 : 0) | ((r & 2) ? x.d1 : 0) | ((r & 4) ? x.d2 : 0) | ((r & 8) ? x.d3 : 0);
    };
    // amplitude at an instance index, as a 32-bit byte offset from the state's uniform base (below
    // 2^26), so loads and stores take one offset register each (scalar base + vector offset)
    auto amp = [](c2d *base, uint32_t g) __attribute__((always_inline)) -> c2d & {
        return *reinterpret_cast<c2d *>(reinterpret_cast<unsigned char *>(base) + (g << 4));
    };
    using L0 = std::integral_constant<int, 0>;
    using L2 = std::integral_constant<int, 2>;

    // load (the first forward pass starts from H^n |0>: every amplitude (1/sqrt2)^n)
    double red[4] = {0, 0, 0, 0};  // energy, post-layer overlap, pre-layer overlap, cost overlap
    if (!REV && (flags & Q_INIT)) {
        double a = 1.0;
        for (int q = 0; q < n; q++) a *= 0.70710678118654752440;
#pragma unroll
        for (int r = 0; r < 16; r++) P[r] = {a, 0.0};
    } else {
        const Idx ix = make_idx(L0{});
#pragma unroll
        for (int r = 0; r < 16; r++) P[r] = amp(psi, at(ix, r));
    }
    if constexpr (REV) {
        const Idx ix = make_idx(L0{});
        if (flags & Q_ENERGY) {  // lambda = H psi (H = -cut) and the energy
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const double h = -(double)cut[at(ix, r)];
                red[0] += h * (P[r].x * P[r].x + P[r].y * P[r].y);
                L[r] = {h * P[r].x, h * P[r].y};
            }
        } else {
#pragma unroll
            for (int r = 0; r < 16; r++) L[r] = amp(lam, at(ix, r));
        }
    }

    // One layer's mixer half over the register groups: layout 0 holds local bits 8-11, layout 1
    // bits 0-3, layout 2 bits 4-7.  b0 <= 3: all three groups; b0 in 4..7: layouts 0 and 2;
    // otherwise layout 0 alone.  mix_out starts in layout 0 and ends in layout 2 (0 when b0 >= 8),
    // mix_back runs the groups the other way round, back to layout 0 (coalesced stores).  REV: the
    // overlaps and the uncompute of both states.
    auto mix_out = [&](int k, double &wy) __attribute__((always_inline)) {
        double cb, sb;
        beta_shear(I, tabs, k, cb, sb);
        mix_group<REV, 0>(P, L, b0, cb, sb, wy);
        if (b0 <= 3) {
            relayout_states<REV, 0, 1>(P, L, buf, lane, wave);
            mix_group<REV, 1>(P, L, b0, cb, sb, wy);
            relayout_states<REV, 1, 2>(P, L, buf, lane, wave);
            mix_group<REV, 2>(P, L, b0, cb, sb, wy);
        } else if (b0 <= 7) {
            relayout_states<REV, 0, 2>(P, L, buf, lane, wave);
            mix_group<REV, 2>(P, L, b0, cb, sb, wy);
        }
    };
    auto mix_back = [&](int k, double &wy) __attribute__((always_inline)) {
        double cb, sb;
        beta_shear(I, tabs, k, cb, sb);
        if (b0 <= 3) {
            mix_group<REV, 2>(P, L, b0, cb, sb, wy);
            relayout_states<REV, 2, 1>(P, L, buf, lane, wave);
            mix_group<REV, 1>(P, L, b0, cb, sb, wy);
            relayout_states<REV, 1, 0>(P, L, buf, lane, wave);
        } else if (b0 <= 7) {
            mix_group<REV, 2>(P, L, b0, cb, sb, wy);
            relayout_states<REV, 2, 0>(P, L, buf, lane, wave);
        }
        mix_group<REV, 0>(P, L, b0, cb, sb, wy);
    };
    // cost layer k in layout LAY: forward exp(-i gamma/2 S(z)) (the first pass computes and keeps
    // cut(z)); reverse the derivative Im <lambda| S |psi> and, unless nothing is stored, the
    // uncompute of both states
    auto cost = [&](auto layc, int k) __attribute__((always_inline)) {
        load_tab(tab, I, tabs, k);
        const Idx ix = make_idx(layc);
        if constexpr (!REV) {
            int c[16];
            if (flags & Q_INIT) {
                // cut(z) of the 16 registers: register 0 in full, every other from the register
                // without its lowest set bit j, by flipping vertex 8 + j (this pass runs on A tiles
                // in layout 0, whose register bits are qubits 8..11):
                //   cut(z ^ e_v) = cut(z) + deg(v) - 2 |{u adjacent to v : z_u != z_v}|
                c[0] = cut_of(sup, n, at(ix, 0));
#pragma unroll
                for (int r = 1; r < 16; r++) {
                    const int j = __builtin_ctz(r);
                    const uint32_t z = at(ix, r ^ (1 << j));
                    const uint32_t zv = (z >> (8 + j)) & 1u;
                    const int diff = __popc(adj[j] & (zv ? ~z : z));
                    c[r] = c[r ^ (1 << j)] + deg[j] - 2 * diff;
                }
#pragma unroll
                for (int r = 0; r < 16; r++) cut[at(ix, r)] = (uint8_t)c[r];
            } else {
#pragma unroll
                for (int r = 0; r < 16; r++) c[r] = cut[at(ix, r)];
            }
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const c2d t = tab[c[r]];
                cmul_in<false>(P[r], t.x, t.y);
            }
        } else {
            // The cut bytes of the next four registers are fetched while the current four are
            // processed: loaded between the in-place register updates (inline asm, which memory
            // operations do not cross) each load's latency stood alone, and fetching all 16 up front
            // costs the kernel its third wave per SIMD.
            const bool undo = !(flags & Q_NOSTORE);
            int cur[4], nxt[4];
#pragma unroll
            for (int i = 0; i < 4; i++) nxt[i] = cut[at(ix, i)];
#pragma unroll
            for (int g = 0; g < 4; g++) {
#pragma unroll
                for (int i = 0; i < 4; i++) cur[i] = nxt[i];
                if (g < 3) {
#pragma unroll
                    for (int i = 0; i < 4; i++) nxt[i] = cut[at(ix, 4 * (g + 1) + i)];
                }
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int r = 4 * g + i;
                    cost_overlap(red[3], (double)(ne - 2 * cur[i]), L[r], P[r]);
                    if (undo) {
                        const c2d t = tab[cur[i]];
                        cmul_in<true>(P[r], t.x, t.y);
                        cmul_in<true>(L[r], t.x, t.y);
                    }
                }
                // table reads four registers at a time (hoisted, all 16 would hold 64 VGPRs)
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        __syncthreads();  // the table and masks share LDS with the next relayout
    };

    // forward: mixer half `pre`, cost `cost`, mixer half `post`; reverse: the same inverted, in the
    // opposite order
    const int kfirst = REV ? kpost : kpre, klast = REV ? kpre : kpost;
    bool in2 = false;  // the registers are in layout 2
    if (kfirst >= 0) {
        mix_out(kfirst, red[1]);
        in2 = b0 <= 7;
    }
    if (kcost >= 0) {
        if (in2) cost(L2{}, kcost);
        else cost(L0{}, kcost);
    }
    if (klast >= 0) {
        if (in2) {
            mix_back(klast, red[2]);
            in2 = false;
        } else {
            mix_out(klast, red[2]);
            in2 = b0 <= 7;
        }
    }
    const bool store = !(REV && (flags & Q_NOSTORE));
    if (store) {
        if (in2) relayout_states<REV, 2, 0>(P, L, buf, lane, wave);
        const Idx ix = make_idx(L0{});
        if constexpr (REV) {
#pragma unroll
            for (int r = 0; r < 16; r++) amp(lam, at(ix, r)) = L[r];
        }
#pragma unroll
        for (int r = 0; r < 16; r++) amp(psi, at(ix, r)) = P[r];
    }
    if constexpr (REV) {
        block_sum4(red, sh);
        if (threadIdx.x == 0) {
            if (flags & Q_ENERGY) part[0] = red[0];
            if (kpost >= 0) part[1 + 3 * kpost + 1 + half] = red[1];
            if (kpre >= 0) part[1 + 3 * kpre + 1 + half] = red[2];
            if (kcost >= 0) part[1 + 3 * kcost] = red[3];
        }
    }
}

// energies / gradients of the large instances from their per-tile partials: one wave per instance,
// lane l summing tiles l, l + 64, ... and a fixed butterfly across the lanes (the same order every run)
__global__ void __launch_bounds__(64) k_qaoa_finish(const QInst *__restrict__ insts, const int32_t *__restrict__ ids,
                                                    const double *__restrict__ parts, double *__restrict__ energies,
                                                    const int64_t *__restrict__ goff, double *__restrict__ grads) {
    const int id = ids[blockIdx.x];
    const QInst &I = insts[id];
    const double *pt = parts + I.part;
    const int ns = I.slot_stride;  // 1 + 3p: energy, then (gamma, beta A half, beta B half) per layer
    auto total = [&](int j) {
        double v = 0;
        for (int t = threadIdx.x; t < I.ntiles; t += 64) v += pt[(int64_t)t * ns + j];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
        return v;
    };
    const double e = total(0);
    if (threadIdx.x == 0) energies[id] = e;
    for (int k = 0; k < I.p; k++) {
        const double gg = total(1 + 3 * k), ga = total(2 + 3 * k), gb = total(3 + 3 * k);
        if (threadIdx.x == 0) {
            grads[goff[id] + 2 * k] = gg;
            grads[goff[id] + 2 * k + 1] = 2.0 * (ga + gb);
        }
    }
}

}  // namespace goqu

using namespace goqu;

extern "C" int goqu_qaoa_batch(int32_t num_instances, const int32_t *num_qubits, const int32_t *layers,
                               const int64_t *edge_offsets, const int32_t *edges, const int64_t *param_offsets,
                               const double *params, double *energies, double *grads) {
    GOQU_GUARD_BEGIN
    if (num_instances < 0 || (num_instances > 0 && (!num_qubits || !layers || !edge_offsets || !edges ||
                                                     !param_offsets || !params || !energies || !grads)))
        return fail(GOQU_ERR_INVALID_ARG, "bad argument");
    if (num_instances == 0) return GOQU_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(GOQU_ERR_NO_DEVICE, "no HIP device visible");
    // host-side instance table
    std::vector<QInst> insts(num_instances);
    std::vector<int32_t> small_ids, large_ids;
    uint64_t pool_elems = 0, cut_bytes = 0;
    int64_t parts = 0, tab_elems = 0;
    for (int i = 0; i < num_instances; i++) {
        QInst &I = insts[i];
        std::memset(&I, 0, sizeof(I));
        I.n = num_qubits[i];
        I.p = layers[i];
        if (I.n < 2 || I.n > kQMaxQubits || I.p < 1 || I.p > kQMaxLayers)
            return fail(GOQU_ERR_INVALID_ARG, "instance " + std::to_string(i) + ": qubits must be in [2, 22], layers in [1, 16]");
        if (param_offsets[i + 1] - param_offsets[i] != 2 * I.p)
            return fail(GOQU_ERR_INVALID_ARG, "instance " + std::to_string(i) + ": needs 2p parameters");
        for (int k = 0; k < I.p; k++) {
            I.gamma[k] = params[param_offsets[i] + 2 * k];
            I.beta[k] = params[param_offsets[i] + 2 * k + 1];
        }
        const int64_t e0 = edge_offsets[i], e1 = edge_offsets[i + 1];
        if (e1 <= e0 || e1 - e0 > kQMaxEdges)
            return fail(GOQU_ERR_INVALID_ARG, "instance " + std::to_string(i) + ": edge count out of range");
        I.nedges = (int32_t)(e1 - e0);
        for (int64_t e = e0; e < e1; e++) {
            int a = edges[2 * e], b = edges[2 * e + 1];
            if (a < 0 || b < 0 || a >= I.n || b >= I.n || a == b)
                return fail(GOQU_ERR_INVALID_ARG, "instance " + std::to_string(i) + ": bad edge");
            if (a > b) std::swap(a, b);
            if (I.up[a] >> b & 1)
                return fail(GOQU_ERR_INVALID_ARG, "instance " + std::to_string(i) + ": repeated edge");
            I.up[a] |= 1u << b;
        }
        I.tab = tab_elems;
        tab_elems += (int64_t)I.p * (I.nedges + kQTabExtra);
        I.small = I.n <= kQTileBits;
        if (I.small) {
            small_ids.push_back(i);
        } else {
            I.ntiles = 1 << (I.n - kQTileBits);
            I.slot_stride = 1 + 3 * I.p;
            I.psi = pool_elems;
            pool_elems += 1ull << I.n;
            I.lam = pool_elems;
            pool_elems += 1ull << I.n;
            I.cut = cut_bytes;
            cut_bytes += 1ull << I.n;
            I.part = parts;
            parts += (int64_t)I.ntiles * I.slot_stride;
            large_ids.push_back(i);
        }
    }
    std::vector<int64_t> goff(num_instances + 1);
    for (int i = 0; i <= num_instances; i++) goff[i] = param_offsets[i] - param_offsets[0];
    // pass schedule of the large instances: forward passes f = 0..p, then the same passes inverted
    // from f = p down to 0; every instance advances one pass per step
    int max_steps = 0;
    for (int id : large_ids) max_steps = std::max(max_steps, 2 * (insts[id].p + 1));
    // launches[2 s]: the forward work of step s, launches[2 s + 1]: its reverse work
    std::vector<std::vector<QWork>> launches(2 * max_steps);
    for (int s = 0; s < max_steps; s++) {
        int blocks[2] = {0, 0};
        for (int id : large_ids) {
            const int p = insts[id].p;
            if (s >= 2 * (p + 1)) continue;
            const int rv = s > p;
            const int f = rv ? 2 * p + 1 - s : s;  // the forward pass this step runs (inverted when rv)
            QWork w;
            w.inst = id;
            w.half = (int8_t)(f & 1);
            w.pre = (int8_t)(f >= 1 ? f - 1 : -1);
            w.cost = (int8_t)(f < p ? f : -1);
            w.post = (int8_t)(f < p ? f : -1);
            w.flags = (!rv && f == 0 ? Q_INIT : 0) | (rv && f == p ? Q_ENERGY : 0) | (rv && f == 0 ? Q_NOSTORE : 0);
            w.first_block = blocks[rv];
            blocks[rv] += insts[id].ntiles;
            launches[2 * s + rv].push_back(w);
        }
    }
    // Every device buffer of the call is one carve of a single retained allocation: the instance
    // table, id lists, parameter offsets and work items (uploaded with one copy), then the results,
    // the layer tables and, for the large instances, the pool (psi, lambda), the per-tile partials
    // and the cut bytes.
    const size_t ng = (size_t)(goff[num_instances]);
    size_t wtotal = 0;
    std::vector<size_t> woff(launches.size());
    for (size_t s = 0; s < launches.size(); s++) {
        woff[s] = wtotal;
        wtotal += launches[s].size();
    }
    size_t carved = 0;
    auto carve = [&](size_t bytes) {
        const size_t o = carved;
        carved += (std::max<size_t>(bytes, 1) + 255) & ~size_t(255);
        return o;
    };
    const size_t o_inst = carve(insts.size() * sizeof(QInst));
    const size_t o_small = carve(small_ids.size() * 4);
    const size_t o_large = carve(large_ids.size() * 4);
    const size_t o_goff = carve(goff.size() * 8);
    const size_t o_work = carve(wtotal * sizeof(QWork));
    const size_t up_bytes = carved;
    const size_t o_e = carve((size_t)num_instances * 8);
    const size_t o_g = carve(ng * 8);
    const size_t o_tabs = carve((size_t)tab_elems * sizeof(c2d));
    const size_t o_pool = carve((size_t)pool_elems * sizeof(c2d));
    const size_t o_parts = carve((size_t)parts * 8);
    const size_t o_cut = carve((size_t)cut_bytes);
    CachedDevBuf dbuf(CACHE_QAOA);
    if (dbuf.take(carved) != hipSuccess) {
        (void)hipGetLastError();
        return fail(GOQU_ERR_OUT_OF_MEMORY, "qaoa batch: state pool of " + std::to_string(pool_elems * 16) +
                                                " bytes does not fit; split the batch");
    }
    unsigned char *const base = static_cast<unsigned char *>(dbuf.p);
    const QInst *d_inst = reinterpret_cast<const QInst *>(base + o_inst);
    const int32_t *d_small = reinterpret_cast<const int32_t *>(base + o_small);
    const int32_t *d_large = reinterpret_cast<const int32_t *>(base + o_large);
    const int64_t *d_goff = reinterpret_cast<const int64_t *>(base + o_goff);
    const QWork *d_work = reinterpret_cast<const QWork *>(base + o_work);
    double *d_e = reinterpret_cast<double *>(base + o_e);
    double *d_g = reinterpret_cast<double *>(base + o_g);
    c2d *d_tabs = reinterpret_cast<c2d *>(base + o_tabs);
    c2d *pool = reinterpret_cast<c2d *>(base + o_pool);
    double *partials = reinterpret_cast<double *>(base + o_parts);
    uint8_t *cutb = base + o_cut;
    std::vector<unsigned char> up(up_bytes);
    std::memcpy(up.data() + o_inst, insts.data(), insts.size() * sizeof(QInst));
    std::memcpy(up.data() + o_small, small_ids.data(), small_ids.size() * 4);
    std::memcpy(up.data() + o_large, large_ids.data(), large_ids.size() * 4);
    std::memcpy(up.data() + o_goff, goff.data(), goff.size() * 8);
    for (size_t s = 0; s < launches.size(); s++)
        if (!launches[s].empty())
            std::memcpy(up.data() + o_work + woff[s] * sizeof(QWork), launches[s].data(),
                        launches[s].size() * sizeof(QWork));
    hipStream_t s;
    GOQU_HIP(hipStreamCreate(&s));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{s};
    GOQU_HIP(hipMemcpyAsync(base, up.data(), up_bytes, hipMemcpyHostToDevice, s));
    GOQU_HIP(hipStreamSynchronize(s));

    size_t lds_small = 0;
    for (int id : small_ids)
        lds_small = std::max<size_t>(lds_small, (size_t(2) << insts[id].n) * sizeof(c2d) + (kQMaxEdges + 1) * sizeof(c2d));
    // dynamic LDS above the 64 KiB default (static LDS of the kernel comes on top)
    if (lds_small)
        GOQU_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_qaoa_small),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_small));
    // capture every launch in one graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    GOQU_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const c2d *tabs = d_tabs;
    hipLaunchKernelGGL(k_qaoa_tables, dim3((unsigned)num_instances), dim3(128), 0, s, d_inst, d_tabs);
    if (!small_ids.empty())
        hipLaunchKernelGGL(k_qaoa_small, dim3((unsigned)small_ids.size()), dim3(kQThreads), lds_small, s,
                           d_inst, tabs, d_small, d_e, d_goff, d_g);
    for (size_t ph = 0; ph < launches.size(); ph++) {
        const auto &v = launches[ph];
        if (v.empty()) continue;
        const unsigned blocks = (unsigned)(v.back().first_block + insts[v.back().inst].ntiles);
        if (ph % 2)
            hipLaunchKernelGGL(k_qaoa_pass<true>, dim3(blocks), dim3(kQThreads), 0, s,
                               d_inst, tabs, d_work + woff[ph],
                               (int)v.size(), pool, cutb, partials);
        else
            hipLaunchKernelGGL(k_qaoa_pass<false>, dim3(blocks), dim3(kQThreads), 0, s,
                               d_inst, tabs, d_work + woff[ph],
                               (int)v.size(), pool, cutb, partials);
    }
    if (!large_ids.empty())
        hipLaunchKernelGGL(k_qaoa_finish, dim3((unsigned)large_ids.size()), dim3(64), 0, s, d_inst, d_large, partials,
                           d_e, d_goff, d_g);
    hipError_t ce = hipStreamEndCapture(s, &graph);
    if (ce != hipSuccess) return fail(GOQU_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(ce));
    ce = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    hipEvent_t ev0 = prof_begin(s);
    if (ce == hipSuccess) ce = hipGraphLaunch(exec, s);
    prof_end(ev0, s, PROF_QAOA, 0.0);
    if (ce == hipSuccess) ce = hipStreamSynchronize(s);
    if (ce == hipSuccess) dbuf.idle();
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (ce != hipSuccess) return fail(GOQU_ERR_DEVICE, std::string("qaoa graph: ") + hipGetErrorString(ce));
    GOQU_HIP(hipMemcpy(energies, d_e, num_instances * 8, hipMemcpyDeviceToHost));
    if (ng) GOQU_HIP(hipMemcpy(grads + param_offsets[0], d_g, ng * 8, hipMemcpyDeviceToHost));
    return GOQU_OK;
    GOQU_GUARD_END
}
