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
// instances are swept in tiles of 4096 amplitudes: phase A holds qubits 0..11, phase B the
// qubits 0..L-1 and 12..n-1; the phases of all instances advance together, one launch per
// phase, and the whole sequence is captured in a single hipGraph.  A tile lives in registers, 16
// amplitudes per thread: the mixer runs on register bits, and two or three LDS relayouts bring
// the other tile bits into registers.  cut(z) is computed once per instance (first cost layer)
// and kept as one byte per amplitude; the energy and lambda = H psi are formed when the reverse
// sweep first loads psi.
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

// phase codes for large instances
enum : int32_t { PH_FWD_A = 0, PH_FWD_B = 1, PH_REV_B = 2, PH_REV_A = 3 };

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

// phase table: tab[c] = exp(-i gamma/2 (|E| - 2c)), c = 0..|E|
// The layer tables of every instance, computed once per batch: for layer k at I.tab + k (ne + 2),
// the cost phases exp(-i gamma_k / 2 (ne - 2c)) for c = 0..ne, then (cos beta_k, sin beta_k).  The
// tile kernels copy a layer's phases into LDS instead of evaluating ne + 2 sines and cosines per
// workgroup.
__global__ void k_qaoa_tables(const QInst *__restrict__ insts, c2d *__restrict__ tabs) {
    const QInst &I = insts[blockIdx.x];
    const int ne = I.nedges, T = ne + 2;
    for (int e = threadIdx.x; e < I.p * T; e += blockDim.x) {
        const int k = e / T, c = e % T;
        double sn, cs;
        if (c <= ne) sincos(-0.5 * I.gamma[k] * (double)(ne - 2 * c), &sn, &cs);
        else sincos(I.beta[k], &sn, &cs);
        tabs[I.tab + e] = {cs, sn};
    }
}
// layer k's cost phases into LDS
__device__ __forceinline__ void load_tab(c2d *tab, const QInst &I, const c2d *__restrict__ tabs, int k) {
    const c2d *src = tabs + I.tab + (int64_t)k * (I.nedges + 2);
    for (int c = threadIdx.x; c <= I.nedges; c += blockDim.x) tab[c] = src[c];
    __syncthreads();
}
// (cos beta_k, sin beta_k) of layer k
__device__ __forceinline__ void beta_cs(const QInst &I, const c2d *__restrict__ tabs, int k, double &cb, double &sb) {
    const c2d v = tabs[I.tab + (int64_t)k * (I.nedges + 2) + I.nedges + 1];
    cb = v.x;
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

// ---- large instances: one tile of 4096 amplitudes per workgroup, one launch per phase ----------
struct QWork {
    int32_t inst;
    int32_t phase;  // PH_*
    int32_t layer;
    int32_t first_block;  // prefix over the launch
};

__device__ __forceinline__ uint64_t tile_global(const QInst &I, int phase, uint64_t tile, uint32_t l) {
    if (phase == PH_FWD_A || phase == PH_REV_A) return (tile << kQTileBits) | l;
    // B: local bits 0..L-1 -> qubits 0..L-1, local bits L..11 -> qubits 12..n-1; the tile index
    // supplies qubits L..11
    const int Lb = 2 * kQTileBits - I.n;
    const uint64_t lo = l & ((1u << Lb) - 1), hi = l >> Lb;
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

// RX on register bit J of a 16-amplitude tile: [a0,a1] <- [c a0 - i s a1, -i s a0 + c a1]
template <int J>
__device__ __forceinline__ void rx_reg(c2d (&A)[16], double c, double s) {
#pragma unroll
    for (int r = 0; r < 16; r++) {
        if (r & (1 << J)) continue;
        const c2d a0 = A[r], a1 = A[r | (1 << J)];
        A[r] = {c * a0.x + s * a1.y, c * a0.y - s * a1.x};
        A[r | (1 << J)] = {c * a1.x + s * a0.y, c * a1.y - s * a0.x};
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
        w += l0.x * p1.y - l0.y * p1.x + l1.x * p0.y - l1.y * p0.x;
    }
    return w;
}

// moves a register tile from layout FROM to layout TO through LDS (real parts, then imaginary)
template <int FROM, int TO>
__device__ __forceinline__ void relayout(c2d (&A)[16], double *buf, uint32_t lane, uint32_t wave) {
    const uint32_t wt = qswz(qthread<FROM>(lane, wave)), rt = qswz(qthread<TO>(lane, wave));
#pragma unroll
    for (int part = 0; part < 2; part++) {
#pragma unroll
        for (int r = 0; r < 16; r++) buf[wt ^ qswz(qreg<FROM>(r))] = part ? A[r].y : A[r].x;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const double v = buf[rt ^ qswz(qreg<TO>(r))];
            if (part)
                A[r].y = v;
            else
                A[r].x = v;
        }
        __syncthreads();
    }
}

// the mixer on the register bits of layout LAY that carry local bits >= b0 (forward), or the
// overlaps and the uncompute of both states (reverse)
template <bool REV, int LAY>
__device__ __forceinline__ void mix_group(c2d (&P)[16], c2d (&L)[16], int b0, double cb, double sb, double &wy) {
    constexpr int base = LAY == 0 ? 8 : LAY == 1 ? 0 : 4;
    auto bit = [&](auto jc) __attribute__((always_inline)) {
        constexpr int J = decltype(jc)::value;
        if (base + J < b0) return;
        if constexpr (REV) {
            wy += xov_reg<J>(P, L);
            rx_reg<J>(P, cb, -sb);
            rx_reg<J>(L, cb, -sb);
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

// REV: the launch holds only reverse phases (psi and lambda), otherwise only forward phases (psi):
// the forward kernel keeps half the registers and runs more waves.
template <bool REV>
__global__ void __launch_bounds__(kQThreads) k_qaoa_phase(const QInst *__restrict__ insts,
                                                          const c2d *__restrict__ tabs,
                                                          const QWork *__restrict__ work, int nwork,
                                                          c2d *__restrict__ pool, uint8_t *__restrict__ cuts,
                                                          double *__restrict__ parts) {
    __shared__ double buf[1 << kQTileBits];
    __shared__ c2d tab[kQMaxEdges + 1];
    __shared__ double sh[16];
    __shared__ uint32_t sup[kQMaxQubits];
    // which work item: last with first_block <= blockIdx.x
    int lo = 0, hi = nwork;
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (work[mid].first_block <= (int)blockIdx.x) lo = mid;
        else hi = mid;
    }
    const QWork W = work[lo];
    const QInst &I = insts[W.inst];
    const uint64_t tile = blockIdx.x - W.first_block;
    const int n = I.n, ne = I.nedges, k = W.layer, ph = W.phase;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    c2d *psi = pool + I.psi, *lam = pool + I.lam;
    uint8_t *cut = cuts + I.cut;
    double *part = parts + I.part + (int64_t)tile * I.slot_stride;
    stage_up(sup, I);
    constexpr bool rev = REV;
    const bool isA = ph == PH_FWD_A || ph == PH_REV_A;
    const int Lb = 2 * kQTileBits - n;
    const int b0 = isA ? 0 : Lb;  // local bits b0..11 carry this phase's mixer qubits
    c2d P[16], L[16];
    // global index of register r in layout 0: the tile mapping deposits bits, so it splits into
    // the thread's part and the register index's part
    const uint64_t g0 = tile_global(I, ph, tile, qthread<0>(lane, wave));
    const uint64_t d0 = tile_global(I, ph, 0, 1u << 8), d1 = tile_global(I, ph, 0, 2u << 8);
    const uint64_t d2 = tile_global(I, ph, 0, 4u << 8), d3 = tile_global(I, ph, 0, 8u << 8);
    auto gidx = [&](int r) __attribute__((always_inline)) {
        return g0 | ((r & 1) ? d0 : 0) | ((r & 2) ? d1 : 0) | ((r & 4) ? d2 : 0) | ((r & 8) ? d3 : 0);
    };
    // load (layer 0 starts from H^n |0>: every amplitude (1/sqrt2)^n)
    if (ph == PH_FWD_A && k == 0) {
        double a = 1.0;
        for (int q = 0; q < n; q++) a *= 0.70710678118654752440;
#pragma unroll
        for (int r = 0; r < 16; r++) P[r] = {a, 0.0};
    } else {
#pragma unroll
        for (int r = 0; r < 16; r++) P[r] = psi[gidx(r)];
    }
    if constexpr (REV) {
        if (ph == PH_REV_B && k == I.p - 1) {  // lambda = H psi (H = -cut) and the energy
            double e = 0;
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const double h = -(double)cut[gidx(r)];
                e += h * (P[r].x * P[r].x + P[r].y * P[r].y);
                L[r] = {h * P[r].x, h * P[r].y};
            }
            e = block_sum1(e, sh);
            if (threadIdx.x == 0) part[0] = e;
        } else {
#pragma unroll
            for (int r = 0; r < 16; r++) L[r] = lam[gidx(r)];
        }
    }
    double sb, cb;
    beta_cs(I, tabs, k, cb, sb);
    if (!rev && isA) {  // cost layer: exp(-i gamma/2 S(z)); layer 0 computes and keeps cut(z)
        load_tab(tab, I, tabs, k);
#pragma unroll
        for (int r = 0; r < 16; r++) {
            int c;
            if (k == 0) {
                c = cut_of(sup, n, (uint32_t)gidx(r));
                cut[gidx(r)] = (uint8_t)c;
            } else {
                c = cut[gidx(r)];
            }
            P[r] = cmul(P[r], tab[c]);
        }
    }
    // mixer bits, register group by register group: layout 0 holds local bits 8-11, layout 1
    // bits 0-3, layout 2 bits 4-7.  b0 <= 3: all three groups; b0 in 4..7: groups of layouts 0
    // and 2; otherwise layout 0 alone.  Each path ends back in layout 0 (coalesced stores).
    double wy = 0;
    mix_group<REV, 0>(P, L, b0, cb, sb, wy);
    if (b0 <= 3) {
        relayout_states<REV, 0, 1>(P, L, buf, lane, wave);
        mix_group<REV, 1>(P, L, b0, cb, sb, wy);
        relayout_states<REV, 1, 2>(P, L, buf, lane, wave);
        mix_group<REV, 2>(P, L, b0, cb, sb, wy);
        relayout_states<REV, 2, 0>(P, L, buf, lane, wave);
    } else if (b0 <= 7) {
        relayout_states<REV, 0, 2>(P, L, buf, lane, wave);
        mix_group<REV, 2>(P, L, b0, cb, sb, wy);
        relayout_states<REV, 2, 0>(P, L, buf, lane, wave);
    }
    if constexpr (REV) {
        wy = block_sum1(wy, sh);
        if (threadIdx.x == 0) part[1 + 3 * k + (isA ? 1 : 2)] = wy;
        if (isA) {  // cost derivative and uncompute (all mixer qubits of the layer are undone now)
            load_tab(tab, I, tabs, k);
            double gy = 0;
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int c = cut[gidx(r)];
                const double S = (double)(ne - 2 * c);
                const c2d pz = P[r], lz = L[r];
                gy += S * (lz.x * pz.y - lz.y * pz.x);  // Im(conj(l) S p)
                const c2d t = {tab[c].x, -tab[c].y};
                P[r] = cmul(pz, t);
                L[r] = cmul(lz, t);
            }
            gy = block_sum1(gy, sh);
            if (threadIdx.x == 0) part[1 + 3 * k] = gy;
        }
#pragma unroll
        for (int r = 0; r < 16; r++) lam[gidx(r)] = L[r];
    }
#pragma unroll
    for (int r = 0; r < 16; r++) psi[gidx(r)] = P[r];
}

// energies / gradients of the large instances from their per-tile partials (fixed order)
__global__ void k_qaoa_finish(const QInst *__restrict__ insts, const int32_t *__restrict__ ids, int nids,
                              const double *__restrict__ parts, double *__restrict__ energies,
                              const int64_t *__restrict__ goff, double *__restrict__ grads) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nids) return;
    const int id = ids[t];
    const QInst &I = insts[id];
    const double *pt = parts + I.part;
    double e = 0;
    for (int tl = 0; tl < I.ntiles; tl++) e += pt[(int64_t)tl * I.slot_stride];
    energies[id] = e;
    for (int k = 0; k < I.p; k++) {
        double gg = 0, ga = 0, gb = 0;
        for (int tl = 0; tl < I.ntiles; tl++) {
            const double *s = pt + (int64_t)tl * I.slot_stride + 1 + 3 * k;
            gg += s[0];
            ga += s[1];
            gb += s[2];
        }
        grads[goff[id] + 2 * k] = gg;
        grads[goff[id] + 2 * k + 1] = 2.0 * (ga + gb);
    }
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

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
        tab_elems += (int64_t)I.p * (I.nedges + 2);
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
    // phase schedule of the large instances: forward A,B per layer, then reverse B,A per layer
    int max_phases = 0;
    for (int id : large_ids) max_phases = std::max(max_phases, 4 * insts[id].p);
    // launches[2 s]: the forward work of step s, launches[2 s + 1]: its reverse work
    std::vector<std::vector<QWork>> launches(2 * max_phases);
    for (int s = 0; s < max_phases; s++) {
        int blocks[2] = {0, 0};
        for (int id : large_ids) {
            const int p = insts[id].p;
            if (s >= 4 * p) continue;
            QWork w;
            w.inst = id;
            int rv = 0;
            if (s < 2 * p) {
                w.layer = s / 2;
                w.phase = (s % 2) ? PH_FWD_B : PH_FWD_A;
            } else {
                const int r = s - 2 * p;
                w.layer = p - 1 - r / 2;
                w.phase = (r % 2) ? PH_REV_A : PH_REV_B;
                rv = 1;
            }
            w.first_block = blocks[rv];
            blocks[rv] += insts[id].ntiles;
            launches[2 * s + rv].push_back(w);
        }
    }
    // device buffers
    DevBuf d_inst, d_small, d_large, d_goff, d_e, d_g, d_pool, d_parts, d_work, d_cut, d_tabs;
    const size_t ng = (size_t)(goff[num_instances]);
    size_t wtotal = 0;
    std::vector<size_t> woff(launches.size());
    for (size_t s = 0; s < launches.size(); s++) {
        woff[s] = wtotal;
        wtotal += launches[s].size();
    }
    GOQU_HIP(dmalloc((void **)&d_inst.p, insts.size() * sizeof(QInst)));
    GOQU_HIP(dmalloc((void **)&d_small.p, std::max<size_t>(1, small_ids.size()) * 4));
    GOQU_HIP(dmalloc((void **)&d_large.p, std::max<size_t>(1, large_ids.size()) * 4));
    GOQU_HIP(dmalloc((void **)&d_goff.p, goff.size() * 8));
    GOQU_HIP(dmalloc((void **)&d_e.p, num_instances * 8));
    GOQU_HIP(dmalloc((void **)&d_g.p, std::max<size_t>(1, ng) * 8));
    GOQU_HIP(dmalloc((void **)&d_work.p, std::max<size_t>(1, wtotal) * sizeof(QWork)));
    GOQU_HIP(dmalloc((void **)&d_tabs.p, (size_t)std::max<int64_t>(1, tab_elems) * sizeof(c2d)));
    if (pool_elems) {
        hipError_t e = dmalloc((void **)&d_pool.p, pool_elems * sizeof(c2d));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(GOQU_ERR_OUT_OF_MEMORY, "qaoa batch: state pool of " + std::to_string(pool_elems * 16) +
                                                    " bytes does not fit; split the batch");
        }
        GOQU_HIP(dmalloc((void **)&d_parts.p, std::max<int64_t>(1, parts) * 8));
        GOQU_HIP(dmalloc((void **)&d_cut.p, cut_bytes));
    }
    hipStream_t s;
    GOQU_HIP(hipStreamCreate(&s));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{s};
    GOQU_HIP(hipMemcpyAsync(d_inst.p, insts.data(), insts.size() * sizeof(QInst), hipMemcpyHostToDevice, s));
    if (!small_ids.empty())
        GOQU_HIP(hipMemcpyAsync(d_small.p, small_ids.data(), small_ids.size() * 4, hipMemcpyHostToDevice, s));
    if (!large_ids.empty())
        GOQU_HIP(hipMemcpyAsync(d_large.p, large_ids.data(), large_ids.size() * 4, hipMemcpyHostToDevice, s));
    GOQU_HIP(hipMemcpyAsync(d_goff.p, goff.data(), goff.size() * 8, hipMemcpyHostToDevice, s));
    std::vector<QWork> wflat;
    for (auto &v : launches) wflat.insert(wflat.end(), v.begin(), v.end());
    if (!wflat.empty())
        GOQU_HIP(hipMemcpyAsync(d_work.p, wflat.data(), wflat.size() * sizeof(QWork), hipMemcpyHostToDevice, s));
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
    const c2d *tabs = static_cast<const c2d *>(d_tabs.p);
    hipLaunchKernelGGL(k_qaoa_tables, dim3((unsigned)num_instances), dim3(128), 0, s,
                       static_cast<const QInst *>(d_inst.p), static_cast<c2d *>(d_tabs.p));
    if (!small_ids.empty())
        hipLaunchKernelGGL(k_qaoa_small, dim3((unsigned)small_ids.size()), dim3(kQThreads), lds_small, s,
                           static_cast<const QInst *>(d_inst.p), tabs, static_cast<const int32_t *>(d_small.p),
                           static_cast<double *>(d_e.p), static_cast<const int64_t *>(d_goff.p),
                           static_cast<double *>(d_g.p));
    for (size_t ph = 0; ph < launches.size(); ph++) {
        const auto &v = launches[ph];
        if (v.empty()) continue;
        const unsigned blocks = (unsigned)(v.back().first_block + insts[v.back().inst].ntiles);
        if (ph % 2)
            hipLaunchKernelGGL(k_qaoa_phase<true>, dim3(blocks), dim3(kQThreads), 0, s,
                               static_cast<const QInst *>(d_inst.p), tabs, static_cast<const QWork *>(d_work.p) + woff[ph],
                               (int)v.size(), static_cast<c2d *>(d_pool.p), static_cast<uint8_t *>(d_cut.p),
                               static_cast<double *>(d_parts.p));
        else
            hipLaunchKernelGGL(k_qaoa_phase<false>, dim3(blocks), dim3(kQThreads), 0, s,
                               static_cast<const QInst *>(d_inst.p), tabs, static_cast<const QWork *>(d_work.p) + woff[ph],
                               (int)v.size(), static_cast<c2d *>(d_pool.p), static_cast<uint8_t *>(d_cut.p),
                               static_cast<double *>(d_parts.p));
    }
    if (!large_ids.empty())
        hipLaunchKernelGGL(k_qaoa_finish, dim3((unsigned)((large_ids.size() + 127) / 128)), dim3(128), 0, s,
                           static_cast<const QInst *>(d_inst.p), static_cast<const int32_t *>(d_large.p),
                           (int)large_ids.size(), static_cast<const double *>(d_parts.p),
                           static_cast<double *>(d_e.p), static_cast<const int64_t *>(d_goff.p),
                           static_cast<double *>(d_g.p));
    hipError_t ce = hipStreamEndCapture(s, &graph);
    if (ce != hipSuccess) return fail(GOQU_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(ce));
    ce = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    hipEvent_t ev0 = prof_begin(s);
    if (ce == hipSuccess) ce = hipGraphLaunch(exec, s);
    prof_end(ev0, s, PROF_QAOA, 0.0);
    if (ce == hipSuccess) ce = hipStreamSynchronize(s);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (ce != hipSuccess) return fail(GOQU_ERR_DEVICE, std::string("qaoa graph: ") + hipGetErrorString(ce));
    GOQU_HIP(hipMemcpy(energies, d_e.p, num_instances * 8, hipMemcpyDeviceToHost));
    if (ng) GOQU_HIP(hipMemcpy(grads + param_offsets[0], d_g.p, ng * 8, hipMemcpyDeviceToHost));
    return GOQU_OK;
    GOQU_GUARD_END
}
