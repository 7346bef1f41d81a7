// reduce.hip — FP64 reductions and small sweeps: Pauli expectations, inner products,
// measurement probabilities, collapse, inverse-CDF sampling, and density-matrix observables.
//
// All reductions are deterministic: each workgroup writes one partial per quantity, and a
// single-workgroup finisher adds the partials in a fixed order.
#include <cstdio>
#include <cstring>
#include <mutex>

#include "runtime.h"

namespace goqu {

constexpr int kRedThreads = 256;
constexpr int kRedMaxBlocks = 2048;

int Scratch::ensure(size_t need) {
    if (need <= bytes) return GOQU_OK;
    if (ptr) GOQU_HIP(hipFree(ptr));
    ptr = nullptr;
    bytes = 0;
    GOQU_HIP(dmalloc((void **)&ptr, need));
    bytes = need;
    return GOQU_OK;
}
Scratch::~Scratch() {
    if (ptr) (void)hipFree(ptr);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum `nv` values per thread across the workgroup; thread 0 gets the totals in out[].
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *sh /* >= 4*NV */) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = wave_sum(v[k]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) sh[wave * NV + k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double s = 0;
            for (int w = 0; w < nw; w++) s += sh[w * NV + k];
            v[k] = s;
        }
    }
    __syncthreads();
}

// partials[b*nv + k] -> out[k]: one workgroup per value k, each thread a strided share of the
// blocks, then a fixed-shape tree (deterministic)
__global__ void __launch_bounds__(256) k_finish(const double *__restrict__ partials, int nblocks, int nv,
                                                double *__restrict__ out) {
    __shared__ double sh[256];
    const int k = blockIdx.x;
    double s = 0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) s += partials[(size_t)b * nv + k];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = sh[0];
}

static unsigned red_grid(uint64_t work, unsigned per_block) {
    uint64_t g = (work + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > kRedMaxBlocks) g = kRedMaxBlocks;
    return (unsigned)g;
}

// ---- Pauli sums ------------------------------------------------------------------------------
// 16-point Walsh-Hadamard transform in registers: W[s] = sum_j v[j] (-1)^{|j&s|}
template <typename V>
__device__ __forceinline__ void wht16(V (&v)[16]) {
#pragma unroll
    for (int h = 1; h < 16; h <<= 1)
#pragma unroll
        for (int j = 0; j < 16; j++)
            if (!(j & h)) {
                V a = v[j], b = v[j | h];
                v[j] = a + b;
                v[j | h] = a - b;
            }
}

struct c2 {
    double x, y;
    __device__ c2 operator+(const c2 &o) const { return {x + o.x, y + o.y}; }
    __device__ c2 operator-(const c2 &o) const { return {x - o.x, y - o.y}; }
};

template <typename V>
__device__ __forceinline__ V pick16(const V (&W)[16], int zl) {
    switch (zl) {  // zl is wave-uniform: a scalar branch
    case 0: return W[0];
    case 1: return W[1];
    case 2: return W[2];
    case 3: return W[3];
    case 4: return W[4];
    case 5: return W[5];
    case 6: return W[6];
    case 7: return W[7];
    case 8: return W[8];
    case 9: return W[9];
    case 10: return W[10];
    case 11: return W[11];
    case 12: return W[12];
    case 13: return W[13];
    case 14: return W[14];
    default: return W[15];
    }
}

__device__ __forceinline__ uint64_t insert_zero(uint64_t v, int pos) {
    const uint64_t lo = v & ((1ull << pos) - 1);
    return ((v >> pos) << (pos + 1)) | lo;
}

// x == 0: S_t = sum_i |psi_i|^2 (-1)^{|i & z_t|}; NT terms per pass (real accumulators)
// Both Pauli kernels take states batched along gridDim.y (state y at y * shot_stride amplitudes,
// its partials after those of states 0..y-1).
template <typename R, int NT>
__global__ void __launch_bounds__(kRedThreads) k_pauli_z(const cx<R> *__restrict__ st, uint64_t nbases,
                                                         const uint64_t *__restrict__ z, int nt,
                                                         double *__restrict__ partials, uint64_t shot_stride) {
    __shared__ double sh[4 * NT];
    st += blockIdx.y * shot_stride;
    partials += (size_t)blockIdx.y * gridDim.x * NT;
    double acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = 0;
    for (uint64_t b = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; b < nbases;
         b += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i0 = b << 4;
        double W[16];
#pragma unroll
        for (int j = 0; j < 16; j++) {
            cx<R> a = st[i0 + j];
            W[j] = (double)a.x * (double)a.x + (double)a.y * (double)a.y;
        }
        wht16(W);
#pragma unroll
        for (int t = 0; t < NT; t++) {
            if (t < nt) {
                const uint64_t zt = z[t];
                const double w = pick16(W, (int)(zt & 15));
                acc[t] += (__popcll(i0 & zt) & 1) ? -w : w;
            }
        }
    }
    block_sum<NT>(acc, sh);
    if (threadIdx.x == 0)
        for (int t = 0; t < NT; t++) partials[(size_t)blockIdx.x * NT + t] = acc[t];
}

// x != 0: S_t = sum_i conj(psi[i^x]) (-1)^{|i&z_t|} psi[i]
template <typename R, int NT>
__global__ void __launch_bounds__(kRedThreads) k_pauli_x(const cx<R> *__restrict__ st, uint64_t nbases, uint64_t x,
                                                         int hb, const uint64_t *__restrict__ z, int nt,
                                                         double *__restrict__ partials, uint64_t shot_stride) {
    __shared__ double sh[4 * 2 * NT];
    st += blockIdx.y * shot_stride;
    partials += (size_t)blockIdx.y * gridDim.x * 2 * NT;
    double acc[2 * NT];
#pragma unroll
    for (int t = 0; t < 2 * NT; t++) acc[t] = 0;
    const uint64_t xhi = x & ~15ull;
    const int xlo = (int)(x & 15);
    for (uint64_t b = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; b < nbases;
         b += (uint64_t)gridDim.x * blockDim.x) {
        // with xhi != 0 only blocks with bit hb clear are visited; the partner block
        // (i0 ^ xhi) is handled through conj(W)
        const uint64_t i0 = (xhi ? insert_zero(b, hb) : b) << 4;
        const uint64_t p0 = i0 ^ xhi;
        c2 W[16];
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const cx<R> a = st[i0 + j];
            const cx<R> c = st[p0 + (j ^ xlo)];
            // conj(c) * a
            W[j] = c2{(double)c.x * a.x + (double)c.y * a.y, (double)c.x * a.y - (double)c.y * a.x};
        }
        wht16(W);
#pragma unroll
        for (int t = 0; t < NT; t++) {
            if (t < nt) {
                const uint64_t zt = z[t];
                const int zl = (int)(zt & 15);
                const c2 w = pick16(W, zl);
                double re = (__popcll(i0 & zt) & 1) ? -w.x : w.x;
                double im = (__popcll(i0 & zt) & 1) ? -w.y : w.y;
                if (xhi) {  // partner block: sign((i0^xhi)&z) * sign(xlo & zl) * conj(W[zl])
                    const bool neg = ((__popcll(p0 & zt) + __popc(xlo & zl)) & 1);
                    re += neg ? -w.x : w.x;
                    im += neg ? w.y : -w.y;
                }
                acc[2 * t] += re;
                acc[2 * t + 1] += im;
            }
        }
    }
    block_sum<2 * NT>(acc, sh);
    if (threadIdx.x == 0)
        for (int t = 0; t < 2 * NT; t++) partials[(size_t)blockIdx.x * 2 * NT + t] = acc[t];
}

template <typename R>
static int pauli_sums_t(State &st, uint64_t x, const uint64_t *z, int nt, double *out, Scratch &scr, hipStream_t s) {
    constexpr int NTZ = 32, NTX = 16;
    const int per = x == 0 ? NTZ : NTX;
    const int nv = x == 0 ? per : 2 * per;
    const uint64_t N = st.size();
    const uint64_t xhi = x & ~15ull;
    const uint64_t nbases = xhi ? (N >> 5) : (N >> 4);
    const int hb = xhi ? 63 - __builtin_clzll(xhi) - 4 : 0;  // bit of xhi in base units
    const unsigned grid = red_grid(nbases, kRedThreads * 4);
    const size_t zbytes = 256 * 8, pbytes = (size_t)grid * nv * 8;
    int rc = scr.ensure(zbytes + pbytes + nv * 8 + 256);
    if (rc) return rc;
    uint64_t *dz = static_cast<uint64_t *>(scr.ptr);
    double *dpart = reinterpret_cast<double *>(static_cast<char *>(scr.ptr) + zbytes);
    double *dout = dpart + (size_t)grid * nv;
    std::vector<double> host(nv);
    for (int t0 = 0; t0 < nt; t0 += per) {
        const int cnt = std::min(per, nt - t0);
        GOQU_HIP(hipMemcpyAsync(dz, z + t0, cnt * sizeof(uint64_t), hipMemcpyHostToDevice, s));
        if (x == 0)
            hipLaunchKernelGGL((k_pauli_z<R, NTZ>), dim3(grid), dim3(kRedThreads), 0, s,
                               static_cast<const cx<R> *>(st.amps), nbases, dz, cnt, dpart, 0);
        else
            hipLaunchKernelGGL((k_pauli_x<R, NTX>), dim3(grid), dim3(kRedThreads), 0, s,
                               static_cast<const cx<R> *>(st.amps), nbases, x, hb, dz, cnt, dpart, 0);
        GOQU_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_finish, dim3(nv), dim3(256), 0, s, dpart, (int)grid, nv, dout);
        GOQU_HIP(hipGetLastError());
        GOQU_HIP(hipMemcpyAsync(host.data(), dout, nv * sizeof(double), hipMemcpyDeviceToHost, s));
        GOQU_HIP(hipStreamSynchronize(s));
        for (int t = 0; t < cnt; t++) {
            if (x == 0) {
                out[2 * (t0 + t)] = host[t];
                out[2 * (t0 + t) + 1] = 0.0;
            } else {
                out[2 * (t0 + t)] = host[2 * t];
                out[2 * (t0 + t) + 1] = host[2 * t + 1];
            }
        }
    }
    return GOQU_OK;
}

// Batched finish: for state y, the first nv of the pstride values each of its nblocks blocks
// wrote, summed over the blocks in a fixed order, as complex pairs out[y * out_stride +
// 2 * (out_off + t)] (x == 0: real sums, imaginary parts 0).
__global__ void k_finish_batch(const double *__restrict__ partials, int nblocks, int pstride, int nv, int xmode,
                               double *__restrict__ out, int out_stride, int out_off) {
    const int y = blockIdx.x;
    const double *p = partials + (size_t)y * nblocks * pstride;
    for (int k = threadIdx.x; k < nv; k += blockDim.x) {
        double acc = 0;
        for (int b = 0; b < nblocks; b++) acc += p[(size_t)b * pstride + k];
        double *o = out + (size_t)y * out_stride + 2 * out_off;
        if (xmode) {
            o[k] = acc;
        } else {
            o[2 * k] = acc;
            o[2 * k + 1] = 0.0;
        }
    }
}

template <typename R>
static int pauli_sums_batch_t(const State &view, int nstates, uint64_t stride, uint64_t x, const uint64_t *dz, int nt,
                              double *dout, int out_stride, int out_off, Scratch &scr, hipStream_t s) {
    constexpr int NTZ = 32, NTX = 16;
    const int per = x == 0 ? NTZ : NTX;
    const int nv = x == 0 ? per : 2 * per;
    const uint64_t N = view.size();
    const uint64_t xhi = x & ~15ull;
    const uint64_t nbases = xhi ? (N >> 5) : (N >> 4);
    const int hb = xhi ? 63 - __builtin_clzll(xhi) - 4 : 0;
    // fewer blocks per state when many states share the launch
    const unsigned g1 = red_grid(nbases, kRedThreads * 4);
    const unsigned grid = std::max(1u, std::min(g1, (unsigned)std::max<uint64_t>(1, 4096 / (uint64_t)nstates)));
    int rc = scr.ensure((size_t)grid * nv * 8 * nstates);
    if (rc) return rc;
    double *dpart = static_cast<double *>(scr.ptr);
    for (int t0 = 0; t0 < nt; t0 += per) {
        const int cnt = std::min(per, nt - t0);
        if (x == 0)
            hipLaunchKernelGGL((k_pauli_z<R, NTZ>), dim3(grid, nstates), dim3(kRedThreads), 0, s,
                               static_cast<const cx<R> *>(view.amps), nbases, dz + t0, cnt, dpart, stride);
        else
            hipLaunchKernelGGL((k_pauli_x<R, NTX>), dim3(grid, nstates), dim3(kRedThreads), 0, s,
                               static_cast<const cx<R> *>(view.amps), nbases, x, hb, dz + t0, cnt, dpart, stride);
        GOQU_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_finish_batch, dim3(nstates), dim3(64), 0, s, dpart, (int)grid, nv,
                           x == 0 ? cnt : 2 * cnt, x == 0 ? 0 : 1, dout, out_stride, out_off + t0);
        GOQU_HIP(hipGetLastError());
    }
    return GOQU_OK;
}

int pauli_sums_batch(const State &view, int nstates, uint64_t stride, uint64_t x, const uint64_t *dz, int nt,
                     double *dout, int out_stride, int out_off, Scratch &scr, hipStream_t s) {
    return view.prec == GOQU_C128
               ? pauli_sums_batch_t<double>(view, nstates, stride, x, dz, nt, dout, out_stride, out_off, scr, s)
               : pauli_sums_batch_t<float>(view, nstates, stride, x, dz, nt, dout, out_stride, out_off, scr, s);
}

int pauli_sums(State &st, uint64_t x, const uint64_t *z, int nt, double *out, Scratch &scr, hipStream_t s) {
    return st.prec == GOQU_C128 ? pauli_sums_t<double>(st, x, z, nt, out, scr, s)
                                : pauli_sums_t<float>(st, x, z, nt, out, scr, s);
}

// ---- inner product, probabilities, collapse ---------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(kRedThreads) k_inner(const cx<R> *__restrict__ a, const cx<R> *__restrict__ b,
                                                       uint64_t n, double *__restrict__ partials) {
    __shared__ double sh[8];
    double v[2] = {0, 0};
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const cx<R> p = a[i], q = b[i];
        v[0] += (double)p.x * q.x + (double)p.y * q.y;
        v[1] += (double)p.x * q.y - (double)p.y * q.x;
    }
    block_sum<2>(v, sh);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = v[0];
        partials[2 * blockIdx.x + 1] = v[1];
    }
}

template <typename R>
__global__ void __launch_bounds__(kRedThreads) k_prob0(const cx<R> *__restrict__ a, uint64_t n, int q,
                                                       double *__restrict__ partials) {
    __shared__ double sh[4];
    double v[1] = {0};
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t idx = q >= 0 ? insert_zero(i, q) : i;
        const cx<R> p = a[idx];
        v[0] += (double)p.x * p.x + (double)p.y * p.y;
    }
    block_sum<1>(v, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
}

static int finish_to_host(const double *dpart, int grid, int nv, double *host, Scratch &scr, size_t out_off,
                          hipStream_t s) {
    double *dout = reinterpret_cast<double *>(static_cast<char *>(scr.ptr) + out_off);
    hipLaunchKernelGGL(k_finish, dim3(nv), dim3(256), 0, s, dpart, grid, nv, dout);
    GOQU_HIP(hipGetLastError());
    GOQU_HIP(hipMemcpyAsync(host, dout, nv * sizeof(double), hipMemcpyDeviceToHost, s));
    GOQU_HIP(hipStreamSynchronize(s));
    return GOQU_OK;
}

int inner_product(const State &a, const State &b, double out[2], Scratch &scr, hipStream_t s) {
    const uint64_t n = a.size();
    const unsigned grid = red_grid(n, kRedThreads * 16);
    const size_t pbytes = (size_t)grid * 2 * 8;
    int rc = scr.ensure(pbytes + 64);
    if (rc) return rc;
    double *dpart = static_cast<double *>(scr.ptr);
    if (a.prec == GOQU_C128)
        hipLaunchKernelGGL(k_inner<double>, dim3(grid), dim3(kRedThreads), 0, s,
                           static_cast<const cx<double> *>(a.amps), static_cast<const cx<double> *>(b.amps), n, dpart);
    else
        hipLaunchKernelGGL(k_inner<float>, dim3(grid), dim3(kRedThreads), 0, s,
                           static_cast<const cx<float> *>(a.amps), static_cast<const cx<float> *>(b.amps), n, dpart);
    GOQU_HIP(hipGetLastError());
    return finish_to_host(dpart, grid, 2, out, scr, pbytes, s);
}

int prob_zero(State &st, int q, double *out, Scratch &scr, hipStream_t s) {
    const uint64_t n = q >= 0 ? st.size() / 2 : st.size();
    const unsigned grid = red_grid(n, kRedThreads * 16);
    const size_t pbytes = (size_t)grid * 8;
    int rc = scr.ensure(pbytes + 64);
    if (rc) return rc;
    double *dpart = static_cast<double *>(scr.ptr);
    if (st.prec == GOQU_C128)
        hipLaunchKernelGGL(k_prob0<double>, dim3(grid), dim3(kRedThreads), 0, s,
                           static_cast<const cx<double> *>(st.amps), n, q, dpart);
    else
        hipLaunchKernelGGL(k_prob0<float>, dim3(grid), dim3(kRedThreads), 0, s,
                           static_cast<const cx<float> *>(st.amps), n, q, dpart);
    GOQU_HIP(hipGetLastError());
    return finish_to_host(dpart, grid, 1, out, scr, pbytes, s);
}

template <typename R>
__global__ void k_collapse(cx<R> *__restrict__ a, uint64_t npairs, int q, int outcome, double scale) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < npairs;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i0 = insert_zero(i, q), i1 = i0 | (1ull << q);
        const uint64_t keep = outcome ? i1 : i0, drop = outcome ? i0 : i1;
        cx<R> v = a[keep];
        a[keep] = cx<R>{(R)(v.x * scale), (R)(v.y * scale)};
        a[drop] = cx<R>{R(0), R(0)};
    }
}

int collapse(State &st, int q, int outcome, double scale, hipStream_t s) {
    const uint64_t np = st.size() / 2;
    const unsigned grid = red_grid(np, 256 * 8) * 4;
    if (st.prec == GOQU_C128)
        hipLaunchKernelGGL(k_collapse<double>, dim3(grid), dim3(256), 0, s, static_cast<cx<double> *>(st.amps), np, q,
                           outcome, scale);
    else
        hipLaunchKernelGGL(k_collapse<float>, dim3(grid), dim3(256), 0, s, static_cast<cx<float> *>(st.amps), np, q,
                           outcome, scale);
    GOQU_HIP(hipGetLastError());
    return GOQU_OK;
}

// ---- fill / copy / scale / axpy -----------------------------------------------------------------
template <typename R>
__global__ void k_fill_basis(cx<R> *__restrict__ a, uint64_t n, uint64_t index) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        a[i] = cx<R>{i == index ? R(1) : R(0), R(0)};
}

// nstates consecutive states of view.size() amplitudes each, every one |0...0>
template <typename R>
__global__ void k_fill_zero_batch(cx<R> *__restrict__ a, uint64_t n, uint64_t stride) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        a[i] = cx<R>{(i & (stride - 1)) == 0 ? R(1) : R(0), R(0)};
}

int fill_basis_batch(const State &view, int nstates, hipStream_t s) {
    const uint64_t n = view.size() * (uint64_t)nstates;
    const unsigned grid = (unsigned)std::min<uint64_t>((n + 255) / 256, 1u << 14);
    if (view.prec == GOQU_C128)
        hipLaunchKernelGGL(k_fill_zero_batch<double>, dim3(grid), dim3(256), 0, s,
                           static_cast<cx<double> *>(view.amps), n, view.size());
    else
        hipLaunchKernelGGL(k_fill_zero_batch<float>, dim3(grid), dim3(256), 0, s,
                           static_cast<cx<float> *>(view.amps), n, view.size());
    GOQU_HIP(hipGetLastError());
    return GOQU_OK;
}

int fill_basis(State &st, uint64_t index, hipStream_t s) {
    const uint64_t n = st.size();
    const unsigned grid = (unsigned)std::min<uint64_t>((n + 255) / 256, 1u << 16);
    if (st.prec == GOQU_C128)
        hipLaunchKernelGGL(k_fill_basis<double>, dim3(grid), dim3(256), 0, s, static_cast<cx<double> *>(st.amps), n,
                           index);
    else
        hipLaunchKernelGGL(k_fill_basis<float>, dim3(grid), dim3(256), 0, s, static_cast<cx<float> *>(st.amps), n,
                           index);
    GOQU_HIP(hipGetLastError());
    return GOQU_OK;
}

int copy_state(State &dst, const State &src, hipStream_t s) {
    GOQU_HIP(hipMemcpyAsync(dst.amps, src.amps, src.size() * src.elem, hipMemcpyDeviceToDevice, s));
    return GOQU_OK;
}

template <typename R>
__global__ void k_axpy(cx<R> *__restrict__ y, const cx<R> *__restrict__ x, uint64_t n, double cr, double ci,
                       int scale_only) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        if (scale_only) {
            const cx<R> v = y[i];
            y[i] = cx<R>{(R)(cr * v.x - ci * v.y), (R)(cr * v.y + ci * v.x)};
        } else {
            const cx<R> v = x[i], w = y[i];
            y[i] = cx<R>{(R)(w.x + cr * v.x - ci * v.y), (R)(w.y + cr * v.y + ci * v.x)};
        }
    }
}

int scale_state(State &st, double re, double im, hipStream_t s) {
    const uint64_t n = st.size();
    const unsigned grid = (unsigned)std::min<uint64_t>((n + 255) / 256, 1u << 16);
    if (st.prec == GOQU_C128)
        hipLaunchKernelGGL(k_axpy<double>, dim3(grid), dim3(256), 0, s, static_cast<cx<double> *>(st.amps), nullptr, n,
                           re, im, 1);
    else
        hipLaunchKernelGGL(k_axpy<float>, dim3(grid), dim3(256), 0, s, static_cast<cx<float> *>(st.amps), nullptr, n,
                           re, im, 1);
    GOQU_HIP(hipGetLastError());
    return GOQU_OK;
}

int axpy_state(State &y, const State &x, double re, double im, hipStream_t s) {
    const uint64_t n = y.size();
    const unsigned grid = (unsigned)std::min<uint64_t>((n + 255) / 256, 1u << 16);
    if (y.prec == GOQU_C128)
        hipLaunchKernelGGL(k_axpy<double>, dim3(grid), dim3(256), 0, s, static_cast<cx<double> *>(y.amps),
                           static_cast<const cx<double> *>(x.amps), n, re, im, 0);
    else
        hipLaunchKernelGGL(k_axpy<float>, dim3(grid), dim3(256), 0, s, static_cast<cx<float> *>(y.amps),
                           static_cast<const cx<float> *>(x.amps), n, re, im, 0);
    GOQU_HIP(hipGetLastError());
    return GOQU_OK;
}

// ---- H|psi> accumulation for one x-group -----------------------------------------------------------
// lam[i] += sum_t c_t (-1)^{|(i^x) & z_t|} psi[i^x],   c_t = coeff_t * i^{ny_t}
template <typename R>
__global__ void k_pauli_apply(const cx<R> *__restrict__ psi, cx<R> *__restrict__ lam, uint64_t n, uint64_t x,
                              const uint64_t *__restrict__ z, const double *__restrict__ c, int nt) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t j = i ^ x;
        double fr = 0, fi = 0;
        for (int t = 0; t < nt; t++) {
            const bool neg = __popcll(j & z[t]) & 1;
            fr += neg ? -c[2 * t] : c[2 * t];
            fi += neg ? -c[2 * t + 1] : c[2 * t + 1];
        }
        const cx<R> p = psi[j];
        cx<R> l = lam[i];
        l.x += (R)(fr * p.x - fi * p.y);
        l.y += (R)(fr * p.y + fi * p.x);
        lam[i] = l;
    }
}

int apply_pauli_group(const State &psi, State &lam, uint64_t x, const uint64_t *z, const double *c, int nt,
                      Scratch &scr, hipStream_t s) {
    const size_t zb = (size_t)nt * 8, cb = (size_t)nt * 16;
    int rc = scr.ensure(zb + cb + 64);
    if (rc) return rc;
    uint64_t *dz = static_cast<uint64_t *>(scr.ptr);
    double *dc = reinterpret_cast<double *>(static_cast<char *>(scr.ptr) + ((zb + 15) & ~size_t(15)));
    GOQU_HIP(hipMemcpyAsync(dz, z, zb, hipMemcpyHostToDevice, s));
    GOQU_HIP(hipMemcpyAsync(dc, c, cb, hipMemcpyHostToDevice, s));
    const uint64_t n = psi.size();
    const unsigned grid = (unsigned)std::min<uint64_t>((n + 255) / 256, 1u << 16);
    if (psi.prec == GOQU_C128)
        hipLaunchKernelGGL(k_pauli_apply<double>, dim3(grid), dim3(256), 0, s,
                           static_cast<const cx<double> *>(psi.amps), static_cast<cx<double> *>(lam.amps), n, x, dz, dc,
                           nt);
    else
        hipLaunchKernelGGL(k_pauli_apply<float>, dim3(grid), dim3(256), 0, s,
                           static_cast<const cx<float> *>(psi.amps), static_cast<cx<float> *>(lam.amps), n, x, dz, dc,
                           nt);
    GOQU_HIP(hipGetLastError());
    GOQU_HIP(hipStreamSynchronize(s));  // z/c staging is reused by the next call
    return GOQU_OK;
}

// ---- H|psi> in passes over tiles -------------------------------------------------------------------
// lam (+)= sum over a pass's x-groups g of D_g(i ^ x_g) psi[i ^ x_g], D_g(j) = sum_t c_t (-1)^{|j & z_t|}.
// A pass holds groups whose x-masks all lie inside 12 tile qubits tq[] (ascending, the 3 lowest
// qubits included); one workgroup per tile stages psi's tile in LDS and reads every flipped
// partner from there, so a pass moves psi once and lam once instead of once per group.
constexpr int kHTileBits = 12;
struct HGroup {
    uint64_t x;   // global x-mask
    uint32_t xl;  // x-mask on the tile's local bits
    int32_t t0;   // first term (index into z[] and c[])
    int32_t nt;   // terms
    int32_t pad;
};

template <typename R>
__global__ void __launch_bounds__(256) k_pauli_apply_tile(const cx<R> *__restrict__ psi, cx<R> *__restrict__ lam,
                                                          const uint8_t *__restrict__ tq, const HGroup *__restrict__ grp,
                                                          int ng, const uint64_t *__restrict__ z,
                                                          const double *__restrict__ c, int accumulate) {
    __shared__ cx<R> sm[1 << kHTileBits];
    const uint32_t t = threadIdx.x;
    uint8_t T[kHTileBits];
#pragma unroll
    for (int j = 0; j < kHTileBits; j++) T[j] = tq[j];
    // the tile's base: blockIdx deposited into the qubits outside T
    uint64_t base = blockIdx.x;
#pragma unroll
    for (int j = 0; j < kHTileBits; j++) {
        const uint64_t lo = base & ((1ull << T[j]) - 1);
        base = ((base >> T[j]) << (T[j] + 1)) | lo;
    }
    // local index j = t + 256 r: thread bits on T[0..7], register bits on T[8..11]
    uint64_t gt = base;
#pragma unroll
    for (int u = 0; u < 8; u++)
        if (t >> u & 1) gt |= 1ull << T[u];
    uint64_t gr[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        uint64_t g = 0;
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (r >> u & 1) g |= 1ull << T[8 + u];
        gr[r] = g;
    }
#pragma unroll
    for (int r = 0; r < 16; r++) sm[t + 256 * r] = psi[gt | gr[r]];
    __syncthreads();
    double ar[16], ai[16];
#pragma unroll
    for (int r = 0; r < 16; r++) ar[r] = ai[r] = 0.0;
    for (int g = 0; g < ng; g++) {
        const HGroup G = grp[g];
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const uint64_t j = (gt | gr[r]) ^ G.x;  // global index of the partner
            const cx<R> p = sm[(t + 256 * r) ^ G.xl];
            double fr = 0, fi = 0;
            for (int k = 0; k < G.nt; k++) {
                const bool neg = __popcll(j & z[G.t0 + k]) & 1;
                fr += neg ? -c[2 * (G.t0 + k)] : c[2 * (G.t0 + k)];
                fi += neg ? -c[2 * (G.t0 + k) + 1] : c[2 * (G.t0 + k) + 1];
            }
            ar[r] += fr * (double)p.x - fi * (double)p.y;
            ai[r] += fr * (double)p.y + fi * (double)p.x;
        }
    }
#pragma unroll
    for (int r = 0; r < 16; r++) {
        cx<R> *d = lam + (gt | gr[r]);
        if (accumulate) {
            const cx<R> o = *d;
            *d = cx<R>{(R)(o.x + ar[r]), (R)(o.y + ai[r])};
        } else {
            *d = cx<R>{(R)ar[r], (R)ai[r]};
        }
    }
}

int apply_hamiltonian_tiled(const State &psi, State &lam, int nt, const uint64_t *x, const uint64_t *z,
                            const double *cfold, Scratch &scr, hipStream_t s, bool &done) {
    done = false;
    if (psi.nq < kHTileBits || nt == 0) return GOQU_OK;
    // groups by x-mask, in first-appearance order of their x-mask
    std::vector<uint64_t> gx;
    std::vector<std::vector<int>> gterms;
    for (int t = 0; t < nt; t++) {
        size_t g = 0;
        while (g < gx.size() && gx[g] != x[t]) g++;
        if (g == gx.size()) {
            gx.push_back(x[t]);
            gterms.emplace_back();
        }
        gterms[g].push_back(t);
    }
    const uint64_t low = 7;  // qubits 0..2: runs of 8 amplitudes per global access
    for (uint64_t m : gx)
        if (__builtin_popcountll(m | low) > kHTileBits) return GOQU_OK;  // a group needs more tile qubits
    // greedy passes: a group joins the first pass whose qubit set stays within 12 qubits
    std::vector<uint64_t> pass_q;
    std::vector<std::vector<int>> pass_g;
    for (size_t g = 0; g < gx.size(); g++) {
        size_t p = 0;
        while (p < pass_q.size() && __builtin_popcountll(pass_q[p] | gx[g]) > kHTileBits) p++;
        if (p == pass_q.size()) {
            pass_q.push_back(low);
            pass_g.emplace_back();
        }
        pass_q[p] |= gx[g];
        pass_g[p].push_back((int)g);
    }
    // device staging: per pass 16 bytes of tile qubits + groups, then all terms
    const size_t np = pass_q.size(), ngt = gx.size();
    std::vector<unsigned char> host;
    auto put = [&](const void *p, size_t b) {
        const size_t at = (host.size() + 15) & ~size_t(15);
        host.resize(at + b);
        std::memcpy(host.data() + at, p, b);
        return at;
    };
    std::vector<uint64_t> zt(nt);
    std::vector<double> ct(2 * (size_t)nt);
    std::vector<size_t> off_q(np), off_g(np);
    int32_t k = 0;
    std::vector<int32_t> first(ngt);
    for (size_t g = 0; g < ngt; g++) {
        first[g] = k;
        for (int t : gterms[g]) {
            zt[k] = z[t];
            ct[2 * k] = cfold[2 * t];
            ct[2 * k + 1] = cfold[2 * t + 1];
            k++;
        }
    }
    const size_t off_z = put(zt.data(), zt.size() * 8), off_c = put(ct.data(), ct.size() * 8);
    for (size_t p = 0; p < np; p++) {
        uint64_t q = pass_q[p];
        for (int b = 0; b < psi.nq && __builtin_popcountll(q) < kHTileBits; b++) q |= 1ull << b;
        uint8_t tq[16] = {0};
        int loc[64];
        for (int b = 0, j = 0; b < psi.nq; b++)
            if (q >> b & 1) {
                loc[b] = j;
                tq[j++] = (uint8_t)b;
            }
        std::vector<HGroup> hg;
        for (int g : pass_g[p]) {
            HGroup G{};
            G.x = gx[g];
            uint32_t xl = 0;
            for (uint64_t m = gx[g]; m; m &= m - 1) {
                const int j = loc[__builtin_ctzll(m)];
                // local index bits: thread bits T[0..7] -> index bits 0..7, register bits T[8..11] -> 8..11
                xl |= 1u << j;
            }
            G.xl = xl;
            G.t0 = first[g];
            G.nt = (int32_t)gterms[g].size();
            hg.push_back(G);
        }
        off_q[p] = put(tq, 16);
        off_g[p] = put(hg.data(), hg.size() * sizeof(HGroup));
    }
    int rc = scr.ensure(host.size() + 64);
    if (rc) return rc;
    char *dev = static_cast<char *>(scr.ptr);
    GOQU_HIP(hipMemcpyAsync(dev, host.data(), host.size(), hipMemcpyHostToDevice, s));
    const unsigned grid = (unsigned)(psi.size() >> kHTileBits);
    for (size_t p = 0; p < np; p++) {
        const int ng = (int)pass_g[p].size();
        const uint8_t *dq = reinterpret_cast<const uint8_t *>(dev + off_q[p]);
        const HGroup *dg = reinterpret_cast<const HGroup *>(dev + off_g[p]);
        const uint64_t *dz = reinterpret_cast<const uint64_t *>(dev + off_z);
        const double *dc = reinterpret_cast<const double *>(dev + off_c);
        hipEvent_t ev0 = prof_begin(s);
        if (psi.prec == GOQU_C128)
            hipLaunchKernelGGL(k_pauli_apply_tile<double>, dim3(grid), dim3(256), 0, s,
                               static_cast<const cx<double> *>(psi.amps), static_cast<cx<double> *>(lam.amps), dq, dg,
                               ng, dz, dc, p > 0 ? 1 : 0);
        else
            hipLaunchKernelGGL(k_pauli_apply_tile<float>, dim3(grid), dim3(256), 0, s,
                               static_cast<const cx<float> *>(psi.amps), static_cast<cx<float> *>(lam.amps), dq, dg, ng,
                               dz, dc, p > 0 ? 1 : 0);
        GOQU_HIP(hipGetLastError());
        // psi read, lambda written (and read back after the first pass)
        prof_end(ev0, s, PROF_HAMILTONIAN, (p > 0 ? 3.0 : 2.0) * psi.size() * psi.elem);
    }
    GOQU_HIP(hipStreamSynchronize(s));  // the host staging buffer is released on return
    done = true;
    return GOQU_OK;
}

// ---- sandwich <lam| dM |psi> ---------------------------------------------------------------------
template <typename R, int K>
__global__ void __launch_bounds__(kRedThreads) k_sandwich(const cx<R> *__restrict__ lam, const cx<R> *__restrict__ psi,
                                                          uint64_t ngroups, const int *__restrict__ qs,
                                                          const int *__restrict__ sorted, uint64_t cmask,
                                                          const double *__restrict__ M, double *__restrict__ partials) {
    constexpr int D = 1 << K;
    __shared__ double sh[8];
    double v[2] = {0, 0};
    for (uint64_t gidx = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; gidx < ngroups;
         gidx += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t base = gidx;
#pragma unroll
        for (int j = 0; j < K; j++) base = insert_zero(base, sorted[j]);
        if ((base & cmask) != cmask) continue;
        cx<R> p[D];
        uint64_t idx[D];
#pragma unroll
        for (int r = 0; r < D; r++) {
            uint64_t i = base;
#pragma unroll
            for (int j = 0; j < K; j++)
                if (r >> (K - 1 - j) & 1) i |= 1ull << qs[j];
            idx[r] = i;
            p[r] = psi[i];
        }
#pragma unroll
        for (int r = 0; r < D; r++) {
            double sr = 0, si = 0;
#pragma unroll
            for (int c = 0; c < D; c++) {
                const double mr = M[2 * (r * D + c)], mi = M[2 * (r * D + c) + 1];
                sr += mr * p[c].x - mi * p[c].y;
                si += mr * p[c].y + mi * p[c].x;
            }
            const cx<R> l = lam[idx[r]];
            v[0] += (double)l.x * sr + (double)l.y * si;
            v[1] += (double)l.x * si - (double)l.y * sr;
        }
    }
    block_sum<2>(v, sh);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = v[0];
        partials[2 * blockIdx.x + 1] = v[1];
    }
}

template <typename R>
static int sandwich_t(const State &lam, const State &psi, const EOp &e, double out[2], Scratch &scr, hipStream_t s) {
    const int K = e.k, D = 1 << K;
    const uint64_t ng = psi.size() >> K;
    const unsigned grid = red_grid(ng, kRedThreads * 4);
    const size_t hdr = 512;
    const size_t mbytes = (size_t)2 * D * D * 8;
    const size_t pbytes = (size_t)grid * 2 * 8;
    int rc = scr.ensure(hdr + mbytes + pbytes + 64);
    if (rc) return rc;
    char *base = static_cast<char *>(scr.ptr);
    std::vector<unsigned char> hbuf(hdr + mbytes, 0);
    int *qs = reinterpret_cast<int *>(hbuf.data()), *sorted = qs + 8;
    for (int j = 0; j < K; j++) qs[j] = sorted[j] = e.q[j];
    std::sort(sorted, sorted + K);
    std::memcpy(hbuf.data() + hdr, e.m.data(), mbytes);
    GOQU_HIP(hipMemcpyAsync(base, hbuf.data(), hdr + mbytes, hipMemcpyHostToDevice, s));
    const int *dqs = reinterpret_cast<const int *>(base), *dsorted = dqs + 8;
    const double *dM = reinterpret_cast<const double *>(base + hdr);
    double *dpart = reinterpret_cast<double *>(base + hdr + mbytes);
    const cx<R> *L = static_cast<const cx<R> *>(lam.amps), *P = static_cast<const cx<R> *>(psi.amps);
    switch (K) {
    case 1: hipLaunchKernelGGL((k_sandwich<R, 1>), dim3(grid), dim3(kRedThreads), 0, s, L, P, ng, dqs, dsorted, e.cmask, dM, dpart); break;
    case 2: hipLaunchKernelGGL((k_sandwich<R, 2>), dim3(grid), dim3(kRedThreads), 0, s, L, P, ng, dqs, dsorted, e.cmask, dM, dpart); break;
    case 3: hipLaunchKernelGGL((k_sandwich<R, 3>), dim3(grid), dim3(kRedThreads), 0, s, L, P, ng, dqs, dsorted, e.cmask, dM, dpart); break;
    case 4: hipLaunchKernelGGL((k_sandwich<R, 4>), dim3(grid), dim3(kRedThreads), 0, s, L, P, ng, dqs, dsorted, e.cmask, dM, dpart); break;
    default: return fail(GOQU_ERR_UNSUPPORTED, "derivative on more than 4 target qubits");
    }
    GOQU_HIP(hipGetLastError());
    return finish_to_host(dpart, grid, 2, out, scr, hdr + mbytes + pbytes, s);
}

int sandwich(const State &lam, const State &psi, const EOp &e, double out[2], Scratch &scr, hipStream_t s) {
    return psi.prec == GOQU_C128 ? sandwich_t<double>(lam, psi, e, out, scr, s)
                                 : sandwich_t<float>(lam, psi, e, out, scr, s);
}

// Several <lam| M_j |psi> at once, the M_j on the same qubits: one read of lam and psi for all.
template <typename R, int K>
__global__ void __launch_bounds__(kRedThreads) k_sandwich_multi(const cx<R> *__restrict__ lam,
                                                                const cx<R> *__restrict__ psi, uint64_t ngroups,
                                                                const int *__restrict__ qs,
                                                                const int *__restrict__ sorted, uint64_t cmask,
                                                                const double *__restrict__ M, int nm,
                                                                double *__restrict__ partials) {
    constexpr int D = 1 << K;
    __shared__ double sh[4 * 2 * kMaxSandwich];
    double v[2 * kMaxSandwich];
#pragma unroll
    for (int j = 0; j < 2 * kMaxSandwich; j++) v[j] = 0;
    for (uint64_t gidx = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; gidx < ngroups;
         gidx += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t base = gidx;
#pragma unroll
        for (int j = 0; j < K; j++) base = insert_zero(base, sorted[j]);
        if ((base & cmask) != cmask) continue;
        cx<R> p[D], l[D];
#pragma unroll
        for (int r = 0; r < D; r++) {
            uint64_t i = base;
#pragma unroll
            for (int j = 0; j < K; j++)
                if (r >> (K - 1 - j) & 1) i |= 1ull << qs[j];
            p[r] = psi[i];
            l[r] = lam[i];
        }
#pragma unroll
        for (int m = 0; m < kMaxSandwich; m++) {
            if (m >= nm) break;
            const double *Mm = M + (size_t)m * 2 * D * D;
#pragma unroll
            for (int r = 0; r < D; r++) {
                double sr = 0, si = 0;
#pragma unroll
                for (int c = 0; c < D; c++) {
                    const double mr = Mm[2 * (r * D + c)], mi = Mm[2 * (r * D + c) + 1];
                    sr += mr * p[c].x - mi * p[c].y;
                    si += mr * p[c].y + mi * p[c].x;
                }
                v[2 * m] += (double)l[r].x * sr + (double)l[r].y * si;
                v[2 * m + 1] += (double)l[r].x * si - (double)l[r].y * sr;
            }
        }
    }
    block_sum<2 * kMaxSandwich>(v, sh);
    if (threadIdx.x == 0)
        for (int j = 0; j < 2 * nm; j++) partials[(size_t)blockIdx.x * 2 * kMaxSandwich + j] = v[j];
}

template <typename R>
static int sandwich_multi_t(const State &lam, const State &psi, const std::vector<EOp> &ms, double *out,
                            Scratch &scr, hipStream_t s) {
    const EOp &e = ms[0];
    const int K = e.k, D = 1 << K, nm = (int)ms.size();
    const uint64_t ng = psi.size() >> K;
    const unsigned grid = red_grid(ng, kRedThreads * 4);
    const size_t hdr = 512;
    const size_t mbytes = (size_t)nm * 2 * D * D * 8;
    const size_t pbytes = (size_t)grid * 2 * kMaxSandwich * 8;
    int rc = scr.ensure(hdr + mbytes + pbytes + 2 * kMaxSandwich * 8 + 64);
    if (rc) return rc;
    char *base = static_cast<char *>(scr.ptr);
    std::vector<unsigned char> hbuf(hdr + mbytes, 0);
    int *qs = reinterpret_cast<int *>(hbuf.data()), *sorted = qs + 8;
    for (int j = 0; j < K; j++) qs[j] = sorted[j] = e.q[j];
    std::sort(sorted, sorted + K);
    for (int m = 0; m < nm; m++)
        std::memcpy(hbuf.data() + hdr + (size_t)m * 2 * D * D * 8, ms[m].m.data(), (size_t)2 * D * D * 8);
    GOQU_HIP(hipMemcpyAsync(base, hbuf.data(), hbuf.size(), hipMemcpyHostToDevice, s));
    const int *dqs = reinterpret_cast<const int *>(base), *dsorted = dqs + 8;
    const double *dM = reinterpret_cast<const double *>(base + hdr);
    double *dpart = reinterpret_cast<double *>(base + hdr + mbytes);
    const cx<R> *L = static_cast<const cx<R> *>(lam.amps), *P = static_cast<const cx<R> *>(psi.amps);
    switch (K) {
    case 1: hipLaunchKernelGGL((k_sandwich_multi<R, 1>), dim3(grid), dim3(kRedThreads), 0, s, L, P, ng, dqs, dsorted, e.cmask, dM, nm, dpart); break;
    case 2: hipLaunchKernelGGL((k_sandwich_multi<R, 2>), dim3(grid), dim3(kRedThreads), 0, s, L, P, ng, dqs, dsorted, e.cmask, dM, nm, dpart); break;
    case 3: hipLaunchKernelGGL((k_sandwich_multi<R, 3>), dim3(grid), dim3(kRedThreads), 0, s, L, P, ng, dqs, dsorted, e.cmask, dM, nm, dpart); break;
    case 4: hipLaunchKernelGGL((k_sandwich_multi<R, 4>), dim3(grid), dim3(kRedThreads), 0, s, L, P, ng, dqs, dsorted, e.cmask, dM, nm, dpart); break;
    default: return fail(GOQU_ERR_UNSUPPORTED, "derivative on more than 4 target qubits");
    }
    GOQU_HIP(hipGetLastError());
    std::vector<double> host(2 * kMaxSandwich);
    rc = finish_to_host(dpart, grid, 2 * kMaxSandwich, host.data(), scr, hdr + mbytes + pbytes, s);
    if (rc) return rc;
    for (int j = 0; j < 2 * nm; j++) out[j] = host[j];
    return GOQU_OK;
}

int sandwich_multi(const State &lam, const State &psi, const std::vector<EOp> &ms, double *out, Scratch &scr,
                   hipStream_t s) {
    if (ms.empty()) return GOQU_OK;
    if ((int)ms.size() > kMaxSandwich) return fail(GOQU_ERR_INVALID_ARG, "too many matrices for one sandwich");
    for (const EOp &e : ms)
        if (e.k != ms[0].k || e.cmask != ms[0].cmask || std::memcmp(e.q, ms[0].q, sizeof(e.q)) != 0)
            return fail(GOQU_ERR_INVALID_ARG, "sandwich matrices on different qubits");
    return psi.prec == GOQU_C128 ? sandwich_multi_t<double>(lam, psi, ms, out, scr, s)
                                 : sandwich_multi_t<float>(lam, psi, ms, out, scr, s);
}

// ---- sampling: FP64 inverse CDF ---------------------------------------------------------------------
// Chunk sums in double-double, exclusive scan of the chunk sums, then one wavefront per shot
// walks its chunk.  The prefix is carried in double-double, so the CDF value compared with
// each uniform is the exactly rounded prefix sum.
constexpr int kChunk = 4096;

struct dd {
    double hi, lo;
};
__device__ __forceinline__ dd dd_add(dd a, double b) {
    const double s = a.hi + b;
    const double bb = s - a.hi;
    const double err = (a.hi - (s - bb)) + (b - bb);
    const double lo = a.lo + err;
    const double hi = s + lo;
    return {hi, lo - (hi - s)};
}
__device__ __forceinline__ dd dd_add2(dd a, dd b) {
    dd r = dd_add(a, b.hi);
    return dd_add(r, b.lo);
}

template <bool FROM_AMPS, typename R>
__device__ __forceinline__ double prob_at(const void *src, uint64_t i) {
    if constexpr (FROM_AMPS) {
        const cx<R> a = static_cast<const cx<R> *>(src)[i];
        return (double)a.x * a.x + (double)a.y * a.y;
    } else {
        return static_cast<const double *>(src)[i];
    }
}

// one workgroup per chunk: dd sum of the chunk
template <bool FROM_AMPS, typename R>
__global__ void __launch_bounds__(256) k_chunk_sums(const void *src, uint64_t n, dd *__restrict__ sums) {
    __shared__ dd sh[256];
    const uint64_t c0 = (uint64_t)blockIdx.x * kChunk;
    dd acc = {0, 0};
    for (int j = threadIdx.x; j < kChunk; j += 256) {
        const uint64_t i = c0 + j;
        if (i < n) acc = dd_add(acc, prob_at<FROM_AMPS, R>(src, i));
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        dd t = {0, 0};
        for (int k = 0; k < 256; k++) t = dd_add2(t, sh[k]);
        sums[blockIdx.x] = t;
    }
}

// single workgroup: exclusive prefix of nchunks dd values (in place), total at [nchunks]
__global__ void k_scan_chunks(dd *__restrict__ sums, uint64_t nchunks) {
    __shared__ dd part[256];
    const uint64_t per = (nchunks + 255) / 256;
    const uint64_t b0 = threadIdx.x * per, b1 = min(nchunks, b0 + per);
    dd s = {0, 0};
    for (uint64_t i = b0; i < b1; i++) s = dd_add2(s, sums[i]);
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        dd run = {0, 0};
        for (int k = 0; k < 256; k++) {
            const dd v = part[k];
            part[k] = run;
            run = dd_add2(run, v);
        }
        sums[nchunks] = run;
    }
    __syncthreads();
    dd run = part[threadIdx.x];
    for (uint64_t i = b0; i < b1; i++) {
        const dd v = sums[i];
        sums[i] = run;
        run = dd_add2(run, v);
    }
}

// one wavefront per shot
template <bool FROM_AMPS, typename R>
__global__ void __launch_bounds__(256) k_sample(const void *src, uint64_t n, const dd *__restrict__ prefix,
                                                uint64_t nchunks, const double *__restrict__ u, int64_t shots,
                                                uint64_t *__restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t shot = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (shot >= shots) return;
    const double r = u[shot];
    // last chunk whose exclusive prefix is <= r
    uint64_t lo = 0, hi = nchunks;  // invariant: prefix[lo] <= r (prefix[0] = 0)
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) / 2;
        if (prefix[mid].hi <= r) lo = mid;
        else hi = mid;
    }
    if (!(r < prefix[nchunks].hi)) {  // beyond the total: the last index (sim.go:297)
        if (lane == 0) out[shot] = n - 1;
        return;
    }
    // walk chunk `lo`: each lane owns kChunk/64 consecutive entries
    constexpr int per = kChunk / 64;
    const uint64_t c0 = lo * kChunk + (uint64_t)lane * per;
    dd mysum = {0, 0};
    for (int j = 0; j < per; j++) {
        const uint64_t i = c0 + j;
        if (i < n) mysum = dd_add(mysum, prob_at<FROM_AMPS, R>(src, i));
    }
    // exclusive scan across lanes (sequential via shuffles, 64 steps)
    dd run = prefix[lo];
    dd excl = run;
    for (int l = 0; l < 64; l++) {
        const double sh_hi = __shfl(mysum.hi, l, 64);
        const double sh_lo = __shfl(mysum.lo, l, 64);
        if (l == lane) excl = run;
        run = dd_add2(run, dd{sh_hi, sh_lo});
    }
    const dd incl = dd_add2(excl, mysum);
    // the lane whose [excl, incl) contains r
    const bool mine = (excl.hi <= r) && (r < incl.hi);
    const uint64_t ball = __ballot(mine);
    int owner = ball ? __ffsll((unsigned long long)ball) - 1 : 63;
    if (lane == owner) {
        dd c = excl;
        uint64_t hit = c0 + per - 1 < n ? c0 + per - 1 : n - 1;
        for (int j = 0; j < per; j++) {
            const uint64_t i = c0 + j;
            if (i >= n) break;
            c = dd_add(c, prob_at<FROM_AMPS, R>(src, i));
            if (r < c.hi) {
                hit = i;
                break;
            }
        }
        out[shot] = hit;
    }
}

// Bytes of device buffer sample_into needs: chunk prefixes, staged uniforms, sampled indices.
static size_t sample_buffer_bytes(uint64_t n, int64_t shots, size_t *ub_off, size_t *ob_off) {
    const uint64_t nchunks = (n + kChunk - 1) / kChunk;
    const size_t sb = (nchunks + 1) * sizeof(dd), ub = shots * sizeof(double), ob = shots * sizeof(uint64_t);
    const size_t uoff = (sb + 255) & ~size_t(255), ooff = uoff + ((ub + 255) & ~size_t(255));
    if (ub_off) *ub_off = uoff;
    if (ob_off) *ob_off = ooff;
    return ooff + ((ob + 255) & ~size_t(255));
}

// The three sampling kernels over a device buffer of sample_buffer_bytes(n, shots) bytes.  Every
// byte the kernels read is written earlier in the same call (chunk sums by k_chunk_sums, the
// total by k_scan_chunks, uniforms by the copy, indices by k_sample), so the buffer needs no
// clearing between calls.
template <bool FROM_AMPS, typename R>
static int sample_into(const void *src, uint64_t n, const double *u, int64_t shots, uint64_t *out, char *buf,
                       hipStream_t s) {
    if (shots <= 0) return GOQU_OK;
    const uint64_t nchunks = (n + kChunk - 1) / kChunk;
    size_t uoff = 0, ooff = 0;
    sample_buffer_bytes(n, shots, &uoff, &ooff);
    dd *sums = reinterpret_cast<dd *>(buf);
    double *du = reinterpret_cast<double *>(buf + uoff);
    uint64_t *dout = reinterpret_cast<uint64_t *>(buf + ooff);
    GOQU_HIP(hipMemcpyAsync(du, u, shots * sizeof(double), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL((k_chunk_sums<FROM_AMPS, R>), dim3((unsigned)nchunks), dim3(256), 0, s, src, n, sums);
    GOQU_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_scan_chunks, dim3(1), dim3(256), 0, s, sums, nchunks);
    GOQU_HIP(hipGetLastError());
    const unsigned grid = (unsigned)((shots + 3) / 4);
    hipLaunchKernelGGL((k_sample<FROM_AMPS, R>), dim3(grid), dim3(256), 0, s, src, n, sums, nchunks, du, shots, dout);
    GOQU_HIP(hipGetLastError());
    GOQU_HIP(hipMemcpyAsync(out, dout, shots * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    GOQU_HIP(hipStreamSynchronize(s));
    return GOQU_OK;
}

template <bool FROM_AMPS, typename R>
static int sample_impl(const void *src, uint64_t n, const double *u, int64_t shots, uint64_t *out, Scratch &scr,
                       hipStream_t s) {
    if (shots <= 0) return GOQU_OK;
    int rc = scr.ensure(sample_buffer_bytes(n, shots, nullptr, nullptr));
    if (rc) return rc;
    return sample_into<FROM_AMPS, R>(src, n, u, shots, out, static_cast<char *>(scr.ptr), s);
}

int sample_state(State &st, uint64_t nvalid, const double *u, int64_t shots, uint64_t *out, Scratch &scr,
                 hipStream_t s) {
    return st.prec == GOQU_C128 ? sample_impl<true, double>(st.amps, nvalid, u, shots, out, scr, s)
                                : sample_impl<true, float>(st.amps, nvalid, u, shots, out, scr, s);
}

int sample_probs(const double *dprobs, uint64_t n, const double *u, int64_t shots, uint64_t *out, Scratch &scr,
                 hipStream_t s) {
    return sample_impl<false, double>(dprobs, n, u, shots, out, scr, s);
}

// ---- allocation-free sampling ------------------------------------------------------------------------
// A file-scope device array stands in for the Scratch buffer of sample_state.  Each device holds its
// own copy until the process ends; host code reaches it only through the address that
// hipGetSymbolAddress returns on that device (the array's name is not a device address in host
// code).  Calls that overlap on one device share the array and race.
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

static std::mutex g_static_mu;
static void *g_static_ptr[kMaxDevices];

// this device's address of `scratch`, resolved on the first call on the device
static int static_scratch(int dev, char **out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(GOQU_ERR_DEVICE, "device ordinal beyond the supported range");
    std::lock_guard<std::mutex> lk(g_static_mu);
    if (!g_static_ptr[dev]) {
        void *p = nullptr;
        GOQU_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(GOQU_ERR_DEVICE, "hipGetSymbolAddress returned no address for the static scratch");
        g_static_ptr[dev] = p;
    }
    *out = static_cast<char *>(g_static_ptr[dev]);
    return GOQU_OK;
}

int sample_state_static(State &st, int dev, uint64_t nvalid, const double *u, int64_t shots, uint64_t *out,
                        hipStream_t s) {
    if (shots <= 0) return GOQU_OK;
    const size_t need = sample_buffer_bytes(nvalid, shots, nullptr, nullptr);
    if (need > kStaticScratchBytes) {
        char msg[256];
        std::snprintf(msg, sizeof(msg),
                      "sample without allocation: %lld shots over 2^%d amplitudes need %zu bytes of scratch, "
                      "the static scratch holds %zu",
                      (long long)shots, 63 - __builtin_clzll(nvalid), need, kStaticScratchBytes);
        return fail_cstr(GOQU_ERR_INVALID_ARG, msg);
    }
    char *buf = nullptr;
    int rc = static_scratch(dev, &buf);
    if (rc) return rc;
    return st.prec == GOQU_C128 ? sample_into<true, double>(st.amps, nvalid, u, shots, out, buf, s)
                                : sample_into<true, float>(st.amps, nvalid, u, shots, out, buf, s);
}

// ---- density matrix ---------------------------------------------------------------------------------
__global__ void k_dm_diag(const cx<double> *__restrict__ rho, uint64_t dim, double *__restrict__ p,
                          uint64_t shot_stride) {
    rho += blockIdx.y * shot_stride;
    p += blockIdx.y * dim;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < dim; i += (uint64_t)gridDim.x * blockDim.x) {
        const double v = rho[i * dim + i].x;
        p[i] = v < 0 ? 0.0 : v;  // measure.go:11-21
    }
}

int dm_diag(const State &rho, int n, double *dprobs, hipStream_t s) {
    const uint64_t dim = 1ull << n;
    const unsigned grid = (unsigned)std::min<uint64_t>((dim + 255) / 256, 4096);
    hipLaunchKernelGGL(k_dm_diag, dim3(grid), dim3(256), 0, s, static_cast<const cx<double> *>(rho.amps), dim, dprobs,
                       0);
    GOQU_HIP(hipGetLastError());
    return GOQU_OK;
}

int dm_diag_batch(const State &view, int n, int nstates, uint64_t stride, double *dprobs, hipStream_t s) {
    const uint64_t dim = 1ull << n;
    const unsigned grid = (unsigned)std::min<uint64_t>((dim + 255) / 256, 1024);
    hipLaunchKernelGGL(k_dm_diag, dim3(grid, nstates), dim3(256), 0, s, static_cast<const cx<double> *>(view.amps),
                       dim, dprobs, stride);
    GOQU_HIP(hipGetLastError());
    return GOQU_OK;
}

// one ApplyReadoutError qubit pass (measure.go:48-68): out[i] = own share + flipped-in share
__global__ void k_readout(const double *__restrict__ in, double *__restrict__ out, uint64_t dim, int q, double p01,
                          double p10) {
#pragma clang fp contract(off)
    const uint64_t mask = 1ull << q;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < dim; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t f = i ^ mask;
        if (!(i & mask))
            out[i] = in[i] * (1 - p01) + in[f] * p10;
        else
            out[i] = in[f] * p01 + in[i] * (1 - p10);
    }
}

int dm_readout(double *dprobs, double *tmp, int n, int q, double p01, double p10, hipStream_t s) {
    const uint64_t dim = 1ull << n;
    const unsigned grid = (unsigned)std::min<uint64_t>((dim + 255) / 256, 4096);
    hipLaunchKernelGGL(k_readout, dim3(grid), dim3(256), 0, s, dprobs, tmp, dim, q, p01, p10);
    GOQU_HIP(hipGetLastError());
    GOQU_HIP(hipMemcpyAsync(dprobs, tmp, dim * sizeof(double), hipMemcpyDeviceToDevice, s));
    return GOQU_OK;
}

// sum_i rho[i, i^x] (-1)^{|i&z|}  (expect_dm.go:9-21, without coeff and i^ny)
// density matrices batched along gridDim.y as in k_pauli_z
__global__ void __launch_bounds__(kRedThreads) k_dm_pauli(const cx<double> *__restrict__ rho, uint64_t dim, uint64_t x,
                                                          uint64_t z, double *__restrict__ partials,
                                                          uint64_t shot_stride) {
    __shared__ double sh[8];
    rho += blockIdx.y * shot_stride;
    partials += (size_t)blockIdx.y * gridDim.x * 2;
    double v[2] = {0, 0};
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < dim; i += (uint64_t)gridDim.x * blockDim.x) {
        const cx<double> e = rho[i * dim + (i ^ x)];
        const bool neg = __popcll(i & z) & 1;
        v[0] += neg ? -e.x : e.x;
        v[1] += neg ? -e.y : e.y;
    }
    block_sum<2>(v, sh);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = v[0];
        partials[2 * blockIdx.x + 1] = v[1];
    }
}

int dm_pauli(const State &rho, int n, uint64_t x, uint64_t z, double out[2], Scratch &scr, hipStream_t s) {
    const uint64_t dim = 1ull << n;
    const unsigned grid = red_grid(dim, kRedThreads * 4);
    const size_t pbytes = (size_t)grid * 16;
    int rc = scr.ensure(pbytes + 64);
    if (rc) return rc;
    double *dpart = static_cast<double *>(scr.ptr);
    hipLaunchKernelGGL(k_dm_pauli, dim3(grid), dim3(kRedThreads), 0, s, static_cast<const cx<double> *>(rho.amps), dim,
                       x, z, dpart, 0);
    GOQU_HIP(hipGetLastError());
    return finish_to_host(dpart, grid, 2, out, scr, pbytes, s);
}

int dm_pauli_batch(const State &view, int n, int nstates, uint64_t stride, uint64_t x, uint64_t z, double *dout,
                   int out_stride, int out_off, Scratch &scr, hipStream_t s) {
    const uint64_t dim = 1ull << n;
    const unsigned g1 = red_grid(dim, kRedThreads * 4);
    const unsigned grid = std::max(1u, std::min(g1, (unsigned)std::max<uint64_t>(1, 4096 / (uint64_t)nstates)));
    int rc = scr.ensure((size_t)grid * 16 * nstates);
    if (rc) return rc;
    double *dpart = static_cast<double *>(scr.ptr);
    hipLaunchKernelGGL(k_dm_pauli, dim3(grid, nstates), dim3(kRedThreads), 0, s,
                       static_cast<const cx<double> *>(view.amps), dim, x, z, dpart, stride);
    GOQU_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_finish_batch, dim3(nstates), dim3(64), 0, s, dpart, (int)grid, 2, 2, 1, dout, out_stride,
                       out_off);
    GOQU_HIP(hipGetLastError());
    return GOQU_OK;
}

// Purity (sim.go:265-276): Re sum_ij rho_ij conj(rho_ji), tiled 32x32 transposes
__global__ void __launch_bounds__(256) k_dm_purity(const cx<double> *__restrict__ rho, uint64_t dim,
                                                   double *__restrict__ partials) {
    __shared__ cx<double> tile[32][33];
    __shared__ double sh[8];
    double v[2] = {0, 0};
    const uint64_t nt = (dim + 31) / 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (uint64_t t = blockIdx.x; t < nt * nt; t += gridDim.x) {
        const uint64_t bi = t / nt, bj = t % nt;
        __syncthreads();
        for (int r = ty; r < 32; r += 8) {  // tile (bj, bi) transposed into LDS
            const uint64_t gi = bj * 32 + r, gj = bi * 32 + tx;
            tile[r][tx] = (gi < dim && gj < dim) ? rho[gi * dim + gj] : cx<double>{0, 0};
        }
        __syncthreads();
        for (int r = ty; r < 32; r += 8) {
            const uint64_t gi = bi * 32 + r, gj = bj * 32 + tx;
            if (gi < dim && gj < dim) {
                const cx<double> a = rho[gi * dim + gj];
                const cx<double> b = tile[tx][r];  // rho[gj, gi]
                v[0] += a.x * b.x + a.y * b.y;       // a * conj(b)
                v[1] += a.y * b.x - a.x * b.y;
            }
        }
    }
    block_sum<2>(v, sh);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = v[0];
        partials[2 * blockIdx.x + 1] = v[1];
    }
}

int dm_purity(const State &rho, int n, double *out, Scratch &scr, hipStream_t s) {
    const uint64_t dim = 1ull << n;
    const uint64_t nt = (dim + 31) / 32;
    const unsigned grid = (unsigned)std::min<uint64_t>(nt * nt, kRedMaxBlocks);
    const size_t pbytes = (size_t)grid * 16;
    int rc = scr.ensure(pbytes + 64);
    if (rc) return rc;
    double *dpart = static_cast<double *>(scr.ptr);
    hipLaunchKernelGGL(k_dm_purity, dim3(grid), dim3(256), 0, s, static_cast<const cx<double> *>(rho.amps), dim, dpart);
    GOQU_HIP(hipGetLastError());
    double h[2];
    rc = finish_to_host(dpart, grid, 2, h, scr, pbytes, s);
    if (rc) return rc;
    *out = std::fabs(h[0]);
    return GOQU_OK;
}

// Fidelity (sim.go:279-291): |Re sum_ij conj(psi_i) rho_ij psi_j|; one wavefront per row
__global__ void __launch_bounds__(256) k_dm_fid(const cx<double> *__restrict__ rho, uint64_t dim,
                                                const cx<double> *__restrict__ psi, double *__restrict__ partials) {
    __shared__ double sh[8];
    double v[2] = {0, 0};
    const int lane = threadIdx.x & 63;
    const uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t i = wave; i < dim; i += nwaves) {
        double sr = 0, si = 0;
        for (uint64_t j = lane; j < dim; j += 64) {
            const cx<double> a = rho[i * dim + j], p = psi[j];
            sr += a.x * p.x - a.y * p.y;
            si += a.x * p.y + a.y * p.x;
        }
        const cx<double> c = psi[i];  // conj(c) * (sr + i si)
        v[0] += c.x * sr + c.y * si;
        v[1] += c.x * si - c.y * sr;
    }
    block_sum<2>(v, sh);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = v[0];
        partials[2 * blockIdx.x + 1] = v[1];
    }
}

int dm_fidelity(const State &rho, int n, const void *dpsi, double *out, Scratch &scr, hipStream_t s) {
    const uint64_t dim = 1ull << n;
    const unsigned grid = (unsigned)std::min<uint64_t>((dim + 3) / 4, kRedMaxBlocks);
    const size_t pbytes = (size_t)grid * 16;
    int rc = scr.ensure(pbytes + 64);
    if (rc) return rc;
    double *dpart = static_cast<double *>(scr.ptr);
    hipLaunchKernelGGL(k_dm_fid, dim3(grid), dim3(256), 0, s, static_cast<const cx<double> *>(rho.amps), dim,
                       static_cast<const cx<double> *>(dpsi), dpart);
    GOQU_HIP(hipGetLastError());
    double h[2];
    rc = finish_to_host(dpart, grid, 2, h, scr, pbytes, s);
    if (rc) return rc;
    *out = std::fabs(h[0]);
    return GOQU_OK;
}

}  // namespace goqu
