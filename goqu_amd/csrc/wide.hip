// wide.hip — dense ops on 5..9 target qubits as FP64 MFMA contractions.
//
// A wide op U (2^k x 2^k, k = 5..9, qubits qw[0] = MSB ... qw[k-1] = LSB of the matrix index)
// under a control mask acts on every column of the state viewed as a 2^k x 2^f matrix: the row
// index runs over the target qubits, the column index over the f free qubits (neither target
// nor control), and only columns whose controls are 1 take part.  This is the reference's
// controlled-gate path on the full controlled matrix (sim/statevector/kernel_controlled.go:
// 249-292), whose cost per amplitude is 2^k complex multiply-adds: at k >= 5 the contraction,
// not HBM, bounds it.
//
// One workgroup takes 16 columns: it stages their 2^k x 16 amplitudes in LDS (real and imaginary
// parts as separate fp64 matrices), each wave computes 16-row tiles of U Psi with
// v_mfma_f64_16x16x4f64 (four real products per complex step: Re = Ur Pr - Ui Pi, Im = Ur Pi +
// Ui Pr), and writes its rows back in place (every input column is in LDS before the first
// write).
#include "runtime.h"

namespace goqu {

typedef double d4 __attribute__((ext_vector_type(4)));

// Operand layouts of v_mfma_f64_16x16x4f64 (one wave, lane l):
//   A (16 x 4): A[l % 16][l / 16];  B (4 x 16): B[l / 16][l % 16];
//   D (16 x 16): register v of lane l holds D[4 v + l / 16][l % 16].
template <typename R>
__global__ void __launch_bounds__(256) k_wide_mfma(cx<R> *__restrict__ st, int k, const uint64_t *__restrict__ rowbit,
                                                   const uint8_t *__restrict__ freeq, int nfree, uint64_t cmask,
                                                   uint64_t ncols, const double *__restrict__ Ur,
                                                   const double *__restrict__ Ui, const double *__restrict__ Uin,
                                                   const uint8_t *__restrict__ act, uint64_t shot_stride) {
    if (act && !act[blockIdx.y]) return;  // a shot that this op does not reach
    st += blockIdx.y * shot_stride;
    extern __shared__ double lds[];
    const int D = 1 << k;
    double *Br = lds, *Bi = lds + (size_t)D * 16;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint64_t ngroups = (ncols + 15) / 16;
    for (uint64_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
        const uint64_t c0 = grp * 16;
        // column c -> its state index: c deposited into the free qubits, controls set
        auto col_index = [&](uint64_t c) {
            uint64_t g = cmask;
            for (int b = 0; b < nfree; b++)
                if (c >> b & 1) g |= 1ull << freeq[b];
            return g;
        };
        auto row_index = [&](int r) {
            uint64_t g = 0;
            for (int b = 0; b < k; b++)
                if (r >> b & 1) g |= rowbit[b];
            return g;
        };
        // stage the 2^k x 16 block (columns beyond ncols are zero and never written back)
        for (int e = t; e < D * 16; e += blockDim.x) {
            const int r = e >> 4, c = e & 15;
            double vr = 0, vi = 0;
            if (c0 + c < ncols) {
                const cx<R> a = st[col_index(c0 + c) | row_index(r)];
                vr = (double)a.x;
                vi = (double)a.y;
            }
            Br[e] = vr;
            Bi[e] = vi;
        }
        __syncthreads();
        const int nwaves = blockDim.x >> 6;
        const int ac = lane / 16, ar = lane % 16;  // A and B operand coordinates of this lane
        for (int rt = wave; rt < D / 16; rt += nwaves) {
            d4 accr = {0, 0, 0, 0}, acci = {0, 0, 0, 0};
            const size_t arow = (size_t)(rt * 16 + ar) * D;
            for (int kk = 0; kk < D / 4; kk++) {
                const size_t ai = arow + kk * 4 + ac;
                const double a_r = Ur[ai], a_i = Ui[ai], a_in = Uin[ai];
                const int bi = (kk * 4 + ac) * 16 + ar;
                const double b_r = Br[bi], b_i = Bi[bi];
                accr = __builtin_amdgcn_mfma_f64_16x16x4f64(a_r, b_r, accr, 0, 0, 0);
                accr = __builtin_amdgcn_mfma_f64_16x16x4f64(a_in, b_i, accr, 0, 0, 0);
                acci = __builtin_amdgcn_mfma_f64_16x16x4f64(a_r, b_i, acci, 0, 0, 0);
                acci = __builtin_amdgcn_mfma_f64_16x16x4f64(a_i, b_r, acci, 0, 0, 0);
            }
            const uint64_t c = c0 + (lane & 15);
            if (c < ncols) {
                const uint64_t gc = col_index(c);
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int r = rt * 16 + 4 * v + (lane >> 4);
                    st[gc | row_index(r)] = cx<R>{(R)accr[v], (R)acci[v]};
                }
            }
        }
        __syncthreads();  // the next group restages LDS
    }
}

int prepare_wide(int nq, const EOp &e, hipStream_t s, WideOp &w) {
    const int k = e.k;
    if (e.kind != EOP_WIDE || k < 5 || k > kMaxWideTargets || (int)e.qw.size() != k)
        return fail(GOQU_ERR_INVALID_ARG, "malformed wide op");
    const int D = 1 << k;
    // row bit b (LSB = 0) <-> qubit qw[k-1-b]; free qubits: neither targets nor controls
    std::vector<uint64_t> rowbit(k);
    uint64_t used = e.cmask;
    for (int b = 0; b < k; b++) {
        const int q = e.qw[k - 1 - b];
        if (q < 0 || q >= nq) return fail(GOQU_ERR_INVALID_ARG, "wide op qubit out of range");
        rowbit[b] = 1ull << q;
        used |= rowbit[b];
    }
    if (e.cmask >> nq) return fail(GOQU_ERR_INVALID_ARG, "wide op control out of range");
    std::vector<uint8_t> freeq;
    for (int q = 0; q < nq; q++)
        if (!(used >> q & 1)) freeq.push_back((uint8_t)q);
    // device staging: row bits | free qubits | Ur | Ui | -Ui
    const size_t mat = (size_t)D * D;
    const size_t off_f = (size_t)k * 8, off_m = (off_f + 64 + 255) & ~size_t(255);
    const size_t bytes = off_m + 3 * mat * 8;
    std::vector<unsigned char> host(bytes, 0);
    std::memcpy(host.data(), rowbit.data(), (size_t)k * 8);
    if (!freeq.empty()) std::memcpy(host.data() + off_f, freeq.data(), freeq.size());
    double *ur = reinterpret_cast<double *>(host.data() + off_m), *ui = ur + mat, *uin = ui + mat;
    for (size_t i = 0; i < mat; i++) {
        ur[i] = e.m[2 * i];
        ui[i] = e.m[2 * i + 1];
        uin[i] = -e.m[2 * i + 1];
    }
    w.release();
    GOQU_HIP(dmalloc(&w.dev, bytes));
    w.k = k;
    w.nfree = (int)freeq.size();
    w.cmask = e.cmask;
    w.ncols = 1ull << freeq.size();
    w.off_f = off_f;
    w.off_m = off_m;
    GOQU_HIP(hipMemcpyAsync(w.dev, host.data(), bytes, hipMemcpyHostToDevice, s));
    GOQU_HIP(hipStreamSynchronize(s));  // the host staging buffer is released on return
    return GOQU_OK;
}

void WideOp::release() {
    if (dev) (void)hipFree(dev);
    dev = nullptr;
}

int launch_wide(void *amps, int prec, const WideOp &w, hipStream_t s, const ShotBatch *b) {
    const int D = 1 << w.k;
    const size_t mat = (size_t)D * D;
    const char *dev = static_cast<const char *>(w.dev);
    const double *dur = reinterpret_cast<const double *>(dev + w.off_m);
    const uint64_t ngroups = (w.ncols + 15) / 16;
    const unsigned grid = (unsigned)std::min<uint64_t>(ngroups, 1u << 20);
    const unsigned shots = b ? (unsigned)b->nshots : 1u;
    const uint8_t *act = b ? b->act : nullptr;
    const uint64_t stride = b ? b->stride : 0;
    const size_t lds = (size_t)D * 16 * 2 * sizeof(double);
    const uint64_t *rb = reinterpret_cast<const uint64_t *>(dev);
    const uint8_t *fq = reinterpret_cast<const uint8_t *>(dev + w.off_f);
    hipEvent_t ev0 = prof_begin(s);
    if (prec == GOQU_C128)
        hipLaunchKernelGGL(k_wide_mfma<double>, dim3(grid, shots), dim3(256), lds, s, static_cast<cx<double> *>(amps),
                           w.k, rb, fq, w.nfree, w.cmask, w.ncols, dur, dur + mat, dur + 2 * mat, act, stride);
    else
        hipLaunchKernelGGL(k_wide_mfma<float>, dim3(grid, shots), dim3(256), lds, s, static_cast<cx<float> *>(amps),
                           w.k, rb, fq, w.nfree, w.cmask, w.ncols, dur, dur + mat, dur + 2 * mat, act, stride);
    GOQU_HIP(hipGetLastError());
    // FP64 flops of the contraction: 2^k x 2^k complex times 2^k x ncols, 8 real flops per complex
    // multiply-add, for every shot of the batch
    prof_end(ev0, s, PROF_WIDE, 8.0 * D * D * (double)w.ncols * shots);
    return GOQU_OK;
}

int run_wide(State &st, const EOp &e, hipStream_t s) {
    WideOp w;
    int rc = prepare_wide(st.nq, e, s, w);
    if (!rc) rc = launch_wide(st.amps, st.prec, w, s, nullptr);
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = fail(GOQU_ERR_DEVICE, "wide op failed");
    w.release();
    return rc;
}

}  // namespace goqu
