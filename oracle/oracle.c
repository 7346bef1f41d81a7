/*
 * oracle.c — CPU restatement of goqu's complex128 simulators.  TEST INFRASTRUCTURE ONLY.
 *
 * This file is the parity oracle: tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg are its only users, and only as the checker (or as the timed CPU
 * baseline).  The product library (goqu_amd/libgoqu_hip.so) never links or calls it.
 *
 * It restates the reference algorithms loop for loop, with the same complex arithmetic
 * order ((a*b) as (ar*br - ai*bi, ar*bi + ai*br), left-to-right sums, no FMA contraction:
 * build with -ffp-contract=off), so that the GPU results can be held to 1e-12 against it.
 * Each function cites the reference file:line it follows.  Parallel loops follow goqu's
 * worker policy (optimalWorkers, sim/statevector/sim.go:327-341): they split disjoint
 * index ranges, so results do not depend on the worker count, except the chunked partial
 * sums of expectParallel, which are restated with the same chunking.
 *
 * Pinning (what holds this restatement to the reference, besides reading): the reference's
 * own known-answer tests (tests/test_oracle_kats.py) and brute-force Kronecker-product
 * unitaries built independently in numpy (tests/test_oracle_bruteforce.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/goqu_hip.h"

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    double re, im;
} cx;

static inline cx C(double re, double im) {
    cx r = {re, im};
    return r;
}
static inline cx cmul(cx a, cx b) { return C(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline cx cadd(cx a, cx b) { return C(a.re + b.re, a.im + b.im); }
static inline cx cconj(cx a) { return C(a.re, -a.im); }
static inline cx cneg(cx a) { return C(-a.re, -a.im); }
static inline cx mat(const double *m, int i) { return C(m[2 * i], m[2 * i + 1]); }

static int g_workers = 1; /* GOMAXPROCS analogue */

void orc_set_workers(int w) { g_workers = w < 1 ? 1 : w; }

/* optimalWorkers — sim/statevector/sim.go:327-341 */
static int optimal_workers(int nq) {
    if (nq <= 16) return 1;
    long long amps = 1LL << nq;
    long long by_work = amps / 8192;
    if (by_work < 1) by_work = 1;
    return g_workers < by_work ? g_workers : (int)by_work;
}

/* ======================================================================================
 * Statevector
 * ====================================================================================== */

/* resetQubit — sim/statevector/sim.go:168-186 */
static void sv_reset_qubit(cx *s, int nq, int q) {
    const int64_t N = 1LL << nq, half = 1LL << q, block = half << 1;
    const int w = optimal_workers(nq);
#pragma omp parallel for num_threads(w) schedule(static)
    for (int64_t b0 = 0; b0 < N; b0 += block)
        for (int64_t off = 0; off < half; off++) {
            int64_t i0 = b0 + off, i1 = i0 + half;
            cx a0 = s[i0], a1 = s[i1];
            double norm = sqrt(a0.re * a0.re + a0.im * a0.im + a1.re * a1.re + a1.im * a1.im);
            s[i0] = norm > 1e-15 ? C(norm, 0) : C(0, 0);
            s[i1] = C(0, 0);
        }
}

/* applyGate1 / applyGate1Parallel — sim/statevector/sim.go:211-272 */
static void sv_gate1(cx *s, int nq, int q, const double *m) {
    const int64_t N = 1LL << nq, half = 1LL << q, block = half << 1;
    const cx m0 = mat(m, 0), m1 = mat(m, 1), m2 = mat(m, 2), m3 = mat(m, 3);
    const int w = optimal_workers(nq);
#pragma omp parallel for num_threads(w) schedule(static)
    for (int64_t b0 = 0; b0 < N; b0 += block)
        for (int64_t off = 0; off < half; off++) {
            int64_t i0 = b0 + off, i1 = i0 + half;
            cx a0 = s[i0], a1 = s[i1];
            s[i0] = cadd(cmul(m0, a0), cmul(m1, a1));
            s[i1] = cadd(cmul(m2, a0), cmul(m3, a1));
        }
}

/* The 3-level block-stride loop of blockStride2 (kernel2q.go:116-148) visits every offset
 * with both qubit bits clear, exactly once; order is irrelevant for these in-place bodies. */
#define FOR_PAIRS2(N, lo, hi, ...)                                                           \
    do {                                                                                     \
        const int64_t loM = 1LL << (lo), hiM = 1LL << (hi);                                  \
        const int w_ = optimal_workers(nq);                                                  \
        _Pragma("omp parallel for num_threads(w_) schedule(static)") for (int64_t hi0 = 0;  \
                                                                           hi0 < (N);       \
                                                                           hi0 += hiM << 1) \
            for (int64_t lo0 = hi0; lo0 < hi0 + hiM; lo0 += loM << 1)                        \
                for (int64_t offset = lo0; offset < lo0 + loM; offset++) { __VA_ARGS__ }             \
    } while (0)

/* dispatchGate2 — sim/statevector/kernel2q.go:24-114 (three tiers; nil matrix = identity) */
static void sv_gate2(cx *s, int nq, int gate_id, int q0, int q1, const double *m) {
    const int64_t N = 1LL << nq;
    const int64_t mask0 = 1LL << q0, mask1 = 1LL << q1;
    const int lo = q0 < q1 ? q0 : q1, hi = q0 < q1 ? q1 : q0;
    switch (gate_id) {
    case GOQU_GATE_CNOT: /* kernel2qCNOT kernel2q.go:151-165 */
        FOR_PAIRS2(N, lo, hi, {
            int64_t i10 = offset | mask0, i11 = offset | mask0 | mask1;
            cx t = s[i10];
            s[i10] = s[i11];
            s[i11] = t;
        });
        return;
    case GOQU_GATE_CZ: /* kernel2qCZ :168-180 */
        FOR_PAIRS2(N, lo, hi, {
            int64_t i11 = offset | mask0 | mask1;
            s[i11] = cneg(s[i11]);
        });
        return;
    case GOQU_GATE_SWAP: /* kernel2qSWAP :183-197 */
        FOR_PAIRS2(N, lo, hi, {
            int64_t i01 = offset | mask1, i10 = offset | mask0;
            cx t = s[i01];
            s[i01] = s[i10];
            s[i10] = t;
        });
        return;
    case GOQU_GATE_CY: /* kernel2qCY :200-215: |10> <- -i a3, |11> <- i a2 */
        FOR_PAIRS2(N, lo, hi, {
            int64_t i10 = offset | mask0, i11 = offset | mask0 | mask1;
            cx a2 = s[i10], a3 = s[i11];
            s[i10] = cmul(C(0, -1), a3);
            s[i11] = cmul(C(0, 1), a2);
        });
        return;
    case GOQU_GATE_ISWAP: /* kernel2qISWAP :307-322 */
        FOR_PAIRS2(N, lo, hi, {
            int64_t i01 = offset | mask1, i10 = offset | mask0;
            cx a1 = s[i01], a2 = s[i10];
            s[i01] = cmul(C(0, 1), a2);
            s[i10] = cmul(C(0, 1), a1);
        });
        return;
    case GOQU_GATE_SYCAMORE: { /* kernel2qSycamore :325-346, d11 = matrix[15] */
        const cx d11 = m ? mat(m, 15) : C(cos(M_PI / 6), -sin(M_PI / 6));
        FOR_PAIRS2(N, lo, hi, {
            int64_t i01 = offset | mask1, i10 = offset | mask0, i11 = offset | mask0 | mask1;
            cx a1 = s[i01], a2 = s[i10];
            s[i01] = cmul(C(0, -1), a2);
            s[i10] = cmul(C(0, -1), a1);
            s[i11] = cmul(s[i11], d11);
        });
        return;
    }
    default:
        break;
    }
    if (!m) return; /* opaque gate: identity (kernel2q.go:106-108) */
    if (gate_id == GOQU_GATE_CTRL_DIAG) { /* kernel2qDiagonal :222-237 (s *= d) */
        const cx d2 = mat(m, 10), d3 = mat(m, 15);
        FOR_PAIRS2(N, lo, hi, {
            int64_t i10 = offset | mask0, i11 = offset | mask0 | mask1;
            s[i10] = cmul(s[i10], d2);
            s[i11] = cmul(s[i11], d3);
        });
        return;
    }
    if (gate_id == GOQU_GATE_DIAG2) { /* kernel2qFullDiagonal :241-256 */
        const cx d0 = mat(m, 0), d1 = mat(m, 5), d2 = mat(m, 10), d3 = mat(m, 15);
        FOR_PAIRS2(N, lo, hi, {
            s[offset] = cmul(s[offset], d0);
            s[offset | mask1] = cmul(s[offset | mask1], d1);
            s[offset | mask0] = cmul(s[offset | mask0], d2);
            s[offset | mask0 | mask1] = cmul(s[offset | mask0 | mask1], d3);
        });
        return;
    }
    if (gate_id == GOQU_GATE_CTRL_U) { /* kernel2qControlled :260-276 */
        const cx u00 = mat(m, 10), u01 = mat(m, 11), u10 = mat(m, 14), u11 = mat(m, 15);
        FOR_PAIRS2(N, lo, hi, {
            int64_t i10 = offset | mask0, i11 = offset | mask0 | mask1;
            cx a2 = s[i10], a3 = s[i11];
            s[i10] = cadd(cmul(u00, a2), cmul(u01, a3));
            s[i11] = cadd(cmul(u10, a2), cmul(u11, a3));
        });
        return;
    }
    /* kernel2qGeneric :284-304 */
    cx M[16];
    for (int i = 0; i < 16; i++) M[i] = mat(m, i);
    FOR_PAIRS2(N, lo, hi, {
        int64_t i00 = offset, i01 = offset | mask1, i10 = offset | mask0, i11 = offset | mask0 | mask1;
        cx a0 = s[i00], a1 = s[i01], a2 = s[i10], a3 = s[i11];
        s[i00] = cadd(cadd(cadd(cmul(M[0], a0), cmul(M[1], a1)), cmul(M[2], a2)), cmul(M[3], a3));
        s[i01] = cadd(cadd(cadd(cmul(M[4], a0), cmul(M[5], a1)), cmul(M[6], a2)), cmul(M[7], a3));
        s[i10] = cadd(cadd(cadd(cmul(M[8], a0), cmul(M[9], a1)), cmul(M[10], a2)), cmul(M[11], a3));
        s[i11] = cadd(cadd(cadd(cmul(M[12], a0), cmul(M[13], a1)), cmul(M[14], a2)), cmul(M[15], a3));
    });
}

/* dispatchGate3 — sim/statevector/kernel3q.go:10-45 and the kernels :90-280.  The block-stride
 * loops enumerate the offsets with all three bits clear; here they are enumerated directly. */
static void sv_gate3(cx *s, int nq, int gate_id, int q0, int q1, int q2, const double *m) {
    const int64_t N = 1LL << nq;
    const int64_t m0 = 1LL << q0, m1 = 1LL << q1, m2 = 1LL << q2, all = m0 | m1 | m2;
    const int w = optimal_workers(nq);
    if (gate_id == GOQU_GATE_CCX) { /* kernel3qCCX: swap |110> <-> |111> */
#pragma omp parallel for num_threads(w) schedule(static)
        for (int64_t off = 0; off < N; off++) {
            if (off & all) continue;
            int64_t a = off | m0 | m1, b = off | m0 | m1 | m2;
            cx t = s[a];
            s[a] = s[b];
            s[b] = t;
        }
        return;
    }
    if (gate_id == GOQU_GATE_CSWAP) { /* kernel3qCSWAP: swap |101> <-> |110> */
#pragma omp parallel for num_threads(w) schedule(static)
        for (int64_t off = 0; off < N; off++) {
            if (off & all) continue;
            int64_t a = off | m0 | m2, b = off | m0 | m1;
            cx t = s[a];
            s[a] = s[b];
            s[b] = t;
        }
        return;
    }
    if (gate_id == GOQU_GATE_CCZ) { /* kernel3qCCZ: negate |111> */
#pragma omp parallel for num_threads(w) schedule(static)
        for (int64_t off = 0; off < N; off++) {
            if (off & all) continue;
            s[off | all] = cneg(s[off | all]);
        }
        return;
    }
    if (!m) return; /* opaque: identity (kernel3q.go:38-41) */
    cx M[64];
    for (int i = 0; i < 64; i++) M[i] = mat(m, i);
    /* kernel3qGeneric :145-175: local index r: bit2 <-> q0, bit1 <-> q1, bit0 <-> q2 */
#pragma omp parallel for num_threads(w) schedule(static)
    for (int64_t off = 0; off < N; off++) {
        if (off & all) continue;
        int64_t idx[8];
        cx a[8];
        for (int r = 0; r < 8; r++) {
            int64_t i = off;
            if (r & 4) i |= m0;
            if (r & 2) i |= m1;
            if (r & 1) i |= m2;
            idx[r] = i;
            a[r] = s[i];
        }
        for (int r = 0; r < 8; r++) {
            cx sum = C(0, 0);
            for (int c = 0; c < 8; c++) sum = cadd(sum, cmul(M[r * 8 + c], a[c]));
            s[idx[r]] = sum;
        }
    }
}

/* Full controlled matrix of a ControlledGate — circuit/gate/controlled.go:84-110: identity
 * except the bottom-right inner block. */
static void controlled_full_matrix(const double *inner, int nc, int nt, cx *out) {
    const int dim = 1 << (nc + nt), idim = 1 << nt, off = dim - idim;
    for (int i = 0; i < dim * dim; i++) out[i] = C(0, 0);
    for (int i = 0; i < off; i++) out[i * dim + i] = C(1, 0);
    for (int r = 0; r < idim; r++)
        for (int c = 0; c < idim; c++) out[(off + r) * dim + off + c] = mat(inner, r * idim + c);
}

/* dispatchControlled — sim/statevector/kernel_controlled.go:11-292 */
static int sv_controlled(cx *s, int nq, const goqu_op_t *op) {
    const int nc = op->num_controls, nt = op->num_qubits - nc;
    const int64_t N = 1LL << nq;
    const double *m = op->matrix;
    const int w = optimal_workers(nq);
    if (!m) return GOQU_OK; /* opaque inner: identity */
    int64_t cmask = 0;
    for (int i = 0; i < nc; i++) cmask |= 1LL << op->qubits[i];
    if (nt == 1) { /* applyControlledGate1 :52-91 */
        const int64_t tb = 1LL << op->qubits[nc];
        const cx m0 = mat(m, 0), m1 = mat(m, 1), m2 = mat(m, 2), m3 = mat(m, 3);
#pragma omp parallel for num_threads(w) schedule(static)
        for (int64_t i = 0; i < N; i++) {
            if ((i & cmask) != cmask || (i & tb)) continue;
            cx a0 = s[i], a1 = s[i | tb];
            s[i] = cadd(cmul(m0, a0), cmul(m1, a1));
            s[i | tb] = cadd(cmul(m2, a0), cmul(m3, a1));
        }
        return GOQU_OK;
    }
    if (nt == 2) { /* applyControlledGate2 :138-175 */
        const int64_t mk0 = 1LL << op->qubits[nc], mk1 = 1LL << op->qubits[nc + 1], tm = mk0 | mk1;
        cx M[16];
        for (int i = 0; i < 16; i++) M[i] = mat(m, i);
#pragma omp parallel for num_threads(w) schedule(static)
        for (int64_t i = 0; i < N; i++) {
            if ((i & cmask) != cmask || (i & tm)) continue;
            int64_t idx[4] = {i, i | mk1, i | mk0, i | mk0 | mk1};
            cx a[4];
            for (int j = 0; j < 4; j++) a[j] = s[idx[j]];
            for (int r = 0; r < 4; r++) {
                cx sum = C(0, 0);
                for (int c = 0; c < 4; c++) sum = cadd(sum, cmul(M[r * 4 + c], a[c]));
                s[idx[r]] = sum;
            }
        }
        return GOQU_OK;
    }
    /* applyControlledGateN :249-292 (full controlled matrix, at most 10 qubits:
     * circuit/gate/controlled.go:87-89) */
    const int tot = op->num_qubits;
    if (tot > 10) return GOQU_ERR_UNSUPPORTED;
    const int dim = 1 << tot;
    cx *M = (cx *)malloc(sizeof(cx) * dim * dim);
    controlled_full_matrix(m, nc, nt, M);
    int64_t masks[16], all = 0;
    for (int i = 0; i < tot; i++) {
        masks[i] = 1LL << op->qubits[i];
        all |= masks[i];
    }
#pragma omp parallel num_threads(w)
    {
        int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * dim);
        cx *a = (cx *)malloc(sizeof(cx) * dim);
#pragma omp for schedule(static)
        for (int64_t base = 0; base < N; base++) {
            if (base & all) continue;
            for (int r = 0; r < dim; r++) {
                int64_t i = base;
                for (int b = 0; b < tot; b++)
                    if (r & (1 << (tot - 1 - b))) i |= masks[b];
                idx[r] = i;
            }
            for (int j = 0; j < dim; j++) a[j] = s[idx[j]];
            for (int r = 0; r < dim; r++) {
                cx sum = C(0, 0);
                for (int c = 0; c < dim; c++) sum = cadd(sum, cmul(M[r * dim + c], a[c]));
                s[idx[r]] = sum;
            }
        }
        free(idx);
        free(a);
    }
    free(M);
    return GOQU_OK;
}

/* Apply — sim/statevector/sim.go:92-157 (StatePrep is resolved by the host beforehand). */
int orc_sv_apply(double *state, int nq, const goqu_op_t *ops, int nops) {
    cx *s = (cx *)state;
    for (int k = 0; k < nops; k++) {
        const goqu_op_t *op = &ops[k];
        switch (op->kind) {
        case GOQU_OP_SKIP:
            continue;
        case GOQU_OP_RESET:
            sv_reset_qubit(s, nq, op->qubits[0]);
            continue;
        case GOQU_OP_CONTROLLED: {
            int rc = sv_controlled(s, nq, op);
            if (rc) return rc;
            continue;
        }
        case GOQU_OP_GATE:
            break;
        default:
            return GOQU_ERR_INVALID_ARG;
        }
        switch (op->num_qubits) {
        case 1:
            if (!op->matrix) return GOQU_ERR_NO_MATRIX;
            sv_gate1(s, nq, op->qubits[0], op->matrix);
            break;
        case 2:
            sv_gate2(s, nq, op->gate_id, op->qubits[0], op->qubits[1], op->matrix);
            break;
        case 3:
            sv_gate3(s, nq, op->gate_id, op->qubits[0], op->qubits[1], op->qubits[2], op->matrix);
            break;
        default:
            return GOQU_ERR_UNSUPPORTED; /* sim.go:152 */
        }
    }
    return GOQU_OK;
}

/* probabilities + sampleIndex — sim/statevector/sim.go:274-298: sequential cumulative sum,
 * first index with r < cum, else the last index. */
void orc_sample_probs(const double *probs, int64_t N, const double *u, int64_t shots, uint64_t *out) {
    for (int64_t k = 0; k < shots; k++) {
        double r = u[k], cum = 0.0;
        int64_t hit = N - 1;
        for (int64_t i = 0; i < N; i++) {
            cum += probs[i];
            if (r < cum) {
                hit = i;
                break;
            }
        }
        out[k] = (uint64_t)hit;
    }
}

void orc_sv_probabilities(const double *state, int nq, double *probs) {
    const int64_t N = 1LL << nq;
    for (int64_t i = 0; i < N; i++) probs[i] = state[2 * i] * state[2 * i] + state[2 * i + 1] * state[2 * i + 1];
}

void orc_sv_sample(const double *state, int nq, const double *u, int64_t shots, uint64_t *out) {
    const int64_t N = 1LL << nq;
    double *p = (double *)malloc(sizeof(double) * N);
    orc_sv_probabilities(state, nq, p);
    orc_sample_probs(p, N, u, shots, out);
    free(p);
}

/* iPow — sim/pauli/pauli.go:212-224 */
static cx ipow(int n) {
    switch (((n % 4) + 4) % 4) {
    case 1: return C(0, 1);
    case 2: return C(-1, 0);
    case 3: return C(0, -1);
    default: return C(1, 0);
    }
}

/* sign — sim/pauli/pauli.go:226-232 */
static inline double psign(uint64_t v) { return (__builtin_popcountll(v) & 1) ? -1.0 : 1.0; }

/* pauli.Expect + expectParallel — sim/pauli/expect_sv.go:21-88 */
void orc_sv_expect(const double *state, int nq, uint64_t xm, uint64_t zm, int ny, double cre, double cim,
                   double *out) {
    const cx *s = (const cx *)state;
    const int64_t N = 1LL << nq;
    const cx yph = ipow(ny);
    cx total = C(0, 0);
    if (nq >= 17) {
        int nw = optimal_workers(nq);
        if (nw < 1) nw = 1;
        int64_t chunk = N / nw;
        cx *partials = (cx *)calloc(nw, sizeof(cx));
#pragma omp parallel for num_threads(nw) schedule(static, 1)
        for (int w = 0; w < nw; w++) {
            int64_t start = w * chunk, end = (w == nw - 1) ? N : start + chunk;
            cx acc = C(0, 0);
            for (int64_t i = start; i < end; i++) {
                cx sj = s[i ^ xm];
                acc = cadd(acc, cmul(cmul(cmul(cconj(sj), C(psign(i & zm), 0)), yph), s[i]));
            }
            partials[w] = acc;
        }
        for (int w = 0; w < nw; w++) total = cadd(total, partials[w]);
        free(partials);
    } else {
        for (int64_t i = 0; i < N; i++) {
            cx sj = s[i ^ xm];
            total = cadd(total, cmul(cmul(cmul(cconj(sj), C(psign(i & zm), 0)), yph), s[i]));
        }
    }
    cx r = cmul(C(cre, cim), total);
    out[0] = r.re;
    out[1] = r.im;
}

/* ExpectationValue — sim/statevector/sim.go:368-383 (rounded to 1e-14) */
double orc_sv_expectation_value(const double *state, int nq, uint64_t mask, int round14) {
    const int64_t N = 1LL << nq;
    double ev = 0;
    for (int64_t i = 0; i < N; i++) {
        double p = state[2 * i] * state[2 * i] + state[2 * i + 1] * state[2 * i + 1];
        if (__builtin_popcountll((uint64_t)i & mask) % 2 == 0)
            ev += p;
        else
            ev -= p;
    }
    return round14 ? round(ev * 1e14) / 1e14 : ev;
}

/* measureQubit — sim/statevector/dynamic.go:186-230 */
int orc_sv_measure_qubit(double *state, int nq, int q, double r) {
    cx *s = (cx *)state;
    const int64_t N = 1LL << nq, half = 1LL << q, block = half << 1;
    double prob0 = 0.0;
    for (int64_t b0 = 0; b0 < N; b0 += block)
        for (int64_t off = 0; off < half; off++) {
            cx a = s[b0 + off];
            prob0 += a.re * a.re + a.im * a.im;
        }
    int outcome = (r >= prob0) ? 1 : 0;
    double norm = outcome == 0 ? sqrt(prob0) : sqrt(1 - prob0);
    if (norm < 1e-15) return outcome;
    double inv = 1.0 / norm;
    cx scale = C(inv, 0);
    for (int64_t b0 = 0; b0 < N; b0 += block)
        for (int64_t off = 0; off < half; off++) {
            int64_t i0 = b0 + off, i1 = i0 + half;
            if (outcome == 0) {
                s[i0] = cmul(s[i0], scale);
                s[i1] = C(0, 0);
            } else {
                s[i0] = C(0, 0);
                s[i1] = cmul(s[i1], scale);
            }
        }
    return outcome;
}

/* ======================================================================================
 * Density matrix
 * ====================================================================================== */

/* applyGate1 = gate1Rows then gate1Cols — sim/densitymatrix/kernel1q.go:10-67 */
static void dm_gate1(cx *rho, int nq, int q, const double *m) {
    const int64_t dim = 1LL << nq, half = 1LL << q, block = half << 1;
    const cx m0 = mat(m, 0), m1 = mat(m, 1), m2 = mat(m, 2), m3 = mat(m, 3);
    for (int64_t b0 = 0; b0 < dim; b0 += block)
        for (int64_t off = 0; off < half; off++) {
            int64_t r0 = b0 + off, r1 = r0 + half;
            for (int64_t c = 0; c < dim; c++) {
                cx a0 = rho[r0 * dim + c], a1 = rho[r1 * dim + c];
                rho[r0 * dim + c] = cadd(cmul(m0, a0), cmul(m1, a1));
                rho[r1 * dim + c] = cadd(cmul(m2, a0), cmul(m3, a1));
            }
        }
    const cx u0 = cconj(m0), u1 = cconj(m1), u2 = cconj(m2), u3 = cconj(m3);
    for (int64_t r = 0; r < dim; r++) {
        cx *row = rho + r * dim;
        for (int64_t b0 = 0; b0 < dim; b0 += block)
            for (int64_t off = 0; off < half; off++) {
                int64_t c0 = b0 + off, c1 = c0 + half;
                cx a0 = row[c0], a1 = row[c1];
                row[c0] = cadd(cmul(u0, a0), cmul(u1, a1));
                row[c1] = cadd(cmul(u2, a0), cmul(u3, a1));
            }
    }
}

/* applyGate2 = gate2Rows then gate2Cols — sim/densitymatrix/kernel2q.go:7-110 */
static void dm_gate2(cx *rho, int nq, int q0, int q1, const double *m) {
    const int64_t dim = 1LL << nq, mask0 = 1LL << q0, mask1 = 1LL << q1, both = mask0 | mask1;
    cx M[16], U[16];
    for (int i = 0; i < 16; i++) {
        M[i] = mat(m, i);
        U[i] = cconj(M[i]);
    }
    for (int64_t off = 0; off < dim; off++) {
        if (off & both) continue;
        int64_t r[4] = {off, off | mask1, off | mask0, off | mask0 | mask1};
        for (int64_t c = 0; c < dim; c++) {
            cx a[4];
            for (int j = 0; j < 4; j++) a[j] = rho[r[j] * dim + c];
            for (int i = 0; i < 4; i++)
                rho[r[i] * dim + c] = cadd(cadd(cadd(cmul(M[4 * i], a[0]), cmul(M[4 * i + 1], a[1])),
                                                cmul(M[4 * i + 2], a[2])),
                                           cmul(M[4 * i + 3], a[3]));
        }
    }
    for (int64_t row = 0; row < dim; row++) {
        cx *rw = rho + row * dim;
        for (int64_t off = 0; off < dim; off++) {
            if (off & both) continue;
            int64_t c[4] = {off, off | mask1, off | mask0, off | mask0 | mask1};
            cx a[4];
            for (int j = 0; j < 4; j++) a[j] = rw[c[j]];
            for (int i = 0; i < 4; i++)
                rw[c[i]] = cadd(cadd(cadd(cmul(U[4 * i], a[0]), cmul(U[4 * i + 1], a[1])), cmul(U[4 * i + 2], a[2])),
                                cmul(U[4 * i + 3], a[3]));
        }
    }
}

/* resetQubit — sim/densitymatrix/sim.go:195-217 */
static void dm_reset_qubit(cx *rho, int nq, int q) {
    const int64_t dim = 1LL << nq, mask = 1LL << q;
    for (int64_t r = 0; r < dim; r++) {
        if ((r >> q) & 1) continue;
        for (int64_t c = 0; c < dim; c++) {
            if ((c >> q) & 1) continue;
            rho[r * dim + c] = cadd(rho[r * dim + c], rho[(r | mask) * dim + (c | mask)]);
        }
    }
    for (int64_t r = 0; r < dim; r++)
        for (int64_t c = 0; c < dim; c++)
            if (((r >> q) & 1) || ((c >> q) & 1)) rho[r * dim + c] = C(0, 0);
}

/* measureQubitDM — sim/densitymatrix/dynamic.go:211-255 */
int orc_dm_measure_qubit(double *rho_d, int nq, int q, double r) {
    cx *rho = (cx *)rho_d;
    const int64_t dim = 1LL << nq, mask = 1LL << q;
    double prob0 = 0.0;
    for (int64_t i = 0; i < dim; i++)
        if ((i & mask) == 0) prob0 += rho[i * dim + i].re;
    int outcome = (r >= prob0) ? 1 : 0;
    double p = outcome == 0 ? prob0 : 1 - prob0;
    if (p < 1e-15) return outcome;
    cx scale = C(1.0 / p, 0);
    for (int64_t rr = 0; rr < dim; rr++) {
        int rb = (int)((rr >> q) & 1);
        for (int64_t c = 0; c < dim; c++) {
            int cb = (int)((c >> q) & 1);
            int64_t idx = rr * dim + c;
            if (rb == outcome && cb == outcome)
                rho[idx] = cmul(rho[idx], scale);
            else
                rho[idx] = C(0, 0);
        }
    }
    return outcome;
}

/* applyChannel — sim/densitymatrix/sim.go:220-255 */
static void dm_channel(cx *rho, int nq, const goqu_op_t *op) {
    const int k = op->num_qubits, d = 1 << k;
    const int64_t n = 1LL << (2 * nq);
    if (op->num_kraus == 1) {
        if (k == 1) dm_gate1(rho, nq, op->qubits[0], op->matrix);
        if (k == 2) dm_gate2(rho, nq, op->qubits[0], op->qubits[1], op->matrix);
        return;
    }
    cx *temp = (cx *)calloc(n, sizeof(cx));
    cx *saved = (cx *)malloc(n * sizeof(cx));
    memcpy(saved, rho, n * sizeof(cx));
    for (int e = 0; e < op->num_kraus; e++) {
        const double *ek = op->matrix + (size_t)2 * d * d * e;
        memcpy(rho, saved, n * sizeof(cx));
        if (k == 1) dm_gate1(rho, nq, op->qubits[0], ek);
        if (k == 2) dm_gate2(rho, nq, op->qubits[0], op->qubits[1], ek);
        for (int64_t i = 0; i < n; i++) temp[i] = cadd(temp[i], rho[i]);
    }
    memcpy(rho, temp, n * sizeof(cx));
    free(temp);
    free(saved);
}

/* DM Apply — sim/densitymatrix/sim.go:87-191 for the op kinds the host hands down. */
int orc_dm_apply(double *rho_d, int nq, const goqu_op_t *ops, int nops) {
    cx *rho = (cx *)rho_d;
    for (int k = 0; k < nops; k++) {
        const goqu_op_t *op = &ops[k];
        switch (op->kind) {
        case GOQU_OP_SKIP:
            break;
        case GOQU_OP_RESET:
            dm_reset_qubit(rho, nq, op->qubits[0]);
            break;
        case GOQU_OP_CHANNEL:
            if (op->num_qubits < 1 || op->num_qubits > 2) return GOQU_ERR_UNSUPPORTED;
            dm_channel(rho, nq, op);
            break;
        case GOQU_OP_GATE:
            if (!op->matrix) break; /* sim.go:147-150: nil matrix skipped */
            if (op->num_qubits == 1)
                dm_gate1(rho, nq, op->qubits[0], op->matrix);
            else if (op->num_qubits == 2)
                dm_gate2(rho, nq, op->qubits[0], op->qubits[1], op->matrix);
            else
                return GOQU_ERR_UNSUPPORTED;
            break;
        default:
            return GOQU_ERR_INVALID_ARG;
        }
    }
    return GOQU_OK;
}

/* diagonalProbs — sim/densitymatrix/measure.go:11-21 */
void orc_dm_diag_probs(const double *rho, int nq, double *probs) {
    const int64_t dim = 1LL << nq;
    for (int64_t i = 0; i < dim; i++) {
        double p = rho[2 * (i * dim + i)];
        probs[i] = p < 0 ? 0 : p;
    }
}

/* ApplyReadoutError — sim/densitymatrix/measure.go:29-71 (p01/p10 NULL => no error on any
 * qubit; entries < 0 mark a qubit without readout error). */
void orc_dm_readout(const double *probs, int nq, const double *p01, const double *p10, double *out) {
    const int64_t dim = 1LL << nq;
    double *res = (double *)calloc(dim, sizeof(double));
    for (int64_t i = 0; i < dim; i++)
        if (probs[i] != 0) res[i] += probs[i];
    if (p01 && p10) {
        double *nr = (double *)malloc(dim * sizeof(double));
        for (int q = 0; q < nq; q++) {
            if (p01[q] < 0) continue;
            memset(nr, 0, dim * sizeof(double));
            const int64_t mask = 1LL << q;
            for (int64_t i = 0; i < dim; i++) {
                if (res[i] == 0) continue;
                int64_t f = i ^ mask;
                if (!(i & mask)) {
                    nr[i] += res[i] * (1 - p01[q]);
                    nr[f] += res[i] * p01[q];
                } else {
                    nr[i] += res[i] * (1 - p10[q]);
                    nr[f] += res[i] * p10[q];
                }
            }
            memcpy(res, nr, dim * sizeof(double));
        }
        free(nr);
    }
    memcpy(out, res, dim * sizeof(double));
    free(res);
}

/* pauli.ExpectDM — sim/pauli/expect_dm.go:9-21 */
void orc_dm_expect(const double *rho_d, int nq, uint64_t xm, uint64_t zm, int ny, double cre, double cim,
                   double *out) {
    const cx *rho = (const cx *)rho_d;
    const int64_t dim = 1LL << nq;
    const cx yph = ipow(ny);
    cx sum = C(0, 0);
    for (int64_t i = 0; i < dim; i++) {
        int64_t j = i ^ (int64_t)xm;
        sum = cadd(sum, cmul(cmul(rho[i * dim + j], C(psign((uint64_t)i & zm), 0)), yph));
    }
    cx r = cmul(C(cre, cim), sum);
    out[0] = r.re;
    out[1] = r.im;
}

/* ExpectationValue — sim/densitymatrix/measure.go:76-91 */
double orc_dm_expectation_value(const double *rho, int nq, uint64_t mask, int round14) {
    const int64_t dim = 1LL << nq;
    double ev = 0;
    for (int64_t i = 0; i < dim; i++) {
        double p = rho[2 * (i * dim + i)];
        if (__builtin_popcountll((uint64_t)i & mask) % 2 == 0)
            ev += p;
        else
            ev -= p;
    }
    return round14 ? round(ev * 1e14) / 1e14 : ev;
}

/* Purity — sim/densitymatrix/sim.go:265-276: |Re sum_ij rho_ij conj(rho_ji)| */
double orc_dm_purity(const double *rho_d, int nq) {
    const cx *rho = (const cx *)rho_d;
    const int64_t dim = 1LL << nq;
    cx tr = C(0, 0);
    for (int64_t i = 0; i < dim; i++)
        for (int64_t j = 0; j < dim; j++) tr = cadd(tr, cmul(rho[i * dim + j], cconj(rho[j * dim + i])));
    return fabs(tr.re);
}

/* Fidelity — sim/densitymatrix/sim.go:279-291 */
double orc_dm_fidelity(const double *rho_d, int nq, const double *pure_d) {
    const cx *rho = (const cx *)rho_d, *ps = (const cx *)pure_d;
    const int64_t dim = 1LL << nq;
    cx f = C(0, 0);
    for (int64_t i = 0; i < dim; i++)
        for (int64_t j = 0; j < dim; j++) f = cadd(f, cmul(cmul(cconj(ps[i]), rho[i * dim + j]), ps[j]));
    return fabs(f.re);
}

/* ======================================================================================
 * Host memory bandwidth, to label the CPU baseline: a STREAM-style copy b[i] = a[i] of n
 * doubles by `threads` OpenMP threads (first touch in parallel), best of `reps`, in GB/s of
 * bytes read + written.  Not part of the reference.
 * ====================================================================================== */
double orc_stream_copy(int64_t n, int threads, int reps) {
    double *a = (double *)malloc((size_t)n * sizeof(double));
    double *b = (double *)malloc((size_t)n * sizeof(double));
    if (!a || !b) {
        free(a);
        free(b);
        return 0.0;
    }
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t i = 0; i < n; i++) {
        a[i] = (double)i;
        b[i] = 0.0;
    }
    double best = 1e30;
    for (int r = 0; r < reps; r++) {
        double t0 = omp_get_wtime();
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int64_t i = 0; i < n; i++) b[i] = a[i];
        double dt = omp_get_wtime() - t0;
        if (dt < best) best = dt;
    }
    double sink = b[n / 2];
    free(a);
    free(b);
    return sink >= 0 ? 2.0 * (double)n * sizeof(double) / best / 1e9 : 0.0;
}
