/*
 * goqu_hip.h — C ABI of libgoqu_hip, the MI355X (gfx950) simulation backend for goqu.
 *
 * Plain C types only (pointers, sizes, status codes): this is what a Go cgo package,
 * Python ctypes or C++ binds.  Every entry point returns a status (GOQU_OK == 0);
 * goqu_last_error() returns the calling thread's last message.
 *
 * Conventions (identical to goqu's CPU simulators):
 *   - basis index bit k <-> qubit k (qubit 0 = least significant bit)
 *     (sim/statevector/doc.go:4-14, sim/statevector/sim.go:188-230);
 *   - gate matrices are flat row-major (2^k x 2^k) with qubits[0] as the MOST significant
 *     bit of the gate-local index (circuit/gate/fixed.go:169-171,316-317);
 *   - complex numbers are interleaved doubles {re, im};
 *   - density matrices are row-major rho[r*dim + c] (sim/densitymatrix/sim.go:24-38).
 *
 * Handles are not thread-safe: one handle is used from one thread at a time, but any
 * number of handles may live concurrently (each owns its own HIP stream), matching the
 * per-goroutine simulator instances of goqu (algorithm/gradient/gradient.go:21-24).
 */
#ifndef GOQU_HIP_H
#define GOQU_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------- */
#define GOQU_OK 0
#define GOQU_ERR_INVALID_ARG 1   /* bad qubit index, size mismatch, null pointer ...     */
#define GOQU_ERR_OUT_OF_MEMORY 2 /* device allocation failed                             */
#define GOQU_ERR_UNSUPPORTED 3   /* e.g. >=4-qubit non-controlled gate (sim.go:152)      */
#define GOQU_ERR_DEVICE 4        /* HIP runtime error                                    */
#define GOQU_ERR_NO_DEVICE 5     /* no HIP device visible                                */
#define GOQU_ERR_NO_MATRIX 6     /* 1-qubit gate without a matrix (sim.go:139-142)       */

/* ---- precisions ----------------------------------------------------------------------- */
#define GOQU_C128 0 /* complex128 state (fp64 arithmetic)                                  */
#define GOQU_C64 1  /* complex64 state, fp32 arithmetic, fp64 reductions                  */

/* ---- op kinds --------------------------------------------------------------------------- */
#define GOQU_OP_GATE 0       /* unitary on qubits[0..k-1]; matrix (2^k)^2, may be NULL (opaque) */
#define GOQU_OP_CONTROLLED 1 /* qubits = controls ++ targets; matrix = inner target matrix      */
#define GOQU_OP_RESET 2      /* Evolve-path reset of qubits[0] (sim/statevector/sim.go:168-186) */
#define GOQU_OP_SKIP 3       /* measurement / barrier / delay: no effect on the state           */
#define GOQU_OP_CHANNEL 4    /* density matrix only: num_kraus Kraus matrices, (2^k)^2 each     */

/* ---- gate identity tier (mirrors dispatchGate2/dispatchGate3's pointer and interface
 *      tiers, sim/statevector/kernel2q.go:24-114, kernel3q.go:10-45) ----------------------- */
#define GOQU_GATE_GENERIC 0
#define GOQU_GATE_CNOT 1
#define GOQU_GATE_CZ 2
#define GOQU_GATE_SWAP 3
#define GOQU_GATE_CY 4
#define GOQU_GATE_ISWAP 5
#define GOQU_GATE_SYCAMORE 6
#define GOQU_GATE_CCX 7
#define GOQU_GATE_CSWAP 8
#define GOQU_GATE_CCZ 9
#define GOQU_GATE_CTRL_DIAG 10 /* gate.ControlDiagonal2Q: diag(1,1,m[10],m[15])  (gate.go:56-62) */
#define GOQU_GATE_DIAG2 11     /* gate.Diagonal2Q: diag(m[0],m[5],m[10],m[15])   (gate.go:64-69) */
#define GOQU_GATE_CTRL_U 12    /* gate.ControlU2Q: 2x2 block m[10,11,14,15]      (gate.go:71-76) */

#define GOQU_MAX_OP_QUBITS 16

/* One operation of a compiled circuit.  Mirrors ir.Operation (circuit/ir/circuit.go:79-85)
 * after the host has fetched gate.Matrix() and classified the gate. */
typedef struct goqu_op {
    int32_t kind;         /* GOQU_OP_*                                               */
    int32_t gate_id;      /* GOQU_GATE_*                                             */
    int32_t num_qubits;   /* entries of qubits[] in use                               */
    int32_t num_controls; /* GOQU_OP_CONTROLLED: leading control qubits               */
    int32_t qubits[GOQU_MAX_OP_QUBITS];
    const double *matrix; /* interleaved complex; borrowed for the call only          */
    int32_t num_kraus;    /* GOQU_OP_CHANNEL: number of Kraus matrices in matrix[]     */
    int32_t reserved;
} goqu_op_t;

/* ---- library ------------------------------------------------------------------------------ */
const char *goqu_last_error(void);
int goqu_device_count(int32_t *out);
/* Architecture name of the current device (e.g. "gfx950:sramecc+:xnack-"), NUL-terminated in
 * buf[0 .. len); the autotuner's cache key. */
int goqu_device_name(char *buf, int32_t len);

/* ---- statevector (replaces sim/statevector.Sim, sim/statevector/sim.go:21-383) ------------ */
typedef struct goqu_sv goqu_sv_t;

/* statevector.New (sim.go:27-35): state |0...0>.  num_qubits in [1, 34] (device memory
 * permitting; the CPU reference stops at 28). */
int goqu_sv_create(int32_t num_qubits, int32_t precision, goqu_sv_t **out);
int goqu_sv_destroy(goqu_sv_t *h);                                   /* Close  sim.go:68    */
int goqu_sv_reset(goqu_sv_t *h);                                     /* Reset  sim.go:71-77 */
/* Apply (sim.go:92-157): ops in order, without resetting.  Evolve = reset + apply. */
int goqu_sv_apply(goqu_sv_t *h, const goqu_op_t *ops, int32_t num_ops);
/* Tile shape of this handle's Apply programs (no counterpart in the reference; the autotuner
 * sets it): 2^tile_bits amplitudes per tile, of which the low_bits lowest qubits are in every
 * tile.  tile_bits in 8..13, low_bits in 2..min(6, tile_bits - 4); (0, 0) restores the default. */
int goqu_sv_set_plan(goqu_sv_t *h, int32_t tile_bits, int32_t low_bits);
/* StateVector (sim.go:61-65) as a chunked copy-out; set_state is the StatePrep fast path
 * (sim.go:105-120). amps holds 2*count doubles. */
int goqu_sv_get_state(goqu_sv_t *h, double *amps, uint64_t offset, uint64_t count);
int goqu_sv_set_state(goqu_sv_t *h, const double *amps, uint64_t offset, uint64_t count);
/* Run's sampling (sim.go:46-57, 274-298): out_idx[s] = first i with u[s] < cum_i, where cum
 * is the FP64 prefix sum of |a_i|^2; falls back to the last index. */
int goqu_sv_sample(goqu_sv_t *h, const double *uniforms, int64_t shots, uint64_t *out_idx);
/* goqu_sv_sample without any device allocation.  No counterpart in the reference: for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same result as goqu_sv_sample for the same uniforms.  Its buffers (chunk prefix
 * sums, staged uniforms, indices) live in an 8 MiB static device array, one per device, whose
 * address is looked up once per device; a call needing more than 8 MiB (16 * 2^n / 4096 +
 * 16 * shots bytes, each part rounded up to 256) returns GOQU_ERR_INVALID_ARG.  The handle's
 * device must be current.  Calls that overlap on one device share the array and race. */
int goqu_sv_sample_noalloc(goqu_sv_t *h, const double *uniforms, int64_t shots, uint64_t *out_idx);
/* pauli.Expect per term (sim/pauli/expect_sv.go:21-47): out[2t..2t+1] = coeff_t *
 * sum_i conj(psi[i^x_t]) (-1)^{|i&z_t|} i^{ny_t} psi[i].  Terms sharing an x-mask share a
 * sweep of the state. */
int goqu_sv_expect_paulis(goqu_sv_t *h, int32_t num_terms, const uint64_t *xmask,
                          const uint64_t *zmask, const int32_t *num_y, const double *coeffs,
                          double *out);
/* ExpectationValue before rounding (sim.go:368-383): sum_i |a_i|^2 (-1)^{|i&mask|}. */
int goqu_sv_expectation_z(goqu_sv_t *h, uint64_t mask, double *out);
/* measureQubit (sim/statevector/dynamic.go:186-230): P0, outcome = (u >= P0), collapse and
 * renormalise; norm < 1e-15 leaves the state untouched. */
int goqu_sv_measure_qubit(goqu_sv_t *h, int32_t qubit, double u, int32_t *outcome);
/* Squared norm sum_i |a_i|^2 in FP64. */
int goqu_sv_norm2(goqu_sv_t *h, double *out);

/* ---- density matrix (replaces sim/densitymatrix.Sim, sim/densitymatrix/sim.go:24-313) ----- */
typedef struct goqu_dm goqu_dm_t;

/* densitymatrix.New (sim.go:32-38): rho = |0><0|.  num_qubits in [1, 16]. */
int goqu_dm_create(int32_t num_qubits, goqu_dm_t **out);
int goqu_dm_destroy(goqu_dm_t *h);
int goqu_dm_reset(goqu_dm_t *h);
/* Apply (sim.go:87-191) over GATE (1 or 2 qubits), CHANNEL (applyChannel sim.go:220-255) and
 * RESET (resetQubit sim.go:195-217) ops; the host resolves noise lookups and decomposes
 * >=3-qubit gates beforehand, exactly as the reference does on the host. */
int goqu_dm_apply(goqu_dm_t *h, const goqu_op_t *ops, int32_t num_ops);
int goqu_dm_get(goqu_dm_t *h, double *rho, uint64_t offset, uint64_t count);
int goqu_dm_set(goqu_dm_t *h, const double *rho, uint64_t offset, uint64_t count);
/* diagonalProbs (measure.go:11-21) followed by ApplyReadoutError (measure.go:29-71):
 * p01/p10 are per-qubit readout flip probabilities (NULL = no readout error). */
int goqu_dm_probs(goqu_dm_t *h, const double *p01, const double *p10, double *out);
int goqu_dm_expect_paulis(goqu_dm_t *h, int32_t num_terms, const uint64_t *xmask,
                          const uint64_t *zmask, const int32_t *num_y, const double *coeffs,
                          double *out);                               /* expect_dm.go:9-30 */
int goqu_dm_expectation_z(goqu_dm_t *h, uint64_t mask, double *out); /* measure.go:76-91  */
int goqu_dm_purity(goqu_dm_t *h, double *out);                       /* sim.go:265-276    */
int goqu_dm_fidelity(goqu_dm_t *h, const double *pure, double *out); /* sim.go:279-291    */
/* Inverse-CDF sampling over given probabilities (DM Run, sim.go:51-67). */
int goqu_dm_sample(goqu_dm_t *h, const double *probs, const double *uniforms, int64_t shots,
                   uint64_t *out_idx);

/* ---- adjoint differentiation (new; beside algorithm/gradient/gradient.go:25-86) ----------- */
/* A derivative entry says: parameter `param` enters op `op` with d(matrix)/d(param) =
 * dmatrix * scale (chain rule factor).  Several entries may name the same op (U3) or the same
 * parameter (QAOA reuse). */
typedef struct goqu_deriv {
    int32_t op;
    int32_t param;
    double scale;
    const double *dmatrix; /* same shape as the op's matrix (inner matrix for CONTROLLED) */
} goqu_deriv_t;

/* One forward + one reverse sweep (Jones & Gacon 2020): energy = Re<psi|H|psi> and
 * grads[p] = sum over entries of scale * 2 Re<lambda_k| dU_k |psi_{k-1}>.  The state is left
 * at |psi(theta)> after the call.  grads has num_params entries. */
int goqu_sv_adjoint(goqu_sv_t *h, const goqu_op_t *ops, int32_t num_ops,
                    const goqu_deriv_t *derivs, int32_t num_derivs, int32_t num_params,
                    int32_t num_terms, const uint64_t *xmask, const uint64_t *zmask,
                    const int32_t *num_y, const double *coeffs, double *energy,
                    double *grads);

/* Noisy adjoint on the density matrix: the forward pass stores rho every `checkpoint_every`
 * ops and the reverse pass back-propagates the observable through adjoint channels,
 * recomputing rho from the nearest checkpoint.  grads[p] = d Tr(H rho)/d theta_p.
 * checkpoint_every is the densest spacing: the library widens it so that the checkpoints fit
 * in free device memory. */
int goqu_dm_adjoint(goqu_dm_t *h, const goqu_op_t *ops, int32_t num_ops,
                    const goqu_deriv_t *derivs, int32_t num_derivs, int32_t num_params,
                    int32_t num_terms, const uint64_t *xmask, const uint64_t *zmask,
                    const int32_t *num_y, const double *coeffs, int32_t checkpoint_every,
                    double *energy, double *grads);

/* ---- batched QAOA MaxCut (new; beside sweep.RunSim sweep/run.go:27-73 and
 *      algorithm/qaoa/qaoa.go:61-196) ------------------------------------------------------- */
/* Instance i has num_qubits[i] qubits, edges edge_offsets[i]..edge_offsets[i+1] of the flat
 * edges[] array (pairs), p = layers[i] and parameters params[param_offsets[i] ...] in
 * ir.FreeParameters order (gamma_0, beta_0, gamma_1, ...).  The circuit is H on every qubit,
 * then per layer RZZ(gamma_k) on every edge and RX(2 beta_k) on every qubit; H is
 * sum_edges (-0.5 I + 0.5 Z_i Z_j) (algorithm/qaoa/maxcut.go:13-36).  energies[i] and
 * grads[param_offsets[i] ...] are written.  All instances run in one captured launch graph.
 * The call's device memory stays reserved for the next batch call until the process ends (a
 * failed device allocation anywhere in the library releases it first). */
int goqu_qaoa_batch(int32_t num_instances, const int32_t *num_qubits, const int32_t *layers,
                    const int64_t *edge_offsets, const int32_t *edges,
                    const int64_t *param_offsets, const double *params, double *energies,
                    double *grads);

/* ---- dynamic circuits with the shots batched on the device (replaces RunDynamic /
 *      executeOps / executeControlFlow / measureQubit, sim/statevector/dynamic.go:14-230 and
 *      sim/densitymatrix/dynamic.go:13-255) ------------------------------------------------ */
/* A dynamic circuit is a flat program of instructions over an op array (goqu_op_t, as for Apply;
 * for the density matrix with the noise channels already resolved, as for goqu_dm_apply).
 * Control flow is structured: BRANCH, WHILE and FOR each open a group that END closes. */
#define GOQU_DYN_GATES 0   /* ops[a .. a+b) on every shot that reaches the instruction         */
#define GOQU_DYN_MEASURE 1 /* measureQubit(a) into clbit b: one uniform u, outcome = (u >= P0) */
#define GOQU_DYN_RESET 2   /* measure qubit a (one uniform), then X on outcome 1               */
#define GOQU_DYN_BRANCH 3  /* opens an if/else or switch group of CASE / DEFAULT bodies        */
#define GOQU_DYN_CASE 4    /* body for shots whose clbits cond_bits[a .. a+b) read, little-endian,
                              as `value` and that took no earlier CASE of the group            */
#define GOQU_DYN_DEFAULT 5 /* body for shots that took no CASE (the group's last body)         */
#define GOQU_DYN_END 6     /* closes BRANCH, WHILE or FOR                                       */
#define GOQU_DYN_WHILE 7   /* body repeats while cond_bits[a .. a+b) read `value`, at most 1000
                              times (ir.MaxControlFlowIterations, circuit/ir/circuit.go:129)   */
#define GOQU_DYN_FOR 8     /* body repeats a times                                              */
typedef struct goqu_dyn_instr {
    int32_t kind; /* GOQU_DYN_* */
    int32_t a;
    int32_t b;
    int32_t value;
} goqu_dyn_instr_t;

/* RunDynamic: every shot starts from |0...0> (|0><0|), runs the program and leaves its clbits in
 * clbits_out[shot * num_clbits + c] (0 or 1).  Shot s draws its uniforms in order from
 * uniforms[s * draws_per_shot ...] (the reference draws from a fresh generator per shot); a shot
 * that needs more returns GOQU_ERR_INVALID_ARG.  Shots run in batches side by side on the device;
 * the handle is left holding the last shot's final state.  P0 is an FP64 sum of the same terms as
 * the reference's in another order, so outcomes agree except where u - P0 is within rounding. */
int goqu_sv_run_dynamic(goqu_sv_t *h, const goqu_op_t *ops, int32_t num_ops, const goqu_dyn_instr_t *prog,
                        int32_t num_instr, const int32_t *cond_bits, int32_t num_cond_bits, int32_t num_clbits,
                        int64_t shots, const double *uniforms, int32_t draws_per_shot, uint8_t *clbits_out);
int goqu_dm_run_dynamic(goqu_dm_t *h, const goqu_op_t *ops, int32_t num_ops, const goqu_dyn_instr_t *prog,
                        int32_t num_instr, const int32_t *cond_bits, int32_t num_cond_bits, int32_t num_clbits,
                        int64_t shots, const double *uniforms, int32_t draws_per_shot, uint8_t *clbits_out);

/* ---- batches of circuits on one register size (beside sweep.RunSim / RunDensitySim,
 *      sweep/run.go:27-171, and mitigation.StatevectorExecutor / DensityMatrixExecutor,
 *      algorithm/mitigation/executor.go:20-41, which evaluate circuits one fresh simulator at a
 *      time) ------------------------------------------------------------------------------- */
/* Circuit i is ops[op_offsets[i] .. op_offsets[i+1]) (op_offsets[0] = 0, num_instances + 1
 * entries), run from |0...0> on num_qubits qubits.  Circuits whose programs launch alike run
 * together, one launch per stage sweep over all their states.  tile_bits / low_bits choose the
 * tile shape as goqu_sv_set_plan does ((0, 0) = the default).
 * batch_expect: out[i * 2 * num_terms + 2 t ...] = coeff_t <psi_i|P_t|psi_i> (as
 * goqu_sv_expect_paulis per circuit).  batch_sample: out_idx[i * shots + s] for
 * uniforms[i * shots + s] (as goqu_sv_sample per circuit). */
int goqu_sv_batch_expect(int32_t num_qubits, int32_t precision, int32_t tile_bits, int32_t low_bits,
                         int32_t num_instances, const goqu_op_t *ops, const int64_t *op_offsets, int32_t num_terms,
                         const uint64_t *xmask, const uint64_t *zmask, const int32_t *num_y, const double *coeffs,
                         double *out);
int goqu_sv_batch_sample(int32_t num_qubits, int32_t precision, int32_t tile_bits, int32_t low_bits,
                         int32_t num_instances, const goqu_op_t *ops, const int64_t *op_offsets, int64_t shots,
                         const double *uniforms, uint64_t *out_idx);
/* Density-matrix circuits (ops as for goqu_dm_apply, noise resolved): Tr(rho_i P_t) * coeff_t
 * (as goqu_dm_expect_paulis), or samples of the diagonal after the per-qubit readout error
 * p01 / p10 (NULL: none; as goqu_dm_probs + goqu_dm_sample). */
int goqu_dm_batch_expect(int32_t num_qubits, int32_t num_instances, const goqu_op_t *ops, const int64_t *op_offsets,
                         int32_t num_terms, const uint64_t *xmask, const uint64_t *zmask, const int32_t *num_y,
                         const double *coeffs, double *out);
int goqu_dm_batch_sample(int32_t num_qubits, int32_t num_instances, const goqu_op_t *ops, const int64_t *op_offsets,
                         const double *p01, const double *p10, int64_t shots, const double *uniforms,
                         uint64_t *out_idx);

/* ---- host-side planning and test hooks (no counterpart in the reference) ------------------ */
/* Lowers and plans an op list on the host only (no device is touched): *stages receives the
 * number of state sweeps (stage-kernel launches) and *instrs the instruction count.  density != 0
 * plans the density-matrix lowering of num_qubits qubits.  tile_bits 0 = the default tile. */
int goqu_plan_stats(int32_t num_qubits, int32_t density, const goqu_op_t *ops, int32_t num_ops,
                    int32_t tile_bits, int64_t *stages, int64_t *instrs);
/* Fault injection for tests of the error paths: kind 1 makes the countdown-th device
 * allocation from now on fail (GOQU_ERR_OUT_OF_MEMORY), kind 2 makes the countdown-th op-list
 * lowering throw std::bad_alloc inside the library (caught: GOQU_ERR_OUT_OF_MEMORY); kind 0 or
 * countdown 0 disarms.  GOQU_FAULT_INJECT=device_alloc:<n> / host_alloc:<n> in the environment
 * arms it when the library loads.  No C++ exception ever crosses this ABI. */
#define GOQU_FAULT_DEVICE_ALLOC 1
#define GOQU_FAULT_HOST_ALLOC 2
int goqu_debug_inject_fault(int32_t kind, int32_t countdown);

/* ---- profiling (no counterpart in the reference; for benchmarks) -------------------------- */
/* When enabled, every launch of the profiled kernels is bracketed by HIP events on its own stream;
 * the totals accumulate process-wide until goqu_profile_reset. */
int goqu_profile_enable(int32_t on);
int goqu_profile_reset(void);
/* kernel_ms: summed event time of the stage kernels; launches: their count; bytes: their
 * algorithmic HBM bytes (read + write of the whole state per launch). */
int goqu_profile_read(double *kernel_ms, int64_t *launches, double *bytes);
/* The same per kernel: 0 k_stage / k_stage_big, 1 k_stage_dm, 2 k_stage_pair (bytes: read + write
 * of psi and lambda), 3 k_pauli_apply_tile (psi read, lambda written, and read after a group's first
 * pass), 4 k_wide_mfma (bytes holds its FP64 flops), 5 one batched-QAOA launch graph (no bytes).
 * Reading resolves the recorded events, which synchronises the device. */
#define GOQU_PROF_STAGE 0
#define GOQU_PROF_STAGE_DM 1
#define GOQU_PROF_STAGE_PAIR 2
#define GOQU_PROF_HAMILTONIAN 3
#define GOQU_PROF_WIDE 4
#define GOQU_PROF_QAOA 5
int goqu_profile_read_kernel(int32_t kernel, double *kernel_ms, int64_t *launches, double *bytes);

#ifdef __cplusplus
}
#endif
#endif /* GOQU_HIP_H */
