// engine.h — shared host/device structures of the tiled statevector engine.
//
// A circuit is lowered to elementary ops (EOp), which the planner groups into stages.
// A stage is one sweep over the state: every workgroup loads one tile of 2^m amplitudes
// (the tile qubits T = the L lowest qubits plus up to m-L chosen qubits), runs the
// stage's instruction list on it in registers (re-distributing amplitudes between
// threads through LDS when an op needs other qubits in registers), and stores it back.
#pragma once
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/goqu_hip.h"

namespace goqu {

// Register slots per thread: each thread keeps 2^kRegBits amplitudes in registers.
constexpr int kRegBits = 4;
constexpr int kRegAmps = 1 << kRegBits;
// Smallest internal qubit count: tiny circuits are padded with |0> qubits so that every
// tile has at least kRegBits register qubits and a few thread qubits.
constexpr int kMinInternalQubits = 6;
constexpr int kMaxTileBits = 13;

// ---- elementary ops ------------------------------------------------------------------
enum EOpKind : int32_t {
    EOP_DENSE = 0,  // k targets (k <= 4), dense 2^k x 2^k matrix, qubits[0] = MSB
    EOP_DIAG = 1,   // k qubits anywhere (k <= 4), 2^k diagonal entries, qubits[0] = MSB
    EOP_ANTI1 = 2,  // 1 target: [a0, a1] <- [p0 * a1, p1 * a0]   (X, Y, CNOT, CY, CCX ...)
    EOP_SWAP2 = 3,  // 2 targets: |01> <- p * |10>, |10> <- p * |01>  (SWAP, iSWAP, CSWAP)
    EOP_RESET = 4,  // 1 target: deterministic reset (a0 <- |(a0,a1)|, a1 <- 0)
    EOP_GRAD = 5,   // paired states only: Re <lambda| D |psi> on k targets into gradient slot `slot`
    EOP_PARITY = 6, // density-matrix channel superoperator that preserves each qubit's row^column
                    // parity: k = 2 (q = row, col) or 4 (q = row a, row b, col a, col b); m holds
                    // the full 2^k x 2^k superoperator (q[0] = MSB), zero across parity classes
    EOP_WIDE = 7,   // dense on k = 5..kMaxWideTargets targets listed in qw (qw[0] = MSB), under
                    // cmask: not part of any stage; run by the MFMA kernel between stage programs
};
constexpr int kMaxWideTargets = 9;

// Gradient slots a single stage may hold (per-wave partials live in LDS).
constexpr int kMaxGradPerStage = 64;

struct EOp {
    int32_t kind = EOP_DENSE;
    int32_t k = 0;        // number of qubits in q[]
    int32_t q[4] = {0, 0, 0, 0};
    uint64_t cmask = 0;   // control qubits (all must be 1)
    std::vector<double> m;  // interleaved complex payload (dense matrix / diag / phases)
    std::vector<int> qw;    // EOP_WIDE: target qubits, qw[0] = MSB of the matrix index
    bool pure_swap = false; // ANTI1 / SWAP2 with unit phases: move only
    int32_t slot = -1;      // EOP_GRAD: global gradient slot
    // paired programs: 0 = the op acts on both states, 1 = on the first only, 2 = on the second
    // only (EOP_PARITY ops alone: a density-matrix reverse sweep uncomputes rho through the
    // inverse of a channel while h moves through its adjoint)
    int32_t side = 0;
    // which of the q[] are targets that must sit in registers (DENSE, ANTI1, SWAP2, RESET)
    bool needs_regs() const { return kind != EOP_DIAG; }
    bool is_grad() const { return kind == EOP_GRAD; }
    // The statevector reset a0 <- |(a0, a1)| is not linear: it discards any phase the two
    // amplitudes of a pair share, so it commutes with no op on any qubit (not even a diagonal
    // one on another qubit).  Every reordering treats it as touching all qubits non-diagonally.
    bool is_barrier() const { return kind == EOP_RESET; }
    uint64_t touch() const { return cmask | targets_of(); }
    uint64_t targets() const { return needs_regs() ? targets_of() : 0; }
    uint64_t targets_of() const {
        uint64_t t = 0;
        if (kind == EOP_WIDE)
            for (int x : qw) t |= 1ull << x;
        else
            for (int i = 0; i < k; i++) t |= 1ull << q[i];
        return t;
    }
    // qubits acted on non-diagonally (targets) and only diagonally (controls, diagonal gates):
    // two ops commute when every qubit they share is diagonal for both
    uint64_t nondiag() const { return targets(); }
    uint64_t diag() const { return touch() & ~targets(); }
};

// ---- device instructions ---------------------------------------------------------------
enum InstrOp : int32_t {
    I_DENSE = 0,  // dense on register slot mask `smask`; matrix at pool[mat] in slot order
    I_DIAG = 1,   // phase table at pool[mat], see Instr::smask / k / ls
    I_ANTI1 = 2,  // slot s[0]: [a0, a1] <- [p0 a1, p1 a0], phases at pool[mat]
    I_SWAP2 = 3,  // slots s[0] < s[1]: |01> <- p |10>, |10> <- p |01>, phase at pool[mat]
    I_RESET = 4,  // slot s[0]
    I_REMAP = 5,  // new layout lay[0..m-1] (register slots 0..3, then thread bits), LDS offsets kr
    I_GRAD = 6,   // like I_DENSE, but accumulates Re <b| M |a> into local gradient slot `pad`
    I_PARITY = 7, // parity-block superoperator on 2 or 4 register slots (canonical form below)
    I_SHEAR = 8,  // slot s[0]: k in-place shears, a0 += c0 a1, a1 += c1 a0, a0 += c2 a1 (the
                  // first k of them); c0, c1, c2 are doubles stored in the instruction at byte
                  // offsets 40, 48, 56 (set_shear_coefs), so the op needs no pool load
    I_XSWAP = 9,  // register slot s[0] trades places with lane bit k (0..5) by cross-lane moves
                  // (permlane swaps for lane bits 4, 5; row-shift DPP for 0..3); lay / kr give
                  // the resulting layout and its LDS offsets, as for I_REMAP
    I_END = 10,   // closes every stage's instruction list, right after its ninstr instructions (not
                  // counted in ninstr): the kernel runs a stage's ops up to it without a count test
    I_SHEARS = 11, // I_SHEAR on several register slots at once (consecutive shears on distinct
                   // slots commute): bits 2s..2s+1 of k = the shear count on slot s (0, 2 or 3),
                   // the coefficients c0, c1, c2 of slot s at pool[mat] + 3 s (twelve reals)
};

// Instr::flags, bits 0..7
constexpr int32_t F_PURE = 1;   // ANTI1 / SWAP2 with unit phases: a move
constexpr int32_t F_TCTRL = 2;  // controls on thread-id bits (tmask)
constexpr int32_t F_PAT = 4;    // controls on register slots (pat)
constexpr int32_t F_TILE = 8;   // controls on qubits outside the tile (omask), uniform per tile
constexpr int32_t F_OUTER = 16; // run by the kernel outside its inner op loop (REMAP, XSWAP, RESET, dense on 3-4 slots)
constexpr int32_t F_REAL = 32;  // I_PARITY, I_DENSE: every matrix entry is real (only real parts are read)
constexpr int32_t F_SIDE_A = 64;  // I_PARITY in a paired program: acts on the first state only
constexpr int32_t F_SIDE_B = 128; // I_PARITY in a paired program: acts on the second state only

// Instr::flags, bits 8..15: the kernel code path of the instruction (assign_codes), so that the
// stage kernel dispatches every op with one switch on a small integer.  Ranges:
enum KCode : int32_t {
    K_SHEAR = 0,     // 8: 2 * slot + (three shears)
    K_SHEARS = 8,    // I_SHEARS
    K_DIAG_FULL = 12, // uniform phase table applied to registers 1..15 without tests (and to
                      // register 0 if bit 0 of `pat` is set): at least 11 of them nontrivial
    K_DIAG_MASK = 13, // uniform phase table, the registers of `pat`
    K_DIAG_TILE = 14, // phase table selected by bits of the tile base (uniform per workgroup)
    K_DIAG_LANE = 15, // phase table indexed by thread-id bits, or under thread-bit controls
    K_DIAG_SLOT = 16, // 4: slot; one uniform phase (table entry 1 << slot) on the 8 registers
                      // with that slot's bit set, every other entry 1
    K_DIAG_TSLOT = 20, // 4: slot; as K_DIAG_SLOT, on the lanes whose thread-bit controls hold
    K_ANTI1 = 24,    // 16: 4 * slot + 2 * pure + thread controls
    K_SWAP2 = 40,    // 24: 4 * slot pair (01 02 03 12 13 23) + 2 * pure + thread controls
    K_DENSE1 = 64,   // 12: 3 * slot + (0 controlled, 1 real, 2 complex)
    K_DENSE2 = 76,   // 12: 2 * slot pair + (0 controlled, 1 complex)
    K_PARITY2 = 88,  // 12: 2 * slot pair + (0 real, 1 complex)
    K_PARITY4 = 100, // 6: 2 * (partner of slot 0 - 1) + (0 real, 1 complex)
    K_GRAD = 106,    // 10: index of the target slot mask in {1,2,3,4,5,6,8,9,10,12}
    K_OUTER = 255,   // F_OUTER instructions (the kernel's outer loop)
};
constexpr int kCodeShift = 8;
inline int instr_code(int32_t flags) { return (flags >> kCodeShift) & 255; }

// I_PARITY canonical form (the kernel's view).  k = 2: slots lo < hi (smask); class
// p = bit lo ^ bit hi; element i of class p has bit lo = i, bit hi = i ^ p; the pool holds
// 2 blocks of 2x2 (class-major, row-major).  k = 4: all four slots, paired as
// s[0]-s[1] and s[2]-s[3] (s[0] = 0); class c = 2 pa + pb with pa = bit s0 ^ bit s1,
// pb = bit s2 ^ bit s3; element i = 2 ia + ib has bit s0 = ia, bit s1 = ia ^ pa, bit s2 = ib,
// bit s3 = ib ^ pb; the pool holds 4 blocks of 4x4.

// One instruction.  The planner resolves every qubit an op reads against the register layout in
// force when the op runs, so the kernel never maps global qubits itself:
//   register slots      -> compile-time register indices (smask / s[] select the code path),
//                          controls there -> `pat`, bit r = 1 iff they hold at register index r;
//   thread-id bits      -> `tmask` (controls), ls[] < 16 (diagonal index bits);
//   qubits off the tile -> `omask` (controls) and ls[] >= 16, uniform over a workgroup.
// Byte arrays are read by the kernel as aligned 32-bit words.
struct alignas(16) Instr {
    int32_t op;
    int32_t k;        // DENSE / GRAD: target count; DIAG: lane index bits (entries of ls[])
    int32_t mat;      // offset (complex elements) into the program pool
    int32_t flags;
    uint32_t pat;     // DENSE / ANTI1 / SWAP2 / GRAD: register control pattern (0xFFFF: none);
                      // DIAG: bit v = 1 iff column v of the phase table holds an entry != 1
    uint32_t tmask;   // thread-id bits that must be 1
    uint64_t omask;   // qubits outside the tile that must be 1
    uint8_t s[4];     // register slots
    int32_t smask;    // DENSE / GRAD: target slot mask (matrix in slot-bit order);
                      // DIAG: register slots indexing the table (diagonal qubits and controls there)
    int32_t pad;      // GRAD: gradient slot within the stage
    uint8_t ls[4];    // DIAG: lane index bits, MSB first: < 16 thread-id bit, >= 16 qubit (ls-16) off the tile
    uint8_t lay[16];  // REMAP: tile-local bit of register slots 0..3, then of thread-id bits 0..m-5
    uint16_t kr[16];  // REMAP: LDS byte offset (8-byte elements) of register index r in the new layout
};
static_assert(sizeof(Instr) == 96, "Instr must stay 96 bytes");

inline void set_shear_coefs(Instr &I, double c0, double c1, double c2) {
    unsigned char *b = reinterpret_cast<unsigned char *>(&I);
    std::memcpy(b + 40, &c0, 8);
    std::memcpy(b + 48, &c1, 8);
    std::memcpy(b + 56, &c2, 8);
}
inline double shear_coef(const Instr &I, int i) {
    double c;
    std::memcpy(&c, reinterpret_cast<const unsigned char *>(&I) + 40 + 8 * i, 8);
    return c;
}

// DIAG phase table: 16 entries per lane index, entry (lane_index << 4) | r for register index r
// (lane_index: the ls[] bits, MSB first); pat bit r = 1 iff some entry of register r is != 1.
// smask keeps the register slots the op's qubits and controls occupy (informative only).

struct alignas(16) StageHeader {
    int32_t m;            // tile bits
    int32_t ninstr;
    int32_t instr_off;    // index of first Instr in the program's instruction array
    int32_t pool_off;     // unused (instructions address the program pool directly)
    uint8_t T[16];        // tile qubits, ascending (local bit j <-> global T[j])
    uint8_t lay0[16];     // initial layout (local bits: 4 register slots then thread bits)
    uint8_t gq0[16];      // global qubits of lay0
    uint8_t layN[16];     // final layout (the one in force after the last instruction)
    uint8_t gqN[16];      // global qubits of layN
    uint16_t k0[16];      // LDS byte offsets (8-byte elements) of register index r in lay0
    int32_t grad_base;    // global gradient slot of local slot 0
    int32_t ngrad;        // gradient slots of this stage
    int32_t flags;        // SF_BIG: holds a dense op on 3-4 register slots
    int32_t pad;
};
static_assert(sizeof(StageHeader) == 144, "StageHeader layout");
// StageHeader::flags
constexpr int32_t SF_BIG = 1;
// the stage holds density-matrix channel ops (I_PARITY): run by the density-matrix kernel
constexpr int32_t SF_DM = 4;
// The stage stores through a different global mapping than it loads with (uncontrolled SWAPs
// relabelled its local bits, or lane exchanges moved amplitudes between threads) and holds no LDS
// remap: a thread may then store to a position another wave (or lane) has still to load, so the
// kernel waits for every load of the tile (a workgroup barrier) before the first store.
constexpr int32_t SF_SYNC_STORE = 2;

// LDS element index swizzle of the stage kernel's remaps, for 8-byte elements (a complex64
// amplitude, or one component of a complex128 amplitude: those move real parts and imaginary
// parts in two passes).  XOR-linear, swz(a ^ b) = swz(a) ^ swz(b), so offsets split into a
// per-thread and a per-register part.  Lanes that differ in any three tile bits whose positions
// are distinct mod 4 land in distinct 8-byte bank pairs.
inline uint32_t lds_swizzle(uint32_t l) {
    const uint32_t h = l >> 4;
    return l ^ ((h ^ (h >> 4) ^ (h >> 8)) & 15u);
}
struct Program {
    int nq = 0;           // internal qubit count
    std::vector<StageHeader> stages;
    std::vector<Instr> instrs;
    std::vector<double> pool;  // interleaved complex
    int64_t num_eops = 0;
    int32_t num_grad = 0;      // gradient slots over all stages
    std::vector<int32_t> grad_slots;  // program gradient slot -> the EOp's `slot`
};

struct PlanConfig {
    int tile_bits = 12;   // m
    int low_bits = 4;     // L: lowest qubits in every tile (contiguous runs of 2^L amps)
};

// Lowering of goqu_op_t records to elementary ops (statevector semantics).  Returns a
// GOQU_* status and an error message.
struct goqu_op_view;
int lower_sv_ops(const void *ops /* goqu_op_t* */, int nops, int nq_user,
                 std::vector<EOp> &out, std::string &err);

// Lowering for a density matrix of n qubits stored as a 2n-qubit vector.
int lower_dm_ops(const void *ops, int nops, int n, std::vector<EOp> &out, std::string &err);
// folds each parity-block channel into the last channel on a superset of its qubits when no op in
// between touches those qubits (lower_dm_ops applies it to every list it lowers)
void merge_channels(std::vector<EOp> &ops);

// Greedy stage planner + register allocation (no EOP_WIDE ops: run_eops splits at them).
Program plan_program(const std::vector<EOp> &eops, int nq_internal, const PlanConfig &cfg);
// Fills Instr::kr and StageHeader::k0: LDS byte offsets of 8-byte elements.
void set_lds_offsets(Program &P);
// The kernel code path of every instruction (flags bits 8..15) and the stage kernel flags.
void assign_codes(Program &P);
// The register slot of a single-slot uniform phase table (K_DIAG_SLOT / K_DIAG_TSLOT), or -1.
int diag_single_slot(const Instr &I, const std::vector<double> &pool);

// ---- adjoint differentiation programs (adjoint_plan.cpp) ----------------------------------
// conjugate transpose of a unitary EOp (RESET has none)
int invert_eop(const EOp &e, EOp &inv, std::string &err);
// dense derivative op dU of a goqu_op_t on its targets (controls kept), from `dm`
int deriv_eop(const goqu_op_t &op, const double *dm, EOp &out, std::string &err);
// EOP_GRAD op D = dU U^dag of one derivative (false: more than two targets or no matrix)
bool grad_eop(const goqu_op_t &op, const double *dm, int slot, EOp &out);
// The reverse sweep of a statevector adjoint as one program on (psi, lam): for k = nops-1..0,
// the EOP_GRAD ops of op k's derivatives (slot = derivative index), then the inverse of op k.
// fused = false (and rev empty) when some derivative does not fit an EOP_GRAD op.
int build_sv_reverse(const goqu_op_t *ops, int nops, int nq_user, const goqu_deriv_t *derivs, int nd,
                     std::vector<EOp> &rev, bool &fused, std::string &err);
// The inverse of a channel superoperator (EOP_PARITY or EOP_DENSE) by Gaussian elimination with
// partial pivoting; false when it is singular or its largest entry exceeds `bound` (uncomputing
// rho through it would amplify rounding errors too much).
bool inverse_superop(const EOp &e, EOp &inv, double bound);
// The reverse sweep of a density-matrix adjoint over ops [k0, k1) as one program on (rho, h), rho
// the first state and h = vec(H) adjoint-propagated the second: for k = k1-1..k0, an EOP_GRAD per
// derivative of op k (D = dU U^dag on its row qubits: 2 Re <h_k| D rho_k> is the derivative
// term), then the inverse of op k: unitary pieces on both states, a channel superoperator S as
// S^-1 on rho (side 1) and S^dag on h (side 2).  fused = false when some op cannot be run this way
// (a reset, a channel that is not parity-block or badly conditioned, a wide or >2-target gate).
int build_dm_reverse(const goqu_op_t *ops, int k0, int k1, int n, const goqu_deriv_t *derivs, int nd,
                     std::vector<EOp> &rev, bool &fused, std::string &err);
// A bound on how much uncomputing rho through the inverse of op k can grow errors (the infinity
// norm of its inverse superoperator; 1 for gates, infinity for a reset or a singular channel).
double dm_uncompute_growth(const goqu_op_t &op, int n);

}  // namespace goqu
