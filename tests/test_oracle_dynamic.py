"""Pins the oracle's measurement, density-matrix reset and density-matrix ExpectationValue:

  * orc_sv_measure_qubit (measureQubit, sim/statevector/dynamic.go:186-230) through closed forms
    on product states (outcome rule r >= P0, collapse, renormalisation, the norm < 1e-15 skip),
    and through the reference's dynamic-circuit tests (dynamic_test.go:14-162: conditional,
    conditional not fired, reset and reuse, 3-qubit teleportation) run by a shot loop over the
    oracle that follows dynamic.go's executeOps;
  * the oracle's density-matrix reset (resetQubit, sim/densitymatrix/sim.go:195-217) through
    the reference's TestEvolve_Reset (densitymatrix/sim_test.go:550-620) and the partial-trace
    closed form on a random two-qubit state;
  * orc_dm_expectation_value (measure.go:76-91) through closed forms on mixed states, its
    1e-14 rounding, and agreement with Tr(rho Z...) built in numpy.
"""
import math

import numpy as np
import pytest

from goqu_amd import densitymatrix as D
from goqu_amd import gate as G
from goqu_amd import noise as NZ
from goqu_amd import _native as N
from goqu_amd.circuit import NewBuilder
from oracle import oracle as O

MAX_ITER = 1000  # ir.MaxControlFlowIterations, circuit/ir/circuit.go:129


class OracleDynamic:
    """The reference's RunDynamic shot loop (dynamic.go:14-120) over the oracle's kernels."""

    def __init__(self, n):
        self.n = n

    def apply(self, g, qubits):
        buf = N.OpBuffer()
        N.add_sv_gate(buf, g, qubits)
        assert O.sv_apply_ops(self.psi, self.n, buf) == 0

    def run(self, c, shots, rng):
        counts = {}
        nc = c.NumClbits()
        for _ in range(shots):
            self.psi = O.zero_state(self.n)
            cl = [0] * nc
            self.exec(c.Ops(), cl, rng)
            bs = "".join("1" if cl[nc - 1 - i] else "0" for i in range(nc))
            counts[bs] = counts.get(bs, 0) + 1
        return counts

    def exec(self, ops, cl, rng):
        for op in ops:
            cf = op.ControlFlow
            if cf is not None:
                if cf.Type == "if_else":
                    if cl[cf.Condition.Clbit] == cf.Condition.Value:
                        self.exec(cf.Bodies[0], cl, rng)
                    elif len(cf.Bodies) > 1:
                        self.exec(cf.Bodies[1], cl, rng)
                elif cf.Type == "while":
                    for _ in range(MAX_ITER):
                        if cl[cf.Condition.Clbit] != cf.Condition.Value:
                            break
                        self.exec(cf.Bodies[0], cl, rng)
                elif cf.Type == "for":
                    for _ in range(*cf.ForRange):
                        self.exec(cf.Bodies[0], cl, rng)
                continue
            g = op.Gate
            if g is None:
                if op.Clbits and op.Qubits:
                    cl[op.Clbits[0]] = O.sv_measure_qubit(self.psi, op.Qubits[0], rng.random())
                continue
            if g.Name() in ("barrier", "delay"):
                continue
            if g.Name() == "reset":
                if O.sv_measure_qubit(self.psi, op.Qubits[0], rng.random()) == 1:
                    self.apply(G.X, op.Qubits)
                continue
            if op.Condition is not None and cl[op.Condition.Clbit] != op.Condition.Value:
                continue
            self.apply(g, op.Qubits)


# ---- measureQubit closed forms ---------------------------------------------------------------------
def _product(t, phi, alpha, beta):
    """(cos t |0> + e^{i phi} sin t |1>) on qubit 1 (x) (alpha |0> + beta |1>) on qubit 0."""
    q1 = np.array([math.cos(t), np.exp(1j * phi) * math.sin(t)])
    q0 = np.array([alpha, beta])
    return np.kron(q1, q0)


@pytest.mark.parametrize("t", [0.3, 1.1])
def test_measure_qubit_outcomes_and_collapse(t):
    a, b = 0.6, 0.8j
    p0 = math.cos(t) ** 2
    for r, outcome in [(0.0, 0), (p0 * 0.999, 0), (p0, 1), (min(0.999999, p0 + 1e-9), 1), (0.999999, 1)]:
        psi = _product(t, 0.7, a, b)
        assert O.sv_measure_qubit(psi, 1, r) == outcome
        if outcome == 0:
            want = np.kron([1, 0], [a, b])
        else:
            want = np.kron([0, np.exp(0.7j)], [a, b])
        assert np.max(np.abs(psi - want)) < 1e-15, (r, psi)


def test_measure_qubit_skips_zero_norm_branch():
    """Outcome 0 on a branch of probability 1e-40: norm 1e-20 < 1e-15 leaves the state as it is."""
    psi = np.array([1e-20, 0, 1, 0], dtype=np.complex128)  # not normalised on purpose
    ref = psi.copy()
    assert O.sv_measure_qubit(psi, 1, 0.0) == 0
    assert np.array_equal(psi, ref)


def test_measure_qubit_probability_is_the_marginal():
    """P0 of a random 3-qubit state on each qubit: the outcome flips exactly at the marginal."""
    rng = np.random.default_rng(4)
    s = rng.normal(size=8) + 1j * rng.normal(size=8)
    s /= np.linalg.norm(s)
    for q in range(3):
        p0 = sum(abs(s[i]) ** 2 for i in range(8) if not (i >> q) & 1)
        lo, hi = s.copy(), s.copy()
        assert O.sv_measure_qubit(lo, q, np.nextafter(p0, 0)) == 0
        assert O.sv_measure_qubit(hi, q, p0) == 1
        keep = np.array([0 if (i >> q) & 1 else 1 for i in range(8)])
        assert np.max(np.abs(lo - s * keep / math.sqrt(p0))) < 1e-15
        assert np.max(np.abs(hi - s * (1 - keep) / math.sqrt(1 - p0))) < 1e-15


# ---- the reference's dynamic-circuit tests (dynamic_test.go) ----------------------------------------
def test_dynamic_simple_conditional():  # dynamic_test.go:14-40
    c = NewBuilder("cond", 2).WithClbits(1).X(0).Measure(0, 0).If(0, 1, G.X, 1).Build()
    assert c.IsDynamic()
    assert OracleDynamic(2).run(c, 1000, np.random.default_rng(1)) == {"1": 1000}


def test_dynamic_conditional_not_fired():  # dynamic_test.go:42-63
    c = NewBuilder("cond_no", 2).WithClbits(1).Measure(0, 0).If(0, 1, G.X, 1).Build()
    assert OracleDynamic(2).run(c, 1000, np.random.default_rng(2)) == {"0": 1000}


def test_dynamic_reset_and_reuse():  # dynamic_test.go:65-104
    c = NewBuilder("reset_reuse", 1).WithClbits(2).H(0).Measure(0, 0).Reset(0).H(0).Measure(0, 1).Build()
    assert c.IsDynamic()
    counts = OracleDynamic(1).run(c, 10000, np.random.default_rng(3))
    one = sum(v for k, v in counts.items() if k[0] == "1")
    assert abs(one / 10000 - 0.5) < 0.05


def teleport_circuit():  # dynamic_test.go:106-162 (the verifying variant)
    b = NewBuilder("teleport_verify", 3).WithClbits(3)
    b.X(0).H(1).CNOT(1, 2).CNOT(0, 1).H(0).Measure(0, 0).Measure(1, 1)
    b.If(1, 1, G.X, 2).If(0, 1, G.Z, 2).Measure(2, 2)
    return b.Build()


def test_dynamic_teleportation():
    counts = OracleDynamic(3).run(teleport_circuit(), 4000, np.random.default_rng(4))
    assert all(len(k) == 3 and k[0] == "1" for k in counts)
    assert len(counts) == 4  # the Bell-measurement outcomes c0 c1 are uniform


# ---- density-matrix reset (densitymatrix/sim_test.go:550-620) ---------------------------------------
def _dm_evolve(c, n):
    rho = O.dm_zero(n)
    assert O.dm_apply_ops(rho, n, D.marshal(c)) == 0
    d = 1 << n
    return rho.reshape(d, d)


def test_dm_reset_known_answers():
    zero = np.zeros((2, 2))
    zero[0, 0] = 1
    assert np.max(np.abs(_dm_evolve(NewBuilder("xreset", 1).X(0).Reset(0).Build(), 1) - zero)) < 1e-10
    assert np.max(np.abs(_dm_evolve(NewBuilder("hreset", 1).H(0).Reset(0).Build(), 1) - zero)) < 1e-10
    rho = _dm_evolve(NewBuilder("bellreset", 2).H(0).CNOT(0, 1).Reset(0).Build(), 2)
    assert np.max(np.abs(rho - np.diag([0.5, 0, 0.5, 0]))) < 1e-10
    assert abs(np.trace(rho) - 1) < 1e-10


@pytest.mark.parametrize("q", [0, 1])
def test_dm_reset_is_partial_trace(q):
    """rho -> |0><0|_q (x) Tr_q(rho) on a random (complex) two-qubit mixed state."""
    rng = np.random.default_rng(10 + q)
    A = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    rho = A @ A.conj().T
    rho /= np.trace(rho)
    flat = np.ascontiguousarray(rho.reshape(-1))
    buf = N.OpBuffer()
    buf.add(N.OP_RESET, [q])
    assert O.dm_apply_ops(flat, 2, buf) == 0
    t = rho.reshape(2, 2, 2, 2)  # [b1, b0, b1', b0']
    if q == 0:
        red = np.einsum("aibi->ab", t)  # trace out qubit 0 (the low bit)
        want = np.kron(red, np.diag([1, 0]))
    else:
        red = np.einsum("iaib->ab", t)
        want = np.kron(np.diag([1, 0]), red)
    assert np.max(np.abs(flat.reshape(4, 4) - want)) < 1e-15


# ---- density-matrix ExpectationValue (measure.go:76-91) ---------------------------------------------
def test_dm_expectation_value_closed_forms():
    p = 0.3
    rho = np.zeros((4, 4), dtype=np.complex128)
    rho[0, 0], rho[3, 3] = p, 1 - p  # p|00><00| + (1-p)|11><11|
    rho[0, 3] = rho[3, 0] = 0.2  # coherences do not enter a Z observable
    flat = np.ascontiguousarray(rho.reshape(-1))
    assert O.dm_expectation_value(flat, 2, [0, 1]) == 1.0
    assert abs(O.dm_expectation_value(flat, 2, [0]) - (2 * p - 1)) < 1e-15
    assert abs(O.dm_expectation_value(flat, 2, [1]) - (2 * p - 1)) < 1e-15
    assert O.dm_expectation_value(flat, 2, []) == 1.0


def test_dm_expectation_value_depolarized_and_rounding():
    """|0> through Depolarizing1Q(p): <Z> = 1 - 4p/3; values within 1e-14 of an integer round to
    it, others keep 14 decimals."""
    for p in [0.05, 0.3, 0.75]:
        b = NewBuilder("dep", 1).Apply(G.I, 0)
        nm = NZ.New()
        nm.AddGateError("I", NZ.Depolarizing1Q(p))
        rho = O.dm_zero(1)
        assert O.dm_apply_ops(rho, 1, D.marshal(b.Build(), nm)) == 0
        ev = O.dm_expectation_value(rho, 1, [0])
        assert ev == round((1 - 4 * p / 3) * 1e14) / 1e14
    rho = O.dm_zero(1)
    rho[0] = 1 - 3e-16
    rho[3] = 3e-16
    assert O.dm_expectation_value(rho, 1, [0]) == 1.0
    assert O.dm_expectation_value(rho, 1, [0], round14=False) != 1.0


def test_dm_expectation_value_matches_trace_formula():
    rng = np.random.default_rng(12)
    n = 3
    A = rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8))
    rho = A @ A.conj().T
    rho /= np.trace(rho)
    flat = np.ascontiguousarray(rho.reshape(-1))
    Zm = np.diag([1.0, -1.0])
    for qs in [[0], [2], [0, 1], [0, 1, 2]]:
        op = np.array([[1.0]])
        for q in reversed(range(n)):  # qubit n-1 is the most significant factor
            op = np.kron(op, Zm if q in qs else np.eye(2))
        want = np.trace(rho @ op).real
        assert abs(O.dm_expectation_value(flat, n, qs, round14=False) - want) < 1e-15
