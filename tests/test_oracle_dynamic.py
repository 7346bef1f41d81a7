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
    """The reference's RunDynamic shot loop over the oracle's kernels: executeOps and
    executeControlFlow of sim/statevector/dynamic.go:55-146 (density=False) or
    sim/densitymatrix/dynamic.go:54-157 (density=True, with the noise model's lookups).  Shot s
    draws its uniforms in order from draw(s), as the reference draws from a fresh generator per
    shot."""

    def __init__(self, n, density=False, noise=None):
        self.n = n
        self.density = density
        self.noise = noise

    def _apply_buf(self, buf):
        if not len(buf):
            return
        if self.density:
            assert O.dm_apply_ops(self.psi, self.n, buf) == 0
        else:
            assert O.sv_apply_ops(self.psi, self.n, buf) == 0

    def apply(self, g, qubits):
        buf = N.OpBuffer()
        N.add_sv_gate(buf, g, qubits) if not self.density else buf.add(N.OP_GATE, qubits, g.Matrix())
        self._apply_buf(buf)

    def _channel(self, name, qubits):
        if self.noise is None:
            return
        ch = self.noise.Lookup(name, qubits)
        if ch is not None:
            buf = N.OpBuffer()
            D._add_channel(buf, ch, qubits)
            self._apply_buf(buf)

    def measure(self, q, r):
        if self.density:
            return O.dm_measure_qubit(self.psi, self.n, q, r)
        return O.sv_measure_qubit(self.psi, q, r)

    def shot(self, c, draws):
        self.psi = O.dm_zero(self.n) if self.density else O.zero_state(self.n)
        cl = [0] * c.NumClbits()
        self.draw = draws
        self.exec(c.Ops(), cl)
        return cl

    def run_shots(self, c, U):
        """clbits of every shot (shot s draws from U[s] in order) and the last shot's state."""
        rows = []
        for s in range(len(U)):
            it = iter(U[s])
            rows.append(self.shot(c, lambda: next(it)))
        return np.array(rows, dtype=np.uint8).reshape(len(U), c.NumClbits()), self.psi

    def run(self, c, shots, rng):
        counts = {}
        nc = c.NumClbits()
        for _ in range(shots):
            cl = self.shot(c, rng.random)
            bs = "".join("1" if cl[nc - 1 - i] else "0" for i in range(nc))
            counts[bs] = counts.get(bs, 0) + 1
        return counts

    def exec(self, ops, cl):
        for op in ops:
            cf = op.ControlFlow
            if cf is not None:
                if cf.Type == "if_else":
                    if cl[cf.Condition.Clbit] == cf.Condition.Value:
                        self.exec(cf.Bodies[0], cl)
                    elif len(cf.Bodies) > 1:
                        self.exec(cf.Bodies[1], cl)
                elif cf.Type == "while":
                    for _ in range(MAX_ITER):
                        if cl[cf.Condition.Clbit] != cf.Condition.Value:
                            break
                        self.exec(cf.Bodies[0], cl)
                elif cf.Type == "for":
                    for _ in range(*cf.ForRange):
                        self.exec(cf.Bodies[0], cl)
                elif cf.Type == "switch":  # readClassicalValue: little-endian (dynamic.go:150-157)
                    val = sum(cl[b] << i for i, b in enumerate(cf.SwitchClbits))
                    for i, v in enumerate(cf.SwitchCases):
                        if val == v and i < len(cf.Bodies):
                            self.exec(cf.Bodies[i], cl)
                            break
                    else:
                        if len(cf.Bodies) > len(cf.SwitchCases):
                            self.exec(cf.Bodies[-1], cl)
                continue
            g = op.Gate
            if g is None:
                if op.Clbits and op.Qubits:
                    cl[op.Clbits[0]] = self.measure(op.Qubits[0], self.draw())
                continue
            name = g.Name()
            if name == "barrier" or (name == "delay" and not self.density):
                continue
            if name == "delay":
                self._channel("delay", op.Qubits)
                continue
            if name == "reset":
                if self.measure(op.Qubits[0], self.draw()) == 1:
                    self.apply(G.X, op.Qubits[:1])
                if self.density:
                    self._channel("reset", op.Qubits)
                continue
            if op.Condition is not None and cl[op.Condition.Clbit] != op.Condition.Value:
                continue
            if self.density:
                buf = N.OpBuffer()
                D._marshal_op(buf, op, self.noise)
                self._apply_buf(buf)
            else:
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


# ---- control flow: the reference's simulation tests (circuit/builder/controlflow_test.go) -------------
def test_control_flow_kats_on_the_oracle():
    """TestWhileSimulation, TestForSimulation, TestIfElseSimulation, TestSwitchSimulation
    (controlflow_test.go:184-300) through the oracle shot loop."""
    rng = np.random.default_rng(5)
    c = NewBuilder("while-sim", 1).WithClbits(1).X(0).Measure(0, 0).While(
        0, 1, lambda b: b.X(0).Measure(0, 0)).Build()
    assert OracleDynamic(1).run(c, 100, rng) == {"0": 100}
    c = NewBuilder("for-sim", 1).WithClbits(1).For(0, 0, 3, 1, lambda b: b.X(0)).MeasureAll().Build()
    assert OracleDynamic(1).run(c, 100, rng) == {"1": 100}
    c = NewBuilder("ifelse-sim", 1).WithClbits(1).X(0).Measure(0, 0).IfElseBlock(
        0, 1, lambda b: None, lambda b: b.X(0)).Measure(0, 0).Build()
    assert OracleDynamic(1).run(c, 100, rng) == {"1": 100}
    c = NewBuilder("switch-sim", 1).WithClbits(2).X(0).Measure(0, 0).Switch(
        [0, 1], {0: lambda b: b.X(0), 1: lambda b: None}, None).Measure(0, 0).Build()
    assert OracleDynamic(1).run(c, 100, rng) == {"01": 100}


def test_switch_reads_clbits_little_endian_and_falls_to_default():
    """readClassicalValue (dynamic.go:150-157): clbits[0] is bit 0; no matching case -> default."""
    def circ(c0, c1):
        b = NewBuilder("sw", 2).WithClbits(4)
        if c0:
            b.X(0)
        if c1:
            b.X(1)
        b.Measure(0, 0).Measure(1, 1)
        b.Switch([0, 1], {1: lambda x: x.X(0), 2: lambda x: x.X(1)}, lambda x: (x.X(0), x.X(1)))
        return b.Measure(0, 2).Measure(1, 3).Build()

    rng = np.random.default_rng(6)
    # value 1 (c0=1): X(0) -> q0 = 0, q1 = 0;  value 2 (c1=1): X(1) -> q1 = 0;  value 0 / 3: default
    assert OracleDynamic(2).run(circ(1, 0), 10, rng) == {"0001": 10}
    assert OracleDynamic(2).run(circ(0, 1), 10, rng) == {"0010": 10}
    assert OracleDynamic(2).run(circ(0, 0), 10, rng) == {"1100": 10}
    assert OracleDynamic(2).run(circ(1, 1), 10, rng) == {"0011": 10}


def test_builder_rejects_bad_clbits_and_zero_step():
    """TestWhileInvalidClbit (controlflow_test.go:302-311) and the For step check."""
    with pytest.raises(ValueError):
        NewBuilder("bad-while", 1).WithClbits(1).While(5, 1, lambda b: b.X(0))
    with pytest.raises(ValueError):
        NewBuilder("bad-for", 1).WithClbits(1).For(0, 0, 3, 0, lambda b: b.X(0))
    with pytest.raises(ValueError):
        NewBuilder("bad-switch", 1).WithClbits(1).Switch([0, 1], {0: lambda b: None})


# ---- measureQubitDM (sim/densitymatrix/dynamic.go:211-255) ----------------------------------------------
def _mixed_1q(p, coh):
    return np.array([[p, coh], [np.conj(coh), 1 - p]])


@pytest.mark.parametrize("q", [0, 1, 2])
def test_dm_measure_qubit_closed_forms(q):
    """rho = rho_2 (x) rho_1 (x) rho_0: P0 is qubit q's own p; the outcome rule r >= P0; the kept
    block is Pi rho Pi / p, the rest zero; a certain outcome (p < 1e-15) leaves rho as it is."""
    singles = [_mixed_1q(0.3, 0.1 + 0.2j), _mixed_1q(0.8, -0.05j), _mixed_1q(0.55, 0.3)]
    rho = np.kron(singles[2], np.kron(singles[1], singles[0]))
    p0 = singles[q][0, 0].real
    for r, outcome in [(0.0, 0), (p0 - 1e-9, 0), (p0, 1), (0.999, 1)]:
        work = rho.reshape(-1).copy()
        assert O.dm_measure_qubit(work, 3, q, r) == outcome
        proj = np.diag([1.0 if ((i >> q) & 1) == outcome else 0.0 for i in range(8)])
        want = proj @ rho @ proj / (p0 if outcome == 0 else 1 - p0)
        assert np.max(np.abs(work.reshape(8, 8) - want)) < 1e-15
    # P(1) = 2^-51 < 1e-15 and r >= P0: outcome 1 with rho untouched
    eps = 2.0 ** -51
    near = np.zeros((8, 8), dtype=complex)
    near[0, 0], near[1 << q, 1 << q] = 1 - eps, eps
    work = near.reshape(-1).copy()
    assert O.dm_measure_qubit(work, 3, q, 1 - 2.0 ** -52) == 1
    assert np.array_equal(work.reshape(8, 8), near)


def test_dm_measure_matches_statevector_measure_on_pure_states():
    """On rho = |psi><psi|, the DM measurement gives the statevector measurement's outcome and
    |psi'><psi'| of its collapsed state (orc_sv_measure_qubit is pinned above)."""
    rng = np.random.default_rng(8)
    for trial in range(20):
        psi = rng.normal(size=16) + 1j * rng.normal(size=16)
        psi /= np.linalg.norm(psi)
        q, r = int(rng.integers(4)), float(rng.random())
        rho = np.outer(psi, psi.conj()).reshape(-1)
        o1 = O.dm_measure_qubit(rho, 4, q, r)
        o2 = O.sv_measure_qubit(psi, q, r)
        assert o1 == o2
        assert np.max(np.abs(rho - np.outer(psi, psi.conj()).reshape(-1))) < 1e-14


def test_oracle_dm_dynamic_equals_statevector_dynamic_without_noise():
    """Noiseless dynamic circuits: the DM shot loop and the statevector shot loop give the same
    clbits for the same uniforms, and the DM's final state is |psi><psi| of the statevector's."""
    import sys
    sys.path.insert(0, str(__import__("pathlib").Path(__file__).parent))
    from circuits import random_dynamic_circuit

    for seed in range(4):
        rng = np.random.default_rng(900 + seed)
        c = random_dynamic_circuit(3, 3, 4, rng, density=True)
        U = rng.random((20, 3000))
        cs, psi = OracleDynamic(3).run_shots(c, U)
        cd, rho = OracleDynamic(3, density=True).run_shots(c, U)
        assert np.array_equal(cs, cd)
        assert np.max(np.abs(rho - np.outer(psi, psi.conj()).reshape(-1))) < 1e-12


# ---- flattening into the device program (goqu_amd.dynamic) ---------------------------------------------
def test_flatten_lists_the_reference_cases_in_order():
    from goqu_amd import dynamic as DY

    b = NewBuilder("flat", 2).WithClbits(2)
    b.H(0).CNOT(0, 1).Measure(0, 0).If(0, 1, G.X, 1).Reset(1).Barrier()
    b.IfElseBlock(1, 0, lambda x: x.Z(0), lambda x: x.Measure(1, 1))
    b.Switch([0, 1], {2: lambda x: x.X(0), 0: lambda x: None}, lambda x: x.Y(1))
    b.For(0, 0, 5, 2, lambda x: x.S(0))
    b.While(1, 1, lambda x: x.Measure(1, 1))
    p = DY.flatten(b.Build(), lambda buf, op: N.add_sv_gate(buf, op.Gate, op.Qubits))
    want = [
        (N.DYN_GATES, 0, 2, 0), (N.DYN_MEASURE, 0, 0, 0),
        (N.DYN_BRANCH, 0, 0, 0), (N.DYN_CASE, 0, 1, 1), (N.DYN_GATES, 2, 1, 0), (N.DYN_END, 0, 0, 0),
        (N.DYN_RESET, 1, 0, 0),
        (N.DYN_BRANCH, 0, 0, 0), (N.DYN_CASE, 1, 1, 0), (N.DYN_GATES, 3, 1, 0), (N.DYN_DEFAULT, 0, 0, 0),
        (N.DYN_MEASURE, 1, 1, 0), (N.DYN_END, 0, 0, 0),
        (N.DYN_BRANCH, 0, 0, 0), (N.DYN_CASE, 2, 2, 0), (N.DYN_CASE, 2, 2, 2), (N.DYN_GATES, 4, 1, 0),
        (N.DYN_DEFAULT, 0, 0, 0), (N.DYN_GATES, 5, 1, 0), (N.DYN_END, 0, 0, 0),
        (N.DYN_FOR, 3, 0, 0), (N.DYN_GATES, 6, 1, 0), (N.DYN_END, 0, 0, 0),
        (N.DYN_WHILE, 4, 1, 1), (N.DYN_MEASURE, 1, 1, 0), (N.DYN_END, 0, 0, 0),
    ]
    assert p.instrs == want
    assert p.cond_bits == [0, 1, 0, 1, 1]
    assert p.draws == 1 + 1 + 1 + 1000  # measure, reset, the else-branch measure, the while body
