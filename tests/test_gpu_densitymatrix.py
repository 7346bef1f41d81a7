"""Density-matrix parity on the GPU against the oracle (goqu's two-pass / Kraus-loop
algorithm restated in C), max |delta rho| <= 1e-12, plus the reference's channel KATs."""
import math

import numpy as np
import pytest

from circuits import random_circuit
from goqu_amd import densitymatrix as D
from goqu_amd import gate as G
from goqu_amd import noise as NZ
from goqu_amd.circuit import NewBuilder
from goqu_amd.pauli import NewPauliString, X, Y, Z
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-12


def noise_model():
    nm = NZ.New()
    nm.AddDefaultError(1, NZ.Depolarizing1Q(1e-2))
    nm.AddDefaultError(2, NZ.Depolarizing2Q(2e-2))
    nm.AddGateError("H", NZ.AmplitudeDamping(0.05))
    nm.AddGateError("delay", NZ.PhaseDamping(0.03))
    nm.AddGateQubitError("X", [0], NZ.GeneralizedAmplitudeDamping(0.3, 0.1))
    nm.AddReadoutError(0, NZ.NewReadoutError(0.02, 0.05))
    return nm


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 9])
def test_noisy_dm_parity(n):
    rng = np.random.default_rng(500 + n)
    c = random_circuit(n, 4, rng, three_q=False, controlled=False)
    b = NewBuilder("withdelay", n)
    b.Compose(c)
    for q in range(n):
        b.Delay(q, 1e-6)
    if n > 1:
        b.Reset(1)
    c = b.Build()
    nm = noise_model()
    dm = D.New(n).WithNoise(nm)
    dm.Evolve(c)
    got = dm.DensityMatrix()
    want = O.dm_zero(n)
    assert O.dm_apply_ops(want, n, D.marshal(c, nm)) == 0
    assert np.max(np.abs(got - want)) <= TOL
    p01 = np.full(n, -1.0)
    p10 = np.full(n, -1.0)
    p01[0], p10[0] = 0.02, 0.05
    assert np.max(np.abs(dm.Probabilities() - O.dm_probs(want, n, p01, p10))) <= TOL
    for _ in range(6):
        ops = {int(q): [X, Y, Z][rng.integers(3)] for q in rng.choice(n, int(rng.integers(1, n + 1)), replace=False)}
        ps = NewPauliString(rng.normal(), ops, n)
        assert abs(dm.ExpectPauliString(ps) - O.dm_expect(want, n, ps).real) <= TOL
    assert abs(dm.Purity() - O.dm_purity(want, n)) <= 1e-12
    # both sides are rounded to 14 decimals (the reference rounds); sums that differ in the last
    # bits may fall on either side of a rounding boundary, one unit of the 14th decimal apart
    assert abs(dm.ExpectationValue([0]) - O.dm_expectation_value(want, n, [0])) <= 1.5e-14
    u = rng.random(200)
    counts = dm.Run(c, 200, uniforms=u)
    assert sum(counts.values()) == 200


def test_dm_kats():
    nm = NZ.New()
    nm.AddDefaultError(1, NZ.Depolarizing1Q(0.75))
    dm = D.New(1).WithNoise(nm)
    dm.Evolve(NewBuilder("x", 1).X(0).Build())
    rho = dm.DensityMatrix()
    assert abs(rho[0].real - 0.5) < 1e-10 and abs(rho[3].real - 0.5) < 1e-10
    nm = NZ.New()
    nm.AddDefaultError(1, NZ.Depolarizing1Q(0.5))
    dm = D.New(1).WithNoise(nm)
    dm.Evolve(NewBuilder("h", 1).H(0).Build())
    assert round(dm.Purity(), 4) == 0.5556


@pytest.mark.parametrize("n", [4, 6])
def test_dm_multiqubit_gates_exact(n):
    """3+ qubit and multi-controlled gates act exactly: rho = |psi><psi| of the statevector."""
    rng = np.random.default_rng(n)
    c = random_circuit(n, 3, rng, three_q=True, controlled=True)
    dm = D.New(n)
    dm.Evolve(c)
    psi = O.sv_evolve(c)
    assert np.max(np.abs(dm.DensityMatrix() - np.outer(psi, psi.conj()).reshape(-1))) <= TOL
    assert abs(dm.Fidelity(psi) - 1) < 1e-12


def test_dm_dynamic_feed_forward_and_reset():
    """densitymatrix RunDynamic (dynamic.go): measurement collapse, feed-forward and reset."""
    b = NewBuilder("ff", 2).WithClbits(2)
    b.X(0).Measure(0, 0).If(0, 1, G.X, 1).Measure(1, 1)
    c = b.Build()
    dm = D.New(2)
    assert dm.Run(c, 10, rng=np.random.default_rng(1)) == {"11": 10}
    b = NewBuilder("rs", 1).WithClbits(1)
    b.H(0).Reset(0).Measure(0, 0)
    assert D.New(1).Run(b.Build(), 10, rng=np.random.default_rng(2)) == {"0": 10}
    # collapse renormalises: after measuring qubit 0 of a Bell pair, qubit 1 agrees
    b = NewBuilder("bell", 2).WithClbits(2)
    b.H(0).CNOT(0, 1).Measure(0, 0).H(1).H(1).Measure(1, 1)
    counts = D.New(2).Run(b.Build(), 200, rng=np.random.default_rng(3))
    assert set(counts) <= {"00", "11"} and 60 <= counts.get("00", 0) <= 140


def test_dm_14q_trace_and_hermitian():
    """At the 14-qubit configuration size: trace 1 and Hermiticity through noisy layers."""
    n = 14
    rng = np.random.default_rng(14)
    b = NewBuilder("c5", n)
    for _ in range(2):
        for q in range(n):
            b.U3(*rng.uniform(0, 2 * math.pi, 3), q)
        perm = rng.permutation(n)
        for i in range(0, n - 1, 2):
            b.CNOT(int(perm[i]), int(perm[i + 1]))
    nm = NZ.New()
    nm.AddDefaultError(1, NZ.Depolarizing1Q(1e-3))
    nm.AddDefaultError(2, NZ.Depolarizing2Q(1e-2))
    dm = D.New(n).WithNoise(nm)
    dm.Evolve(b.Build())
    p = dm.DiagonalProbs()
    assert abs(p.sum() - 1) < 1e-10
    blk = np.empty((64, 64), dtype=complex)
    full = dm.DensityMatrix()
    dim = 1 << n
    for r in range(64):
        blk[r] = full[r * dim: r * dim + 64]
    assert np.max(np.abs(blk - blk.conj().T)) < 1e-13


@pytest.mark.parametrize("n", [2, 4, 7])
def test_dm_complex_parity_channels_and_dense_fallback(n):
    """Channels with complex parity blocks (one and two qubits) and a parity-mixing channel
    (dense superoperator), on the GPU against the oracle's Kraus loop."""
    p = 0.3
    k1 = np.sqrt(p) * np.array([[0, 1], [1j, 0]])
    c1 = NZ.Channel("cplx1", 1, [np.sqrt(1 - p) * np.eye(2), k1])
    k2 = np.sqrt(p) * np.kron(np.array([[0, 1], [1j, 0]]), np.array([[np.exp(0.4j), 0], [0, 1]]))
    c2 = NZ.Channel("cplx2", 2, [np.sqrt(1 - p) * np.eye(4), k2])
    th = 0.7
    r = np.array([[np.cos(th), -1j * np.sin(th)], [-1j * np.sin(th), np.cos(th)]])
    c3 = NZ.Channel("mix", 1, [np.sqrt(1 - p) * np.eye(2), np.sqrt(p) * r])
    nm = NZ.New()
    nm.AddDefaultError(1, c1)
    nm.AddDefaultError(2, c2)
    nm.AddGateError("H", c3)
    rng = np.random.default_rng(90 + n)
    b = NewBuilder("dmc", n)
    for q in range(n):
        b.U3(*rng.uniform(0, 2 * np.pi, 3), q)
    b.CNOT(0, 1).H(n - 1).CZ(n - 2, n - 1)
    if n > 3:
        b.CNOT(3, 0).H(2)
    c = b.Build()
    dm = D.New(n).WithNoise(nm)
    dm.Evolve(c)
    got = dm.DensityMatrix().reshape(-1)
    dm.Close()
    want = O.dm_zero(n)
    assert O.dm_apply_ops(want, n, D.marshal(c, nm)) == 0
    assert np.max(np.abs(got - want)) <= TOL


def test_dm_reset_known_answers_gpu():
    """densitymatrix/sim_test.go:550-620 (TestEvolve_Reset) on the device."""
    zero = np.zeros(4)
    zero[0] = 1
    for c in [NewBuilder("xreset", 1).X(0).Reset(0).Build(), NewBuilder("hreset", 1).H(0).Reset(0).Build()]:
        dm = D.New(1)
        dm.Evolve(c)
        assert np.max(np.abs(dm.DensityMatrix() - zero)) < 1e-10
    dm = D.New(2)
    dm.Evolve(NewBuilder("bellreset", 2).H(0).CNOT(0, 1).Reset(0).Build())
    assert np.max(np.abs(dm.DensityMatrix().reshape(4, 4) - np.diag([0.5, 0, 0.5, 0]))) < 1e-10


def test_dm_expectation_value_closed_forms_gpu():
    p = 0.3
    rho = np.zeros((4, 4), dtype=np.complex128)
    rho[0, 0], rho[3, 3] = p, 1 - p
    rho[0, 3] = rho[3, 0] = 0.2
    dm = D.New(2)
    dm.SetDensityMatrix(rho.reshape(-1))
    assert dm.ExpectationValue([0, 1]) == 1.0
    assert abs(dm.ExpectationValue([0]) - (2 * p - 1)) < 1e-14
    nm = NZ.New()
    nm.AddGateError("I", NZ.Depolarizing1Q(0.3))
    dm = D.New(1).WithNoise(nm)
    dm.Evolve(NewBuilder("dep", 1).Apply(G.I, 0).Build())
    assert dm.ExpectationValue([0]) == round((1 - 0.4) * 1e14) / 1e14


@pytest.mark.parametrize("n", [3, 9])
def test_dm_sampled_indices_match_oracle(n):
    """DM Run's sampling (measure.go:93-103) on the device against the oracle's inverse CDF over
    the same readout-corrected probabilities and the same uniforms: identical indices."""
    rng = np.random.default_rng(700 + n)
    c = random_circuit(n, 3, rng, three_q=False, controlled=False)
    nm = noise_model()
    dm = D.New(n).WithNoise(nm)
    dm.Evolve(c)
    probs = dm.Probabilities()
    u = rng.random(5000)
    out = np.empty(len(u), dtype=np.uint64)
    from goqu_amd import _native as N

    N.check(N.lib().goqu_dm_sample(dm._h, N.dptr(probs), N.dptr(u), len(u), N.u64ptr(out)))
    assert np.array_equal(out, O.sample_probs(probs, u))
    dm.Close()


def test_dm_wide_controlled_gate():
    """A controlled gate on 5 targets acts on rho exactly: U on the row qubits and conj(U) on the
    column qubits, each a wide MFMA op; rho = |psi><psi| of the statevector."""
    from scipy.stats import unitary_group

    n = 7
    rng = np.random.default_rng(75)
    U = unitary_group.rvs(32, random_state=5)
    b = NewBuilder("dmwide", n)
    for q in range(n):
        b.U3(*rng.uniform(0, 2 * np.pi, 3), q)
    b.Apply(G.Controlled(G.Unitary("U", U.reshape(-1)), 1), 6, 0, 2, 3, 4, 5)
    c = b.Build()
    dm = D.New(n)
    dm.Evolve(c)
    psi = O.sv_evolve(c)
    assert np.max(np.abs(dm.DensityMatrix() - np.outer(psi, psi.conj()).reshape(-1))) <= TOL


def test_dm_three_qubit_dense_gates():
    """3-qubit gates act on rho exactly as dense ops on three row and three column register slots
    (stages for the big density-matrix kernel), next to channel-free 1-2 qubit gates: rho equals
    |psi><psi| of the statevector oracle (the density-matrix oracle, like the reference, takes no
    3-qubit gates)."""
    from scipy.stats import unitary_group

    n = 6
    rng = np.random.default_rng(76)
    U = unitary_group.rvs(8, random_state=7)
    b = NewBuilder("dm3", n)
    for q in range(n):
        b.U3(*rng.uniform(0, 2 * np.pi, 3), q)
    b.Apply(G.Unitary("U", U.reshape(-1), 3), 4, 0, 2)
    b.CNOT(1, 3).Apply(G.CCX, 0, 5, 3).Apply(G.CSWAP, 2, 1, 4)
    b.Apply(G.Unitary("U", U.reshape(-1), 3), 1, 5, 3)
    c = b.Build()
    dm = D.New(n)
    dm.Evolve(c)
    psi = O.sv_evolve(c)
    assert np.max(np.abs(dm.DensityMatrix() - np.outer(psi, psi.conj()).reshape(-1))) <= TOL
    dm.Close()
