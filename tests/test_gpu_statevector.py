"""Statevector parity on the GPU: the HIP path against the CPU oracle on the same seeded
circuits.  complex128: max |delta amplitude| <= 1e-12; complex64 (fp32 arithmetic, fp64
reductions) against the fp64 oracle: <= 1e-5.  Sampled indices identical for the same
uniforms."""
import math

import numpy as np
import pytest

from circuits import config1_circuit, config2_circuit, random_circuit
from goqu_amd import _native as N
from goqu_amd import gate as G
from goqu_amd import statevector
from goqu_amd.circuit import NewBuilder
from goqu_amd.pauli import NewPauliString, NewPauliSum, Parse, X, Y, Z
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL128 = 1e-12
TOL64 = 1e-5


def _check(c, tol=TOL128, precision=N.C128):
    sim = statevector.New(c.NumQubits(), precision)
    sim.Evolve(c)
    got = sim.StateVector()
    want = O.sv_evolve(c)
    err = float(np.max(np.abs(got - want)))
    assert err <= tol, f"max |delta amplitude| = {err:.3e}"
    return sim, want


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 9, 11, 13, 14, 16])
@pytest.mark.parametrize("seed", [0, 1])
def test_random_circuits_all_gates(n, seed):
    rng = np.random.default_rng(1000 * n + seed)
    c = random_circuit(n, 6, rng)
    print(f"random circuit n={n} seed={seed} ops={c.NumOps()}", flush=True)
    _check(c)


@pytest.mark.parametrize("n", [5, 12, 17, 20])
def test_random_circuits_c64(n):
    rng = np.random.default_rng(77 + n)
    c = random_circuit(n, 4, rng)
    _check(c, TOL64, N.C64)


@pytest.mark.parametrize("precision,tol", [(N.C128, TOL128), (N.C64, TOL64)])
def test_swap_relabel_stages(precision, tol):
    """SWAPs among high qubits only (a stage without instructions that stores permuted), SWAPs
    that move low qubits, and SWAPs mixed with rotations and CNOTs, on random states."""
    from test_planner_emulated import _swap_circuits, random_state

    n = 14
    for i, c in enumerate(_swap_circuits(n)):
        s0 = random_state(n, 40 + i)
        sim = statevector.New(n, precision)
        sim.SetStateVector(s0)
        sim.Apply(c)
        got = sim.StateVector()
        sim.Close()
        want = O.sv_evolve(c, s0.copy())
        assert np.max(np.abs(got - want)) <= tol, i


def test_config1_parity_and_sampling():
    c = config1_circuit(20, 20, seed=11)
    sim, want = _check(c)
    u = np.random.default_rng(5).random(1024)
    got = sim.Sample(u)
    ref = O.sv_sample(want, u)
    assert np.array_equal(got, ref)
    counts = sim.Run(c, 1024, uniforms=u)
    assert sum(counts.values()) == 1024
    assert all(len(k) == 20 for k in counts)


def test_config2_shape_22q():
    c = config2_circuit(22, 8, seed=3)
    _check(c)


def test_high_qubit_gates_24q_unitarity():
    """U then U^dagger returns |0...0> at a size where the oracle is still quick."""
    n = 24
    c = config2_circuit(n, 6, seed=4)
    sim = statevector.New(n)
    sim.Evolve(c)
    b = NewBuilder("inv", n)
    for op in reversed(c.Ops()):
        if op.Gate is not None:
            b.Apply(op.Gate.Inverse(), *op.Qubits)
    sim.Apply(b.Build())
    sv = sim.StateVectorRange(0, 4)
    assert abs(sv[0] - 1) < 1e-11
    assert abs(sim.Norm2() - 1) < 1e-12


def test_expect_paulis_vs_oracle():
    n = 12
    rng = np.random.default_rng(3)
    c = random_circuit(n, 4, rng)
    sim, psi = _check(c)
    terms = []
    for t in range(60):
        ops = {}
        for q in rng.choice(n, int(rng.integers(0, 6)), replace=False):
            ops[int(q)] = [X, Y, Z][rng.integers(3)]
        terms.append(NewPauliString(complex(rng.normal(), rng.normal() * (t % 3 == 0)), ops, n))
    # terms sharing one x mask (Z strings)
    for t in range(40):
        ops = {int(q): Z for q in rng.choice(n, int(rng.integers(1, n)), replace=False)}
        terms.append(NewPauliString(rng.normal(), ops, n))
    got = sim.ExpectTerms(terms)
    for t, ps in enumerate(terms):
        want = O.sv_expect(psi, ps)
        assert abs(got[t] - want) < 1e-12, (t, str(ps), got[t], want)
    H = NewPauliSum(terms)
    assert abs(sim.ExpectPauliSum(H) - sum(O.sv_expect(psi, p) for p in terms).real) < 1e-11


def test_expectation_value_rounding():
    c = NewBuilder("bell", 2).H(0).CNOT(0, 1).Build()
    sim = statevector.New(2)
    sim.Evolve(c)
    assert sim.ExpectationValue([0, 1]) == 1.0
    assert sim.ExpectationValue([0]) == 0.0


def test_reset_semantics():
    """sim_test.go:666-746: Evolve's reset concentrates probability into |0>."""
    s2 = 1 / math.sqrt(2)
    cases = [
        (NewBuilder("xr", 1).X(0).Reset(0).Build(), [1, 0]),
        (NewBuilder("hr", 1).H(0).Reset(0).Build(), [1, 0]),
        (NewBuilder("br", 2).H(0).CNOT(0, 1).Reset(0).Build(), [s2, 0, s2, 0]),
        (NewBuilder("r0", 1).Reset(0).Build(), [1, 0]),
    ]
    for c, want in cases:
        sim = statevector.New(c.NumQubits())
        sim.Evolve(c)
        assert np.max(np.abs(sim.StateVector() - np.array(want))) < 1e-12


def test_measure_qubit_matches_oracle():
    n = 8
    rng = np.random.default_rng(9)
    c = random_circuit(n, 3, rng, three_q=False, controlled=False)
    sim, psi = _check(c)
    ref = psi.copy()
    for q, u in [(3, 0.3), (0, 0.9), (7, 0.5)]:
        o_gpu = sim.MeasureQubit(q, u)
        o_cpu = O.sv_measure_qubit(ref, q, u)
        assert o_gpu == o_cpu
    assert np.max(np.abs(sim.StateVector() - ref)) < 1e-12


def test_errors_are_status_codes():
    sim = statevector.New(3)
    buf = N.OpBuffer()
    buf.add(N.OP_GATE, [0, 1, 2], np.eye(8))
    buf.add(N.OP_GATE, [0, 1, 5], np.eye(8))  # qubit out of range
    with pytest.raises(N.GoquError) as e:
        sim.ApplyOps(buf)
    assert e.value.code == N.ERR_INVALID_ARG
    buf = N.OpBuffer()
    buf.add(N.OP_GATE, [0], None)  # 1q gate without matrix: sim.go:139-142
    with pytest.raises(N.GoquError) as e:
        sim.ApplyOps(buf)
    assert e.value.code == N.ERR_NO_MATRIX


def test_bell_counts_statistics():
    """sim_test.go:135-167: Bell counts in [4000, 6000] of 10000."""
    c = NewBuilder("bell", 2).H(0).CNOT(0, 1).MeasureAll().Build()
    sim = statevector.New(2)
    counts = sim.Run(c, 10000, rng=np.random.default_rng(1))
    assert set(counts) <= {"00", "11"}
    assert 4000 <= counts.get("00", 0) <= 6000


def test_dynamic_teleportation_like():
    """RunDynamic with mid-circuit measurement and feed-forward (dynamic.go)."""
    b = NewBuilder("ff", 2).WithClbits(2)
    b.X(0).Measure(0, 0).If(0, 1, G.X, 1).Measure(1, 1)
    c = b.Build()
    assert c.IsDynamic()
    sim = statevector.New(2)
    counts = sim.Run(c, 20, rng=np.random.default_rng(2))
    assert counts == {"11": 20}


@pytest.mark.parametrize("k", [0, 1])
def test_reset_known_answers_gpu(k):
    """H0 H1 Reset1 S0 -> [1, i, 0, 0]/sqrt2 and H0 H1 RZ(0.7, 0) Reset1 -> [1, 1, 0, 0]/sqrt2: no
    op moves across the (non-linear) reset."""
    from test_planner_emulated import _reset_kat_circuits

    c, want = _reset_kat_circuits()[k]
    for prec, tol in [(N.C128, 1e-15), (N.C64, 1e-7)]:
        sim = statevector.New(2, prec)
        sim.Evolve(c)
        assert np.max(np.abs(sim.StateVector() - want)) < tol
        sim.Close()


@pytest.mark.parametrize("n", [3, 7, 12])
def test_random_circuits_with_resets(n):
    rng = np.random.default_rng(4200 + n)
    c = random_circuit(n, 5, rng, resets=0.2)
    print(f"reset circuit n={n} ops={c.NumOps()}", flush=True)
    _check(c)


def test_dynamic_reference_kats_gpu():
    """The reference's dynamic-circuit tests (dynamic_test.go:14-162) on the device path:
    conditional, conditional not fired, reset and reuse, 3-qubit teleportation."""
    from test_oracle_dynamic import teleport_circuit

    c = NewBuilder("cond", 2).WithClbits(1).X(0).Measure(0, 0).If(0, 1, G.X, 1).Build()
    assert statevector.New(2).Run(c, 200, rng=np.random.default_rng(1)) == {"1": 200}
    c = NewBuilder("cond_no", 2).WithClbits(1).Measure(0, 0).If(0, 1, G.X, 1).Build()
    assert statevector.New(2).Run(c, 200, rng=np.random.default_rng(2)) == {"0": 200}
    c = NewBuilder("reset_reuse", 1).WithClbits(2).H(0).Measure(0, 0).Reset(0).H(0).Measure(0, 1).Build()
    counts = statevector.New(1).Run(c, 2000, rng=np.random.default_rng(3))
    assert abs(sum(v for k, v in counts.items() if k[0] == "1") / 2000 - 0.5) < 0.05
    counts = statevector.New(3).Run(teleport_circuit(), 1000, rng=np.random.default_rng(4))
    assert all(len(k) == 3 and k[0] == "1" for k in counts) and len(counts) == 4


def test_measure_qubit_closed_forms_gpu():
    """measureQubit on product states: outcome rule r >= P0, collapse and renormalisation exact."""
    from test_oracle_dynamic import _product

    t, a, b = 0.3, 0.6, 0.8j
    p0 = math.cos(t) ** 2
    for prec, tol in [(N.C128, 1e-15), (N.C64, 1e-6)]:  # complex64: a few fp32 ulps
        for r, outcome in [(0.0, 0), (p0 * 0.999, 0), (p0, 1), (0.999999, 1)]:
            sim = statevector.New(2, prec)
            sim.SetStateVector(_product(t, 0.7, a, b))
            assert sim.MeasureQubit(1, r) == outcome
            want = np.kron([1, 0], [a, b]) if outcome == 0 else np.kron([0, np.exp(0.7j)], [a, b])
            assert np.max(np.abs(sim.StateVector() - want)) < tol
            sim.Close()


def test_stateprep_kats_and_subset_parity():
    """StatePrep (sim.go:105-136): the reference's known answers (sim_test.go:775-872) on the
    device, and random StatePrep on qubit subsets (the gate's RY + CNOT cascade) against the
    oracle."""
    from test_oracle_kats import _stateprep_kat_circuits

    for c, want in _stateprep_kat_circuits():
        sim = statevector.New(c.NumQubits())
        sim.Evolve(c)
        assert np.max(np.abs(sim.StateVector() - np.asarray(want))) < 1e-12, c.Name()
        sim.Close()
    rng = np.random.default_rng(12)
    for k, qs in [(1, [3]), (2, [4, 1]), (3, [0, 5, 2]), (4, [6, 1, 3, 0])]:
        v = rng.normal(size=1 << k) + 1j * rng.normal(size=1 << k)
        v /= np.linalg.norm(v)
        b = NewBuilder("sp", 7)
        for q in range(7):
            b.H(q)
        b.Apply(G.StatePrep(v), *qs).CNOT(0, 6)
        c = b.Build()
        _check(c)


@pytest.mark.parametrize("n,nt,nc", [(11, 5, 1), (12, 6, 2), (16, 9, 1), (22, 7, 3)])
def test_wide_controlled_gates_mfma(n, nt, nc):
    """Controlled gates with 5..9 target qubits (the reference's full controlled-matrix path,
    kernel_controlled.go:249-292) run as FP64 MFMA contractions: against the oracle, complex128
    and complex64."""
    from scipy.stats import unitary_group

    rng = np.random.default_rng(100 * nt + n)
    U = unitary_group.rvs(1 << nt, random_state=int(rng.integers(1 << 30)))
    qs = [int(x) for x in rng.choice(n, nt + nc, replace=False)]
    b = NewBuilder("wide", n)
    for q in range(n):
        b.U3(*rng.uniform(0, 2 * np.pi, 3), q)
    b.Apply(G.Controlled(G.Unitary("U", U.reshape(-1)), nc), *qs)
    b.CNOT(0, n - 1).RZ(0.4, 2)
    c = b.Build()
    print(f"wide n={n} targets={nt} controls={nc}", flush=True)
    _check(c)
    _check(c, TOL64, N.C64)


def test_program_cache_repeats_and_misses():
    """Apply keeps the planned programs of recent op lists per handle: repeating an op list
    reuses them (twice without a reset is U.U), an op list that differs in one matrix entry or a
    new tile shape plans afresh, and more distinct lists than the cache holds evict the oldest."""
    n = 12
    rng = np.random.default_rng(91)
    c = random_circuit(n, 5, rng)
    sim = statevector.New(n)
    for _ in range(3):  # Evolve of one circuit, three times
        sim.Evolve(c)
        assert np.max(np.abs(sim.StateVector() - O.sv_evolve(c))) <= TOL128
    sim.Apply(c)  # no reset: the cached program runs on U|0>
    want2 = O.sv_evolve(c, O.sv_evolve(c))
    assert np.max(np.abs(sim.StateVector() - want2)) <= TOL128
    # one rotation angle changed by 1e-9: a different op list
    b = NewBuilder("c1", n)
    b2 = NewBuilder("c2", n)
    for q in range(n):
        b.RY(0.3 + 0.1 * q, q)
        b2.RY(0.3 + 0.1 * q + (1e-9 if q == 5 else 0.0), q)
    for q in range(n - 1):
        b.CNOT(q, q + 1)
        b2.CNOT(q, q + 1)
    c1, c2 = b.Build(), b2.Build()
    for cc in (c1, c2, c1, c2):
        sim.Evolve(cc)
        assert np.max(np.abs(sim.StateVector() - O.sv_evolve(cc))) <= TOL128
    # a new tile shape replans; more lists than the cache holds
    sim.SetPlan(8, 3)
    circuits = [random_circuit(n, 2, np.random.default_rng(200 + i)) for i in range(6)]
    for cc in circuits + circuits[::-1]:
        sim.Evolve(cc)
        assert np.max(np.abs(sim.StateVector() - O.sv_evolve(cc))) <= TOL128
    sim.Close()
