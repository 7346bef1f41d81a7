"""Batches of circuits on the device (goqu_sv_batch_* / goqu_dm_batch_*) and their callers
(sweep.RunSim / RunDensitySim, the mitigation executors, the local backend) against the oracle:
every circuit's per-term expectations within 1e-12 (complex128), sampled indices identical for
the same uniforms."""
import math

import numpy as np
import pytest

from circuits import random_circuit, random_pauli_sum
from goqu_amd import backend as BK
from goqu_amd import batch as B
from goqu_amd import densitymatrix as D
from goqu_amd import gate as G
from goqu_amd import mitigation as MT
from goqu_amd import noise as NZ
from goqu_amd import observe
from goqu_amd import sweep as SW
from goqu_amd import _native as N
from goqu_amd.circuit import Literal, Mul, New, NewBuilder
from goqu_amd.qaoa import EfficientSU2
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _oracle_terms(c, H):
    psi = O.sv_evolve(c)
    return np.array([O.sv_expect(psi, t) for t in H.Terms()])


def _mixed_batch(n, count, seed):
    """Circuits of two structures (several launch groups) plus random one-offs."""
    rng = np.random.default_rng(seed)
    base = [random_circuit(n, 3, np.random.default_rng(seed + k)) for k in range(2)]
    out = []
    for i in range(count):
        if i % 3 == 2:
            out.append(random_circuit(n, 2, rng))
        else:
            out.append(base[i % 3])
    return out


@pytest.mark.parametrize("n", [3, 8, 13])
def test_batch_expect_matches_oracle(n):
    circs = _mixed_batch(n, 9, 100 + n)
    H = random_pauli_sum(n, 24, np.random.default_rng(n))
    print(f"batch expect n={n} circuits={len(circs)}", flush=True)
    got = B.expect_terms(circs, H)
    for i, c in enumerate(circs):
        assert np.max(np.abs(got[i] - _oracle_terms(c, H))) < 1e-12, i
    e = B.energies(circs, H)
    assert np.max(np.abs(e - np.array([_oracle_terms(c, H).sum().real for c in circs]))) < 1e-12


def test_batch_chunks_and_complex64(monkeypatch):
    n = 10
    circs = _mixed_batch(n, 11, 300)
    H = random_pauli_sum(n, 12, np.random.default_rng(3))
    want = np.array([_oracle_terms(c, H) for c in circs])
    monkeypatch.setenv("GOQU_BATCH_MAX_STATES", "2")
    assert np.max(np.abs(B.expect_terms(circs, H) - want)) < 1e-12
    monkeypatch.delenv("GOQU_BATCH_MAX_STATES")
    assert np.max(np.abs(B.expect_terms(circs, H, precision=N.C64) - want)) < 2e-5
    # a tile shape chosen by the caller (the autotuner's knob) changes nothing in the results
    assert np.max(np.abs(B.expect_terms(circs, H, plan=(10, 3)) - want)) < 1e-12


def test_batch_sample_matches_oracle():
    n = 9
    circs = _mixed_batch(n, 6, 400)
    U = np.random.default_rng(4).random((len(circs), 500))
    got = B.sample(circs, 500, uniforms=U)
    for i, c in enumerate(circs):
        assert np.array_equal(got[i], O.sv_sample(O.sv_evolve(c), U[i])), i


def test_batch_with_a_wide_op_and_an_efficientsu2_vqe_batch():
    rng = np.random.default_rng(500)
    inner = np.linalg.qr(rng.normal(size=(32, 32)) + 1j * rng.normal(size=(32, 32)))[0]
    wide = NewBuilder("w", 7)
    for q in range(7):
        wide.H(q)
    wide.Apply(G.Controlled(G.Unitary("U5", inner.reshape(-1)), 1), 0, 1, 2, 3, 4, 5)
    circs = [wide.Build(), random_circuit(7, 2, rng)]
    H = random_pauli_sum(7, 10, rng)
    got = B.expect_terms(circs, H)
    for i, c in enumerate(circs):
        assert np.max(np.abs(got[i] - _oracle_terms(c, H))) < 1e-12
    # VQE: EfficientSU2(10, 2) at 40 parameter sets in one batch
    c, npar = EfficientSU2(10, 2)
    params = rng.uniform(0, 2 * math.pi, (40, npar))
    H = random_pauli_sum(10, 30, rng)
    e = B.vqe_energies(c, H, params)
    for k in (0, 17, 39):
        want = _oracle_terms(B.bind_all(c, params[k:k + 1])[0], H).sum().real
        assert abs(e[k] - want) < 1e-12


def _dm_oracle(c, n, nm):
    rho = O.dm_zero(n)
    assert O.dm_apply_ops(rho, n, D.marshal(c, nm)) == 0
    return rho


@pytest.mark.parametrize("noisy", [False, True])
def test_dm_batch_expect_and_sample_match_oracle(noisy):
    n = 4
    rng = np.random.default_rng(600)
    circs = [random_circuit(n, 3, rng, three_q=False, controlled=False) for _ in range(5)]
    nm = None
    if noisy:
        nm = NZ.New()
        nm.AddDefaultError(1, NZ.Depolarizing1Q(0.01))
        nm.AddDefaultError(2, NZ.Depolarizing2Q(0.03))
        nm.AddReadoutError(1, NZ.NewReadoutError(0.02, 0.05))
    H = random_pauli_sum(n, 15, rng)
    got = B.dm_expect_terms(circs, H, nm)
    U = rng.random((len(circs), 300))
    idx = B.dm_sample(circs, 300, nm, uniforms=U)
    p01, p10 = D._readout_arrays(nm, n)
    for i, c in enumerate(circs):
        rho = _dm_oracle(c, n, nm)
        want = np.array([O.dm_expect(rho, n, t) for t in H.Terms()])
        assert np.max(np.abs(got[i] - want)) < 1e-12
        probs = O.dm_probs(rho, n, p01, p10) if p01 is not None else O.dm_probs(rho, n)
        assert np.array_equal(idx[i], O.sample_probs(probs, U[i]))


def _ry_circuit():
    return NewBuilder("ry", 1).SymRY(New("theta"), 0).MeasureAll().Build()


def test_run_sim_reference_cases():
    """sweep/run_test.go:30-200 on the device batch."""
    res = SW.RunSim(None, _ry_circuit(), 1000, SW.Linspace("theta", 0, math.pi, 5), rng=np.random.default_rng(1))
    assert len(res) == 5 and all(r.Err is None for r in res)
    assert res[0].Counts.get("1", 0) <= 10 and res[4].Counts.get("0", 0) <= 10
    th, ph = New("theta"), New("phi")
    c = NewBuilder("multi", 2).SymRY(th, 0).SymRY(ph, 1).MeasureAll().Build()
    sw = SW.Product(SW.Linspace("theta", 0, math.pi, 3), SW.Linspace("phi", 0, math.pi, 2))
    res = SW.RunSim(None, c, 100, sw, rng=np.random.default_rng(2))
    assert [r.Index for r in res] == list(range(6)) and all(r.Err is None for r in res)
    calls = {"sweep": 0, "sim": 0}

    def wrap_sweep(ctx, info):
        calls["sweep"] += 1
        assert info.NumPoints == 3
        return ctx, lambda e: None

    def wrap_sim(ctx, info):
        calls["sim"] += 1
        return ctx, lambda e: None

    ctx = observe.WithHooks(None, observe.Hooks(WrapSim=wrap_sim, WrapSweep=wrap_sweep))
    assert len(SW.RunSim(ctx, _ry_circuit(), 100, SW.Linspace("theta", 0, math.pi, 3))) == 3
    assert calls == {"sweep": 1, "sim": 3}
    res = SW.RunDensitySim(None, _ry_circuit(), 1000, SW.Linspace("theta", 0, math.pi, 3), None)
    assert res[0].Counts.get("1", 0) <= 10
    nm = NZ.New()
    nm.AddDefaultError(1, NZ.BitFlip(0.1))
    res = SW.RunDensitySim(None, _ry_circuit(), 1000, SW.NewPoints("theta", [math.pi]), nm)
    assert res[0].Counts.get("0", 0) > 0
    fixed = NewBuilder("fixed", 1).X(0).MeasureAll().Build()
    assert SW.RunSim(None, fixed, 100, SW.NewPoints("ignored", [1.0]))[0].Counts == {"1": 100}


def test_run_sim_counts_match_oracle_per_point():
    th = New("theta")
    b = NewBuilder("p", 3).SymRY(th, 0).CNOT(0, 1).SymRX(Mul(Literal(0.5), th), 2).MeasureAll()
    c = b.Build()
    sw = SW.Linspace("theta", 0.1, 2.9, 6)
    U = np.random.default_rng(5).random((6, 200))
    res = SW.RunSim(None, c, 200, sw, uniforms=U)
    from goqu_amd.circuit import Bind

    for i, pt in enumerate(sw.Resolve()):
        idx = O.sv_sample(O.sv_evolve(Bind(c, pt)), U[i])
        vals, cnt = np.unique(idx, return_counts=True)
        assert res[i].Counts == {format(int(v), "03b"): int(k) for v, k in zip(vals, cnt)}


def test_mitigation_executors_and_local_backend():
    rng = np.random.default_rng(700)
    c = random_circuit(5, 3, rng)
    H = random_pauli_sum(5, 8, rng)
    want = _oracle_terms(c, H).sum().real
    assert abs(MT.StatevectorExecutor(H)(None, c) - want) < 1e-12
    assert abs(MT.StatevectorBatchExecutor(H)(None, [c, c])[1] - want) < 1e-12
    cd = random_circuit(3, 2, rng, three_q=False, controlled=False)
    Hd = random_pauli_sum(3, 5, rng)
    nm = NZ.New()
    nm.AddDefaultError(1, NZ.Depolarizing1Q(0.02))
    rho = _dm_oracle(cd, 3, nm)
    wd = sum(O.dm_expect(rho, 3, t) for t in Hd.Terms()).real
    assert abs(MT.DensityMatrixExecutor(Hd, nm)(None, cd) - wd) < 1e-12
    # the plugin wiring: local.New(WithMaxQubits(34), WithSimulator(rocm.NewSimulator))
    be = BK.New(BK.WithMaxQubits(34), BK.WithSimulator(BK.NewSimulator))
    seen = []
    ctx = observe.WithHooks(None, observe.Hooks(WrapSim=lambda ctx, info: (seen.append(info) or ctx,
                                                                           lambda e: seen.append(e))))
    bell = NewBuilder("bell", 2).H(0).CNOT(0, 1).MeasureAll().Build()
    job = be.Submit(ctx, BK.SubmitRequest(bell, 400))
    counts = be.Result(ctx, job.ID).Counts
    assert set(counts) <= {"00", "11"} and sum(counts.values()) == 400
    assert seen[0].NumQubits == 2 and seen[0].Shots == 400 and seen[1] is None
    assert be.Status(ctx, job.ID).State == BK.StateCompleted


def test_autotuner_picks_a_measured_shape_and_results_agree(tmp_path, monkeypatch):
    """tune() times the candidates on the device, persists the fastest, and a simulator built with
    plan="tuned" gives the default shape's amplitudes to rounding."""
    from goqu_amd import autotune as AT
    from goqu_amd import statevector

    monkeypatch.setenv("GOQU_TUNE_CACHE", str(tmp_path / "tune.json"))
    c = random_circuit(16, 6, np.random.default_rng(800), three_q=False)
    best, times = AT.tune(c, candidates=[(10, 3), (11, 4), (12, 4)], reps=1)
    assert best in [(10, 3), (11, 4), (12, 4)] and len(times) == 3
    assert min(times.values()) == times[f"{best[0]},{best[1]}"]
    assert AT.plan_for(16) == best
    a = statevector.New(16, plan="tuned")
    b = statevector.New(16)
    a.Evolve(c)
    b.Evolve(c)
    assert np.max(np.abs(a.StateVector() - b.StateVector())) < 1e-12
    a.Close()
    b.Close()
