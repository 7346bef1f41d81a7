"""Dynamic circuits on the device (goqu_sv_run_dynamic / goqu_dm_run_dynamic) against the oracle's
shot loop (executeOps / executeControlFlow of sim/statevector/dynamic.go:14-230 and
sim/densitymatrix/dynamic.go:13-255): every shot's clbits identical for the same uniform slabs,
and the last shot's final state within 1e-12 (complex128)."""
import math

import numpy as np
import pytest

from circuits import random_dynamic_circuit
from goqu_amd import densitymatrix as D
from goqu_amd import gate as G
from goqu_amd import noise as NZ
from goqu_amd import statevector
from goqu_amd import _native as N
from goqu_amd.circuit import NewBuilder
from test_oracle_dynamic import OracleDynamic, teleport_circuit

pytestmark = pytest.mark.gpu


def _slabs(c, shots, seed, prog):
    return np.random.default_rng(seed).random((shots, max(prog.draws, 1)))


def _check_sv(c, n, shots, seed, prec=N.C128):
    sim = statevector.New(n, prec)
    prog = sim.DynamicProgram(c)
    U = _slabs(c, shots, seed, prog)
    print(f"sv dynamic n={n} shots={shots} seed={seed} instrs={len(prog.instrs)} draws={prog.draws}", flush=True)
    got = sim.RunDynamicShots(c, shots, uniforms=U)
    want, psi = OracleDynamic(n).run_shots(c, U)
    last = sim.StateVector()
    sim.Close()
    return got, want, last, psi


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("n", [2, 5, 9])
def test_random_dynamic_circuits_match_oracle(n, seed):
    rng = np.random.default_rng(7000 + 10 * n + seed)
    c = random_dynamic_circuit(n, 3, 5, rng)
    got, want, last, psi = _check_sv(c, n, 64, seed, N.C128)
    assert np.array_equal(got, want)
    assert np.max(np.abs(last - psi)) < 1e-12


def test_reference_dynamic_and_control_flow_kats():
    """dynamic_test.go:14-162 and controlflow_test.go:184-300 on the device path."""
    rng = np.random.default_rng(1)
    c = NewBuilder("cond", 2).WithClbits(1).X(0).Measure(0, 0).If(0, 1, G.X, 1).Build()
    assert statevector.New(2).RunDynamic(c, 500, rng=rng) == {"1": 500}
    c = NewBuilder("cond_no", 2).WithClbits(1).Measure(0, 0).If(0, 1, G.X, 1).Build()
    assert statevector.New(2).RunDynamic(c, 500, rng=rng) == {"0": 500}
    counts = statevector.New(3).RunDynamic(teleport_circuit(), 2000, rng=rng)
    assert all(len(k) == 3 and k[0] == "1" for k in counts) and len(counts) == 4
    c = NewBuilder("while-sim", 1).WithClbits(1).X(0).Measure(0, 0).While(
        0, 1, lambda b: b.X(0).Measure(0, 0)).Build()
    assert statevector.New(1).RunDynamic(c, 100, rng=rng) == {"0": 100}
    c = NewBuilder("for-sim", 1).WithClbits(1).For(0, 0, 3, 1, lambda b: b.X(0)).MeasureAll().Build()
    assert statevector.New(1).RunDynamic(c, 100, rng=rng) == {"1": 100}
    c = NewBuilder("switch-sim", 1).WithClbits(2).X(0).Measure(0, 0).Switch(
        [0, 1], {0: lambda b: b.X(0), 1: lambda b: None}, None).Measure(0, 0).Build()
    assert statevector.New(1).RunDynamic(c, 100, rng=rng) == {"01": 100}


def test_repeat_until_success_loop_statistics():
    """while (c == 1) { H; measure }: the iteration count is geometric; per-shot parity with the
    oracle, and the final clbit is always 0."""
    b = NewBuilder("rus", 1).WithClbits(2).H(0).Measure(0, 0)
    b.While(0, 1, lambda x: x.H(0).Measure(0, 0))
    c = b.Build()
    got, want, _, _ = _check_sv(c, 1, 2000, 3)
    assert np.array_equal(got, want)
    assert not got[:, 0].any()


def test_complex64_matches_oracle_outcomes():
    """complex64 state, FP64 probabilities: outcomes equal the FP64 oracle's except where a
    uniform falls within complex64 rounding of P0 (none for these seeds)."""
    for seed in range(3):
        rng = np.random.default_rng(7500 + seed)
        c = random_dynamic_circuit(6, 3, 5, rng)
        got, want, last, psi = _check_sv(c, 6, 64, seed, N.C64)
        assert np.array_equal(got, want)
        assert np.max(np.abs(last - psi)) < 1e-5


def test_batches_of_shots_and_a_large_register(monkeypatch):
    """Shots in several device batches (GOQU_DYN_MAX_BATCH caps a batch) and a 16-qubit register
    (1 MiB per shot, several 2^12-amplitude tiles per shot)."""
    rng = np.random.default_rng(7600)
    c = random_dynamic_circuit(5, 3, 4, rng)
    monkeypatch.setenv("GOQU_DYN_MAX_BATCH", "7")
    got, want, last, psi = _check_sv(c, 5, 50, 11)
    assert np.array_equal(got, want)
    assert np.max(np.abs(last - psi)) < 1e-12
    monkeypatch.delenv("GOQU_DYN_MAX_BATCH")
    c = random_dynamic_circuit(16, 4, 3, np.random.default_rng(7601))
    got, want, last, psi = _check_sv(c, 16, 24, 12)
    assert np.array_equal(got, want)
    assert np.max(np.abs(last - psi)) < 1e-12


def test_reset_and_measure_edge_cases():
    """Reset of |1> and of |+> (measure then X), measurement of a certain outcome (the
    degenerate branch leaves the state untouched), a circuit without clbits."""
    b = NewBuilder("edges", 3).WithClbits(3)
    b.X(0).Reset(0).Measure(0, 0)        # reset of |1>: 0
    b.H(1).Reset(1).Measure(1, 1)        # reset of |+>: 0 whatever the draw
    b.X(2).Measure(2, 2).Measure(2, 2)   # certain outcome 1 twice
    got, want, last, psi = _check_sv(b.Build(), 3, 200, 5)
    assert np.array_equal(got, want)
    assert (got == np.array([0, 0, 1])).all()
    assert np.max(np.abs(last - psi)) < 1e-12
    c = NewBuilder("noclbits", 2).H(0).Reset(0).CNOT(0, 1).Build()
    sim = statevector.New(2)
    assert sim.RunDynamic(c, 10, rng=np.random.default_rng(1)) == {"": 10}
    sim.Close()


def test_wide_controlled_gate_inside_a_branch():
    """A 5-target controlled gate (MFMA wide op) under a feed-forward condition."""
    rng = np.random.default_rng(7700)
    inner = np.linalg.qr(rng.normal(size=(32, 32)) + 1j * rng.normal(size=(32, 32)))[0]
    g = G.Controlled(G.Unitary("U5", inner.reshape(-1)), 1)
    b = NewBuilder("wide", 7).WithClbits(2)
    for q in range(7):
        b.H(q)
    b.Measure(0, 0).IfElseBlock(0, 1, lambda x: x.Apply(g, 1, 2, 3, 4, 5, 6), lambda x: x.X(6))
    b.Measure(6, 1)
    got, want, last, psi = _check_sv(b.Build(), 7, 64, 21)
    assert np.array_equal(got, want)
    assert np.max(np.abs(last - psi)) < 1e-12


def test_invalid_programs_are_rejected():
    sim = statevector.New(2)
    ops = N.OpBuffer()
    ops.add(N.OP_GATE, [0], G.H.Matrix())
    bits = np.zeros(1, dtype=np.int32)
    u = np.zeros((4, 4))
    out = np.zeros((4, 1), dtype=np.uint8)
    import ctypes as C

    def call(instrs, nclbits=1, draws=4):
        arr = (N.GoquDynInstr * max(1, len(instrs)))()
        for i, t in enumerate(instrs):
            arr[i].kind, arr[i].a, arr[i].b, arr[i].value = t
        return N.lib().goqu_sv_run_dynamic(sim._h, ops.array(), len(ops), arr, len(instrs), N.i32ptr(bits), 1,
                                           nclbits, 4, N.dptr(u), draws, out.ctypes.data_as(C.POINTER(C.c_uint8)))

    assert call([(N.DYN_BRANCH, 0, 0, 0), (N.DYN_CASE, 0, 1, 1)]) == N.ERR_INVALID_ARG  # not closed
    assert call([(N.DYN_GATES, 0, 2, 0)]) == N.ERR_INVALID_ARG                          # op range
    assert call([(N.DYN_MEASURE, 2, 0, 0)]) == N.ERR_INVALID_ARG                        # qubit
    assert call([(N.DYN_MEASURE, 0, 1, 0)]) == N.ERR_INVALID_ARG                        # clbit
    assert call([(N.DYN_CASE, 0, 1, 0), (N.DYN_END, 0, 0, 0)]) == N.ERR_INVALID_ARG     # outside a group
    assert call([(N.DYN_BRANCH, 0, 0, 0), (N.DYN_GATES, 0, 1, 0), (N.DYN_END, 0, 0, 0)]) == N.ERR_INVALID_ARG
    # a shot needing more uniforms than its slab
    assert call([(N.DYN_MEASURE, 0, 0, 0)] * 3, draws=2) == N.ERR_INVALID_ARG
    assert "draws_per_shot" in N.lib().goqu_last_error().decode()
    assert call([(N.DYN_GATES, 0, 1, 0), (N.DYN_MEASURE, 0, 0, 0)]) == N.OK
    sim.Close()


# ---- density matrix ------------------------------------------------------------------------------------
def _noise():
    nm = NZ.New()
    nm.AddDefaultError(1, NZ.Depolarizing1Q(0.02))
    nm.AddDefaultError(2, NZ.Depolarizing2Q(0.05))
    nm.AddGateError("delay", NZ.AmplitudeDamping(0.1))
    nm.AddGateError("reset", NZ.BitFlip(0.03))
    return nm


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("noisy", [False, True])
def test_dm_dynamic_matches_oracle(noisy, seed):
    n = 3 + seed % 2
    rng = np.random.default_rng(7800 + seed)
    c = random_dynamic_circuit(n, 3, 4, rng, density=True)
    nm = _noise() if noisy else None
    dm = D.New(n)
    if nm is not None:
        dm.WithNoise(nm)
    prog = dm.DynamicProgram(c)
    U = np.random.default_rng(seed).random((48, max(prog.draws, 1)))
    print(f"dm dynamic n={n} noisy={noisy} seed={seed} instrs={len(prog.instrs)}", flush=True)
    got = dm.RunDynamicShots(c, 48, uniforms=U)
    want, rho = OracleDynamic(n, density=True, noise=nm).run_shots(c, U)
    assert np.array_equal(got, want)
    assert np.max(np.abs(dm.DensityMatrix() - rho)) < 1e-12
    dm.Close()


def test_dm_dynamic_reference_cases():
    """densitymatrix RunDynamic: feed-forward, reset, Bell-pair collapse (dynamic.go)."""
    b = NewBuilder("ff", 2).WithClbits(2)
    b.X(0).Measure(0, 0).If(0, 1, G.X, 1).Measure(1, 1)
    assert D.New(2).RunDynamic(b.Build(), 50, rng=np.random.default_rng(1)) == {"11": 50}
    b = NewBuilder("rs", 1).WithClbits(1)
    b.H(0).Reset(0).Measure(0, 0)
    assert D.New(1).RunDynamic(b.Build(), 50, rng=np.random.default_rng(2)) == {"0": 50}
    b = NewBuilder("bell", 2).WithClbits(2)
    b.H(0).CNOT(0, 1).Measure(0, 0).H(1).H(1).Measure(1, 1)
    counts = D.New(2).RunDynamic(b.Build(), 400, rng=np.random.default_rng(3))
    assert set(counts) <= {"00", "11"} and 150 <= counts.get("00", 0) <= 250
