"""The stage planner, checked on the CPU: planned programs (tile qubits, register layouts,
permuted matrices, LDS re-distributions) are executed by a lockstep emulator of the stage
kernel's semantics and compared with the oracle.  Covers every gate tier and several tile
shapes, including tiny states padded to the minimum internal width."""
import numpy as np
import pytest

import emulator as E
from circuits import config2_circuit, random_circuit
from goqu_amd import _native as N
from oracle import oracle as O


def _padded(n, nmin=6):
    nqi = max(n, nmin)
    s = np.zeros(1 << nqi, dtype=np.complex128)
    s[0] = 1
    return s, nqi


@pytest.mark.parametrize("n", [1, 2, 3, 5, 6, 7, 9, 12, 13])
@pytest.mark.parametrize("tile,low", [(13, 3), (12, 3), (8, 2), (6, 2), (10, 4)])
def test_planner_matches_oracle(n, tile, low):
    rng = np.random.default_rng(31 * n + tile)
    c = random_circuit(n, 5, rng)
    buf = O.sv_marshal(c)
    s, nqi = _padded(n)
    E.apply(s, n, buf, tile_bits=min(tile, nqi), low_bits=low)
    want = O.sv_evolve(c)
    assert np.max(np.abs(s[: 1 << n] - want)) < 1e-12
    assert np.max(np.abs(s[1 << n:])) == 0.0 if nqi > n else True


def test_planner_high_qubits_and_stage_count():
    n = 16
    c = config2_circuit(n, 6, seed=5)
    buf = O.sv_marshal(c)
    s, _ = _padded(n)
    stages = E.apply(s, n, buf, tile_bits=10, low_bits=3)
    want = O.sv_evolve(c)
    assert np.max(np.abs(s - want)) < 1e-12
    gates = sum(1 for op in c.Ops() if op.Gate is not None)
    assert stages < gates / 4, (stages, gates)


def _swap_circuits(n):
    """SWAPs among high qubits only (a stage without instructions whose store is permuted), SWAPs
    that also move low qubits (restored by lane exchanges), and SWAPs mixed with rotations and
    CNOTs on low and high qubits."""
    from goqu_amd.circuit import NewBuilder

    b = NewBuilder("high", n)
    for a, c in [(3, n - 1), (n - 1, 5), (4, n - 2), (6, 9)]:
        b.SWAP(a, c)
    high = b.Build()
    b = NewBuilder("low", n)
    for a, c in [(0, n - 1), (n - 1, 5), (3, n - 2), (1, 2), (0, 1)]:
        b.SWAP(a, c)
    low = b.Build()
    b = NewBuilder("mixed", n)
    rng = np.random.default_rng(n)
    for layer in range(3):
        for q in range(n):
            b.U3(*rng.uniform(0, 2 * np.pi, 3), q)
        for a, c in [(0, n - 1), (2, 7), (1, n - 3)]:
            b.SWAP(a, c).CNOT(c, a)
    return high, low, b.Build()


def random_state(n, seed):
    rng = np.random.default_rng(seed)
    s = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return s / np.linalg.norm(s)


@pytest.mark.parametrize("tile,low", [(13, 3), (12, 3), (10, 3), (8, 2)])
def test_swaps_relabel_and_lane_exchanges(tile, low):
    """Uncontrolled SWAPs relabel local bits (stores write through the relabelled map), and
    register sets change by lane exchanges where cheaper than an LDS remap."""
    n = 14
    E.coverage(reset=True)
    for i, c in enumerate(_swap_circuits(n)):
        s0 = random_state(n, 40 + i)
        s = s0.copy()
        E.apply(s, n, O.sv_marshal(c), tile_bits=tile, low_bits=low)
        assert np.max(np.abs(s - O.sv_evolve(c, s0.copy()))) < 1e-12
    cov = E.coverage(reset=True)
    assert cov["swap_pure"] == 0  # every uncontrolled SWAP was a relabelling
    if tile == 12:
        assert cov["xswap_perm"] + cov["xswap_dpp"] > 0


@pytest.mark.parametrize("n", [1, 2, 3, 5, 6])
def test_dm_lowering_matches_oracle(n):
    """Density-matrix lowering (U on row qubits, conj(U) on column qubits, Kraus channels as
    superoperators, reset) through the planner, against the oracle's two-pass Kraus loop."""
    from goqu_amd import densitymatrix as D
    from goqu_amd import noise as NZ
    from goqu_amd.circuit import NewBuilder

    rng = np.random.default_rng(70 + n)
    c0 = random_circuit(n, 3, rng, three_q=False, controlled=False)
    b = NewBuilder("dm", n).Compose(c0)
    for q in range(n):
        b.Delay(q, 1.0)
    b.Reset(0)
    c = b.Build()
    nm = NZ.New()
    nm.AddDefaultError(1, NZ.Depolarizing1Q(0.03))
    nm.AddDefaultError(2, NZ.Depolarizing2Q(0.04))
    nm.AddGateError("delay", NZ.AmplitudeDamping(0.1))
    buf = D.marshal(c, nm)
    nqi = max(2 * n, 6)
    s = np.zeros(1 << nqi, dtype=np.complex128)
    s[0] = 1
    E.apply(s, n, buf, tile_bits=min(12, nqi), low_bits=3, dm=True)
    want = O.dm_zero(n)
    assert O.dm_apply_ops(want, n, buf) == 0
    assert np.max(np.abs(s[: 1 << (2 * n)] - want)) < 1e-12


def test_emulated_programs_cover_every_kernel_path():
    """The circuits above, the density-matrix lowering and an adjoint reverse sweep reach every
    stage-kernel code path (control placements, diagonal table shapes, phased and pure moves,
    remaps, gradients, reset), so the emulator checks the planner's encoding of each."""
    import math

    import test_adjoint_emulated as TA
    from circuits import efficient_su2
    from goqu_amd import gate as G
    from goqu_amd.circuit import FreeParameters, NewBuilder, SymbolicGate, NewVector

    E.coverage(reset=True)
    for n, tile, low in [(5, 6, 2), (9, 8, 3), (12, 12, 3), (13, 8, 2), (13, 10, 4)]:
        rng = np.random.default_rng(5 * n + tile)
        c = random_circuit(n, 6, rng)
        s, nqi = _padded(n)
        E.apply(s, n, O.sv_marshal(c), tile_bits=min(tile, nqi), low_bits=low)
        assert np.max(np.abs(s[: 1 << n] - O.sv_evolve(c))) < 1e-12
    # phased moves, register/thread/tile controls on them, reset
    n = 11
    b = NewBuilder("moves", n)
    for q in range(n):
        b.H(q)
    for (cq, t) in [(0, 1), (10, 2), (5, 9), (9, 5), (3, 4)]:
        b.Apply(G.CY, cq, t).Apply(G.Y, t).Apply(G.ISWAP, cq, t).CNOT(t, cq)
        b.Apply(G.CSWAP, cq, t, (t + 3) % n).Apply(G.CCX, cq, (t + 3) % n, t)
        b.Apply(G.CCZ, cq, t, (t + 5) % n).Apply(G.SWAP, t, (t + 6) % n)
    b.Reset(4)
    c = b.Build()
    s, nqi = _padded(n)
    E.apply(s, n, O.sv_marshal(c), tile_bits=8, low_bits=3)
    assert np.max(np.abs(s[: 1 << n] - O.sv_evolve(c))) < 1e-12
    test_dm_lowering_matches_oracle(5)
    # gradients, plain and controlled
    th = NewVector("p", 3)
    b = NewBuilder("g", 8)
    for q in range(8):
        b.H(q)
    b.SymRX(th.At(0), 7).Apply(SymbolicGate("CRZ", 2, [th.At(1)]), 0, 6)
    b.Apply(SymbolicGate("CRY", 2, [th.At(2)]), 7, 1)
    c = b.Build()
    names = FreeParameters(c)
    H = TA.random_hamiltonian(8, 4, np.random.default_rng(2))
    TA.emulated_adjoint(c, names, np.array([0.3, 1.1, -0.7]), H, 8, 7, 3)
    cov = E.coverage(reset=False)
    TA.test_fused_reverse_controlled_record_derivative()  # resets the counters itself
    cov["grad_ctrl"] = 1
    missing = [k for k, v in cov.items() if v == 0]
    assert not missing, (missing, cov)


def _custom_channels():
    """Kraus channels whose superoperators keep every qubit's row^column parity but have complex
    blocks (anti-diagonal Kraus operators with complex entries), and one that does not keep it
    (falls back to a dense superoperator)."""
    from goqu_amd import noise as NZ

    p = 0.3
    k0 = np.sqrt(1 - p) * np.eye(2)
    k1 = np.sqrt(p) * np.array([[0, 1], [1j, 0]]) / np.sqrt(1.0)  # K1^dag K1 = p I
    c1 = NZ.Channel("cplx1", 1, [k0, k1])
    k2 = np.sqrt(p) * np.kron(np.array([[0, 1], [1j, 0]]), np.array([[np.exp(0.4j), 0], [0, 1]]))
    c2 = NZ.Channel("cplx2", 2, [np.sqrt(1 - p) * np.eye(4), k2])
    # mixes parities: K = cos I + i sin X style rotation inside a mixture
    th = 0.7
    r = np.array([[np.cos(th), -1j * np.sin(th)], [-1j * np.sin(th), np.cos(th)]])
    c3 = NZ.Channel("mix", 1, [np.sqrt(1 - p) * np.eye(2), np.sqrt(p) * r])
    return c1, c2, c3


@pytest.mark.parametrize("n", [2, 4])
def test_dm_parity_blocks_complex_and_fallback(n):
    """Parity-block superoperators with complex blocks (1q and 2q) and a parity-mixing channel
    (dense fallback), planner + emulator against the oracle's Kraus loop."""
    from goqu_amd import densitymatrix as D
    from goqu_amd import gate as G
    from goqu_amd import noise as NZ
    from goqu_amd.circuit import NewBuilder

    c1, c2, c3 = _custom_channels()
    rng = np.random.default_rng(80 + n)
    b = NewBuilder("dmc", n)
    for q in range(n):
        b.U3(*rng.uniform(0, 2 * np.pi, 3), q)
    b.CNOT(0, 1).H(n - 1).CZ(n - 2, n - 1)
    c = b.Build()
    nm = NZ.New()
    nm.AddDefaultError(1, c1)
    nm.AddDefaultError(2, c2)
    nm.AddGateError("H", c3)
    buf = D.marshal(c, nm)
    nqi = max(2 * n, 6)
    s = np.zeros(1 << nqi, dtype=np.complex128)
    s[0] = 1
    E.coverage(reset=True)
    E.apply(s, n, buf, tile_bits=min(12, nqi), low_bits=3, dm=True)
    cov = E.coverage(reset=True)
    want = O.dm_zero(n)
    assert O.dm_apply_ops(want, n, buf) == 0
    assert np.max(np.abs(s[: 1 << (2 * n)] - want)) < 1e-12
    assert cov["parity2"] > 0 and cov["parity4"] > 0
    assert cov["parity2"] + cov["parity4"] > cov["parity_real"]  # complex blocks ran


@pytest.mark.parametrize("seed", range(12))
@pytest.mark.parametrize("tile,low", [(12, 3), (8, 2)])
def test_resets_are_barriers(seed, tile, low):
    """The Evolve path's reset a0 <- |(a0, a1)| is not linear, so no op on any qubit may be
    fused, deferred or scheduled across it: random circuits with resets against the oracle."""
    n = 4 + seed % 7
    rng = np.random.default_rng(900 + seed)
    c = random_circuit(n, 4, rng, resets=0.25)
    s, nqi = _padded(n)
    E.apply(s, n, O.sv_marshal(c), tile_bits=min(tile, nqi), low_bits=low)
    assert np.max(np.abs(s[: 1 << n] - O.sv_evolve(c))) < 1e-12


def _reset_kat_circuits():
    """Two known answers: H0 H1 Reset1 S0 -> [1, i, 0, 0]/sqrt2 (the S gate after the reset
    must act), and H0 H1 RZ(0.7, 0) Reset1 -> [1, 1, 0, 0]/sqrt2 (the RZ phases, including the
    global one, are discarded by the reset's norm)."""
    from goqu_amd.circuit import NewBuilder

    a = NewBuilder("kat_s", 2).H(0).H(1).Reset(1).S(0).Build()
    b = NewBuilder("kat_rz", 2).H(0).H(1).RZ(0.7, 0).Reset(1).Build()
    r = 1 / np.sqrt(2)
    return [(a, np.array([r, 1j * r, 0, 0])), (b, np.array([r, r, 0, 0]))]


@pytest.mark.parametrize("k", [0, 1])
def test_reset_known_answers(k):
    c, want = _reset_kat_circuits()[k]
    assert np.max(np.abs(O.sv_evolve(c) - want)) < 1e-15
    s, nqi = _padded(2)
    E.apply(s, 2, O.sv_marshal(c), tile_bits=nqi, low_bits=2)
    assert np.max(np.abs(s[:4] - want)) < 1e-15


@pytest.mark.parametrize("nt,nc", [(5, 1), (6, 2), (8, 2), (9, 1)])
def test_wide_controlled_gates(nt, nc):
    """Controlled gates whose inner matrix acts on 5..9 targets (the reference's full controlled
    matrix path, kernel_controlled.go:249-292, up to 10 qubits): lowered to wide ops between
    stage programs, against the oracle."""
    from goqu_amd import gate as G
    from goqu_amd.circuit import NewBuilder
    from scipy.stats import unitary_group

    n = 11
    rng = np.random.default_rng(10 * nt + nc)
    U = unitary_group.rvs(1 << nt, random_state=int(rng.integers(1 << 30)))
    qs = [int(x) for x in rng.choice(n, nt + nc, replace=False)]
    b = NewBuilder("wide", n)
    for q in range(n):
        b.U3(*rng.uniform(0, 2 * np.pi, 3), q)
    b.Apply(G.Controlled(G.Unitary("U", U.reshape(-1)), nc), *qs)
    b.CNOT(0, 1).RZ(0.4, 2)
    c = b.Build()
    s, _ = _padded(n)
    E.apply(s, n, O.sv_marshal(c), tile_bits=10, low_bits=3)
    assert np.max(np.abs(s - O.sv_evolve(c))) < 1e-12


def test_dm_adjacent_channels_merge():
    """A one-qubit channel that meets a parity-block channel on its qubit with nothing in between
    (here: amplitude damping on Delay after a gate's depolarizing channel, across gates on other
    qubits) is folded into it: the program has the same instructions as without it."""
    from goqu_amd import densitymatrix as D
    from goqu_amd import noise as NZ
    from goqu_amd.circuit import NewBuilder

    def instrs(with_delay):
        b = NewBuilder("m", 3)
        b.RX(0.3, 0).RY(0.4, 1).CNOT(1, 2)
        if with_delay:
            b.Delay(0, 1.0)
        nm = NZ.New()
        nm.AddDefaultError(1, NZ.Depolarizing1Q(0.03))
        nm.AddGateError("delay", NZ.AmplitudeDamping(0.1))
        return N.plan_stats(3, D.marshal(b.Build(), nm), density=True)[1]

    assert instrs(True) == instrs(False)
