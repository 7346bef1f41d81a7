"""The reference's own known-answer tests, restated against the CPU oracle (and the host
marshalling that feeds both the oracle and the device library).  Each test cites the goqu
test it restates; tolerances are the reference's (eps = 1e-10, sim_test.go:13)."""
import math

import numpy as np
import pytest

from goqu_amd import densitymatrix as D
from goqu_amd import gate as G
from goqu_amd import noise as NZ
from goqu_amd.circuit import NewBuilder
from goqu_amd.pauli import NewPauliString, NewPauliSum, Parse, X, Y, Z
from oracle import oracle as O

EPS = 1e-10
s2 = 1 / math.sqrt(2)


def evolve(b):
    return O.sv_evolve(b.Build())


def close(got, want, eps=EPS):
    return np.max(np.abs(np.asarray(got) - np.asarray(want, dtype=complex))) < eps


def test_bell():  # sim_test.go:15-35
    assert close(evolve(NewBuilder("bell", 2).H(0).CNOT(0, 1)), [s2, 0, 0, s2])


def test_ghz3():  # sim_test.go:37-56
    sv = evolve(NewBuilder("ghz", 3).H(0).CNOT(0, 1).CNOT(1, 2))
    want = np.zeros(8, complex)
    want[0] = want[7] = s2
    assert close(sv, want)


def test_x_h_hh():  # sim_test.go:58-106
    assert close(evolve(NewBuilder("x", 1).X(0)), [0, 1])
    assert close(evolve(NewBuilder("h", 1).H(0)), [s2, s2])
    assert close(evolve(NewBuilder("hh", 1).H(0).H(0)), [1, 0])


def test_qft3_from_zero_uniform():  # sim_test.go:107-133
    b = NewBuilder("qft3", 3).H(0).Apply(G.CP(math.pi / 2), 1, 0).Apply(G.CP(math.pi / 4), 2, 0)
    b.H(1).Apply(G.CP(math.pi / 2), 2, 1).H(2)
    assert np.all(np.abs(np.abs(evolve(b)) - 1 / math.sqrt(8)) < EPS)


def test_ccx():  # sim_test.go:169-200
    sv = evolve(NewBuilder("ccx", 3).X(0).X(1).CCX(0, 1, 2))
    assert abs(sv[7] - 1) < EPS


def test_swap():  # sim_test.go:202-221
    sv = evolve(NewBuilder("swap", 2).X(0).SWAP(0, 1))
    assert abs(sv[2] - 1) < EPS


def test_z_expectations():  # sim_test.go:223-277
    sv = evolve(NewBuilder("z", 2).H(0).CNOT(0, 1))
    assert O.sv_expectation_value(sv, [0]) == 0.0
    assert O.sv_expectation_value(sv, [0, 1]) == 1.0
    sv = evolve(NewBuilder("x", 2).X(0))
    assert O.sv_expectation_value(sv, [0]) == -1.0
    assert O.sv_expectation_value(sv, [1]) == 1.0


def test_non_adjacent_cnot():  # sim_test.go:279-300
    sv = evolve(NewBuilder("cx", 3).X(0).CNOT(0, 2))
    assert abs(sv[5] - 1) < EPS


def test_cz_cy():  # sim_test.go:302-347
    sv = evolve(NewBuilder("cz", 2).X(0).X(1).CZ(0, 1))
    assert abs(sv[3] + 1) < EPS
    sv = evolve(NewBuilder("cy", 2).X(0).Apply(G.CY, 0, 1))
    assert abs(sv[3] - 1j) < EPS


def test_cp_crz_diagonal():  # sim_test.go:348-394
    phi = 0.7
    sv = evolve(NewBuilder("cp", 2).X(0).X(1).Apply(G.CP(phi), 0, 1))
    assert abs(sv[3] - np.exp(1j * phi)) < EPS
    sv = evolve(NewBuilder("crz", 2).X(0).X(1).Apply(G.CRZ(phi), 0, 1))
    assert abs(sv[3] - np.exp(1j * phi / 2)) < EPS
    sv = evolve(NewBuilder("crz0", 2).X(1).Apply(G.CRZ(phi), 0, 1))  # control off
    assert abs(sv[2] - 1) < EPS


def test_crx_cry_controlled():  # sim_test.go:395-438
    th = 1.1
    sv = evolve(NewBuilder("crx", 2).X(0).Apply(G.CRX(th), 0, 1))
    assert close(sv, [0, math.cos(th / 2), 0, -1j * math.sin(th / 2)])
    sv = evolve(NewBuilder("cry", 2).X(0).Apply(G.CRY(th), 0, 1))
    assert close(sv, [0, math.cos(th / 2), 0, math.sin(th / 2)])


def test_ms_generic():  # sim_test.go:439-462: MS(0,0)|00> = (|00> - i|11>)/sqrt2
    sv = evolve(NewBuilder("ms", 2).Apply(G.MS(0, 0), 0, 1))
    assert close(sv, [s2, 0, 0, -1j * s2])


def test_ccx_nonadjacent_cswap():  # sim_test.go:464-513
    sv = evolve(NewBuilder("ccx", 4).X(0).X(3).CCX(0, 3, 1))
    assert abs(sv[0b1011] - 1) < EPS
    sv = evolve(NewBuilder("cswap", 3).X(0).X(1).Apply(G.CSWAP, 0, 1, 2))
    assert abs(sv[0b101] - 1) < EPS


def test_mcx_vs_ccx():  # kernel_controlled_test.go:13-80
    for ctrl in ([0, 1, 2], [3, 1, 0], [1, 4, 2, 0]):
        n = max(ctrl) + 2
        t = n - 1
        b = NewBuilder("mcx", n)
        for q in ctrl:
            b.X(q)
        b.MCX(ctrl, t)
        sv = evolve(b)
        idx = sum(1 << q for q in ctrl) | (1 << t)
        assert abs(sv[idx] - 1) < EPS
        b = NewBuilder("mcx-off", n)
        for q in ctrl[:-1]:
            b.X(q)
        b.MCX(ctrl, t)
        sv = evolve(b)
        assert abs(sv[sum(1 << q for q in ctrl[:-1])] - 1) < EPS


def test_reset_semantics():  # sim_test.go:666-746
    assert close(evolve(NewBuilder("xr", 1).X(0).Reset(0)), [1, 0])
    assert close(evolve(NewBuilder("hr", 1).H(0).Reset(0)), [1, 0])
    assert close(evolve(NewBuilder("br", 2).H(0).CNOT(0, 1).Reset(0)), [s2, 0, s2, 0])
    assert close(evolve(NewBuilder("r0", 1).Reset(0)), [1, 0])


def test_parallel_threshold_17q():  # sim_test.go:748-772 (worker split must not change results)
    b = NewBuilder("par17", 17).H(0)
    for q in range(16):
        b.CNOT(q, q + 1)
    O.set_workers(1)
    a = evolve(b)
    O.set_workers(8)
    c = evolve(b)
    O.set_workers(1)
    assert np.array_equal(a, c)
    assert abs(a[0] - s2) < EPS and abs(a[-1] - s2) < EPS


def test_pauli_bell_expectations():  # expect_sv_test.go:35-80, pauli/example_test.go:10-26
    sv = evolve(NewBuilder("bell", 2).H(0).CNOT(0, 1))
    for s, want in [("ZZ", 1), ("XX", 1), ("YY", -1), ("ZI", 0), ("IZ", 0), ("XY", 0)]:
        assert abs(O.sv_expect(sv, Parse(s)) - want) < EPS
    ps = NewPauliString(0.5, {0: Z, 1: Z}, 2)
    assert abs(O.sv_expect(sv, ps) - 0.5) < EPS


def test_pauli_parse_order():  # pauli.go:66-88: character i acts on qubit i
    sv = evolve(NewBuilder("x0", 2).X(0))
    assert abs(O.sv_expect(sv, Parse("ZI")) + 1) < EPS
    assert abs(O.sv_expect(sv, Parse("IZ")) - 1) < EPS


def _dm(c, n, nm=None):
    rho = O.dm_zero(n)
    assert O.dm_apply_ops(rho, n, D.marshal(c, nm)) == 0
    return rho


def test_dm_depolarizing_maximally_mixed():  # densitymatrix/sim_test.go:164-186
    nm = NZ.New()
    nm.AddDefaultError(1, NZ.Depolarizing1Q(0.75))
    rho = _dm(NewBuilder("x", 1).X(0).Build(), 1, nm)
    assert abs(rho[0].real - 0.5) < EPS and abs(rho[3].real - 0.5) < EPS


def test_dm_amplitude_damping():  # sim_test.go:188-211
    nm = NZ.New()
    nm.AddGateError("X", NZ.AmplitudeDamping(0.5))
    rho = _dm(NewBuilder("x", 1).X(0).Build(), 1, nm)
    assert abs(rho[0].real - 0.5) < EPS and abs(rho[3].real - 0.5) < EPS


def test_dm_phase_damping():  # sim_test.go:213-234
    lam = 0.5
    nm = NZ.New()
    nm.AddGateError("H", NZ.PhaseDamping(lam))
    rho = _dm(NewBuilder("h", 1).H(0).Build(), 1, nm)
    assert abs(rho[1] - 0.5 * math.sqrt(1 - lam)) < EPS


def test_dm_readout():  # sim_test.go:236-253
    rho = O.dm_zero(1)
    p = O.dm_probs(rho, 1, [0.1], [0.2])
    assert abs(p[0] - 0.9) < EPS and abs(p[1] - 0.1) < EPS


def test_dm_trace_and_positivity():  # sim_test.go:255-301
    nm = NZ.New()
    nm.AddDefaultError(1, NZ.Depolarizing1Q(0.1))
    nm.AddDefaultError(2, NZ.Depolarizing2Q(0.1))
    c = NewBuilder("t", 3).H(0).CNOT(0, 1).CNOT(1, 2).Build()
    rho = _dm(c, 3, nm).reshape(8, 8)
    assert abs(np.trace(rho) - 1) < EPS
    assert np.min(np.linalg.eigvalsh((rho + rho.conj().T) / 2)) > -EPS


def test_dm_purity_example():  # densitymatrix/example_test.go:34-52
    c = NewBuilder("h", 1).H(0).Build()
    assert round(O.dm_purity(_dm(c, 1), 1), 4) == 1.0
    nm = NZ.New()
    nm.AddDefaultError(1, NZ.Depolarizing1Q(0.5))
    assert round(O.dm_purity(_dm(c, 1, nm), 1), 4) == 0.5556


def test_dm_expect_sum():  # expect_dm_test.go:127-139: ExpectSumDM = 0.5
    rho = _dm(NewBuilder("bell", 2).H(0).CNOT(0, 1).Build(), 2)
    H = NewPauliSum([NewPauliString(0.25, {0: Z, 1: Z}, 2), NewPauliString(0.25, {0: X, 1: X}, 2)])
    total = sum(O.dm_expect(rho, 2, t) for t in H.Terms())
    assert abs(total - 0.5) < EPS


def test_dm_fidelity_pure():  # kernel2q_test.go:11-84 (fidelity 1 against the statevector)
    c = NewBuilder("f", 4).H(0).CNOT(0, 1).RY(0.3, 2).CNOT(2, 3).Apply(G.RXX(0.4), 1, 3).Build()
    psi = O.sv_evolve(c)
    assert abs(O.dm_fidelity(_dm(c, 4), 4, psi) - 1) < EPS


def _stateprep_kat_circuits():  # sim_test.go:775-872
    from goqu_amd import gate as G

    out = []
    out.append((NewBuilder("sp1", 1).Apply(G.StatePrep([s2, s2]), 0).Build(), [s2, s2]))
    out.append((NewBuilder("sp-bell", 2).Apply(G.StatePrep([s2, 0, 0, s2]), 0, 1).Build(), [s2, 0, 0, s2]))
    out.append((NewBuilder("sp-decompose", 2).Apply(G.StatePrep([s2, s2]), 1).Build(), [s2, 0, s2, 0]))
    out.append((NewBuilder("sp-one", 1).Apply(G.StatePrep([0, 1]), 0).Build(), [0, 1]))
    return out


def test_stateprep_kats():
    for c, want in _stateprep_kat_circuits():
        assert close(O.sv_evolve(c), want), c.Name()
    # TestStatePrep_Normalized: |0> + i|1> over sqrt2 stays normalised
    from goqu_amd import gate as G

    sv = O.sv_evolve(NewBuilder("sp-arb", 2).Apply(G.StatePrep([s2, 1j * s2]), 1).Build())
    assert abs(np.vdot(sv, sv).real - 1) < EPS
    # the one-qubit cascade: RY(2 atan2(|a1|, |a0|)) then RZ(phase difference) (stateprep.go:185-211)
    a = np.array([0.6, 0.8 * np.exp(0.9j)])
    sv = O.sv_evolve(NewBuilder("sp-ph", 2).Apply(G.StatePrep(a), 0).Build())
    assert abs(abs(sv[0]) - 0.6) < EPS and abs(abs(sv[1]) - 0.8) < EPS
    assert abs(np.angle(sv[1] / sv[0]) - 0.9) < EPS
