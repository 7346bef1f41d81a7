"""The statevector adjoint's fused reverse sweep, checked on the CPU: the program that
build_sv_reverse + the planner produce (EOP_GRAD ops interleaved with inverse ops, run on psi
and lambda together) is executed by the stage-kernel emulator, and the gradients are compared
with exact parameter shift (gradient.go:49-66; exact for exp(-i theta P/2) gates) and with
4th-order central differences of oracle energies.  The reverse sweep must also return psi to
|0...0>."""
import math

import numpy as np
import pytest

import emulator as E
from circuits import efficient_su2
from goqu_amd import gradient as GR
from goqu_amd.circuit import Bind, FreeParameters, Literal, Mul, NewBuilder, NewVector, SymbolicGate
from goqu_amd.pauli import NewPauliString, NewPauliSum, X, Y, Z
from oracle import oracle as O


def random_hamiltonian(n, terms, rng, wmax=4):
    out = []
    for _ in range(terms):
        w = int(rng.integers(1, min(wmax, n) + 1))
        ops = {int(q): [X, Y, Z][rng.integers(3)] for q in rng.choice(n, w, replace=False)}
        out.append(NewPauliString(rng.normal(), ops, n))
    return NewPauliSum(out)


def apply_h(H, psi):
    """H psi with P = i^ny X^x Z^z: (P psi)[i ^ x] = i^ny (-1)^{|i & z|} psi[i]."""
    idx = np.arange(len(psi), dtype=np.int64)
    out = np.zeros_like(psi)
    for t in H.Terms():
        x, z, ny = t.xmask(), t.zmask(), t.numY()
        sign = 1 - 2 * (np.vectorize(lambda v: bin(v).count("1") & 1)(idx & z))
        out[idx ^ x] += t.Coeff() * (1j ** ny) * sign * psi
    return out


def energy(c, names, x, H):
    psi = O.sv_evolve(Bind(c, dict(zip(names, x))))
    return float(np.real(np.vdot(psi, apply_h(H, psi))))


def emulated_adjoint(c, names, x, H, n, tile=12, low=3):
    prog = GR.AdjointProgram(c, names)
    buf, derivs, nd, keep = prog.build(x)
    psi = O.sv_evolve(Bind(c, dict(zip(names, x))))
    lam = apply_h(H, psi)
    e = float(np.real(np.vdot(psi, lam)))
    nqi = max(n, 6)
    P = np.zeros(1 << nqi, dtype=np.complex128)
    L = np.zeros(1 << nqi, dtype=np.complex128)
    P[: 1 << n] = psi
    L[: 1 << n] = lam
    vals, stages = E.sv_reverse(P, L, n, buf, derivs, nd, tile_bits=min(tile, nqi), low_bits=low)
    g = np.zeros(len(names))
    for i in range(nd):
        g[derivs[i].param] += derivs[i].scale * 2.0 * vals[i]
    zero = np.zeros(1 << nqi, dtype=np.complex128)
    zero[0] = 1
    assert np.max(np.abs(P - zero)) < 1e-12, "reverse sweep did not uncompute psi"
    return e, g, stages


def fd4(f, x, h=1e-3):
    out = np.zeros(len(x))
    for i in range(len(x)):
        d = np.zeros(len(x))
        d[i] = 1
        out[i] = (-f(x + 2 * h * d) + 8 * f(x + h * d) - 8 * f(x - h * d) + f(x - 2 * h * d)) / (12 * h)
    return out


@pytest.mark.parametrize("n,reps,tile,low", [(3, 2, 6, 2), (5, 3, 12, 3), (7, 2, 6, 2), (8, 2, 7, 3)])
def test_fused_reverse_matches_parameter_shift(n, reps, tile, low):
    rng = np.random.default_rng(10 + n)
    c, npar = efficient_su2(n, reps)
    names = FreeParameters(c)
    H = random_hamiltonian(n, 12, rng)
    x = rng.uniform(0, 2 * math.pi, npar)
    e, g, _ = emulated_adjoint(c, names, x, H, n, tile, low)
    assert abs(e - energy(c, names, x, H)) < 1e-12
    ps = np.zeros(npar)
    for i in range(npar):
        xp, xm = x.copy(), x.copy()
        xp[i] += 0.5 * math.pi
        xm[i] -= 0.5 * math.pi
        ps[i] = (energy(c, names, xp, H) - energy(c, names, xm, H)) / 2
    assert np.max(np.abs(g - ps)) < 1e-11


def test_fused_reverse_more_slots_than_one_stage_holds():
    """96 derivatives on 6 qubits: more than one stage's 64 gradient slots."""
    n, reps = 6, 7
    rng = np.random.default_rng(5)
    c, npar = efficient_su2(n, reps)
    names = FreeParameters(c)
    H = random_hamiltonian(n, 8, rng)
    x = rng.uniform(0, 2 * math.pi, npar)
    _, g, stages = emulated_adjoint(c, names, x, H, n, 6, 2)
    assert stages >= 2
    f = lambda v: energy(c, names, v, H)  # noqa: E731
    assert np.max(np.abs(g - fd4(f, x))) < 1e-8


def test_fused_reverse_all_parameterised_gates_and_controls():
    """Every differentiable constructor plus controlled rotations (CRX/CRY/CRZ, a rotation
    under two controls) and a parameter reused with scale factors."""
    n = 4
    th = NewVector("p", 19)
    b = NewBuilder("mix", n)
    for q in range(n):
        b.H(q)
    b.SymRX(th.At(0), 0).SymRY(th.At(1), 1).SymRZ(th.At(2), 2)
    b.SymPhase(th.At(3), 0).SymU3(th.At(4), th.At(5), th.At(6), 1)
    b.SymCP(th.At(7), 0, 1).SymRXX(th.At(8), 1, 2).SymRYY(th.At(9), 0, 3).SymRZZ(th.At(10), 0, 1)
    b.SymRot(th.At(11), th.At(12), th.At(13), 2).SymFSim(th.At(14), th.At(15), 3, 2)
    b.Apply(SymbolicGate("CRX", 2, [th.At(16)]), 3, 0)
    b.Apply(SymbolicGate("CRY", 2, [Mul(Literal(0.5), th.At(17))]), 1, 3)
    b.Apply(SymbolicGate("CRZ", 2, [th.At(18)]), 2, 1)
    b.SymRZZ(Mul(Literal(-1.5), th.At(0)), 2, 3)
    b.CNOT(0, 3).CZ(1, 2).SWAP(0, 2)
    c = b.Build()
    names = FreeParameters(c)
    rng = np.random.default_rng(4)
    H = random_hamiltonian(n, 10, rng, 4)
    x = rng.uniform(0, 2 * math.pi, len(names))
    for tile, low in [(6, 2), (12, 3)]:
        _, g, _ = emulated_adjoint(c, names, x, H, n, tile, low)
        f = lambda v: energy(c, names, v, H)  # noqa: E731
        assert np.max(np.abs(g - fd4(f, x))) < 1e-8


def test_fused_reverse_wide_state_high_qubits():
    """13 qubits with a 2^8 tile: derivatives on qubits outside the low tile bits, so the
    planner must bring them into registers for the EOP_GRAD ops."""
    n = 13
    rng = np.random.default_rng(13)
    c, npar = efficient_su2(n, 1)
    names = FreeParameters(c)
    H = random_hamiltonian(n, 6, rng)
    x = rng.uniform(0, 2 * math.pi, npar)
    _, g, stages = emulated_adjoint(c, names, x, H, n, 8, 3)
    assert stages >= 2
    ps = np.zeros(npar)
    for i in range(npar):
        xp, xm = x.copy(), x.copy()
        xp[i] += 0.5 * math.pi
        xm[i] -= 0.5 * math.pi
        ps[i] = (energy(c, names, xp, H) - energy(c, names, xm, H)) / 2
    assert np.max(np.abs(g - ps)) < 1e-11


def test_fused_reverse_controlled_record_derivative():
    """A CONTROLLED record (two controls on RX(theta), as Controlled(RX, 2) marshals) with its
    derivative given on the target matrix: the EOP_GRAD op keeps the controls, checked against
    4th-order differences of the oracle's evolution of the same records."""
    from goqu_amd import _native as N

    n = 7
    rng = np.random.default_rng(8)
    H = random_hamiltonian(n, 6, rng)

    def rx(t):
        c, s = math.cos(t / 2), math.sin(t / 2)
        return np.array([[c, -1j * s], [-1j * s, c]])

    def drx(t):
        c, s = math.cos(t / 2), math.sin(t / 2)
        return 0.5 * np.array([[-s, -1j * c], [-1j * c, -s]])

    pre = [(q, rng.uniform(0, 2 * math.pi, 3)) for q in range(n)]

    def records(t):
        buf = N.OpBuffer()
        for q, (a, b_, c_) in pre:
            u3 = np.array([[math.cos(a / 2), -np.exp(1j * c_) * math.sin(a / 2)],
                           [np.exp(1j * b_) * math.sin(a / 2), np.exp(1j * (b_ + c_)) * math.cos(a / 2)]])
            buf.add(N.OP_GATE, [q], u3)
        # controls (0, 6) on target 3, and (2,) on target 5 with the same parameter scaled by -2
        k1 = buf.add(N.OP_CONTROLLED, [0, 6, 3], rx(t), num_controls=2)
        k2 = buf.add(N.OP_CONTROLLED, [2, 5], rx(-2 * t), num_controls=1)
        return buf, k1, k2

    def energy_of(t):
        buf, _, _ = records(t)
        psi = np.zeros(1 << n, dtype=np.complex128)
        psi[0] = 1
        assert O.sv_apply_ops(psi, n, buf) == 0
        return float(np.real(np.vdot(psi, apply_h(H, psi))))

    t = 0.83
    buf, k1, k2 = records(t)
    d1, d2 = drx(t), drx(-2 * t)
    derivs = (N.GoquDeriv * 2)()
    keep = []
    for i, (k, sc, dm) in enumerate([(k1, 1.0, d1), (k2, -2.0, d2)]):
        flat = np.ascontiguousarray(dm, dtype=np.complex128).reshape(-1).view(np.float64)
        keep.append(flat)
        derivs[i].op = k
        derivs[i].param = 0
        derivs[i].scale = sc
        derivs[i].dmatrix = N.dptr(flat)
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1
    assert O.sv_apply_ops(psi, n, buf) == 0
    lam = apply_h(H, psi)
    E.coverage(reset=True)
    for tile, low in [(6, 2), (7, 3)]:
        P, L = psi.copy(), lam.copy()
        vals, _ = E.sv_reverse(P, L, n, buf, derivs, 2, tile_bits=tile, low_bits=low)
        g = sum(derivs[i].scale * 2.0 * vals[i] for i in range(2))
        h = 1e-3
        fd = (-energy_of(t + 2 * h) + 8 * energy_of(t + h) - 8 * energy_of(t - h) + energy_of(t - 2 * h)) / (12 * h)
        assert abs(g - fd) < 1e-9, (tile, g, fd)
    assert E.coverage(reset=True)["grad_ctrl"] > 0


def _dm_noise(p=0.2):
    from goqu_amd import noise as NZ

    k1 = np.sqrt(p) * np.array([[0, 1], [1j, 0]])
    nm = NZ.New()
    nm.AddDefaultError(1, NZ.Depolarizing1Q(0.02))
    nm.AddDefaultError(2, NZ.Depolarizing2Q(0.03))
    nm.AddGateError("H", NZ.Channel("cplx1", 1, [np.sqrt(1 - p) * np.eye(2), k1]))
    nm.AddGateError("RY", NZ.AmplitudeDamping(0.05))
    nm.AddGateError("delay", NZ.PhaseDamping(0.04))
    return nm


def _dm_energy(c, names, x, H, nm, n):
    from goqu_amd import densitymatrix as D

    rho = O.dm_zero(n)
    assert O.dm_apply_ops(rho, n, D.marshal(Bind(c, dict(zip(names, x))), nm)) == 0
    return sum(O.dm_expect(rho, n, t) for t in H.Terms()).real


def _h_matrix(H, n):
    """vec(H): H[r][c] at index r 2^n + c (row bits high, the density matrix's layout)."""
    dim = 1 << n
    M = np.zeros((dim, dim), dtype=np.complex128)
    for col in range(dim):
        e = np.zeros(dim, dtype=np.complex128)
        e[col] = 1
        M[:, col] = apply_h(H, e)
    return M.reshape(-1)


@pytest.mark.parametrize("n,tile,low", [(2, 6, 2), (3, 6, 2), (4, 8, 3)])
def test_dm_fused_reverse_vs_finite_differences(n, tile, low):
    """The density-matrix adjoint's fused reverse sweep (build_dm_reverse: EOP_GRAD ops on the row
    qubits, gates inverted on rho and h together, channels as S^-1 on rho and S^dag on h) run by
    the emulator on (rho, vec(H)): gradients against 4th-order central differences of oracle
    energies (the oracle's Kraus loop), under depolarizing, amplitude-damping, phase-damping and a
    complex parity-block channel; rho is uncomputed back to |0><0|."""
    th = NewVector("p", 6 * n)
    b = NewBuilder("dmf", n)
    k = 0
    for layer in range(2):
        for q in range(n):
            b.SymU3(th.At(k), th.At(k + 1), th.At(k + 2), q)
            k += 3
        for q in range(layer % 2, n - 1, 2):
            b.CNOT(q, q + 1)
        b.H(0).SymRY(th.At(k % (6 * n)), n - 1).Delay(n - 1, 100.0)
        if n > 2:
            b.SymRZZ(th.At((k + 1) % (6 * n)), 0, n - 1)
    c = b.Build()
    names = FreeParameters(c)
    nm = _dm_noise()
    rng = np.random.default_rng(40 + n)
    H = random_hamiltonian(n, 6, rng, 3)
    x = rng.uniform(0, 2 * math.pi, len(names))
    prog = GR.AdjointProgram(c, names, density=True, noise=nm)
    buf, derivs, nd, keep = prog.build(x)
    rho = O.dm_zero(n)
    assert O.dm_apply_ops(rho, n, buf) == 0
    hv = _h_matrix(H, n)
    e = float(np.real(np.vdot(hv, rho)))
    assert abs(e - _dm_energy(c, names, x, H, nm, n)) < 1e-12
    nqi = max(2 * n, 6)
    R = np.zeros(1 << nqi, dtype=np.complex128)
    Hv = np.zeros(1 << nqi, dtype=np.complex128)
    R[: len(rho)] = rho
    Hv[: len(hv)] = hv
    vals, _ = E.dm_reverse(R, Hv, n, buf.array(), 0, len(buf), derivs, nd, tile_bits=tile, low_bits=low)
    g = np.zeros(len(names))
    for i in range(nd):
        g[derivs[i].param] += derivs[i].scale * 2.0 * vals[i]
    fd = fd4(lambda v: _dm_energy(c, names, v, H, nm, n), x)
    assert np.max(np.abs(g - fd)) < 1e-9, np.max(np.abs(g - fd))
    zero = np.zeros(1 << nqi, dtype=np.complex128)
    zero[0] = 1
    assert np.max(np.abs(R - zero)) < 1e-11, "reverse sweep did not uncompute rho"


def test_dm_fused_reverse_declines_resets_and_singular_channels():
    """A reset (not invertible) and a fully depolarizing channel (singular superoperator) leave the
    fused sweep unavailable: the driver then takes the checkpointed method."""
    from goqu_amd import noise as NZ

    n = 2
    th = NewVector("p", 2)
    for variant in range(2):
        b = NewBuilder("dmr", n)
        b.SymRY(th.At(0), 0).CNOT(0, 1).SymRX(th.At(1), 1)
        nm = NZ.New()
        if variant == 0:
            b.Reset(1)
        else:
            nm.AddGateError("CNOT", NZ.Depolarizing2Q(15.0 / 16.0))
        c = b.Build()
        names = FreeParameters(c)
        prog = GR.AdjointProgram(c, names, density=True, noise=nm)
        buf, derivs, nd, keep = prog.build(np.array([0.3, 0.7]))
        R = np.zeros(1 << 6, dtype=np.complex128)
        Hv = np.zeros(1 << 6, dtype=np.complex128)
        with pytest.raises(RuntimeError, match="status 200"):
            E.dm_reverse(R, Hv, n, buf.array(), 0, len(buf), derivs, nd, tile_bits=6, low_bits=2)


def test_dm_fused_reverse_three_qubit_dense():
    """A fixed 3-qubit unitary (dense ops on three row and three column register slots: stages
    for the paired kernels' big variants) in the density-matrix reverse sweep: rho = |psi><psi|
    from the statevector oracle is uncomputed to |0><0| and the gradient matches statevector
    oracle finite differences (the density-matrix oracle, like the reference, takes no 3-qubit
    gates)."""
    from goqu_amd import gate as G

    rng = np.random.default_rng(23)
    a = rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8))
    q, r = np.linalg.qr(a)
    u3q = q * (np.diag(r) / np.abs(np.diag(r)))
    n = 3
    th = NewVector("q", 6)
    b = NewBuilder("dmbig", n)
    for i in range(n):
        b.SymRY(th.At(i), i)
    b.Apply(G.Unitary("u3q", u3q.reshape(-1), 3), 2, 0, 1)
    for i in range(n):
        b.SymRX(th.At(n + i), i)
    c = b.Build()
    names = FreeParameters(c)
    H = random_hamiltonian(n, 6, rng, 3)
    x = rng.uniform(0, 2 * math.pi, len(names))
    prog = GR.AdjointProgram(c, names, density=True)
    buf, derivs, nd, keep = prog.build(x)
    psi = O.sv_evolve(Bind(c, dict(zip(names, x))))
    R = np.outer(psi, psi.conj()).reshape(-1)
    Hv = _h_matrix(H, n)
    vals, _ = E.dm_reverse(R, Hv, n, buf.array(), 0, len(buf), derivs, nd, tile_bits=6, low_bits=2)
    g = np.zeros(len(names))
    for i in range(nd):
        g[derivs[i].param] += derivs[i].scale * 2.0 * vals[i]
    fd = fd4(lambda v: energy(c, names, v, H), x)
    assert np.max(np.abs(g - fd)) < 1e-9
    zero = np.zeros(len(R), dtype=np.complex128)
    zero[0] = 1
    assert np.max(np.abs(R - zero)) < 1e-12
