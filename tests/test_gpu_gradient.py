"""Adjoint differentiation on the GPU, checked against the reference's parameter-shift rule
(exact when each parameter enters once as exp(-i theta P / 2)), central finite differences,
and the oracle-based energy."""
import math

import numpy as np
import pytest

from circuits import efficient_su2
from goqu_amd import gradient as GR
from goqu_amd import statevector
from goqu_amd.circuit import Bind, FreeParameters, NewBuilder, NewVector, Literal, Mul
from goqu_amd.pauli import NewPauliString, NewPauliSum, X, Y, Z
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def random_hamiltonian(n, terms, rng, wmax=4):
    out = []
    for _ in range(terms):
        w = int(rng.integers(1, min(wmax, n) + 1))
        ops = {int(q): [X, Y, Z][rng.integers(3)] for q in rng.choice(n, w, replace=False)}
        out.append(NewPauliString(rng.normal(), ops, n))
    return NewPauliSum(out)


def oracle_energy(c, names, x, H):
    psi = O.sv_evolve(Bind(c, dict(zip(names, x))))
    return sum(O.sv_expect(psi, t) for t in H.Terms()).real


def test_gradient_closed_form():
    """algorithm/gradient/gradient_test.go:15-87: E = cos(theta), dE/dtheta = -sin(theta)."""
    th = NewVector("t", 1)
    b = NewBuilder("ry", 1)
    b.SymRY(th.At(0), 0)
    c = b.Build()
    H = NewPauliSum([NewPauliString(1, {0: Z}, 1)])
    names = FreeParameters(c)
    for x in [0.3, 1.2, -2.0]:
        e, g = GR.Adjoint(c, H, names).value_and_grad([x])
        assert abs(e - math.cos(x)) < 1e-12
        assert abs(g[0] + math.sin(x)) < 1e-12
        ps = GR.ParameterShift(GR.CostFunc(c, H, names))([x])
        assert abs(ps[0] + math.sin(x)) < 1e-10


@pytest.mark.parametrize("n,reps", [(4, 2), (8, 3), (12, 2)])
def test_adjoint_vs_parameter_shift_efficient_su2(n, reps):
    rng = np.random.default_rng(n)
    c, npar = efficient_su2(n, reps)
    names = FreeParameters(c)
    assert len(names) == npar
    H = random_hamiltonian(n, 20, rng)
    x = rng.uniform(0, 2 * math.pi, npar)
    e, g = GR.Adjoint(c, H, names).value_and_grad(x)
    assert abs(e - oracle_energy(c, names, x, H)) < 1e-11
    sim = statevector.New(n)
    ps = GR.ParameterShift(GR.CostFunc(c, H, names, sim=sim))(x)
    assert np.max(np.abs(g - ps)) < 1e-10


def test_adjoint_vs_finite_difference_reused_scaled_params():
    """QAOA-style reuse: the same parameter in many gates with scale factors (parameter shift
    is inexact here; central differences are the check)."""
    n = 6
    gam, bet = NewVector("g", 2), NewVector("b", 2)
    b = NewBuilder("qaoa", n)
    for q in range(n):
        b.H(q)
    edges = [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (0, 5), (1, 4)]
    for k in range(2):
        for (i, j) in edges:
            b.SymRZZ(Mul(Literal(1.0), gam.At(k)), i, j)
        for q in range(n):
            b.SymRX(Mul(Literal(2.0), bet.At(k)), q)
    c = b.Build()
    names = FreeParameters(c)
    terms = []
    for (i, j) in edges:
        terms.append(NewPauliString(-0.5, {}, n))
        terms.append(NewPauliString(0.5, {i: Z, j: Z}, n))
    H = NewPauliSum(terms)
    x = np.array([0.4, 0.9, 1.3, 0.2])
    e, g = GR.Adjoint(c, H, names).value_and_grad(x)
    fd = GR.FiniteDifference(GR.CostFunc(c, H, names), 1e-5)(x)
    assert np.max(np.abs(g - fd)) < 1e-8
    assert abs(e - oracle_energy(c, names, x, H)) < 1e-11


def test_adjoint_all_parameterised_gates():
    """Every differentiable constructor: adjoint against 4th-order central differences."""
    n = 3
    th = NewVector("p", 16)
    b = NewBuilder("mix", n)
    b.H(0).H(1).H(2)
    b.SymRX(th.At(0), 0).SymRY(th.At(1), 1).SymRZ(th.At(2), 2)
    b.SymPhase(th.At(3), 0).SymU3(th.At(4), th.At(5), th.At(6), 1)
    b.SymCP(th.At(7), 0, 1).SymRXX(th.At(8), 1, 2).SymRYY(th.At(9), 0, 2).SymRZZ(th.At(10), 0, 1)
    b.SymRot(th.At(11), th.At(12), th.At(13), 2).SymFSim(th.At(14), th.At(15), 0, 2)
    c = b.Build()
    names = FreeParameters(c)
    rng = np.random.default_rng(3)
    H = random_hamiltonian(n, 10, rng, 3)
    x = rng.uniform(0, 2 * math.pi, len(names))
    _, g = GR.Adjoint(c, H, names).value_and_grad(x)
    f = lambda v: oracle_energy(c, names, v, H)  # noqa: E731
    h = 1e-3
    for i in range(len(x)):
        d = np.zeros(len(x))
        d[i] = 1
        fd = (-f(x + 2 * h * d) + 8 * f(x + h * d) - 8 * f(x - h * d) + f(x - 2 * h * d)) / (12 * h)
        assert abs(g[i] - fd) < 1e-9, (names[i], g[i], fd)


def _dm_energy(c, names, x, H, nm, n):
    from goqu_amd import densitymatrix as D

    rho = O.dm_zero(n)
    assert O.dm_apply_ops(rho, n, D.marshal(Bind(c, dict(zip(names, x))), nm)) == 0
    return sum(O.dm_expect(rho, n, t) for t in H.Terms()).real


@pytest.mark.parametrize("every", [1, 3])
def test_adjoint_density_matrix_vs_finite_differences(every):
    """Checkpointed density-matrix adjoint (h = vec(H) back-propagated through adjoint channel
    superoperators, rho recomputed from checkpoints): energy against the oracle's Kraus-loop
    density matrix, gradient against 4th-order central differences of oracle energies, under
    depolarizing / amplitude-damping channels, a complex parity-block channel and a reset."""
    from goqu_amd import noise as NZ

    n = 3
    th = NewVector("p", 9)
    b = NewBuilder("dmg", n)
    b.SymRX(th.At(0), 0).SymRY(th.At(1), 1).SymU3(th.At(2), th.At(3), th.At(4), 2)
    b.CNOT(0, 1).SymRZZ(th.At(5), 1, 2).H(0)
    b.Reset(2)
    b.SymRY(th.At(6), 2).SymCP(th.At(7), 0, 2).SymRX(th.At(8), 1)
    c = b.Build()
    names = FreeParameters(c)
    p = 0.2
    k1 = np.sqrt(p) * np.array([[0, 1], [1j, 0]])
    nm = NZ.New()
    nm.AddDefaultError(1, NZ.Depolarizing1Q(0.02))
    nm.AddDefaultError(2, NZ.Depolarizing2Q(0.03))
    nm.AddGateError("H", NZ.Channel("cplx1", 1, [np.sqrt(1 - p) * np.eye(2), k1]))
    nm.AddGateError("RY", NZ.AmplitudeDamping(0.05))
    rng = np.random.default_rng(17)
    H = random_hamiltonian(n, 8, rng, 3)
    x = rng.uniform(0, 2 * math.pi, len(names))
    e, g = GR.AdjointDensity(c, H, names, noise=nm, checkpoint_every=every).value_and_grad(x)
    assert abs(e - _dm_energy(c, names, x, H, nm, n)) < 1e-11
    f = lambda v: _dm_energy(c, names, v, H, nm, n)  # noqa: E731
    h = 1e-3
    for i in range(len(x)):
        d = np.zeros(len(x))
        d[i] = 1
        fd = (-f(x + 2 * h * d) + 8 * f(x + h * d) - 8 * f(x - h * d) + f(x - 2 * h * d)) / (12 * h)
        assert abs(g[i] - fd) < 1e-9, (names[i], g[i], fd)


def test_adjoint_paired_big_stages():
    """Reverse sweeps whose stages hold 3-qubit dense ops run on the paired kernels' big variants:
    a statevector circuit with a fixed random 3-qubit unitary between parameterised layers, and a
    noisy density matrix with a Toffoli (U and conj(U) on 3 row / column qubits), against
    4th-order central differences of oracle energies."""
    from goqu_amd import gate as G

    rng = np.random.default_rng(23)
    a = rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8))
    q, r = np.linalg.qr(a)
    u3q = q * (np.diag(r) / np.abs(np.diag(r)))
    n = 4
    th = NewVector("p", 8)
    b = NewBuilder("big", n)
    for i in range(n):
        b.SymRY(th.At(i), i)
    b.Apply(G.Unitary("u3q", u3q.reshape(-1), 3), 0, 2, 3)
    for i in range(n):
        b.SymRZ(th.At(n + i), i)
    b.CNOT(3, 1)
    c = b.Build()
    names = FreeParameters(c)
    H = random_hamiltonian(n, 6, rng, 3)
    x = rng.uniform(0, 2 * math.pi, len(names))
    _, g = GR.Adjoint(c, H, names).value_and_grad(x)
    f = lambda v: oracle_energy(c, names, v, H)  # noqa: E731
    h = 1e-3
    for i in range(len(x)):
        d = np.zeros(len(x))
        d[i] = 1
        fd = (-f(x + 2 * h * d) + 8 * f(x + h * d) - 8 * f(x - h * d) + f(x - 2 * h * d)) / (12 * h)
        assert abs(g[i] - fd) < 1e-9, (names[i], g[i], fd)

    # The density-matrix oracle (like the reference, sim/densitymatrix/sim.go:161-181) takes no
    # 3-qubit gates, so the density-matrix case runs noiseless and is checked against statevector
    # oracle energies: Tr(H |psi><psi|) = <psi|H|psi>.
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
    e, g = GR.AdjointDensity(c, H, names).value_and_grad(x)
    f = lambda v: oracle_energy(c, names, v, H)  # noqa: E731
    assert abs(e - f(x)) < 1e-11
    for i in range(len(x)):
        d = np.zeros(len(x))
        d[i] = 1
        fd = (-f(x + 2 * h * d) + 8 * f(x + h * d) - 8 * f(x - h * d) + f(x - 2 * h * d)) / (12 * h)
        assert abs(g[i] - fd) < 1e-9, (names[i], g[i], fd)
