"""Every benchmark configuration at its stated size, on the GPU, against the oracle where the
oracle can run and through size-independent properties where it cannot.

  C2  30 and 33 qubits, depth 40: exact parity of amplitudes and 64 Z-strings on a full-size
      circuit whose state factorises (two config2 halves on interleaved qubits, so gates span
      low and high qubits of the whole register); the real config2 circuit through norm and
      U then U^dagger = identity.
  C3  EfficientSU2(24, 7), 384 parameters, H = 200 Pauli strings: complex128 energy against the
      oracle, gradients against the oracle's parameter-shift rule (sampled parameters) and the
      GPU parameter-shift rule (all); complex64 within 1e-5 of the fp64 values.
  C4  4,096 MaxCut instances in one batch: one instance per qubit count 10..22 against the
      oracle energy, gradients against central differences of oracle energies (n <= 12) and
      against the generic adjoint.
  C5  14-qubit noisy density matrix, depth 20: trace, Hermiticity, non-negative diagonal,
      adjoint energy equal to the forward energy, adjoint gradient against central differences
      of the GPU forward energy; gradient parity against the oracle at 8 qubits with the same
      layer structure.
"""
import math
import os

import numpy as np
import pytest

from circuits import (config2_circuit, efficient_su2, interleaved_product_circuit, noisy_dm_workload,
                      qaoa_instances, random_pauli_sum, random_z_masks, split_bits)
from goqu_amd import _native as N
from goqu_amd import gradient as GR
from goqu_amd import statevector
from goqu_amd.circuit import Bind, FreeParameters, NewBuilder
from goqu_amd.pauli import NewPauliString, Z
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _oracle_workers():
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    O.set_workers(max(1, min(cores, 16)))


# ---- C2 ------------------------------------------------------------------------------------------
@pytest.mark.parametrize("n", [30, 33])
def test_c2_full_size_product_parity(n):
    """Amplitudes (16 chunks of 2^16 at random offsets) and 64 random Z-strings of the full
    depth-40 factorising circuit against psi_A (x) psi_B from the oracle: <= 1e-12."""
    _say(f"C2 product parity n={n}")
    c, ca, cb = interleaved_product_circuit(n, 40, seed=101 + n)
    na, nb = ca.NumQubits(), cb.NumQubits()
    psi_a, psi_b = O.sv_evolve(ca), O.sv_evolve(cb)
    sim = statevector.New(n)
    try:
        sim.Evolve(c)
        _say("  evolved")
        assert abs(sim.Norm2() - 1) < 1e-12
        rng = np.random.default_rng(n)
        chunk = 1 << 16
        for off in rng.choice(1 << (n - 16), 16, replace=False):
            idx = np.arange(int(off) * chunk, int(off + 1) * chunk, dtype=np.uint64)
            got = sim.StateVectorRange(int(off) * chunk, chunk)
            want = psi_a[split_bits(idx, 0, na)] * psi_b[split_bits(idx, 1, nb)]
            err = float(np.max(np.abs(got - want)))
            assert err < 1e-12, (off, err)
        _say("  amplitudes ok")
        masks = random_z_masks(n, 64, rng)
        terms = [NewPauliString(1.0, {q: Z for q in range(n) if m >> q & 1}, n) for m in masks]
        got = sim.ExpectTerms(terms).real
        for m, g in zip(masks, got):
            qa = [j for j in range(na) if m >> (2 * j) & 1]
            qb = [j for j in range(nb) if m >> (2 * j + 1) & 1]
            want = O.sv_expectation_value(psi_a, qa, False) * O.sv_expectation_value(psi_b, qb, False)
            assert abs(g - want) < 1e-12, (hex(m), g, want)
        _say("  64 Z-strings ok")
    finally:
        sim.Close()


def _inverse(c):
    b = NewBuilder("inv", c.NumQubits())
    for op in reversed(c.Ops()):
        if op.Gate is not None:
            b.Apply(op.Gate.Inverse(), *op.Qubits)
    return b.Build()


@pytest.mark.parametrize("n", [30, 33])
def test_c2_full_size_unitarity(n):
    """The config2 circuit itself (depth 40): the norm stays 1 and the 64 Z-strings are real and
    in [-1, 1]; U then U^dagger returns the state to |0...0> (the first 2^16 amplitudes within
    1e-10)."""
    _say(f"C2 unitarity n={n}")
    c = config2_circuit(n, 40, seed=2)
    sim = statevector.New(n)
    try:
        sim.Evolve(c)
        assert abs(sim.Norm2() - 1) < 1e-11
        rng = np.random.default_rng(7 + n)
        masks = random_z_masks(n, 64, rng)
        terms = [NewPauliString(1.0, {q: Z for q in range(n) if m >> q & 1}, n) for m in masks]
        vals = sim.ExpectTerms(terms)
        assert np.all(np.abs(vals.real) <= 1 + 1e-12) and np.all(np.abs(vals.imag) < 1e-14)
        _say("  forward ok")
        sim.Apply(_inverse(c))
        head = sim.StateVectorRange(0, 1 << 16)
        assert abs(head[0] - 1) < 1e-10, head[0]
        assert np.max(np.abs(head[1:])) < 1e-10
        assert abs(sim.Norm2() - 1) < 1e-11
        _say("  U U^dagger ok")
    finally:
        sim.Close()


# ---- C3 ------------------------------------------------------------------------------------------
def _c3_problem():
    rng = np.random.default_rng(3)
    c, npar = efficient_su2(24, 7)
    names = FreeParameters(c)
    H = random_pauli_sum(24, 200, rng)
    x = rng.uniform(0, 2 * math.pi, npar)
    return c, names, H, x


def _oracle_energy(c, names, x, H):
    psi = O.sv_evolve(Bind(c, dict(zip(names, x))))
    return sum(O.sv_expect(psi, t) for t in H.Terms()).real


def test_c3_adjoint_24q_complex128():
    _say("C3 c128")
    _oracle_workers()
    c, names, H, x = _c3_problem()
    sim = statevector.New(24)
    try:
        e, g = GR.Adjoint(c, H, names, sim=sim).value_and_grad(x)
        assert len(g) == 384
        e_orc = _oracle_energy(c, names, x, H)
        assert abs(e - e_orc) < 1e-12, (e, e_orc)
        _say("  energy ok")
        # the reference's parameter-shift rule on oracle energies, first and last parameter (every
        # parameter enters once as exp(-i theta P / 2), so the rule is exact)
        for i in [0, len(x) - 1]:
            d = np.zeros(len(x))
            d[i] = math.pi / 2
            ps = 0.5 * (_oracle_energy(c, names, x + d, H) - _oracle_energy(c, names, x - d, H))
            assert abs(g[i] - ps) < 1e-10, (i, g[i], ps)
        _say("  oracle parameter shift ok")
        ps_all = GR.ParameterShift(GR.CostFunc(c, H, names, sim=sim))(x)
        assert np.max(np.abs(g - ps_all)) < 1e-10
        _say("  GPU parameter shift (384) ok")
    finally:
        sim.Close()


def test_c3_shape_complex64_vs_oracle_small():
    """The c64 adjoint against oracle energies directly, at 10 qubits (same ansatz and H shape)."""
    rng = np.random.default_rng(31)
    c, npar = efficient_su2(10, 3)
    names = FreeParameters(c)
    H = random_pauli_sum(10, 40, rng)
    x = rng.uniform(0, 2 * math.pi, npar)
    sim = statevector.New(10, N.C64)
    e, g = GR.Adjoint(c, H, names, sim=sim).value_and_grad(x)
    sim.Close()
    assert abs(e - _oracle_energy(c, names, x, H)) < 1e-5
    for i in range(0, npar, 7):
        d = np.zeros(npar)
        d[i] = math.pi / 2
        ps = 0.5 * (_oracle_energy(c, names, x + d, H) - _oracle_energy(c, names, x - d, H))
        assert abs(g[i] - ps) < 1e-5, (i, g[i], ps)


def test_c3_adjoint_24q_complex64():
    """complex64 state, fp32 arithmetic and fp64 accumulators: energy and all 384 gradient
    entries within 1e-5 of the fp64 values (the c128 adjoint, itself pinned to the oracle)."""
    _say("C3 c64")
    c, names, H, x = _c3_problem()
    s64, s128 = statevector.New(24, N.C64), statevector.New(24)
    try:
        e32, g32 = GR.Adjoint(c, H, names, sim=s64).value_and_grad(x)
        e64, g64 = GR.Adjoint(c, H, names, sim=s128).value_and_grad(x)
        assert abs(e32 - e64) < 1e-5, (e32, e64)
        assert np.max(np.abs(g32 - g64)) < 1e-5, np.max(np.abs(g32 - g64))
    finally:
        s64.Close()
        s128.Close()


# ---- C4 ------------------------------------------------------------------------------------------
def _qaoa_oracle_energy(it, x):
    from goqu_amd import qaoa as Q

    H = Q.MaxCutHamiltonian(it.edges, it.num_qubits)
    c = Q.QAOACircuit(H, it.layers)
    psi = O.sv_evolve(Bind(c, dict(zip(FreeParameters(c), x))))
    return sum(O.sv_expect(psi, t) for t in H.Terms()).real


def test_c4_full_batch_4096():
    from goqu_amd import qaoa as Q

    _say("C4 4096 instances")
    _oracle_workers()
    insts = qaoa_instances(4096, seed=5)
    energies, grads = Q.BatchMaxCut(insts)
    assert len(energies) == 4096 and all(np.isfinite(energies))
    _say("  batch done")
    for n in range(10, 23):
        i = next(k for k, it in enumerate(insts) if it.num_qubits == n)
        it = insts[i]
        want = _qaoa_oracle_energy(it, it.params)
        assert abs(energies[i] - want) < 1e-10, (n, energies[i], want)
        if n <= 12:
            h = 1e-3
            for j in range(len(it.params)):
                d = np.zeros(len(it.params))
                d[j] = h
                f = lambda v: _qaoa_oracle_energy(it, v)  # noqa: E731
                p = it.params
                fd = (-f(p + 2 * d) + 8 * f(p + d) - 8 * f(p - d) + f(p - 2 * d)) / (12 * h)
                assert abs(grads[i][j] - fd) < 1e-8, (n, j, grads[i][j], fd)
        elif n in (16, 22):
            H = Q.MaxCutHamiltonian(it.edges, n)
            c = Q.QAOACircuit(H, it.layers)
            e2, g2 = GR.Adjoint(c, H, FreeParameters(c)).value_and_grad(it.params)
            assert abs(e2 - energies[i]) < 1e-10
            assert np.max(np.abs(g2 - grads[i])) < 1e-9
        _say(f"  n={n} ok")


# ---- C5 ------------------------------------------------------------------------------------------
def test_c5_14q_depth20_forward_and_adjoint():
    from goqu_amd import densitymatrix

    _say("C5 14q")
    n = 14
    c, nm, H, names, x = noisy_dm_workload(n, 20, seed=9)
    bound = Bind(c, dict(zip(names, x)))
    dm = densitymatrix.New(n).WithNoise(nm)
    try:
        dm.Evolve(bound)
        e = dm.ExpectPauliSum(H)
        rho = dm.DensityMatrix().reshape(1 << n, 1 << n)
        _say("  forward done")
        tr = np.trace(rho)
        assert abs(tr - 1) < 1e-11, tr
        herm = float(np.max(np.abs(rho - rho.conj().T)))
        assert herm < 1e-13, herm
        assert np.min(np.real(np.diag(rho))) > -1e-14
        zsum = sum(float(np.real(np.sum(np.diag(rho) * (1 - 2 * ((np.arange(1 << n) >> q) & 1))))) for q in range(n))
        assert abs(zsum - e) < 1e-11
        del rho
        _say("  trace / Hermiticity / diagonal ok")
        adj = GR.AdjointDensity(c, H, names, noise=nm, checkpoint_every=16)
        e2, g = adj.value_and_grad(x)
        assert abs(e2 - e) < 1e-11, (e2, e)
        assert len(g) == 3 * n * 20 and np.all(np.isfinite(g))
        _say("  adjoint energy ok")

        def f(v):
            dm.Evolve(Bind(c, dict(zip(names, v))))
            return dm.ExpectPauliSum(H)

        rng = np.random.default_rng(14)
        h = 1e-4
        for i in rng.choice(len(x), 4, replace=False):
            d = np.zeros(len(x))
            d[i] = h
            fd = (f(x + d) - f(x - d)) / (2 * h)
            assert abs(g[i] - fd) < 1e-7, (i, g[i], fd)
        _say("  gradient vs GPU central differences ok")
    finally:
        dm.Close()


def test_c5_structure_8q_gradient_vs_oracle():
    """The C5 layer structure at 8 qubits (depth 20, same noise): energy against the oracle's
    Kraus-loop density matrix, gradient against 4th-order central differences of oracle energies
    on 12 sampled angles."""
    from goqu_amd import densitymatrix as D

    n = 8
    c, nm, H, names, x = noisy_dm_workload(n, 20, seed=19)

    def f(v):
        rho = O.dm_zero(n)
        assert O.dm_apply_ops(rho, n, D.marshal(Bind(c, dict(zip(names, v))), nm)) == 0
        return sum(O.dm_expect(rho, n, t) for t in H.Terms()).real

    e, g = GR.AdjointDensity(c, H, names, noise=nm, checkpoint_every=8).value_and_grad(x)
    assert abs(e - f(x)) < 1e-11
    rng = np.random.default_rng(8)
    h = 1e-3
    for i in rng.choice(len(x), 12, replace=False):
        d = np.zeros(len(x))
        d[i] = h
        fd = (-f(x + 2 * d) + 8 * f(x + d) - 8 * f(x - d) + f(x - 2 * d)) / (12 * h)
        assert abs(g[i] - fd) < 1e-9, (i, g[i], fd)
