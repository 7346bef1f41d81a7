"""Batched QAOA MaxCut on the GPU: energies against the oracle statevector of the same
circuit (maxcut.go / qaoa.go construction) and gradients against central finite differences
of the oracle energy and against the generic adjoint."""
import math

import numpy as np
import pytest

from goqu_amd import gradient as GR
from goqu_amd import qaoa as Q
from goqu_amd.circuit import Bind, FreeParameters
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def random_graph(n, p, rng):
    while True:
        edges = [(i, j) for i in range(n) for j in range(i + 1, n) if rng.random() < p]
        if edges:
            return edges


def oracle_energy(n, edges, layers, x):
    H = Q.MaxCutHamiltonian(edges, n)
    c = Q.QAOACircuit(H, layers)
    psi = O.sv_evolve(Bind(c, dict(zip(FreeParameters(c), x))))
    return sum(O.sv_expect(psi, t) for t in H.Terms()).real


def test_batch_small_and_large_instances():
    rng = np.random.default_rng(21)
    insts = []
    for n in [4, 7, 10, 12, 13, 14, 15]:
        p = int(rng.integers(1, 5))
        insts.append(Q.MaxCutInstance(n, random_graph(n, 0.3, rng), p, rng.uniform(0, math.pi, 2 * p)))
    energies, grads = Q.BatchMaxCut(insts)
    for it, e, g in zip(insts, energies, grads):
        want = oracle_energy(it.num_qubits, it.edges, it.layers, it.params)
        assert abs(e - want) < 1e-11, (it.num_qubits, e, want)
        if it.num_qubits <= 13:
            # 4th-order central differences: truncation error ~h^4 f^(5)/30, ~1e-12 here
            h = 1e-3
            f = lambda v: oracle_energy(it.num_qubits, it.edges, it.layers, v)  # noqa: E731
            for i in range(len(it.params)):
                d = np.zeros(len(it.params))
                d[i] = h
                x = it.params
                fd = (-f(x + 2 * d) + 8 * f(x + d) - 8 * f(x - d) + f(x - 2 * d)) / (12 * h)
                assert abs(g[i] - fd) < 1e-8, (it.num_qubits, i, g[i], fd)


def test_batch_gradient_matches_generic_adjoint():
    rng = np.random.default_rng(5)
    n, p = 14, 3
    edges = random_graph(n, 0.3, rng)
    x = rng.uniform(0, math.pi, 2 * p)
    e, g = Q.BatchMaxCut([Q.MaxCutInstance(n, edges, p, x)])
    H = Q.MaxCutHamiltonian(edges, n)
    c = Q.QAOACircuit(H, p)
    e2, g2 = GR.Adjoint(c, H, FreeParameters(c)).value_and_grad(x)
    assert abs(e[0] - e2) < 1e-11
    assert np.max(np.abs(g[0] - g2)) < 1e-10


def test_batch_22_qubits():
    rng = np.random.default_rng(22)
    n, p = 22, 2
    edges = random_graph(n, 0.3, rng)
    x = rng.uniform(0, math.pi, 2 * p)
    e, g = Q.BatchMaxCut([Q.MaxCutInstance(n, edges, p, x)])
    H = Q.MaxCutHamiltonian(edges, n)
    c = Q.QAOACircuit(H, p)
    e2, g2 = GR.Adjoint(c, H, FreeParameters(c)).value_and_grad(x)
    assert abs(e[0] - e2) < 1e-10
    assert np.max(np.abs(g[0] - g2)) < 1e-9


def test_batch_passes_every_tile_shape_and_depth():
    """Instances whose B tiles run the mixer on one, two or three register groups (13-16, 17-20,
    21-22 qubits) at every depth 1-4 in one batch (p + 1 forward and p + 1 reverse passes each,
    the steps of different depths interleaved), each against the generic adjoint."""
    rng = np.random.default_rng(31)
    shapes = [(13, 4), (16, 1), (17, 3), (18, 2), (20, 4), (21, 1), (22, 3), (19, 2)]
    insts = [Q.MaxCutInstance(n, random_graph(n, 0.3, rng), p, rng.uniform(0, math.pi, 2 * p)) for n, p in shapes]
    energies, grads = Q.BatchMaxCut(insts)
    for it, e, g in zip(insts, energies, grads):
        H = Q.MaxCutHamiltonian(it.edges, it.num_qubits)
        c = Q.QAOACircuit(H, it.layers)
        e2, g2 = GR.Adjoint(c, H, FreeParameters(c)).value_and_grad(it.params)
        assert abs(e - e2) < 1e-10, (it.num_qubits, it.layers, e, e2)
        assert np.max(np.abs(g - g2)) < 1e-9, (it.num_qubits, it.layers, g, g2)
    # a second call of the same batch reuses the retained device pool: identical results
    energies2, grads2 = Q.BatchMaxCut(insts)
    assert np.array_equal(energies, energies2)
    assert all(np.array_equal(a, b) for a, b in zip(grads, grads2))
