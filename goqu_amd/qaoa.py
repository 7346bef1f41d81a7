"""QAOA MaxCut and the hardware-efficient ansatz — host-side mirrors of the circuit
construction in goqu's algorithm/qaoa (maxcut.go:13-36, qaoa.go:61-196) and
algorithm/ansatz (efficientsu2.go:19-50, realamplitudes.go:51-70), plus the batched device
evaluation of many MaxCut instances (energy and adjoint gradient per instance)."""
from __future__ import annotations

import ctypes as C
from typing import List, Sequence, Tuple

import numpy as np

from . import _native as N
from .circuit import Builder, Circuit, Literal, Mul, NewBuilder, NewVector
from .pauli import NewPauliString, NewPauliSum, PauliSum, Z


def MaxCutHamiltonian(edges: Sequence[Tuple[int, int]], num_qubits: int) -> PauliSum:
    """maxcut.go:13-36: H = sum_edges (-0.5 I + 0.5 Z_i Z_j)."""
    if not edges:
        raise ValueError("qaoa: no edges provided")
    if num_qubits < 2:
        raise ValueError("qaoa: numQubits must be >= 2")
    terms = []
    for a, b in edges:
        if not (0 <= a < num_qubits and 0 <= b < num_qubits):
            raise ValueError(f"qaoa: edge ({a},{b}) out of range")
        terms.append(NewPauliString(-0.5, {}, num_qubits))
        terms.append(NewPauliString(0.5, {a: Z, b: Z}, num_qubits))
    return NewPauliSum(terms)


def _apply_cost_layer(b: Builder, h: PauliSum, gamma, nq: int) -> None:
    """qaoa.go:164-196: SymRZ / SymRZZ(2 * coeff * gamma) per all-Z term; identities skipped."""
    for term in h.Terms():
        coeff = term.Coeff().real
        if coeff == 0:
            continue
        zq, all_z = [], True
        for q in range(nq):
            op = term.Op(q)
            if op == Z:
                zq.append(q)
            elif op != 0:
                all_z = False
                break
        if not all_z or not zq:
            continue
        angle = Mul(Literal(2 * coeff), gamma)
        if len(zq) == 1:
            b.SymRZ(angle, zq[0])
        elif len(zq) == 2:
            b.SymRZZ(angle, zq[0], zq[1])


def QAOACircuit(cost: PauliSum, p: int) -> Circuit:
    """The circuit of qaoa.Run (qaoa.go:66-95): H on every qubit, then per layer the cost
    layer and RX(2 beta_k) on every qubit.  Free parameters in order gamma_0, beta_0, ..."""
    nq = cost.NumQubits()
    gammas, betas = NewVector("γ", p), NewVector("β", p)
    b = NewBuilder("QAOA", nq)
    for q in range(nq):
        b.H(q)
    for k in range(p):
        _apply_cost_layer(b, cost, gammas.At(k), nq)
        for q in range(nq):
            b.SymRX(Mul(Literal(2), betas.At(k)), q)
    return b.Build()


def EfficientSU2(num_qubits: int, reps: int) -> Tuple[Circuit, int]:
    """ansatz.NewEfficientSU2(n, reps, Circular).Circuit(): reps+1 layers of SymRY/SymRZ on
    every qubit with vector parameters theta[0..2n(reps+1)), a circular CNOT chain between
    layers."""
    nparams = 2 * num_qubits * (reps + 1)
    th = NewVector("θ", nparams)
    b = NewBuilder("EfficientSU2", num_qubits)
    idx = 0
    for rep in range(reps + 1):
        for q in range(num_qubits):
            b.SymRY(th.At(idx), q)
            b.SymRZ(th.At(idx + 1), q)
            idx += 2
        if rep < reps:
            for i in range(num_qubits - 1):
                b.CNOT(i, i + 1)
            if num_qubits > 1:
                b.CNOT(num_qubits - 1, 0)
    return b.Build(), nparams


class MaxCutInstance:
    def __init__(self, num_qubits: int, edges: Sequence[Tuple[int, int]], layers: int, params: Sequence[float]):
        self.num_qubits = num_qubits
        # immutable, so the flat (a0, b0, a1, b1, ...) int32 copy used by batch marshalling stays valid
        self.edges = tuple((int(a), int(b)) for a, b in edges)
        self._edge_arr = np.array(self.edges, dtype=np.int32).reshape(-1)
        self.layers = layers
        self.params = np.asarray(params, dtype=np.float64)
        if len(self.params) != 2 * layers:
            raise ValueError("params must be (gamma_0, beta_0, gamma_1, beta_1, ...)")


class MarshalledBatch:
    """The flat arrays goqu_qaoa_batch takes (qubits, layers, edge offsets, edges, parameter
    offsets, parameters) for a list of instances."""

    def __init__(self, instances: Sequence[MaxCutInstance]):
        ni = len(instances)
        self.ni = ni
        self.nq = np.fromiter((it.num_qubits for it in instances), dtype=np.int32, count=ni)
        self.lay = np.fromiter((it.layers for it in instances), dtype=np.int32, count=ni)
        arrs = [it._edge_arr for it in instances]
        self.eoff = np.zeros(ni + 1, dtype=np.int64)
        np.cumsum(np.fromiter((len(a) // 2 for a in arrs), dtype=np.int64, count=ni), out=self.eoff[1:])
        self.poff = np.zeros(ni + 1, dtype=np.int64)
        np.cumsum(2 * self.lay.astype(np.int64), out=self.poff[1:])
        self.edges = np.concatenate(arrs) if ni and self.eoff[-1] else np.zeros(2, dtype=np.int32)
        self.params = np.concatenate([it.params for it in instances]) if ni else np.zeros(1)


def RunMarshalled(m: MarshalledBatch) -> Tuple[np.ndarray, np.ndarray]:
    """One goqu_qaoa_batch call on arrays already marshalled: the energies, and every instance's
    gradient concatenated in instance order (m.poff delimits them)."""
    energies = np.zeros(max(m.ni, 1))
    grads = np.zeros(max(int(m.poff[-1]), 1))
    N.check(N.lib().goqu_qaoa_batch(m.ni, N.i32ptr(m.nq), N.i32ptr(m.lay), N.i64ptr(m.eoff), N.i32ptr(m.edges),
                                    N.i64ptr(m.poff), N.dptr(m.params), N.dptr(energies), N.dptr(grads)))
    return energies[:m.ni], grads


def BatchMaxCut(instances: Sequence[MaxCutInstance]) -> Tuple[np.ndarray, List[np.ndarray]]:
    """Energy <H> and its gradient (gamma_0, beta_0, ...) for every instance, all instances in
    one device launch graph (goqu_qaoa_batch)."""
    m = MarshalledBatch(instances)
    energies, grads = RunMarshalled(m)
    po = m.poff.tolist()
    return energies, [grads[a:b] for a, b in zip(po, po[1:])]
