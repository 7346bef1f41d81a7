"""Pauli strings and sums — host-side mirror of goqu's sim/pauli (pauli.go:17-232).

Encoding per qubit: I=00, X=10, Z=01, Y=11 (x-bit, z-bit), ops[q] acts on qubit q.  Parse
maps character i to qubit i (pauli.go:66-88), the reverse of bitstring display order.
"""
from __future__ import annotations

from typing import Dict, List, Sequence

I, X, Z, Y = 0b00, 0b10, 0b01, 0b11
_CH = {"I": I, "X": X, "Y": Y, "Z": Z}
_NAME = {I: "I", X: "X", Y: "Y", Z: "Z"}


class PauliString:
    def __init__(self, coeff: complex, ops: Sequence[int], num_qubits: int):
        self.coeff = complex(coeff)
        self.ops = list(ops)
        self.num_qubits = num_qubits

    def Coeff(self) -> complex:
        return self.coeff

    def Op(self, q: int) -> int:
        return self.ops[q] if 0 <= q < self.num_qubits else I

    def NumQubits(self) -> int:
        return self.num_qubits

    def Scale(self, c: complex) -> "PauliString":
        return PauliString(self.coeff * c, self.ops, self.num_qubits)

    def IsIdentity(self) -> bool:
        return all(p == I for p in self.ops)

    def xmask(self) -> int:
        return sum(1 << i for i, p in enumerate(self.ops) if p & 0b10)

    def zmask(self) -> int:
        return sum(1 << i for i, p in enumerate(self.ops) if p & 0b01)

    def numY(self) -> int:
        return sum(1 for p in self.ops if p == Y)

    def __str__(self):
        s = "".join(_NAME[p] for p in self.ops)
        return s if self.coeff == 1 else f"{self.coeff}*{s}"


def NewPauliString(coeff: complex, ops: Dict[int, int], num_qubits: int) -> PauliString:
    """pauli.go:44-58"""
    if num_qubits < 1:
        raise ValueError("pauli: numQubits must be >= 1")
    paulis = [I] * num_qubits
    for q, p in (ops or {}).items():
        if not 0 <= q < num_qubits:
            raise ValueError(f"pauli: qubit {q} out of range [0, {num_qubits})")
        paulis[q] = p
    return PauliString(coeff, paulis, num_qubits)


def Parse(s: str) -> PauliString:
    if not s:
        raise ValueError("pauli: empty string")
    ops = []
    for i, ch in enumerate(s):
        if ch not in _CH:
            raise ValueError(f"pauli: invalid character {ch!r} at position {i}")
        ops.append(_CH[ch])
    return PauliString(1, ops, len(s))


def ZOn(qubits: Sequence[int], num_qubits: int) -> PauliString:
    return NewPauliString(1, {q: Z for q in qubits}, num_qubits)


class PauliSum:
    def __init__(self, terms: List[PauliString]):
        self.terms = list(terms)
        self.num_qubits = terms[0].num_qubits

    def Terms(self) -> List[PauliString]:
        return list(self.terms)

    def NumQubits(self) -> int:
        return self.num_qubits


def NewPauliSum(terms: List[PauliString]) -> PauliSum:
    if not terms:
        raise ValueError("pauli: empty term list")
    nq = terms[0].num_qubits
    for i, t in enumerate(terms[1:]):
        if t.num_qubits != nq:
            raise ValueError(f"pauli: term {i + 1} has {t.num_qubits} qubits, expected {nq}")
    return PauliSum(terms)


def masks(terms: Sequence[PauliString]):
    """(xmask, zmask, numY, coeff) arrays for the C ABI."""
    import numpy as np

    x = np.array([t.xmask() for t in terms], dtype=np.uint64)
    z = np.array([t.zmask() for t in terms], dtype=np.uint64)
    ny = np.array([t.numY() for t in terms], dtype=np.int32)
    c = np.array([[t.coeff.real, t.coeff.imag] for t in terms], dtype=np.float64).reshape(-1)
    return x, z, ny, c
