"""Gate library — the host-side mirror of goqu's circuit/gate package.

Matrices follow circuit/gate exactly: flat row-major (2^k x 2^k) with qubits[0] as the most
significant bit of the gate-local index (circuit/gate/fixed.go:169-171,316-317).  Fixed gates
are module-level singletons (CNOT, CZ, ... as in fixed.go:60-396) so identity comparisons
select the same kernels as goqu's pointer-equality dispatch tier (kernel2q.go:27-76).
"""
from __future__ import annotations

import cmath
import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

# gate identity tiers shared with the C ABI (include/goqu_hip.h)
GENERIC, G_CNOT, G_CZ, G_SWAP, G_CY, G_ISWAP, G_SYCAMORE, G_CCX, G_CSWAP, G_CCZ = range(10)
CTRL_DIAG, DIAG2, CTRL_U = 10, 11, 12


def _m(rows: Sequence[Sequence[complex]]) -> np.ndarray:
    return np.asarray(rows, dtype=np.complex128).reshape(-1)


class Gate:
    """A quantum gate (circuit/gate/gate.go:5-26)."""

    def __init__(self, name: str, n: int, matrix: Optional[np.ndarray], params: Sequence[float] = (),
                 tier: int = GENERIC):
        self._name = name
        self._n = n
        self._matrix = None if matrix is None else np.asarray(matrix, dtype=np.complex128).reshape(-1)
        self._params = tuple(params)
        self.tier = tier

    def Name(self) -> str:
        return self._name

    def Qubits(self) -> int:
        return self._n

    def Matrix(self) -> Optional[np.ndarray]:
        return self._matrix

    def Params(self) -> tuple:
        return self._params

    def Inverse(self) -> "Gate":
        """fixed.go:20-45 / parameterized.go:19-37: conjugate transpose, negated params."""
        if self._matrix is None:
            return self
        if self._name in _SELF_INVERSE:
            return self
        if self._name in _INV_PAIRS:
            return _INV_PAIRS[self._name]()
        d = 1 << self._n
        inv = self._matrix.reshape(d, d).conj().T.copy()
        return Gate(self._name + "†", self._n, inv, tuple(-p for p in self._params), self.tier)

    def Decompose(self, qubits):
        return None

    def __repr__(self) -> str:
        return f"Gate({self._name})"


class ControlledGate(Gate):
    """gate.Controlled (circuit/gate/controlled.go:16-110): qubits = controls ++ targets."""

    def __init__(self, inner: Gate, n_controls: int):
        super().__init__(f"C{n_controls}-{inner.Name()}", n_controls + inner.Qubits(), None, inner.Params())
        self._inner = inner
        self._nc = n_controls

    def Inner(self) -> Gate:
        return self._inner

    def NumControls(self) -> int:
        return self._nc

    def Matrix(self) -> Optional[np.ndarray]:
        """Full controlled matrix; panics (raises) above 10 qubits (controlled.go:84-110)."""
        total = self.Qubits()
        if total > 10:
            raise ValueError(f"gate: Controlled matrix too large ({total} qubits, max 10)")
        inner = self._inner.Matrix()
        if inner is None:
            return None
        dim, idim = 1 << total, 1 << self._inner.Qubits()
        m = np.eye(dim, dtype=np.complex128)
        off = dim - idim
        m[off:, off:] = inner.reshape(idim, idim)
        return m.reshape(-1)

    def Inverse(self) -> Gate:
        return Controlled(self._inner.Inverse(), self._nc)


class _Pseudo(Gate):
    pass


class StatePrepGate(Gate):
    """gate.StatePrep (circuit/gate/stateprep.go:10-73): prepares the given amplitudes."""

    def __init__(self, amplitudes: Sequence[complex]):
        amps = np.asarray(amplitudes, dtype=np.complex128).reshape(-1)
        n = int(round(math.log2(len(amps))))
        if len(amps) != 1 << n:
            raise ValueError("gate: StatePrep needs 2^n amplitudes")
        nrm = np.linalg.norm(amps)
        if abs(nrm - 1.0) > 1e-6:
            raise ValueError("gate: StatePrep amplitudes must be normalised")
        super().__init__("state_prep", n, None)
        self._amps = amps

    def Amplitudes(self) -> np.ndarray:
        return self._amps

    def Decompose(self, qubits: Sequence[int]) -> List[Tuple["Gate", List[int]]]:
        """The reference's RY + CNOT cascade (circuit/gate/stateprep.go:68-106), the path
        statevector.Apply takes when the qubits are not exactly [0..n-1] (sim.go:121-135)."""
        if len(qubits) != self.Qubits():
            return []
        return mottonen_decompose(self._amps, list(qubits))


def _decompose_1q(amps: np.ndarray, q: int) -> List[Tuple[Gate, List[int]]]:
    """stateprep.go:185-211: RY(2 atan2(|a1|, |a0|)) and an RZ for the relative phase."""
    ops: List[Tuple[Gate, List[int]]] = []
    theta = 2 * math.atan2(abs(amps[1]), abs(amps[0]))
    if abs(theta) > 1e-10:
        ops.append((RY(theta), [q]))
    if abs(amps[0]) > 1e-10 and abs(amps[1]) > 1e-10:
        d = cmath.phase(amps[1]) - cmath.phase(amps[0])
        if abs(d) > 1e-10:
            ops.append((RZ(d), [q]))
    elif abs(amps[1]) > 1e-10:
        ph = cmath.phase(amps[1])
        if abs(ph) > 1e-10:
            ops.append((RZ(ph), [q]))
    return ops


def _uniformly_controlled_ry(angles: Sequence[float], controls: List[int], target: int):
    """stateprep.go:213-259: recursive split into (even, odd) half-sum / half-difference angles
    around two CNOTs from the last control."""
    n = len(angles)
    if n == 0:
        return []
    if n == 1:
        return [(RY(angles[0]), [target])] if abs(angles[0]) > 1e-10 else []
    half = n // 2
    even = [(angles[2 * k] + angles[2 * k + 1]) / 2 for k in range(half)]
    odd = [(angles[2 * k] - angles[2 * k + 1]) / 2 for k in range(half)]
    ops = []
    for part in (even, odd):
        if len(controls) > 1:
            ops += _uniformly_controlled_ry(part, controls[:-1], target)
        elif abs(part[0]) > 1e-10:
            ops.append((RY(part[0]), [target]))
        ops.append((CNOT, [controls[-1], target]))
    return ops


def mottonen_decompose(amplitudes: Sequence[complex], qubits: List[int]) -> List[Tuple[Gate, List[int]]]:
    """stateprep.go:106-181: for level = 0..n-1 (target qubits[n-1-level], controls the qubits
    before it), RY angles 2 atan2(|b|, |a|) of the paired halves, the pairs merged into the
    first half scaled to the pair norm; a final RZ for the phase of the last amplitude."""
    n = len(qubits)
    if n == 0:
        return []
    amps = np.asarray(amplitudes, dtype=np.complex128)
    if n == 1:
        return _decompose_1q(amps, qubits[0])
    ops: List[Tuple[Gate, List[int]]] = []
    for level in range(n):
        target = qubits[n - 1 - level]
        hs = 1 << level
        npairs = len(amps) // (2 * hs)
        angles = []
        new = np.zeros(npairs * hs, dtype=np.complex128)
        for p in range(npairs):
            g0 = amps[p * 2 * hs: p * 2 * hs + hs]
            g1 = amps[p * 2 * hs + hs: p * 2 * hs + 2 * hs]
            n0 = math.sqrt(float(np.sum(g0.real * g0.real + g0.imag * g0.imag)))
            n1 = math.sqrt(float(np.sum(g1.real * g1.real + g1.imag * g1.imag)))
            angles.append(2 * math.atan2(n1, n0))
            tot = math.sqrt(n0 * n0 + n1 * n1)
            if tot > 1e-15:
                if n0 > 1e-15:
                    new[p * hs: (p + 1) * hs] = g0 * (tot / n0)
                else:
                    new[p * hs] = tot
        controls = list(qubits[: n - 1 - level])
        if npairs == 1:
            if abs(angles[0]) > 1e-10:
                ops.append((RY(angles[0]), [target]))
        else:
            ops += _uniformly_controlled_ry(angles, controls, target)
        amps = new
    if len(amps) == 1:
        ph = cmath.phase(amps[0])
        if abs(ph) > 1e-10:
            ops.append((RZ(ph), [qubits[0]]))
    return ops


s2 = 1.0 / math.sqrt(2.0)
I = Gate("I", 1, _m([[1, 0], [0, 1]]))
H = Gate("H", 1, _m([[s2, s2], [s2, -s2]]))
X = Gate("X", 1, _m([[0, 1], [1, 0]]))
Y = Gate("Y", 1, _m([[0, -1j], [1j, 0]]))
Z = Gate("Z", 1, _m([[1, 0], [0, -1]]))
S = Gate("S", 1, _m([[1, 0], [0, 1j]]))
Sdg = Gate("S†", 1, _m([[1, 0], [0, -1j]]))
T = Gate("T", 1, _m([[1, 0], [0, complex(s2, s2)]]))
Tdg = Gate("T†", 1, _m([[1, 0], [0, complex(s2, -s2)]]))
SX = Gate("SX", 1, _m([[0.5 + 0.5j, 0.5 - 0.5j], [0.5 - 0.5j, 0.5 + 0.5j]]))

CNOT = Gate("CNOT", 2, _m([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]]), tier=G_CNOT)
CZ = Gate("CZ", 2, _m([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, -1]]), tier=G_CZ)
SWAP = Gate("SWAP", 2, _m([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]]), tier=G_SWAP)
CY = Gate("CY", 2, _m([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, -1j], [0, 0, 1j, 0]]), tier=G_CY)
ISWAP = Gate("iSWAP", 2, _m([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]]), tier=G_ISWAP)
ECR = Gate("ECR", 2, _m([[0, 0, s2, 1j * s2], [0, 0, 1j * s2, s2], [s2, -1j * s2, 0, 0], [-1j * s2, s2, 0, 0]]))
DCX = Gate("DCX", 2, _m([[1, 0, 0, 0], [0, 0, 0, 1], [0, 1, 0, 0], [0, 0, 1, 0]]))
CH = Gate("CH", 2, _m([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, s2, s2], [0, 0, s2, -s2]]))
CSX = Gate("CSX", 2, _m([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0.5 + 0.5j, 0.5 - 0.5j], [0, 0, 0.5 - 0.5j, 0.5 + 0.5j]]))
_ccx = np.eye(8, dtype=np.complex128)
_ccx[6:, 6:] = [[0, 1], [1, 0]]
CCX = Gate("CCX", 3, _ccx, tier=G_CCX)
_cswap = np.eye(8, dtype=np.complex128)
_cswap[5:7, 5:7] = [[0, 1], [1, 0]]
CSWAP = Gate("CSWAP", 3, _cswap, tier=G_CSWAP)
_ccz = np.eye(8, dtype=np.complex128)
_ccz[7, 7] = -1
CCZ = Gate("CCZ", 3, _ccz, tier=G_CCZ)
Sycamore = Gate("Sycamore", 2, _m([[1, 0, 0, 0], [0, 0, -1j, 0], [0, -1j, 0, 0],
                                    [0, 0, 0, cmath.exp(complex(0, -math.pi / 6))]]), tier=G_SYCAMORE)

Reset = _Pseudo("reset", 1, None)


def Delay(duration: float, unit: str = "s") -> Gate:
    g = _Pseudo("delay", 1, None, (duration,))
    g.unit = unit
    return g


def Barrier(n: int) -> Gate:
    return _Pseudo("barrier", n, None)


_SELF_INVERSE = {"I", "H", "X", "Y", "Z", "CNOT", "CZ", "SWAP", "CCX", "CSWAP", "ECR", "CCZ"}
_INV_PAIRS = {"S": lambda: Sdg, "S†": lambda: S, "T": lambda: Tdg, "T†": lambda: T}


def _e(phi: float) -> complex:
    return cmath.exp(complex(0, phi))


# ---- parameterized gates (circuit/gate/parameterized.go:91-564) ------------------------------------
def RX(theta: float) -> Gate:
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    return Gate(f"RX({theta:.4f})", 1, _m([[c, -1j * s], [-1j * s, c]]), (theta,))


def RY(theta: float) -> Gate:
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    return Gate(f"RY({theta:.4f})", 1, _m([[c, -s], [s, c]]), (theta,))


def RZ(theta: float) -> Gate:
    return Gate(f"RZ({theta:.4f})", 1, _m([[_e(-theta / 2), 0], [0, _e(theta / 2)]]), (theta,))


def Phase(phi: float) -> Gate:
    return Gate(f"P({phi:.4f})", 1, _m([[1, 0], [0, _e(phi)]]), (phi,))


def U3(theta: float, phi: float, lam: float) -> Gate:
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    return Gate(f"U3({theta:.4f},{phi:.4f},{lam:.4f})", 1,
                _m([[c, -_e(lam) * s], [_e(phi) * s, _e(phi + lam) * c]]), (theta, phi, lam))


def CP(phi: float) -> Gate:
    return Gate(f"CP({phi:.4f})", 2, np.diag([1, 1, 1, _e(phi)]).astype(np.complex128), (phi,), CTRL_DIAG)


def CRZ(theta: float) -> Gate:
    return Gate(f"CRZ({theta:.4f})", 2, np.diag([1, 1, _e(-theta / 2), _e(theta / 2)]).astype(np.complex128),
                (theta,), CTRL_DIAG)


def CRX(theta: float) -> Gate:
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    m = np.eye(4, dtype=np.complex128)
    m[2:, 2:] = [[c, -1j * s], [-1j * s, c]]
    return Gate(f"CRX({theta:.4f})", 2, m, (theta,), CTRL_U)


def CRY(theta: float) -> Gate:
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    m = np.eye(4, dtype=np.complex128)
    m[2:, 2:] = [[c, -s], [s, c]]
    return Gate(f"CRY({theta:.4f})", 2, m, (theta,), CTRL_U)


def RXX(theta: float) -> Gate:
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    i = complex(0, -s)
    return Gate(f"RXX({theta:.4f})", 2, _m([[c, 0, 0, i], [0, c, i, 0], [0, i, c, 0], [i, 0, 0, c]]), (theta,))


def RYY(theta: float) -> Gate:
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    ip, im = complex(0, s), complex(0, -s)
    return Gate(f"RYY({theta:.4f})", 2, _m([[c, 0, 0, ip], [0, c, im, 0], [0, im, c, 0], [ip, 0, 0, c]]), (theta,))


def RZZ(theta: float) -> Gate:
    em, ep = _e(-theta / 2), _e(theta / 2)
    return Gate(f"RZZ({theta:.4f})", 2, np.diag([em, ep, ep, em]).astype(np.complex128), (theta,), DIAG2)


def ZZ(theta: float) -> Gate:
    em, ep = _e(-theta / 2), _e(theta / 2)
    return Gate(f"ZZ({theta:.4f})", 2, np.diag([em, ep, ep, em]).astype(np.complex128), (theta,), DIAG2)


def GPI(phi: float) -> Gate:
    return Gate(f"GPI({phi:.4f})", 1, _m([[0, _e(-phi)], [_e(phi), 0]]), (phi,))


def GPI2(phi: float) -> Gate:
    return Gate(f"GPI2({phi:.4f})", 1, _m([[s2, s2 * -1j * _e(-phi)], [s2 * -1j * _e(phi), s2]]), (phi,))


def MS(phi0: float, phi1: float) -> Gate:
    ep, em = _e(phi0 + phi1), _e(phi0 - phi1)
    a = s2
    m = [[a, 0, 0, a * -1j * ep.conjugate()], [0, a, a * -1j * em, 0],
         [0, a * -1j * em.conjugate(), a, 0], [a * -1j * ep, 0, 0, a]]
    return Gate(f"MS({phi0:.4f},{phi1:.4f})", 2, _m(m), (phi0, phi1))


def U1(lam: float) -> Gate:
    return Gate(f"U1({lam:.4f})", 1, _m([[1, 0], [0, _e(lam)]]), (lam,))


def U2(phi: float, lam: float) -> Gate:
    return Gate(f"U2({phi:.4f},{lam:.4f})", 1,
                _m([[s2, -_e(lam) * s2], [_e(phi) * s2, _e(phi + lam) * s2]]), (phi, lam))


def Rot(phi: float, theta: float, omega: float) -> Gate:
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    epo, emo = _e((phi + omega) / 2), _e((phi - omega) / 2)
    return Gate(f"Rot({phi:.4f},{theta:.4f},{omega:.4f})", 1,
                _m([[epo.conjugate() * c, -emo.conjugate() * s], [emo * s, epo * c]]), (phi, theta, omega))


def PhasedXZ(x: float, z: float, a: float) -> Gate:
    cx, sx = math.cos(math.pi * x / 2), math.sin(math.pi * x / 2)
    ez, ea = _e(math.pi * z), _e(math.pi * a)
    return Gate(f"PhasedXZ({x:.4f},{z:.4f},{a:.4f})", 1,
                _m([[cx, ea.conjugate() * complex(0, sx)], [ez * ea * complex(0, sx), ez * cx]]), (x, z, a))


def GlobalPhase(phi: float) -> Gate:
    ep = _e(phi)
    return Gate(f"GlobalPhase({phi:.4f})", 1, _m([[ep, 0], [0, ep]]), (phi,))


def FSim(theta: float, phi: float) -> Gate:
    ct, st = math.cos(theta), math.sin(theta)
    return Gate(f"FSim({theta:.4f},{phi:.4f})", 2,
                _m([[1, 0, 0, 0], [0, ct, -1j * st, 0], [0, -1j * st, ct, 0], [0, 0, 0, _e(-phi)]]), (theta, phi))


def PSwap(phi: float) -> Gate:
    ep = _e(phi)
    return Gate(f"PSwap({phi:.4f})", 2, _m([[1, 0, 0, 0], [0, 0, ep, 0], [0, ep, 0, 0], [0, 0, 0, 1]]), (phi,))


def Unitary(name: str, matrix: Sequence[complex], n: Optional[int] = None) -> Gate:
    m = np.asarray(matrix, dtype=np.complex128).reshape(-1)
    k = n if n is not None else int(round(math.log2(len(m)) / 2))
    if len(m) != 1 << (2 * k):
        raise ValueError("gate: Unitary matrix size is not 4^k")
    return Gate(name, k, m)


# ---- controlled (circuit/gate/controlled.go:27-62) ----------------------------------------------------
def Controlled(inner: Gate, n_controls: int) -> Gate:
    if n_controls < 1:
        raise ValueError("gate: Controlled requires at least 1 control qubit")
    if n_controls == 1:
        if inner is X:
            return CNOT
        if inner is Z:
            return CZ
        if inner is Y:
            return CY
    if n_controls == 2 and inner is X:
        return CCX
    return ControlledGate(inner, n_controls)


def MCX(n_controls: int) -> Gate:
    return Controlled(X, n_controls)


def MCZ(n_controls: int) -> Gate:
    return Controlled(Z, n_controls)


def MCP(phi: float, n_controls: int) -> Gate:
    return Controlled(Phase(phi), n_controls)


def StatePrep(amplitudes: Sequence[complex]) -> Gate:
    return StatePrepGate(amplitudes)
