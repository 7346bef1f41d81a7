"""Density-matrix simulator on MI355X — drop-in mirror of goqu's sim/densitymatrix.Sim
(sim/densitymatrix/sim.go:24-518, measure.go, dynamic.go).

rho lives on the device as a 2n-qubit vector (row bits high): a gate is U on the row qubits
and conj(U) on the column qubits, a Kraus channel is its superoperator on both, each applied
by the tiled engine in fused sweeps.  Noise lookups and the >2-qubit decompositions run on
the host exactly as in the reference (sim.go:182-188, 321-518).
"""
from __future__ import annotations

import cmath
import ctypes as C
import math
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _native as N
from . import gate as G
from . import noise as NZ
from .circuit import Circuit, Operation
from .pauli import PauliString, PauliSum, masks
from .statevector import _uniforms, format_bitstring


# ---- >2-qubit decomposition (sim/densitymatrix/sim.go:321-518) --------------------------------------
def _near_zero(x: float) -> bool:
    return abs(math.remainder(x, 2 * math.pi)) < 1e-10


def _mcx(controls: List[int], target: int) -> List[Operation]:
    n = len(controls)
    if n == 1:
        return [Operation(G.CNOT, [controls[0], target])]
    if n == 2:
        c0, c1 = controls
        seq = [(G.H, [target]), (G.CNOT, [c1, target]), (G.Tdg, [target]), (G.CNOT, [c0, target]),
               (G.T, [target]), (G.CNOT, [c1, target]), (G.Tdg, [target]), (G.CNOT, [c0, target]),
               (G.T, [c1]), (G.T, [target]), (G.CNOT, [c0, c1]), (G.H, [target]), (G.T, [c0]),
               (G.Tdg, [c1]), (G.CNOT, [c0, c1])]
        return [Operation(g, q) for g, q in seq]
    last, rest = controls[-1], controls[:-1]
    ops = _c1q(G.SX, rest, target)
    ops.append(Operation(G.CNOT, [last, target]))
    ops += _c1q(G.SX.Inverse(), rest, target)
    ops.append(Operation(G.CNOT, [last, target]))
    ops += _c1q(G.S, rest, last)
    return ops


def _c1q(u: G.Gate, controls: List[int], target: int) -> List[Operation]:
    if len(controls) == 1:
        return [Operation(G.Controlled(u, 1), [controls[0], target])]
    if u is G.X or u.Name() == "X":
        return _mcx(controls, target)
    return _general(u, controls, target)


def _general(u: G.Gate, controls: List[int], target: int) -> List[Operation]:
    m = u.Matrix()
    det = m[0] * m[3] - m[1] * m[2]
    phase = cmath.log(det).imag / 2
    factor = cmath.exp(complex(0, -phase))
    a, b = m[0] * factor, m[1] * factor
    beta = 2 * math.acos(min(max(abs(a), 0.0), 1.0))
    if abs(b) < 1e-10:
        alpha, gamma = -2 * cmath.phase(a), 0.0
    elif abs(a) < 1e-10:
        alpha, beta, gamma = -2 * cmath.phase(-b), math.pi, 0.0
    else:
        apg, amg = cmath.phase(a), cmath.phase(-b)
        alpha, gamma = -(apg + amg), -(apg - amg)
    ops: List[Operation] = []
    if not _near_zero(gamma - alpha):
        ops.append(Operation(G.RZ((gamma - alpha) / 2), [target]))
    ops += _mcx(controls, target)
    if not _near_zero(alpha + gamma):
        ops.append(Operation(G.RZ(-(alpha + gamma) / 2), [target]))
    if not _near_zero(beta):
        ops.append(Operation(G.RY(-beta / 2), [target]))
    ops += _mcx(controls, target)
    if not _near_zero(beta):
        ops.append(Operation(G.RY(beta / 2), [target]))
    if not _near_zero(alpha):
        ops.append(Operation(G.RZ(alpha), [target]))
    if not _near_zero(phase):
        n = len(controls)
        if n == 1:
            ops.append(Operation(G.Phase(phase), [controls[0]]))
        else:
            ops += _c1q(G.Phase(phase), controls[:-1], controls[-1])
    return ops


def decompose_for_density(op: Operation) -> Optional[List[Operation]]:
    """decomposeForDensity (sim.go:321-352)."""
    g = op.Gate
    applied = g.Decompose(op.Qubits)
    if applied is not None:
        out: List[Operation] = []
        for a in applied:
            if a.Gate.Qubits() <= 2:
                out.append(a)
            else:
                inner = decompose_for_density(a)
                if inner is None:
                    return None
                out += inner
        return out
    if isinstance(g, G.ControlledGate):
        nc = g.NumControls()
        controls, targets = op.Qubits[:nc], op.Qubits[nc:]
        inner = g.Inner()
        if nc == 1 and inner.Qubits() == 1:
            return [op]
        if inner.Qubits() != 1:
            return None
        if inner is G.X or inner.Name() == "X":
            return _mcx(list(controls), targets[0])
        return _general(inner, list(controls), targets[0])
    return None


class Sim:
    def __init__(self, num_qubits: int):
        if num_qubits < 1 or num_qubits > 16:
            raise ValueError(f"densitymatrix: numQubits {num_qubits} out of range [1, 16]")
        self.num_qubits = num_qubits
        self.dim = 1 << num_qubits
        self.noise: Optional[NZ.NoiseModel] = None
        h = C.c_void_p()
        N.check(N.lib().goqu_dm_create(num_qubits, C.byref(h)))
        self._h = h

    def WithNoise(self, nm: NZ.NoiseModel) -> "Sim":
        self.noise = nm
        return self

    def Close(self) -> None:
        if self._h:
            N.lib().goqu_dm_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.Close()
        except Exception:
            pass

    def Reset(self) -> None:
        N.check(N.lib().goqu_dm_reset(self._h))

    def Evolve(self, c: Circuit) -> None:
        self.Reset()
        self.Apply(c)

    # -- op marshalling (sim.go:87-191) --
    def _add_channel(self, buf: "N.OpBuffer", ch: NZ.Channel, qubits: Sequence[int]) -> None:
        ks = ch.Kraus()
        if ch.Qubits() not in (1, 2):
            return  # applyChannel only handles 1q/2q channels (sim.go:228-233)
        buf.add(N.OP_CHANNEL, list(qubits)[: ch.Qubits()], np.concatenate(ks), num_kraus=len(ks))

    def _add_gate(self, buf: "N.OpBuffer", g: G.Gate, qubits: Sequence[int]) -> None:
        m = g.Matrix()
        if m is None:
            return
        buf.add(N.OP_GATE, qubits, m)

    def marshal(self, c: Circuit) -> "N.OpBuffer":
        if c.NumQubits() != self.num_qubits:
            raise ValueError(f"circuit has {c.NumQubits()} qubits, simulator has {self.num_qubits}")
        buf = N.OpBuffer()
        for op in c.Ops():
            self._marshal_op(buf, op)
        return buf

    def _marshal_op(self, buf: "N.OpBuffer", op: Operation) -> None:
        g = op.Gate
        if g is None or op.ControlFlow is not None or g.Name() == "barrier":
            return
        if g.Name() == "delay":
            if self.noise is not None:
                ch = self.noise.Lookup("delay", op.Qubits)
                if ch is not None:
                    self._add_channel(buf, ch, op.Qubits)
            return
        if g.Name() == "reset":
            buf.add(N.OP_RESET, op.Qubits[:1])
            if self.noise is not None:
                ch = self.noise.Lookup("reset", op.Qubits)
                if ch is not None:
                    self._add_channel(buf, ch, op.Qubits)
            return
        if isinstance(g, G.StatePrepGate):
            raise NotImplementedError("StatePrep on a density matrix: use SetDensityMatrix")
        k = g.Qubits()
        if k <= 2:
            if g.Matrix() is None:
                return  # sim.go:147-150 (no noise either)
            self._add_gate(buf, g, op.Qubits)
        else:
            subs = decompose_for_density(op)
            if subs is None:
                raise ValueError(f"densitymatrix: unsupported gate size: {k} qubits")
            for sub in subs:
                if sub.Gate.Qubits() > 2:
                    raise ValueError(f"densitymatrix: decomposition produced {sub.Gate.Qubits()}-qubit gate")
                self._add_gate(buf, sub.Gate, sub.Qubits)
        if self.noise is not None:
            ch = self.noise.Lookup(g.Name(), op.Qubits)
            if ch is not None:
                self._add_channel(buf, ch, op.Qubits)

    def Apply(self, c: Circuit) -> None:
        buf = self.marshal(c)
        if len(buf):
            N.check(N.lib().goqu_dm_apply(self._h, buf.array(), len(buf)))

    def ApplyOps(self, buf: "N.OpBuffer") -> None:
        if len(buf):
            N.check(N.lib().goqu_dm_apply(self._h, buf.array(), len(buf)))

    # -- access & observables --
    def DensityMatrix(self) -> np.ndarray:
        n = self.dim * self.dim
        out = np.empty(n, dtype=np.complex128)
        N.check(N.lib().goqu_dm_get(self._h, N.dptr(out.view(np.float64)), 0, n))
        return out

    def SetDensityMatrix(self, rho) -> None:
        a = np.ascontiguousarray(rho, dtype=np.complex128).reshape(-1)
        N.check(N.lib().goqu_dm_set(self._h, N.dptr(a.view(np.float64)), 0, len(a)))

    def _readout_arrays(self):
        if self.noise is None or not self.noise.readout:
            return None, None
        p01 = np.full(self.num_qubits, -1.0)
        p10 = np.full(self.num_qubits, -1.0)
        for q in range(self.num_qubits):
            re = self.noise.ReadoutFor(q)
            if re is not None:
                p01[q], p10[q] = re.P01, re.P10
        return p01, p10

    def DiagonalProbs(self) -> np.ndarray:
        """diagonalProbs (measure.go:11-21)."""
        out = np.empty(self.dim)
        N.check(N.lib().goqu_dm_probs(self._h, None, None, N.dptr(out)))
        return out

    def Probabilities(self) -> np.ndarray:
        """ApplyReadoutError(diagonalProbs()) as used by Run (sim.go:59)."""
        p01, p10 = self._readout_arrays()
        out = np.empty(self.dim)
        N.check(N.lib().goqu_dm_probs(self._h, N.dptr(p01) if p01 is not None else None,
                                      N.dptr(p10) if p10 is not None else None, N.dptr(out)))
        return out

    def Run(self, c: Circuit, shots: int, uniforms=None, rng=None) -> Dict[str, int]:
        if c.IsDynamic():
            raise NotImplementedError("densitymatrix RunDynamic is not supported on the device path")
        self.Evolve(c)
        probs = self.Probabilities()
        counts: Dict[str, int] = {}
        if shots <= 0:
            return counts
        u = _uniforms(shots, uniforms, rng)
        out = np.empty(shots, dtype=np.uint64)
        N.check(N.lib().goqu_dm_sample(self._h, N.dptr(probs), N.dptr(u), shots, N.u64ptr(out)))
        vals, cnt = np.unique(out, return_counts=True)
        for v, k in zip(vals, cnt):
            counts[format_bitstring(v, self.num_qubits)] = int(k)
        return counts

    def ExpectTerms(self, terms: Sequence[PauliString]) -> np.ndarray:
        x, z, ny, cf = masks(terms)
        out = np.empty(2 * len(terms))
        N.check(N.lib().goqu_dm_expect_paulis(self._h, len(terms), N.u64ptr(x), N.u64ptr(z), N.i32ptr(ny),
                                              N.dptr(cf), N.dptr(out)))
        return out.view(np.complex128)

    def ExpectPauliString(self, ps: PauliString) -> float:
        if ps.NumQubits() != self.num_qubits:
            raise ValueError(f"densitymatrix: PauliString has {ps.NumQubits()} qubits, simulator has {self.num_qubits}")
        return float(self.ExpectTerms([ps])[0].real)

    def ExpectPauliSum(self, ps: PauliSum) -> float:
        if ps.NumQubits() != self.num_qubits:
            raise ValueError(f"densitymatrix: PauliSum has {ps.NumQubits()} qubits, simulator has {self.num_qubits}")
        total = 0j
        for v in self.ExpectTerms(ps.Terms()):
            total += v
        return float(total.real)

    def ExpectationValue(self, qubits: Sequence[int]) -> float:
        mask = 0
        for q in qubits:
            mask |= 1 << q
        out = C.c_double(0)
        N.check(N.lib().goqu_dm_expectation_z(self._h, mask, C.byref(out)))
        return round(out.value * 1e14) / 1e14

    def Purity(self) -> float:
        out = C.c_double(0)
        N.check(N.lib().goqu_dm_purity(self._h, C.byref(out)))
        return out.value

    def Fidelity(self, pure) -> float:
        p = np.ascontiguousarray(pure, dtype=np.complex128)
        if len(p) != self.dim:
            raise ValueError(f"densitymatrix: pure state length {len(p)}, expected {self.dim}")
        out = C.c_double(0)
        N.check(N.lib().goqu_dm_fidelity(self._h, N.dptr(p.view(np.float64)), C.byref(out)))
        return out.value


def New(num_qubits: int) -> Sim:
    return Sim(num_qubits)
