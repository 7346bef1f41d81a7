"""Density-matrix simulator on MI355X — drop-in mirror of goqu's sim/densitymatrix.Sim
(sim/densitymatrix/sim.go:24-518, measure.go, dynamic.go).

rho lives on the device as a 2n-qubit vector (row bits high): a gate is U on the row qubits
and conj(U) on the column qubits, a Kraus channel is its superoperator on both, each applied
by the tiled engine in fused sweeps.  Noise lookups run on the host as in the reference
(sim.go:182-188).  Gates on 3+ qubits are applied exactly, where the reference decomposes
controlled gates (sim.go:321-518) and rejects other 3-qubit gates.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _native as N
from . import dynamic as DY
from . import gate as G
from . import noise as NZ
from .circuit import Circuit, Operation
from .pauli import PauliString, PauliSum, masks
from .statevector import _uniforms, format_bitstring


def _add_channel(buf: "N.OpBuffer", ch: NZ.Channel, qubits: Sequence[int]) -> None:
    if ch.Qubits() not in (1, 2):
        return  # applyChannel only handles 1q/2q channels (sim.go:228-233)
    ks = ch.Kraus()
    buf.add(N.OP_CHANNEL, list(qubits)[: ch.Qubits()], np.concatenate(ks), num_kraus=len(ks))


def _marshal_op(buf: "N.OpBuffer", op: Operation, noise: Optional[NZ.NoiseModel]) -> None:
    """One op of DM Apply (sim.go:87-191) -> op records, with the noise lookups."""
    g = op.Gate
    if g is None or op.ControlFlow is not None or g.Name() == "barrier":
        return
    if g.Name() == "delay":
        if noise is not None:
            ch = noise.Lookup("delay", op.Qubits)
            if ch is not None:
                _add_channel(buf, ch, op.Qubits)
        return
    if g.Name() == "reset":
        buf.add(N.OP_RESET, op.Qubits[:1])
        if noise is not None:
            ch = noise.Lookup("reset", op.Qubits)
            if ch is not None:
                _add_channel(buf, ch, op.Qubits)
        return
    if isinstance(g, G.StatePrepGate):
        raise NotImplementedError("StatePrep on a density matrix: use SetDensityMatrix")
    k = g.Qubits()
    if k >= 3 and isinstance(g, G.ControlledGate):
        # exact controlled gate; goqu decomposes it on the host instead (sim.go:321-518),
        # and its multi-controlled X decomposition for 3+ controls leaves the target unchanged
        inner = g.Inner().Matrix()
        if inner is None:
            return
        buf.add(N.OP_CONTROLLED, op.Qubits, inner, num_controls=g.NumControls())
    else:
        m = g.Matrix()
        if m is None:
            return  # sim.go:147-150 (no noise either)
        buf.add(N.OP_GATE, op.Qubits, m)
    if noise is not None:
        ch = noise.Lookup(g.Name(), op.Qubits)
        if ch is not None:
            _add_channel(buf, ch, op.Qubits)


def marshal(c: Circuit, noise: Optional[NZ.NoiseModel] = None) -> "N.OpBuffer":
    """The whole circuit as density-matrix op records."""
    buf = N.OpBuffer()
    for op in c.Ops():
        _marshal_op(buf, op, noise)
    return buf


def _readout_arrays(noise: Optional[NZ.NoiseModel], n: int):
    """Per-qubit readout flip probabilities (-1: none) for goqu_dm_probs, or (None, None)."""
    if noise is None or not noise.readout:
        return None, None
    p01 = np.full(n, -1.0)
    p10 = np.full(n, -1.0)
    for q in range(n):
        re = noise.ReadoutFor(q)
        if re is not None:
            p01[q], p10[q] = re.P01, re.P10
    return p01, p10


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

    def marshal(self, c: Circuit) -> "N.OpBuffer":
        if c.NumQubits() != self.num_qubits:
            raise ValueError(f"circuit has {c.NumQubits()} qubits, simulator has {self.num_qubits}")
        return marshal(c, self.noise)

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
        return _readout_arrays(self.noise, self.num_qubits)

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

    # -- dynamic circuits (sim/densitymatrix/dynamic.go:13-255) --
    def MeasureQubit(self, q: int, u: float) -> int:
        """measureQubitDM (dynamic.go:211-255): P0 from the diagonal, outcome = (u >= P0),
        rho <- P rho P / p_outcome (skipped when p_outcome < 1e-15)."""
        p = self.DiagonalProbs()
        idx = np.arange(self.dim)
        prob0 = float(p[(idx >> q) & 1 == 0].sum())
        outcome = 1 if u >= prob0 else 0
        pout = prob0 if outcome == 0 else 1 - prob0
        if pout < 1e-15:
            return outcome
        e = np.zeros(4, dtype=np.complex128)
        e[3 if outcome else 0] = 1 / np.sqrt(pout)
        buf = N.OpBuffer()
        buf.add(N.OP_CHANNEL, [q], e, num_kraus=1)
        self.ApplyOps(buf)
        return outcome

    def DynamicProgram(self, c: Circuit) -> "DY.DynamicProgram":
        """The circuit as a device program (dynamic.go's executeOps cases, flattened): gates with
        their noise channels, delay as its channel, the "reset" channel after each reset."""
        nm = self.noise

        def gate(buf, op):
            _marshal_op(buf, op, nm)

        def reset_tail(buf, op):
            if nm is not None:
                ch = nm.Lookup("reset", op.Qubits)
                if ch is not None:
                    _add_channel(buf, ch, op.Qubits)

        return DY.flatten(c, gate, emit_delay=gate, emit_reset_tail=reset_tail)

    def RunDynamicShots(self, c: Circuit, shots: int, rng=None, uniforms=None) -> np.ndarray:
        """All shots batched on the device; each shot's clbits, one row per shot."""
        if c.NumQubits() != self.num_qubits:
            raise ValueError(f"circuit has {c.NumQubits()} qubits, simulator has {self.num_qubits}")
        if shots <= 0:
            raise ValueError(f"shots must be positive, got {shots}")
        return DY.run(N.lib().goqu_dm_run_dynamic, self._h, self.DynamicProgram(c), shots, rng, uniforms)

    def RunDynamic(self, c: Circuit, shots: int, rng=None, uniforms=None) -> Dict[str, int]:
        """RunDynamic (dynamic.go:13-51)."""
        return DY.counts_of(self.RunDynamicShots(c, shots, rng, uniforms))

    def Run(self, c: Circuit, shots: int, uniforms=None, rng=None) -> Dict[str, int]:
        """Run (sim.go:51-67): dynamic circuits go to RunDynamic; otherwise evolve, apply the
        readout error and sample."""
        if c.IsDynamic():
            return self.RunDynamic(c, shots, rng=rng)
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
