"""Statevector simulator on MI355X — drop-in mirror of goqu's sim/statevector.Sim
(sim/statevector/sim.go:21-383, dynamic.go:14-230) over the C ABI of libgoqu_hip.

Differences from the CPU reference, by design:
  * New accepts 1..34 qubits (one MI355X holds 33 complex128 qubits); the CPU reference
    panics above 28 (sim.go:28-30).
  * Sampling takes its uniforms from a caller-supplied array or numpy Generator, so runs are
    reproducible; the reference draws from an unseeded PCG per Run (sim.go:51).
  * RunDynamic runs all shots together on the device (goqu_amd.dynamic); shot s draws from its
    own row of uniforms, as the reference draws from a fresh generator per shot (dynamic.go:26).
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _native as N
from . import dynamic as D
from . import gate as G
from .circuit import Circuit
from .pauli import PauliString, PauliSum, masks


def format_bitstring(idx: int, n: int) -> str:
    """formatBitstring (sim.go:306-316): qubit n-1 leftmost."""
    return format(int(idx), f"0{n}b")


def _uniforms(shots: int, uniforms=None, rng=None) -> np.ndarray:
    if uniforms is not None:
        u = np.ascontiguousarray(uniforms, dtype=np.float64)
        if len(u) < shots:
            raise ValueError("fewer uniforms than shots")
        return u[:shots]
    rng = rng if rng is not None else np.random.default_rng()
    return rng.random(shots)


class Sim:
    """sim.Simulator implementation (sim/sim.go:9-22) backed by a device state vector."""

    def __init__(self, num_qubits: int, precision: int = N.C128, plan=None):
        if num_qubits < 1 or num_qubits > 34:
            raise ValueError(f"statevector: numQubits {num_qubits} out of range [1, 34]")
        self.num_qubits = num_qubits
        self.precision = precision
        h = C.c_void_p()
        N.check(N.lib().goqu_sv_create(num_qubits, precision, C.byref(h)))
        self._h = h
        if plan == "tuned":  # the autotuner's persisted shape for this size, if any
            from . import autotune

            plan = autotune.plan_for(num_qubits, precision)
        if plan is not None:
            self.SetPlan(*plan)

    def SetPlan(self, tile_bits: int, low_bits: int) -> None:
        """Tile shape of this simulator's programs (goqu_sv_set_plan; (0, 0) = the default)."""
        N.check(N.lib().goqu_sv_set_plan(self._h, int(tile_bits), int(low_bits)))

    # -- lifecycle --
    def Close(self) -> None:
        if self._h:
            N.lib().goqu_sv_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.Close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.Close()

    def Reset(self) -> None:
        N.check(N.lib().goqu_sv_reset(self._h))

    # -- evolution --
    def Evolve(self, c: Circuit) -> None:
        """Evolve (sim.go:82-85): reset to |0...0> and apply."""
        self.Reset()
        self.Apply(c)

    def Apply(self, c: Circuit) -> None:
        """Apply (sim.go:92-157)."""
        if c.NumQubits() != self.num_qubits:
            raise ValueError(f"circuit has {c.NumQubits()} qubits, simulator has {self.num_qubits}")
        for kind, a, b in _segments(c):
            if kind == "ops":
                N.check(N.lib().goqu_sv_apply(self._h, a[0], b))
            else:
                self._state_prep(a, b)

    def _flush(self, buf: "N.OpBuffer") -> None:
        if len(buf):
            arr = buf.array()
            N.check(N.lib().goqu_sv_apply(self._h, arr, len(buf)))

    def _state_prep(self, g: "G.StatePrepGate", qubits: Sequence[int]) -> None:
        # fast path of sim.go:105-120: full-state copy on qubits [0..n-1]
        if list(qubits) == list(range(self.num_qubits)):
            self.SetStateVector(g.Amplitudes())
            return
        # slow path (sim.go:121-135): the gate's RY + CNOT cascade; gates without a matrix are skipped
        buf = N.OpBuffer()
        for dg, dq in g.Decompose(qubits):
            if dg.Matrix() is not None:
                N.add_sv_gate(buf, dg, dq)
        self._flush(buf)

    def ApplyOps(self, buf: "N.OpBuffer") -> None:
        """Apply pre-marshalled op records."""
        self._flush(buf)

    # -- state access --
    def StateVector(self) -> np.ndarray:
        """StateVector (sim.go:61-65): a host copy."""
        n = 1 << self.num_qubits
        out = np.empty(n, dtype=np.complex128)
        N.check(N.lib().goqu_sv_get_state(self._h, N.dptr(out.view(np.float64)), 0, n))
        return out

    def StateVectorRange(self, offset: int, count: int) -> np.ndarray:
        out = np.empty(count, dtype=np.complex128)
        N.check(N.lib().goqu_sv_get_state(self._h, N.dptr(out.view(np.float64)), offset, count))
        return out

    def SetStateVector(self, amps, offset: int = 0) -> None:
        a = np.ascontiguousarray(amps, dtype=np.complex128)
        N.check(N.lib().goqu_sv_set_state(self._h, N.dptr(a.view(np.float64)), offset, len(a)))

    # -- measurement --
    def Sample(self, uniforms: np.ndarray) -> np.ndarray:
        """Inverse-CDF sampling (sim.go:274-298) for the given uniforms."""
        u = np.ascontiguousarray(uniforms, dtype=np.float64)
        out = np.empty(len(u), dtype=np.uint64)
        if len(u):
            N.check(N.lib().goqu_sv_sample(self._h, N.dptr(u), len(u), N.u64ptr(out)))
        return out

    def Run(self, c: Circuit, shots: int, uniforms=None, rng=None) -> Dict[str, int]:
        """Run (sim.go:40-58); dynamic circuits go to RunDynamic."""
        if c.IsDynamic():
            return self.RunDynamic(c, shots, rng=rng)
        self.Evolve(c)
        counts: Dict[str, int] = {}
        if shots <= 0:
            return counts
        idx = self.Sample(_uniforms(shots, uniforms, rng))
        vals, cnt = np.unique(idx, return_counts=True)
        fmt = f"0{self.num_qubits}b"  # format_bitstring's format, over Python ints
        return {format(v, fmt): k for v, k in zip(vals.tolist(), cnt.tolist())}

    def MeasureQubit(self, q: int, u: float) -> int:
        out = C.c_int32(0)
        N.check(N.lib().goqu_sv_measure_qubit(self._h, q, float(u), C.byref(out)))
        return out.value

    # -- dynamic circuits (dynamic.go:14-230) --
    def DynamicProgram(self, c: Circuit) -> "D.DynamicProgram":
        """The circuit as a device program (dynamic.go's executeOps cases, flattened)."""
        return D.flatten(c, lambda buf, op: N.add_sv_gate(buf, op.Gate, op.Qubits))

    def RunDynamicShots(self, c: Circuit, shots: int, rng=None, uniforms=None) -> np.ndarray:
        """All shots batched on the device; returns each shot's clbits (one row per shot).  Shot s
        takes its uniforms in order from row s of `uniforms` when given."""
        if c.NumQubits() != self.num_qubits:
            raise ValueError(f"circuit has {c.NumQubits()} qubits, simulator has {self.num_qubits}")
        if shots <= 0:
            raise ValueError(f"shots must be positive, got {shots}")
        return D.run(N.lib().goqu_sv_run_dynamic, self._h, self.DynamicProgram(c), shots, rng, uniforms)

    def RunDynamic(self, c: Circuit, shots: int, rng=None, uniforms=None) -> Dict[str, int]:
        """RunDynamic (dynamic.go:14-52): counts of the clbit strings, clbit n_c-1 leftmost."""
        return D.counts_of(self.RunDynamicShots(c, shots, rng, uniforms))

    # -- observables --
    def _check_pauli(self, nq: int, what: str) -> None:
        if nq != self.num_qubits:
            raise ValueError(f"statevector: {what} has {nq} qubits, simulator has {self.num_qubits}")

    def ExpectTerms(self, terms: Sequence[PauliString]) -> np.ndarray:
        """Complex <psi|P_t|psi> (with coefficients) for each term; terms sharing an x mask
        share one sweep of the state."""
        x, z, ny, cf = masks(terms)
        out = np.empty(2 * len(terms), dtype=np.float64)
        N.check(N.lib().goqu_sv_expect_paulis(self._h, len(terms), N.u64ptr(x), N.u64ptr(z), N.i32ptr(ny),
                                              N.dptr(cf), N.dptr(out)))
        return out.view(np.complex128)

    def ExpectPauliString(self, ps: PauliString) -> float:
        """sim.go:346-353"""
        self._check_pauli(ps.NumQubits(), "PauliString")
        return float(self.ExpectTerms([ps])[0].real)

    def ExpectPauliSum(self, ps: PauliSum) -> float:
        """sim.go:355-363: sum over terms in order."""
        self._check_pauli(ps.NumQubits(), "PauliSum")
        vals = self.ExpectTerms(ps.Terms())
        total = 0j
        for v in vals:
            total += v
        return float(total.real)

    def ExpectationValue(self, qubits: Sequence[int]) -> float:
        """sim.go:368-383 (rounded to 1e-14)."""
        mask = 0
        for q in qubits:
            mask |= 1 << q
        out = C.c_double(0)
        N.check(N.lib().goqu_sv_expectation_z(self._h, mask, C.byref(out)))
        return round(out.value * 1e14) / 1e14

    def Norm2(self) -> float:
        out = C.c_double(0)
        N.check(N.lib().goqu_sv_norm2(self._h, C.byref(out)))
        return out.value


def _segments(c: Circuit) -> list:
    """The circuit's op records for goqu_sv_apply, marshalled once per circuit (a built circuit
    does not change): ("ops", (record array, the buffer owning its matrices), count) runs split at
    StatePrep gates, which appear as ("prep", gate, qubits)."""
    segs = getattr(c, "_sv_segments", None)
    if segs is not None:
        return segs
    segs = []
    buf = N.OpBuffer()

    def close(buf):
        if len(buf):
            segs.append(("ops", (buf.array(), buf), len(buf)))

    for op in c.Ops():
        g = op.Gate
        if g is None or op.ControlFlow is not None:
            continue
        if isinstance(g, G.StatePrepGate):
            close(buf)
            buf = N.OpBuffer()
            segs.append(("prep", g, list(op.Qubits)))
            continue
        N.add_sv_gate(buf, g, op.Qubits)
    close(buf)
    c._sv_segments = segs
    return segs


def New(num_qubits: int, precision: int = N.C128, plan=None) -> Sim:
    """statevector.New (sim.go:27-35); plan = (tile_bits, low_bits), or "tuned" for the
    autotuner's choice."""
    return Sim(num_qubits, precision, plan)
