"""Dynamic circuits on the device: a circuit with mid-circuit measurement, reset, conditioned gates
and structured control flow is flattened into the C ABI's instruction program (goqu_dyn_instr_t)
over an op array, and all shots run batched on the GPU (goqu_sv_run_dynamic /
goqu_dm_run_dynamic).

The flattening follows the reference's executeOps / executeControlFlow case by case
(sim/statevector/dynamic.go:55-146, sim/densitymatrix/dynamic.go:54-157):

  * a measurement (no gate, qubits and clbits) -> MEASURE(qubit, clbit);
  * barrier -> nothing; delay -> nothing (statevector) or its noise channel (density matrix);
  * reset -> RESET(qubit) (measure, X on outcome 1), then the "reset" noise channel (density
    matrix); neither reset nor delay looks at a Condition, as in the reference;
  * a gate with a Condition -> BRANCH, CASE(clbit == value), the gate (and its noise), END;
  * if/else -> BRANCH, CASE(clbit == value), body, [DEFAULT, else body], END;
  * switch -> BRANCH, one CASE per case value that has a body (ascending, as the builder stores
    them), DEFAULT for the extra body, END;
  * while -> WHILE(clbit == value), body, END (at most 1000 iterations on the device);
  * for -> FOR(len(range(start, end, step))), body, END.

Runs of consecutive gates become one GATES instruction, planned once into fused stage sweeps.
Each shot draws its uniforms from its own slab; `draws` bounds how many a shot can need.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _native as N
from .circuit import Circuit, MaxControlFlowIterations, Operation

# host memory for the uniform slabs of one call into the library
_MAX_SLAB_BYTES = 256 << 20


@dataclass
class DynamicProgram:
    buf: "N.OpBuffer"
    instrs: List[Tuple[int, int, int, int]] = field(default_factory=list)
    cond_bits: List[int] = field(default_factory=list)
    draws: int = 0  # most uniforms any shot can draw
    num_clbits: int = 0

    def instr_array(self):
        arr = (N.GoquDynInstr * max(1, len(self.instrs)))()
        for i, (k, a, b, v) in enumerate(self.instrs):
            arr[i].kind, arr[i].a, arr[i].b, arr[i].value = k, a, b, v
        return arr


class _Flattener:
    def __init__(self, emit_gate: Callable[["N.OpBuffer", Operation], None],
                 emit_delay: Optional[Callable[["N.OpBuffer", Operation], None]],
                 emit_reset_tail: Optional[Callable[["N.OpBuffer", Operation], None]]):
        self.emit_gate = emit_gate
        self.emit_delay = emit_delay
        self.emit_reset_tail = emit_reset_tail
        self.p = DynamicProgram(N.OpBuffer())
        self.open_gates: Optional[int] = None  # index of the GATES instruction still growing

    def _ins(self, kind: int, a: int = 0, b: int = 0, value: int = 0) -> None:
        self.open_gates = None
        self.p.instrs.append((kind, a, b, value))

    def _gates(self, emit: Callable[["N.OpBuffer", Operation], None], op: Operation) -> None:
        n0 = len(self.p.buf)
        emit(self.p.buf, op)
        added = len(self.p.buf) - n0
        if not added:
            return
        if self.open_gates is None:
            self.p.instrs.append((N.DYN_GATES, n0, 0, 0))
            self.open_gates = len(self.p.instrs) - 1
        k, a, b, v = self.p.instrs[self.open_gates]
        self.p.instrs[self.open_gates] = (k, a, b + added, v)

    def _cond(self, bits: Sequence[int]) -> Tuple[int, int]:
        a = len(self.p.cond_bits)
        self.p.cond_bits.extend(int(x) for x in bits)
        return a, len(bits)

    def ops(self, ops: Sequence[Operation]) -> int:
        """Flattens ops; returns the most uniforms one pass can draw."""
        draws = 0
        for op in ops:
            cf = op.ControlFlow
            if cf is not None:
                draws += self.control_flow(cf)
                continue
            g = op.Gate
            if g is None:
                if op.Clbits and op.Qubits:
                    self._ins(N.DYN_MEASURE, op.Qubits[0], op.Clbits[0])
                    draws += 1
                continue
            name = g.Name()
            if name == "barrier":
                continue
            if name == "delay":
                if self.emit_delay is not None:
                    self._gates(self.emit_delay, op)
                continue
            if name == "reset":
                self._ins(N.DYN_RESET, op.Qubits[0])
                draws += 1
                if self.emit_reset_tail is not None:
                    self._gates(self.emit_reset_tail, op)
                continue
            if op.Condition is not None:
                a, b = self._cond([op.Condition.Clbit])
                self._ins(N.DYN_BRANCH)
                self._ins(N.DYN_CASE, a, b, op.Condition.Value)
                self._gates(self.emit_gate, op)
                self._ins(N.DYN_END)
                continue
            self._gates(self.emit_gate, op)
        return draws

    def control_flow(self, cf) -> int:
        if cf.Type == "while":
            a, b = self._cond([cf.Condition.Clbit])
            self._ins(N.DYN_WHILE, a, b, cf.Condition.Value)
            d = self.ops(cf.Bodies[0])
            self._ins(N.DYN_END)
            return MaxControlFlowIterations * d
        if cf.Type == "for":
            count = len(range(*cf.ForRange))
            self._ins(N.DYN_FOR, count)
            d = self.ops(cf.Bodies[0])
            self._ins(N.DYN_END)
            return count * d
        if cf.Type == "if_else":
            a, b = self._cond([cf.Condition.Clbit])
            self._ins(N.DYN_BRANCH)
            self._ins(N.DYN_CASE, a, b, cf.Condition.Value)
            d = self.ops(cf.Bodies[0])
            if len(cf.Bodies) > 1:
                self._ins(N.DYN_DEFAULT)
                d = max(d, self.ops(cf.Bodies[1]))
            self._ins(N.DYN_END)
            return d
        if cf.Type == "switch":
            cases, bodies = list(cf.SwitchCases), cf.Bodies
            a, b = self._cond(cf.SwitchClbits)
            self._ins(N.DYN_BRANCH)
            d = 0
            for i, v in enumerate(cases):
                if i < len(bodies):
                    self._ins(N.DYN_CASE, a, b, int(v))
                    d = max(d, self.ops(bodies[i]))
            if len(bodies) > len(cases):
                self._ins(N.DYN_DEFAULT)
                d = max(d, self.ops(bodies[-1]))
            self._ins(N.DYN_END)
            return d
        raise ValueError(f"unknown control flow {cf.Type!r}")


def flatten(c: Circuit, emit_gate, emit_delay=None, emit_reset_tail=None) -> DynamicProgram:
    f = _Flattener(emit_gate, emit_delay, emit_reset_tail)
    f.p.draws = f.ops(c.Ops())
    f.p.num_clbits = c.NumClbits()
    return f.p


def run(fn, handle, prog: DynamicProgram, shots: int, rng=None, uniforms: Optional[np.ndarray] = None) -> np.ndarray:
    """Runs `shots` shots of prog through fn (goqu_sv_run_dynamic or goqu_dm_run_dynamic); returns
    the clbits, one row per shot.  uniforms (shots x prog.draws) replaces the generator's draws."""
    nc, d = prog.num_clbits, prog.draws
    out = np.zeros((shots, max(nc, 1)), dtype=np.uint8)
    if uniforms is not None:
        uniforms = np.ascontiguousarray(uniforms, dtype=np.float64).reshape(shots, -1)
        if uniforms.shape[1] < d:
            raise ValueError(f"uniforms: {uniforms.shape[1]} per shot, the circuit may draw {d}")
        d = uniforms.shape[1]
    rng = rng if rng is not None else np.random.default_rng()
    ops = prog.buf.array()
    ins = prog.instr_array()
    bits = np.ascontiguousarray(prog.cond_bits or [0], dtype=np.int32)
    chunk = max(1, min(shots, _MAX_SLAB_BYTES // max(8 * d, 1)))
    for s0 in range(0, shots, chunk):
        k = min(chunk, shots - s0)
        u = uniforms[s0:s0 + k] if uniforms is not None else rng.random((k, d))
        u = np.ascontiguousarray(u if d else np.zeros((k, 1)), dtype=np.float64)
        o = np.zeros((k, max(nc, 1)), dtype=np.uint8)
        N.check(fn(handle, ops, len(prog.buf), ins, len(prog.instrs), N.i32ptr(bits), len(prog.cond_bits), nc, k,
                   N.dptr(u), d, o.ctypes.data_as(C.POINTER(C.c_uint8))))
        out[s0:s0 + k] = o
    return out[:, :nc]


def counts_of(clbits: np.ndarray) -> Dict[str, int]:
    """Bitstrings with clbit n_c - 1 leftmost (dynamic.go:40-48)."""
    nc = clbits.shape[1]
    counts: Dict[str, int] = {}
    if nc == 0:
        counts[""] = clbits.shape[0]
        return counts
    rows, cnt = np.unique(clbits, axis=0, return_counts=True)
    for r, k in zip(rows, cnt):
        counts["".join("1" if r[nc - 1 - i] else "0" for i in range(nc))] = int(k)
    return counts
