"""Circuit IR, symbolic parameters and the fluent builder — host-side mirror of goqu's
circuit/ir (circuit.go:11-322), circuit/param (param.go, gate.go) and circuit/builder
(builder.go:14-760), restricted to what the simulator path consumes.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

from . import gate as G

MaxControlFlowIterations = 1000  # circuit/ir/circuit.go:129


# ---- symbolic parameters (circuit/param) -------------------------------------------------------------
class Expr:
    def eval(self, b: Dict[str, float]) -> float:
        raise NotImplementedError

    def free(self) -> List[str]:
        return []

    def deriv(self, name: str, b: Dict[str, float]) -> float:
        raise NotImplementedError

    def Expr(self) -> "Expr":
        return self


@dataclass(frozen=True)
class Literal(Expr):
    v: float

    def eval(self, b):
        return self.v

    def deriv(self, name, b):
        return 0.0


@dataclass(frozen=True)
class Parameter(Expr):
    name: str

    def Name(self) -> str:
        return self.name

    def eval(self, b):
        if self.name not in b:
            raise KeyError(f"param: unbound parameter {self.name!r}")
        return float(b[self.name])

    def free(self):
        return [self.name]

    def deriv(self, name, b):
        return 1.0 if name == self.name else 0.0


@dataclass(frozen=True)
class Mul(Expr):
    a: Expr
    b: Expr

    def eval(self, bd):
        return self.a.eval(bd) * self.b.eval(bd)

    def free(self):
        return self.a.free() + self.b.free()

    def deriv(self, name, bd):
        return self.a.deriv(name, bd) * self.b.eval(bd) + self.a.eval(bd) * self.b.deriv(name, bd)


@dataclass(frozen=True)
class Add(Expr):
    a: Expr
    b: Expr

    def eval(self, bd):
        return self.a.eval(bd) + self.b.eval(bd)

    def free(self):
        return self.a.free() + self.b.free()

    def deriv(self, name, bd):
        return self.a.deriv(name, bd) + self.b.deriv(name, bd)


def New(name: str) -> Parameter:
    """param.New"""
    return Parameter(name)


class Vector:
    """param.NewVector (param.go:38): elements named "name[i]"."""

    def __init__(self, name: str, n: int):
        self.name = name
        self.n = n

    def At(self, i: int) -> Parameter:
        if not 0 <= i < self.n:
            raise IndexError(f"param: vector index {i} out of range")
        return Parameter(f"{self.name}[{i}]")

    def Size(self) -> int:
        return self.n


def NewVector(name: str, n: int) -> Vector:
    return Vector(name, n)


def _expr(x) -> Expr:
    return x if isinstance(x, Expr) else Literal(float(x))


class SymbolicGate(G.Gate):
    """A gate whose arguments are expressions (circuit/param/gate.go:15-64).  Matrix() of an
    unbound gate raises, as goqu panics (gate.go:31-33)."""

    def __init__(self, ctor_name: str, n: int, exprs: Sequence[Expr]):
        super().__init__(ctor_name, n, None)
        self.ctor_name = ctor_name
        self.exprs = tuple(_expr(e) for e in exprs)

    def Matrix(self):
        raise RuntimeError(f"param: gate {self.ctor_name} has unbound parameters")

    def FreeParameters(self) -> List[str]:
        out: List[str] = []
        for e in self.exprs:
            for n in e.free():
                if n not in out:
                    out.append(n)
        return out

    def Bind(self, b: Dict[str, float]) -> G.Gate:
        args = [e.eval(b) for e in self.exprs]
        return CONSTRUCTORS[self.ctor_name](*args)


CONSTRUCTORS: Dict[str, Callable[..., G.Gate]] = {
    "RX": G.RX, "RY": G.RY, "RZ": G.RZ, "P": G.Phase, "CP": G.CP, "RXX": G.RXX, "RYY": G.RYY,
    "RZZ": G.RZZ, "Rot": G.Rot, "FSim": G.FSim, "PSwap": G.PSwap, "GlobalPhase": G.GlobalPhase,
    "U3": G.U3, "CRX": G.CRX, "CRY": G.CRY, "CRZ": G.CRZ, "ZZ": G.ZZ, "U1": G.U1, "U2": G.U2,
    "GPI": G.GPI, "GPI2": G.GPI2, "MS": G.MS, "PhasedXZ": G.PhasedXZ,
}


# ---- IR (circuit/ir/circuit.go) ----------------------------------------------------------------------
@dataclass
class Condition:
    Clbit: int
    Value: int


@dataclass
class ControlFlow:
    Type: str  # "if_else" | "while" | "for" | "switch"
    Bodies: List[List["Operation"]]
    Condition: Optional[Condition] = None
    ForRange: Optional[tuple] = None  # (start, end, step)
    SwitchClbits: Optional[List[int]] = None
    SwitchCases: Optional[List[int]] = None


@dataclass
class Operation:
    """ir.Operation (circuit.go:79-85); a measurement is Gate=None with Qubits and Clbits."""

    Gate: Optional[G.Gate]
    Qubits: List[int] = field(default_factory=list)
    Clbits: List[int] = field(default_factory=list)
    Condition: Optional[Condition] = None
    ControlFlow: Optional[ControlFlow] = None


class Circuit:
    def __init__(self, name: str, num_qubits: int, num_clbits: int, ops: List[Operation], metadata=None):
        self._name = name
        self._nq = num_qubits
        self._nc = num_clbits
        self._ops = list(ops)
        self._meta = dict(metadata or {})

    def Name(self) -> str:
        return self._name

    def NumQubits(self) -> int:
        return self._nq

    def NumClbits(self) -> int:
        return self._nc

    def Ops(self) -> List[Operation]:
        return list(self._ops)

    def NumOps(self) -> int:
        return len(self._ops)

    def Op(self, i: int) -> Operation:
        return self._ops[i]

    def IsDynamic(self) -> bool:
        """circuit.go:298-322 (the op list is fixed at construction, so the answer is kept)"""
        d = getattr(self, "_dynamic", None)
        if d is None:
            d = self._dynamic = self._is_dynamic()
        return d

    def _is_dynamic(self) -> bool:
        last = -1
        for i in range(len(self._ops) - 1, -1, -1):
            g = self._ops[i].Gate
            if g is not None and g.Name() not in ("barrier", "reset"):
                last = i
                break
        for i, op in enumerate(self._ops):
            if op.ControlFlow is not None or op.Condition is not None:
                return True
            if op.Gate is not None and op.Gate.Name() == "reset":
                return True
            if op.Gate is None and op.Clbits and i < last:
                return True
        return False


def _bind_ops(ops: List[Operation], b: Dict[str, float]) -> List[Operation]:
    out = []
    for op in ops:
        g = op.Gate
        if isinstance(g, SymbolicGate):
            g = g.Bind(b)
        cf = op.ControlFlow
        if cf is not None:
            cf = ControlFlow(cf.Type, [_bind_ops(body, b) for body in cf.Bodies], cf.Condition, cf.ForRange,
                             cf.SwitchClbits, cf.SwitchCases)
        out.append(Operation(g, list(op.Qubits), list(op.Clbits), op.Condition, cf))
    return out


def Bind(c: Circuit, bindings: Dict[str, float]) -> Circuit:
    """ir.Bind (circuit.go:206-251)"""
    return Circuit(c.Name(), c.NumQubits(), c.NumClbits(), _bind_ops(c.Ops(), bindings), c._meta)


def FreeParameters(c: Circuit) -> List[str]:
    """ir.FreeParameters (circuit.go:253-281): names in order of first appearance."""
    names: List[str] = []

    def walk(ops):
        for op in ops:
            if isinstance(op.Gate, SymbolicGate):
                for n in op.Gate.FreeParameters():
                    if n not in names:
                        names.append(n)
            if op.ControlFlow is not None:
                for body in op.ControlFlow.Bodies:
                    walk(body)

    walk(c.Ops())
    return names


# ---- builder (circuit/builder/builder.go) ---------------------------------------------------------------
class Builder:
    def __init__(self, name: str, num_qubits: int):
        if num_qubits < 1:
            raise ValueError("builder: numQubits must be >= 1")
        self.name = name
        self.nq = num_qubits
        self.nc = 0
        self.ops: List[Operation] = []

    def WithClbits(self, n: int) -> "Builder":
        self.nc = n
        return self

    def _q(self, *qs):
        for q in qs:
            if not 0 <= q < self.nq:
                raise ValueError(f"builder: qubit {q} out of range [0, {self.nq})")
        if len(set(qs)) != len(qs):
            raise ValueError("builder: repeated qubit")

    def _cl(self, *cs):
        """validateClbit (builder.go): a clbit outside [0, numClbits) is an error."""
        for c in cs:
            if not 0 <= c < self.nc:
                raise ValueError(f"builder: clbit {c} out of range [0, {self.nc})")

    def Apply(self, g: G.Gate, *qubits: int) -> "Builder":
        self._q(*qubits)
        if g.Qubits() != len(qubits):
            raise ValueError(f"builder: gate {g.Name()} needs {g.Qubits()} qubits, got {len(qubits)}")
        self.ops.append(Operation(g, list(qubits)))
        return self

    # fixed gates
    def H(self, q): return self.Apply(G.H, q)
    def X(self, q): return self.Apply(G.X, q)
    def Y(self, q): return self.Apply(G.Y, q)
    def Z(self, q): return self.Apply(G.Z, q)
    def S(self, q): return self.Apply(G.S, q)
    def T(self, q): return self.Apply(G.T, q)
    def CNOT(self, c, t): return self.Apply(G.CNOT, c, t)
    def CZ(self, c, t): return self.Apply(G.CZ, c, t)
    def SWAP(self, a, b): return self.Apply(G.SWAP, a, b)
    def ISWAP(self, a, b): return self.Apply(G.ISWAP, a, b)
    def ECR(self, a, b): return self.Apply(G.ECR, a, b)
    def DCX(self, a, b): return self.Apply(G.DCX, a, b)
    def CH(self, c, t): return self.Apply(G.CH, c, t)
    def CSX(self, c, t): return self.Apply(G.CSX, c, t)
    def Sycamore(self, a, b): return self.Apply(G.Sycamore, a, b)
    def CCX(self, c0, c1, t): return self.Apply(G.CCX, c0, c1, t)
    def CCZ(self, c0, c1, t): return self.Apply(G.CCZ, c0, c1, t)
    def MCX(self, controls, t): return self.Apply(G.MCX(len(controls)), *controls, t)
    def MCZ(self, controls, t): return self.Apply(G.MCZ(len(controls)), *controls, t)
    def MCP(self, phi, controls, t): return self.Apply(G.MCP(phi, len(controls)), *controls, t)

    def Ctrl(self, g: G.Gate, controls, *targets):
        return self.Apply(G.Controlled(g, len(controls)), *controls, *targets)

    # parameterized
    def RX(self, th, q): return self.Apply(G.RX(th), q)
    def RY(self, th, q): return self.Apply(G.RY(th), q)
    def RZ(self, th, q): return self.Apply(G.RZ(th), q)
    def Phase(self, phi, q): return self.Apply(G.Phase(phi), q)
    def U3(self, th, phi, lam, q): return self.Apply(G.U3(th, phi, lam), q)
    def Rot(self, phi, th, om, q): return self.Apply(G.Rot(phi, th, om), q)
    def PhasedXZ(self, x, z, a, q): return self.Apply(G.PhasedXZ(x, z, a), q)
    def GlobalPhase(self, phi, q): return self.Apply(G.GlobalPhase(phi), q)
    def RXX(self, th, a, b): return self.Apply(G.RXX(th), a, b)
    def RYY(self, th, a, b): return self.Apply(G.RYY(th), a, b)
    def RZZ(self, th, a, b): return self.Apply(G.RZZ(th), a, b)
    def ZZ(self, th, a, b): return self.Apply(G.ZZ(th), a, b)
    def FSim(self, th, phi, a, b): return self.Apply(G.FSim(th, phi), a, b)
    def PSwap(self, phi, a, b): return self.Apply(G.PSwap(phi), a, b)
    def CP(self, phi, c, t): return self.Apply(G.CP(phi), c, t)
    def CRX(self, th, c, t): return self.Apply(G.CRX(th), c, t)
    def CRY(self, th, c, t): return self.Apply(G.CRY(th), c, t)
    def CRZ(self, th, c, t): return self.Apply(G.CRZ(th), c, t)

    def Unitary(self, name, matrix, *qubits):
        return self.Apply(G.Unitary(name, matrix, len(qubits)), *qubits)

    def StatePrep(self, amplitudes, *qubits):
        g = G.StatePrep(amplitudes)
        if g.Qubits() != len(qubits):
            raise ValueError("builder: StatePrep amplitude count does not match qubits")
        return self.Apply(g, *qubits)

    # symbolic
    def _sym(self, name, n, exprs, *qubits):
        return self.Apply(SymbolicGate(name, n, exprs), *qubits)

    def SymRX(self, th, q): return self._sym("RX", 1, [th], q)
    def SymRY(self, th, q): return self._sym("RY", 1, [th], q)
    def SymRZ(self, th, q): return self._sym("RZ", 1, [th], q)
    def SymPhase(self, phi, q): return self._sym("P", 1, [phi], q)
    def SymU3(self, th, phi, lam, q): return self._sym("U3", 1, [th, phi, lam], q)
    def SymCP(self, phi, a, b): return self._sym("CP", 2, [phi], a, b)
    def SymRXX(self, th, a, b): return self._sym("RXX", 2, [th], a, b)
    def SymRYY(self, th, a, b): return self._sym("RYY", 2, [th], a, b)
    def SymRZZ(self, th, a, b): return self._sym("RZZ", 2, [th], a, b)
    def SymRot(self, phi, th, om, q): return self._sym("Rot", 1, [phi, th, om], q)
    def SymFSim(self, th, phi, a, b): return self._sym("FSim", 2, [th, phi], a, b)
    def SymPSwap(self, phi, a, b): return self._sym("PSwap", 2, [phi], a, b)
    def SymGlobalPhase(self, phi, q): return self._sym("GlobalPhase", 1, [phi], q)

    # non-unitary / classical
    def Measure(self, q, c):
        self._q(q)
        if not 0 <= c < self.nc:
            raise ValueError(f"builder: clbit {c} out of range")
        self.ops.append(Operation(None, [q], [c]))
        return self

    def MeasureAll(self):
        """builder.go:381-395"""
        if self.nc < self.nq:
            self.nc = self.nq
        for i in range(self.nq):
            self.ops.append(Operation(None, [i], [i]))
        return self

    def Reset(self, q):
        return self.Apply(G.Reset, q)

    def Delay(self, q, duration, unit="s"):
        return self.Apply(G.Delay(duration, unit), q)

    def Barrier(self, *qubits):
        qs = list(qubits) if qubits else list(range(self.nq))
        self._q(*qs)
        self.ops.append(Operation(G.Barrier(len(qs)), qs))
        return self

    def If(self, clbit, value, g: G.Gate, *qubits):
        self._q(*qubits)
        self._cl(clbit)
        self.ops.append(Operation(g, list(qubits), [], Condition(clbit, value)))
        return self

    def _sub(self) -> "Builder":
        b = Builder(self.name, self.nq)
        b.nc = self.nc
        return b

    def IfElseBlock(self, clbit, value, if_body, else_body=None):
        self._cl(clbit)
        bi = self._sub()
        if_body(bi)
        bodies = [bi.ops]
        if else_body is not None:
            be = self._sub()
            else_body(be)
            bodies.append(be.ops)
        self.ops.append(Operation(None, [], [], None, ControlFlow("if_else", bodies, Condition(clbit, value))))
        return self

    def IfBlock(self, clbit, value, fn):
        return self.IfElseBlock(clbit, value, fn, None)

    def While(self, clbit, value, body):
        self._cl(clbit)
        b = self._sub()
        body(b)
        self.ops.append(Operation(None, [], [], None, ControlFlow("while", [b.ops], Condition(clbit, value))))
        return self

    def For(self, clbit, start, end, step, body):
        """builder.For (builder.go:537-567): step 0 is an error."""
        self._cl(clbit)
        if step == 0:
            raise ValueError("for loop step must be non-zero")
        b = self._sub()
        body(b)
        self.ops.append(Operation(None, [], [], None, ControlFlow("for", [b.ops], None, (start, end, step))))
        return self

    def Switch(self, clbits, cases: Dict[int, object], default_body=None):
        """builder.Switch (builder.go:570-626): case bodies in ascending order of their values, then
        the default body if given; the switch value reads the clbits little-endian."""
        self._cl(*clbits)
        vals = sorted(cases)
        bodies = []
        for v in vals:
            b = self._sub()
            cases[v](b)
            bodies.append(b.ops)
        if default_body is not None:
            b = self._sub()
            default_body(b)
            bodies.append(b.ops)
        self.ops.append(Operation(None, [], [], None, ControlFlow("switch", bodies, None, None, list(clbits), vals)))
        return self

    def Compose(self, c: Circuit, qubit_map: Optional[Dict[int, int]] = None):
        for op in c.Ops():
            qs = [qubit_map.get(q, q) if qubit_map else q for q in op.Qubits]
            self.ops.append(Operation(op.Gate, qs, list(op.Clbits), op.Condition, op.ControlFlow))
        return self

    def Build(self) -> Circuit:
        return Circuit(self.name, self.nq, self.nc, self.ops)


def NewBuilder(name: str, num_qubits: int) -> Builder:
    """builder.New"""
    return Builder(name, num_qubits)
