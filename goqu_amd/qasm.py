"""An OpenQASM 3 reader for circuits: a restatement of goqu's qasm/parser (qasm/parser/lexer.go,
parser.go) over this package's circuit IR, so that QASM files reach the device simulators.

Covered, as in the reference parser: `OPENQASM`, `include` (stdgates.inc is built in), `qubit` /
`bit` / `qreg` / `creg` declarations, user `gate` definitions (inlined at each call), gate calls
with the ctrl / negctrl / inv / pow modifiers, `measure` in its three forms, `reset`, `barrier`,
`delay`, `if` / `else`, `while`, `for`, `switch`, and constant expressions (pi, tau, euler, + - *
/ % **, sin cos tan sqrt exp log acos asin atan).  The reference's lowering is kept, quirks
included: a condition on a whole register tests its first bit (parser.go:624-673); an `if` without
`else` conditions every op of its body (parser.go:718-745); a measurement without a target bit
writes the bit of the same index (parser.go:458-476); a negative control is an X on each side.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import gate as G
from .circuit import Circuit, Condition, ControlFlow, Operation

KEYWORDS = {
    "OPENQASM", "include", "gate", "qubit", "bit", "qreg", "creg", "measure", "reset", "barrier", "delay",
    "if", "else", "for", "while", "switch", "case", "default", "in", "return", "def", "const", "input",
    "output", "ctrl", "negctrl", "inv", "pow", "gphase", "U", "pi", "euler", "tau", "int", "uint",
}
# multi-character operators first
OPERATORS = ["->", "**", "==", "!=", "<=", ">=", "<<", ">>", "&&", "||",
             "(", ")", "[", "]", "{", "}", ";", ",", ":", ".", "@", "~", "+", "%", "^", "-", "*", "/",
             "=", "!", "<", ">", "&", "|"]


@dataclass
class Tok:
    kind: str  # "IDENT", "INT", "FLOAT", "STRING", "EOF", an operator, or a keyword
    text: str
    line: int
    col: int


class QASMError(ValueError):
    pass


def tokenize(src: str) -> List[Tok]:
    """lexer.tokenize (lexer.go:23-179)."""
    out: List[Tok] = []
    i, line, col = 0, 1, 1
    n = len(src)

    def adv(k=1):
        nonlocal i, line, col
        for _ in range(k):
            if src[i] == "\n":
                line += 1
                col = 1
            else:
                col += 1
            i += 1

    while True:
        while i < n and src[i] in " \t\r\n":
            adv()
        if i >= n:
            out.append(Tok("EOF", "", line, col))
            return out
        if src.startswith("//", i):
            while i < n and src[i] != "\n":
                adv()
            continue
        if src.startswith("/*", i):
            adv(2)
            while i < n and not src.startswith("*/", i):
                adv()
            if i < n:
                adv(2)
            continue
        l0, c0 = line, col
        ch = src[i]
        if ch == '"':
            adv()
            j = i
            while i < n and src[i] != '"':
                adv()
            out.append(Tok("STRING", src[j:i], l0, c0))
            if i < n:
                adv()
            continue
        if ch.isdigit() or (ch == "." and i + 1 < n and src[i + 1].isdigit()):
            j = i
            isf = False
            while i < n and (src[i].isdigit() or src[i] == "."):
                isf = isf or src[i] == "."
                adv()
            if i < n and src[i] in "eE" and (i + 1 < n and (src[i + 1].isdigit() or src[i + 1] in "+-")):
                isf = True
                adv()
                if src[i] in "+-":
                    adv()
                while i < n and src[i].isdigit():
                    adv()
            out.append(Tok("FLOAT" if isf else "INT", src[j:i], l0, c0))
            continue
        if ch.isalpha() or ch == "_" or ord(ch) >= 0x80:
            j = i
            while i < n and (src[i].isalnum() or src[i] == "_" or ord(src[i]) >= 0x80):
                adv()
            word = src[j:i]
            out.append(Tok(word if word in KEYWORDS else "IDENT", word, l0, c0))
            continue
        for op in OPERATORS:
            if src.startswith(op, i):
                adv(len(op))
                out.append(Tok(op, op, l0, c0))
                break
        else:
            raise QASMError(f"line {l0}:{c0}: illegal character {ch!r}")


@dataclass
class _Reg:
    start: int
    size: int


@dataclass
class _GateDef:
    params: List[str]
    qubits: List[str]
    body: List[Tok]


def _pow_gate(g: G.Gate, k: int) -> G.Gate:
    """gate.Pow: the matrix to an integer power (a negative power inverts)."""
    m = np.asarray(g.Matrix(), dtype=np.complex128).reshape(1 << g.Qubits(), 1 << g.Qubits())
    p = np.linalg.matrix_power(m if k >= 0 else m.conj().T, abs(k))
    return G.Unitary(f"pow({k})_{g.Name()}", p.reshape(-1), g.Qubits())


@dataclass
class _Parser:
    toks: List[Tok]
    gates: Dict[str, _GateDef] = field(default_factory=dict)
    qregs: Dict[str, _Reg] = field(default_factory=dict)
    cregs: Dict[str, _Reg] = field(default_factory=dict)
    nq: int = 0
    nc: int = 0
    ops: List[Operation] = field(default_factory=list)
    params: Optional[Dict[str, float]] = None
    pos: int = 0
    meta: Dict[str, str] = field(default_factory=dict)

    # -- token helpers --
    def peek(self) -> str:
        return self.toks[self.pos].kind

    def cur(self) -> Tok:
        return self.toks[self.pos]

    def advance(self) -> Tok:
        t = self.toks[self.pos]
        if t.kind != "EOF":
            self.pos += 1
        return t

    def expect(self, kind: str) -> Tok:
        t = self.cur()
        if t.kind != kind:
            raise QASMError(f"line {t.line}:{t.col}: expected {kind}, got {t.kind} ({t.text!r})")
        self.pos += 1
        return t

    # -- program --
    def program(self) -> None:
        if self.peek() == "OPENQASM":
            self.advance()
            t = self.advance()
            if t.kind not in ("INT", "FLOAT"):
                raise QASMError(f"line {t.line}:{t.col}: expected version number")
            self.meta["openqasm_version"] = t.text
            self.expect(";")
        while self.peek() != "EOF":
            self.statement()

    def statement(self) -> None:
        k = self.peek()
        handler = {
            "include": self.include, "gate": self.gate_decl, "qubit": lambda: self.decl(True, False),
            "bit": lambda: self.decl(False, False), "qreg": lambda: self.decl(True, True),
            "creg": lambda: self.decl(False, True), "measure": self.measure, "reset": self.reset,
            "barrier": self.barrier, "delay": self.delay, "if": self.if_stmt, "while": self.while_stmt,
            "for": self.for_stmt, "switch": self.switch_stmt,
        }.get(k)
        if handler is not None:
            handler()
        elif k == "IDENT":
            if self.is_assignment():
                self.assignment()
            else:
                self.gate_call()
        elif k in ("U", "ctrl", "negctrl", "inv", "pow", "gphase"):
            self.gate_call()
        else:
            t = self.cur()
            raise QASMError(f"line {t.line}:{t.col}: unexpected token {t.kind} ({t.text!r})")

    def include(self) -> None:
        self.advance()
        self.expect("STRING")
        self.expect(";")

    def designator(self) -> int:
        self.expect("[")
        v = self.expr()
        self.expect("]")
        if int(v) != v or v <= 0:
            raise QASMError(f"designator must be a positive integer, got {v}")
        return int(v)

    def decl(self, quantum: bool, legacy: bool) -> None:
        self.advance()
        if legacy:
            name = self.expect("IDENT").text
            size = self.designator()
        else:
            size = self.designator() if self.peek() == "[" else 1
            name = self.expect("IDENT").text
        self.expect(";")
        if quantum:
            self.qregs[name] = _Reg(self.nq, size)
            self.nq += size
        else:
            self.cregs[name] = _Reg(self.nc, size)
            self.nc += size

    def gate_decl(self) -> None:
        self.advance()
        name = self.expect("IDENT").text
        params: List[str] = []
        if self.peek() == "(":
            self.advance()
            while self.peek() not in (")", "EOF"):
                params.append(self.expect("IDENT").text)
                if self.peek() == ",":
                    self.advance()
            self.expect(")")
        qubits: List[str] = []
        while self.peek() not in ("{", "EOF"):
            qubits.append(self.expect("IDENT").text)
            if self.peek() == ",":
                self.advance()
        self.expect("{")
        depth, start = 1, self.pos
        while depth > 0 and self.peek() != "EOF":
            if self.peek() == "{":
                depth += 1
            elif self.peek() == "}":
                depth -= 1
            if depth > 0:
                self.advance()
        body = self.toks[start:self.pos]
        self.expect("}")
        self.gates[name] = _GateDef(params, qubits, body)

    def expand(self, name: str, gd: _GateDef, params: List[float], qubits: List[int]) -> None:
        """expandGateCall (parser.go:383-434): the body re-parsed with the arguments bound."""
        if len(params) != len(gd.params):
            raise QASMError(f"gate {name} requires {len(gd.params)} parameters, got {len(params)}")
        if len(qubits) != len(gd.qubits):
            raise QASMError(f"gate {name} requires {len(gd.qubits)} qubit arguments, got {len(qubits)}")
        if not gd.body:
            return
        sub = _Parser(gd.body + [Tok("EOF", "", 0, 0)], gates=self.gates, cregs=self.cregs, nq=self.nq, nc=self.nc)
        sub.qregs = {qn: _Reg(q, 1) for qn, q in zip(gd.qubits, qubits)}
        if gd.params:
            sub.params = dict(zip(gd.params, params))
        while sub.peek() != "EOF":
            sub.statement()
        self.ops.extend(sub.ops)

    # -- measurement, reset, barrier, delay --
    def measure(self) -> None:
        t = self.advance()
        qubits = self.qubit_args()
        if self.peek() == "->":
            self.advance()
            clbits = self.clbit_args()
            self.expect(";")
            if len(qubits) != len(clbits):
                raise QASMError(f"line {t.line}:{t.col}: measure qubit/clbit count mismatch")
            self.ops.extend(Operation(None, [q], [c]) for q, c in zip(qubits, clbits))
            return
        self.expect(";")
        for q in qubits:  # the bit of the same index (parser.go:466-475)
            if q >= self.nc:
                raise QASMError(f"line {t.line}:{t.col}: implicit measurement on q[{q}] requires c[{q}] but only "
                                f"{self.nc} classical bits declared")
            self.ops.append(Operation(None, [q], [q]))

    def is_assignment(self) -> bool:
        i = self.pos + 1
        if self.toks[i].kind == "[":
            depth, i = 1, i + 1
            while i < len(self.toks) and depth > 0:
                depth += {"[": 1, "]": -1}.get(self.toks[i].kind, 0)
                i += 1
        return i < len(self.toks) and self.toks[i].kind == "="

    def assignment(self) -> None:
        clbits = self.clbit_args()
        self.expect("=")
        self.expect("measure")
        qubits = self.qubit_args()
        self.expect(";")
        if len(qubits) != len(clbits):
            raise QASMError("measure qubit/clbit count mismatch")
        self.ops.extend(Operation(None, [q], [c]) for q, c in zip(qubits, clbits))

    def reset(self) -> None:
        self.advance()
        qubits = self.qubit_args()
        self.expect(";")
        self.ops.extend(Operation(G.Reset, [q]) for q in qubits)

    def barrier(self) -> None:
        self.advance()
        qubits = self.qubit_args() if self.peek() != ";" else list(range(self.nq))
        self.expect(";")
        self.ops.append(Operation(G.Barrier(len(qubits)), qubits))

    def delay(self) -> None:
        self.advance()
        self.expect("[")
        dur = self.expr()
        unit = "dt"
        if self.peek() == "IDENT":
            t = self.advance()
            if t.text not in ("ns", "us", "ms", "s", "dt"):
                raise QASMError(f"line {t.line}:{t.col}: unknown delay unit {t.text!r}")
            unit = t.text
        self.expect("]")
        qubits = self.qubit_args()
        self.expect(";")
        self.ops.extend(Operation(G.Delay(dur, unit), [q]) for q in qubits)

    # -- control flow --
    def condition(self) -> Condition:
        """parseCondition (parser.go:624-673): a whole register tests its first bit."""
        self.expect("(")
        reg = self.expect("IDENT").text
        idx = -1
        if self.peek() == "[":
            self.advance()
            idx = int(self.expr())
            self.expect("]")
        self.expect("==")
        val = self.expr()
        self.expect(")")
        clbit = 0
        if reg in self.cregs:
            clbit = self.cregs[reg].start + (idx if idx >= 0 else 0)
        return Condition(clbit, int(val))

    def block(self) -> List[Operation]:
        self.expect("{")
        before = len(self.ops)
        while self.peek() not in ("}", "EOF"):
            self.statement()
        self.expect("}")
        body = self.ops[before:]
        del self.ops[before:]
        return body

    def if_stmt(self) -> None:
        self.advance()
        cond = self.condition()
        if self.peek() == "{":
            body = self.block()
            if self.peek() == "else":
                self.advance()
                other = self.block()
                self.ops.append(Operation(None, [], [], None, ControlFlow("if_else", [body, other], cond)))
                return
            for op in body:  # no else: every op of the body carries the condition
                op.Condition = cond
            self.ops.extend(body)
            return
        before = len(self.ops)
        self.statement()
        for op in self.ops[before:]:
            op.Condition = cond

    def while_stmt(self) -> None:
        self.advance()
        cond = self.condition()
        body = self.block()
        self.ops.append(Operation(None, [], [], None, ControlFlow("while", [body], cond)))

    def for_stmt(self) -> None:
        self.advance()
        if self.peek() in ("int", "uint"):
            self.advance()
            if self.peek() == "[":
                self.advance()
                self.expr()
                self.expect("]")
        self.expect("IDENT")
        self.expect("in")
        self.expect("[")
        start = self.expr()
        self.expect(":")
        end = self.expr()
        step = 1.0
        if self.peek() == ":":
            self.advance()
            step = self.expr()
        self.expect("]")
        body = self.block()
        self.ops.append(Operation(None, [], [], None,
                                  ControlFlow("for", [body], None, (int(start), int(end), int(step)))))

    def switch_stmt(self) -> None:
        self.advance()
        self.expect("(")
        reg = self.expect("IDENT").text
        clbits: List[int] = []
        if self.peek() == "[":
            self.advance()
            v = int(self.expr())
            self.expect("]")
            if reg in self.cregs:
                clbits = [self.cregs[reg].start + v]
        elif reg in self.cregs:
            r = self.cregs[reg]
            clbits = list(range(r.start, r.start + r.size))
        self.expect(")")
        self.expect("{")
        cases: List[int] = []
        bodies: List[List[Operation]] = []
        while self.peek() not in ("}", "EOF"):
            if self.peek() == "case":
                self.advance()
                v = self.expr()
                if self.peek() == ":":
                    self.advance()
                bodies.append(self.block())
                cases.append(int(v))
            elif self.peek() == "default":
                self.advance()
                if self.peek() == ":":
                    self.advance()
                bodies.append(self.block())
            else:
                raise QASMError(f"line {self.cur().line}: expected 'case' or 'default' in switch")
        self.expect("}")
        self.ops.append(Operation(None, [], [], None, ControlFlow("switch", bodies, None, None, clbits, cases)))

    # -- gate calls --
    def gate_call(self) -> None:
        """parseGateCall (parser.go:942-1102)."""
        ctrl = negctrl = inv = 0
        powk: Optional[int] = None
        neg_pos: List[int] = []
        total = 0
        while self.peek() in ("ctrl", "negctrl", "inv", "pow"):
            m = self.advance()
            if m.kind in ("ctrl", "negctrl"):
                k = 1
                if self.peek() == "(":
                    self.advance()
                    k = int(self.expr())
                    self.expect(")")
                if m.kind == "negctrl":
                    neg_pos.extend(range(total, total + k))
                    negctrl += k
                else:
                    ctrl += k
                total += k
            elif m.kind == "inv":
                inv += 1
            else:
                if self.peek() != "(":
                    raise QASMError(f"line {m.line}: pow modifier requires parenthesized exponent")
                self.advance()
                v = self.expr()
                self.expect(")")
                k = int(round(v))
                if abs(v - k) > 1e-10:
                    raise QASMError(f"line {m.line}: pow modifier requires integer exponent, got {v}")
                powk = k
            self.expect("@")
        name = self.advance().text
        params: List[float] = []
        if self.peek() == "(":
            self.advance()
            while self.peek() not in (")", "EOF"):
                params.append(self.expr())
                if self.peek() == ",":
                    self.advance()
            self.expect(")")
        qubits = self.qubit_args()
        self.expect(";")
        if total == 0 and inv == 0 and powk is None and name in self.gates:
            self.expand(name, self.gates[name], params, qubits)
            return
        g = resolve_gate(name, params)
        if powk is not None:
            g = _pow_gate(g, powk)
        if total:
            g = G.Controlled(g, total)
        for _ in range(inv):
            g = g.Inverse()
        for p in neg_pos:
            self.ops.append(Operation(G.X, [qubits[p]]))
        self.ops.append(Operation(g, qubits))
        for p in neg_pos:
            self.ops.append(Operation(G.X, [qubits[p]]))

    # -- arguments --
    def qubit_args(self) -> List[int]:
        out: List[int] = []
        while True:
            out.extend(self.single_arg(self.qregs, "qubit"))
            if self.peek() != ",":
                return out
            self.advance()

    def clbit_args(self) -> List[int]:
        out: List[int] = []
        while True:
            out.extend(self.single_arg(self.cregs, "classical bit"))
            if self.peek() != ",":
                return out
            self.advance()

    def single_arg(self, regs: Dict[str, _Reg], what: str) -> List[int]:
        t = self.expect("IDENT")
        if t.text not in regs:
            kind = "qubit" if what == "qubit" else "classical"
            raise QASMError(f"line {t.line}:{t.col}: undefined {kind} register {t.text!r}")
        r = regs[t.text]
        if self.peek() == "[":
            self.advance()
            idx = int(self.expr())
            self.expect("]")
            if not 0 <= idx < r.size:
                raise QASMError(f"line {t.line}:{t.col}: {what} index {idx} out of range [0, {r.size})")
            return [r.start + idx]
        return list(range(r.start, r.start + r.size))

    # -- constant expressions (parser.go:1346-1520) --
    def expr(self) -> float:
        left = self.mul()
        while self.peek() in ("+", "-"):
            op = self.advance().kind
            right = self.mul()
            left = left + right if op == "+" else left - right
        return left

    def mul(self) -> float:
        left = self.power()
        while self.peek() in ("*", "/", "%"):
            op = self.advance().kind
            right = self.power()
            if op == "*":
                left *= right
            elif op == "/":
                if right == 0:
                    raise QASMError("division by zero")
                left /= right
            else:
                left = math.fmod(left, right)
        return left

    def power(self) -> float:
        base = self.unary()
        if self.peek() == "**":
            self.advance()
            return math.pow(base, self.unary())
        return base

    def unary(self) -> float:
        if self.peek() == "-":
            self.advance()
            return -self.unary()
        if self.peek() == "+":
            self.advance()
            return self.unary()
        return self.primary()

    def primary(self) -> float:
        t = self.cur()
        if t.kind in ("INT", "FLOAT"):
            self.advance()
            return float(t.text)
        if t.kind == "pi":
            self.advance()
            return math.pi
        if t.kind == "euler":
            self.advance()
            return math.e
        if t.kind == "tau":
            self.advance()
            return 2 * math.pi
        if t.kind == "(":
            self.advance()
            v = self.expr()
            self.expect(")")
            return v
        if t.kind == "IDENT":
            if self.params is not None and t.text in self.params:
                self.advance()
                return self.params[t.text]
            self.advance()
            self.expect("(")
            arg = self.expr()
            self.expect(")")
            fns = {"sin": math.sin, "cos": math.cos, "tan": math.tan, "sqrt": math.sqrt, "exp": math.exp,
                   "log": math.log, "ln": math.log, "acos": math.acos, "arccos": math.acos, "asin": math.asin,
                   "arcsin": math.asin, "atan": math.atan, "arctan": math.atan}
            if t.text not in fns:
                raise QASMError(f"line {t.line}:{t.col}: unknown function {t.text!r}")
            return fns[t.text](arg)
        raise QASMError(f"line {t.line}:{t.col}: unexpected token in expression: {t.kind} ({t.text!r})")


_FIXED = {
    "id": G.I, "I": G.I, "h": G.H, "H": G.H, "x": G.X, "X": G.X, "y": G.Y, "Y": G.Y, "z": G.Z, "Z": G.Z,
    "s": G.S, "S": G.S, "sdg": G.Sdg, "t": G.T, "T": G.T, "tdg": G.Tdg, "sx": G.SX,
    "cx": G.CNOT, "CX": G.CNOT, "cnot": G.CNOT, "CNOT": G.CNOT, "cz": G.CZ, "cy": G.CY, "swap": G.SWAP,
    "ccx": G.CCX, "cswap": G.CSWAP, "iswap": G.ISWAP, "ecr": G.ECR, "dcx": G.DCX, "ch": G.CH, "csx": G.CSX,
    "ccz": G.CCZ, "sycamore": G.Sycamore,
}
_PARAM = {
    "rx": (1, G.RX), "ry": (1, G.RY), "rz": (1, G.RZ), "p": (1, G.Phase), "phase": (1, G.Phase),
    "U": (3, G.U3), "u3": (3, G.U3), "u1": (1, G.U1), "u2": (2, G.U2), "cp": (1, G.CP), "cphase": (1, G.CP),
    "crx": (1, G.CRX), "cry": (1, G.CRY), "crz": (1, G.CRZ), "rxx": (1, G.RXX), "ryy": (1, G.RYY),
    "rzz": (1, G.RZZ), "gphase": (1, G.GlobalPhase),
}


def resolve_gate(name: str, params: List[float]) -> G.Gate:
    """resolveGate (parser.go:1104-1240)."""
    if name in _FIXED:
        return _FIXED[name]
    if name in _PARAM:
        k, ctor = _PARAM[name]
        if len(params) != k:
            raise QASMError(f"{name} requires {k} parameter{'s' if k > 1 else ''}, got {len(params)}")
        return ctor(*params)
    raise QASMError(f"unknown gate: {name}")


def ParseString(src: str, name: str = "") -> Circuit:
    """parser.ParseString: the program as a circuit (qubits and clbits in declaration order)."""
    p = _Parser(tokenize(src))
    p.program()
    return Circuit(name, p.nq, p.nc, p.ops, p.meta)


def ParseFile(path: str) -> Circuit:
    with open(path) as f:
        return ParseString(f.read())
