"""Gradients — host-side mirror of goqu's algorithm/gradient (gradient.go:25-86) plus the
adjoint method on the device.

CostFunc, ParameterShift and FiniteDifference follow the reference exactly (they are
combinators over an objective).  Adjoint computes <H> and every partial derivative with one
forward and one reverse sweep on the GPU (goqu_sv_adjoint / goqu_dm_adjoint): each
parameterised op contributes scale * 2 Re <lambda_k| dU_k |psi_{k-1}>, where dU_k is the
derivative of the gate matrix with respect to its argument and scale the derivative of the
argument expression with respect to the parameter (chain rule for reused or scaled
parameters, e.g. QAOA's 2*gamma*coeff).
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Callable, Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _native as N
from . import gate as G
from .circuit import Bind, Circuit, FreeParameters, SymbolicGate, CONSTRUCTORS
from .pauli import PauliSum, masks

ObjectiveFunc = Callable[[Sequence[float]], float]
GradientFunc = Callable[[Sequence[float]], np.ndarray]

_PX = np.array([[0, 1], [1, 0]], dtype=complex)
_PY = np.array([[0, -1j], [1j, 0]], dtype=complex)
_PZ = np.diag([1, -1]).astype(complex)
_GEN = {"RX": _PX, "RY": _PY, "RZ": _PZ, "RXX": np.kron(_PX, _PX), "RYY": np.kron(_PY, _PY),
        "RZZ": np.kron(_PZ, _PZ), "ZZ": np.kron(_PZ, _PZ)}
_CGEN = {"CRX": _PX, "CRY": _PY, "CRZ": _PZ}


def _mat(name: str, args: Sequence[float]) -> np.ndarray:
    g = CONSTRUCTORS[name](*args)
    d = 1 << g.Qubits()
    return g.Matrix().reshape(d, d)


def gate_derivative(name: str, args: Sequence[float], j: int) -> np.ndarray:
    """d(matrix)/d(args[j]) for the parameterised gate constructors."""
    M = _mat(name, args)
    if name in _GEN:  # exp(-i theta/2 G): dM = -i/2 G M
        return -0.5j * _GEN[name] @ M
    if name in _CGEN:
        R = M[2:, 2:]
        out = np.zeros((4, 4), dtype=complex)
        out[2:, 2:] = -0.5j * _CGEN[name] @ R
        return out
    if name in ("P", "U1", "CP"):
        out = np.zeros_like(M)
        out[-1, -1] = 1j * M[-1, -1]
        return out
    if name == "GlobalPhase":
        return 1j * M
    if name == "U3":
        th, ph, la = args
        c, s = math.cos(th / 2), math.sin(th / 2)
        e = lambda a: complex(math.cos(a), math.sin(a))  # noqa: E731
        if j == 0:
            return np.array([[-s / 2, -e(la) * c / 2], [e(ph) * c / 2, -e(ph + la) * s / 2]])
        if j == 1:
            return np.array([[0, 0], [1j * e(ph) * s, 1j * e(ph + la) * c]])
        return np.array([[0, -1j * e(la) * s], [0, 1j * e(ph + la) * c]])
    return _analytic_derivative(name, args, j)


def _e(a: float) -> complex:
    return complex(math.cos(a), math.sin(a))


def _analytic_derivative(name: str, args: Sequence[float], j: int) -> np.ndarray:
    """Entry-wise derivatives of the remaining constructors' matrices (gate.py, following
    circuit/gate/parameterized.go:91-564): every entry is a product of cos / sin of a half angle
    and a phase e^{i(k . args)}, differentiated in closed form."""
    s2 = 1 / math.sqrt(2)
    if name == "GPI":  # [[0, e(-p)], [e(p), 0]]
        (p,) = args
        return np.array([[0, -1j * _e(-p)], [1j * _e(p), 0]])
    if name == "GPI2":  # s2 [[1, -i e(-p)], [-i e(p), 1]]
        (p,) = args
        return np.array([[0, -s2 * _e(-p)], [s2 * _e(p), 0]])
    if name == "MS":  # off-diagonal -i s2 e(+-(p0 + p1)), -i s2 e(+-(p0 - p1))
        p0, p1 = args
        sg = 1.0 if j == 0 else -1.0  # d(p0 - p1)/d(arg j)
        out = np.zeros((4, 4), dtype=complex)
        out[0, 3] = -s2 * _e(-(p0 + p1))
        out[3, 0] = s2 * _e(p0 + p1)
        out[1, 2] = sg * s2 * _e(p0 - p1)
        out[2, 1] = -sg * s2 * _e(-(p0 - p1))
        return out
    if name == "U2":  # s2 [[1, -e(l)], [e(p), e(p + l)]]
        p, lam = args
        if j == 0:
            return np.array([[0, 0], [1j * s2 * _e(p), 1j * s2 * _e(p + lam)]])
        return np.array([[0, -1j * s2 * _e(lam)], [0, 1j * s2 * _e(p + lam)]])
    if name == "Rot":  # [[e(-(p+w)/2) c, -e(-(p-w)/2) s], [e((p-w)/2) s, e((p+w)/2) c]]
        p, th, w = args
        c, s = math.cos(th / 2), math.sin(th / 2)
        if j == 1:
            return np.array([[-_e(-(p + w) / 2) * s / 2, -_e(-(p - w) / 2) * c / 2],
                             [_e((p - w) / 2) * c / 2, -_e((p + w) / 2) * s / 2]])
        M = _mat(name, args)
        k = np.array([[-0.5, -0.5], [0.5, 0.5]]) if j == 0 else np.array([[-0.5, 0.5], [-0.5, 0.5]])
        return 1j * k * M  # each entry's phase exponent is k * arg
    if name == "PhasedXZ":  # [[cx, i sx e(-pi a)], [i sx e(pi (z + a)), cx e(pi z)]]
        x, z, a = args
        cx, sx = math.cos(math.pi * x / 2), math.sin(math.pi * x / 2)
        if j == 0:
            h = math.pi / 2
            return np.array([[-h * sx, 1j * h * cx * _e(-math.pi * a)],
                             [1j * h * cx * _e(math.pi * (z + a)), -h * sx * _e(math.pi * z)]])
        M = _mat(name, args)
        k = np.array([[0, 0], [1, 1]]) if j == 1 else np.array([[0, -1], [1, 0]])
        return 1j * math.pi * k * M
    if name == "FSim":  # [[1,0,0,0],[0,ct,-i st,0],[0,-i st,ct,0],[0,0,0,e(-phi)]]
        th, phi = args
        out = np.zeros((4, 4), dtype=complex)
        if j == 0:
            ct, st = math.cos(th), math.sin(th)
            out[1, 1] = out[2, 2] = -st
            out[1, 2] = out[2, 1] = -1j * ct
        else:
            out[3, 3] = -1j * _e(-phi)
        return out
    if name == "PSwap":  # e(phi) on |01><10| and |10><01|
        (phi,) = args
        out = np.zeros((4, 4), dtype=complex)
        out[1, 2] = out[2, 1] = 1j * _e(phi)
        return out
    raise ValueError(f"no derivative for parameterised gate {name}")


# ---- reference combinators (algorithm/gradient/gradient.go) ------------------------------------
def CostFunc(circuit: Circuit, hamiltonian: PauliSum, param_names: Sequence[str], sim=None,
             precision: int = N.C128) -> ObjectiveFunc:
    """gradient.go:25-43: x -> Bind -> Evolve -> Re<H>.  A reusable simulator may be passed;
    otherwise each call creates one, as the reference does."""
    from . import statevector

    names = list(param_names)

    def f(x):
        bound = Bind(circuit, {n: float(v) for n, v in zip(names, x)})
        s = sim if sim is not None else statevector.New(bound.NumQubits(), precision)
        try:
            s.Evolve(bound)
            return s.ExpectPauliSum(hamiltonian)
        finally:
            if sim is None:
                s.Close()

    return f


def ParameterShift(f: ObjectiveFunc) -> GradientFunc:
    """gradient.go:49-66: g_i = (f(x + pi/2 e_i) - f(x - pi/2 e_i)) / 2."""

    def g(x):
        x = np.asarray(x, dtype=float)
        out = np.zeros(len(x))
        for i in range(len(x)):
            xp, xm = x.copy(), x.copy()
            xp[i] += 0.5 * math.pi
            xm[i] -= 0.5 * math.pi
            out[i] = (f(xp) - f(xm)) / 2
        return out

    return g


def FiniteDifference(f: ObjectiveFunc, h: float) -> GradientFunc:
    """gradient.go:70-86: central difference."""

    def g(x):
        x = np.asarray(x, dtype=float)
        out = np.zeros(len(x))
        for i in range(len(x)):
            xp, xm = x.copy(), x.copy()
            xp[i] += h
            xm[i] -= h
            out[i] = (f(xp) - f(xm)) / (2 * h)
        return out

    return g


# ---- adjoint -------------------------------------------------------------------------------------
class AdjointProgram:
    """Marshals a parameterised circuit for goqu_sv_adjoint / goqu_dm_adjoint."""

    def __init__(self, circuit: Circuit, param_names: Sequence[str], density: bool = False, noise=None):
        self.circuit = circuit
        self.names = list(param_names)
        self.index = {n: i for i, n in enumerate(self.names)}
        self.density = density
        self.noise = noise
        self._fast = None

    def build(self, x: Sequence[float]):
        """(op buffer, derivative records, count, kept arrays) at parameters x.  Statevector
        circuits whose parameterised gates are all RX / RY / RZ of one expression reuse the records
        of their first build and refill only the angle-dependent matrices, vectorised (the arrays
        returned are overwritten by the next call)."""
        if not self.density:
            if self._fast is None:
                self._fast = _FastRotations.make(self, x) or False
            if self._fast:
                return self._fast.refill(x)
        return self._build_generic(x)

    def _build_generic(self, x: Sequence[float]):
        b = {n: float(v) for n, v in zip(self.names, x)}
        buf = N.OpBuffer()
        derivs: List[Tuple[int, int, float, np.ndarray]] = []
        for op in self.circuit.Ops():
            g = op.Gate
            if g is None or op.ControlFlow is not None:
                continue
            sym = g if isinstance(g, SymbolicGate) else None
            bound = sym.Bind(b) if sym is not None else g
            if self.density:
                from .densitymatrix import _marshal_op
                from .circuit import Operation

                start = len(buf)
                _marshal_op(buf, Operation(bound, op.Qubits), self.noise)
                k = start  # the gate record precedes any noise channel
            else:
                k = N.add_sv_gate(buf, bound, op.Qubits)
            if sym is None:
                continue
            args = [e.eval(b) for e in sym.exprs]
            for j, e in enumerate(sym.exprs):
                free = e.free()
                if not free:
                    continue
                dm = np.ascontiguousarray(gate_derivative(sym.ctor_name, args, j), dtype=np.complex128)
                for p in dict.fromkeys(free):
                    if p not in self.index:
                        continue
                    derivs.append((k, self.index[p], e.deriv(p, b), dm))
        arr = (N.GoquDeriv * max(1, len(derivs)))()
        keep = []
        for i, (k, p, sc, dm) in enumerate(derivs):
            flat = dm.reshape(-1).view(np.float64)
            keep.append(flat)
            arr[i].op = k
            arr[i].param = p
            arr[i].scale = sc
            arr[i].dmatrix = N.dptr(flat)
        return buf, arr, len(derivs), keep


class _FrozenOps:
    """Op records built once, with their ctypes array cached (the matrices they point at are
    refilled in place)."""

    def __init__(self, buf: "N.OpBuffer"):
        self.buf = buf
        self._arr = buf.array()

    def __len__(self):
        return len(self.buf)

    def array(self):
        return self._arr


class _FastRotations:
    """The statevector adjoint program of a circuit whose parameterised gates are RX / RY / RZ of
    one expression each: records, matrix storage and derivative records are laid out once; each
    build evaluates the expressions and writes every rotation matrix and its derivative
    -i/2 G M (G = X, Y, Z) with array arithmetic, the same values gate.RX/RY/RZ and
    gate_derivative give."""

    KINDS = {"RX": 0, "RY": 1, "RZ": 2}

    @classmethod
    def make(cls, prog: AdjointProgram, x: Sequence[float]) -> Optional["_FastRotations"]:
        b = {n: float(v) for n, v in zip(prog.names, x)}
        ops = []
        for op in prog.circuit.Ops():
            g = op.Gate
            if g is None or op.ControlFlow is not None:
                continue
            if isinstance(g, SymbolicGate):
                if g.ctor_name not in cls.KINDS or len(g.exprs) != 1:
                    return None
            ops.append(op)
        self = cls()
        self.prog = prog
        buf = N.OpBuffer()
        rot = []  # (record index, kind, expression)
        for op in ops:
            g = op.Gate
            if isinstance(g, SymbolicGate):
                k = N.add_sv_gate(buf, g.Bind(b), op.Qubits)
                rot.append((k, cls.KINDS[g.ctor_name], g.exprs[0]))
            else:
                N.add_sv_gate(buf, g, op.Qubits)
        nr = len(rot)
        self.exprs = [e for _, _, e in rot]
        self.kind = np.array([kd for _, kd, _ in rot], dtype=np.int64)
        self.mats = np.zeros((max(nr, 1), 2, 2), dtype=np.complex128)
        flat = self.mats.reshape(-1).view(np.float64)
        for j, (k, _, _) in enumerate(rot):  # records point into one contiguous store
            r = buf.records[k]
            view = flat[8 * j:8 * j + 8]
            buf.records[k] = (r[0], r[1], r[2], r[3], view, r[5], view.ctypes.data)
        # derivative records: one per (rotation, free parameter of its expression)
        dref = []
        for j, (k, _, e) in enumerate(rot):
            for p in dict.fromkeys(e.free()):
                if p in prog.index:
                    dref.append((j, k, prog.index[p], p, e))
        self.dref = dref
        self.dmats = np.zeros((max(nr, 1), 2, 2), dtype=np.complex128)
        dflat = self.dmats.reshape(-1).view(np.float64)
        self.darr = (N.GoquDeriv * max(1, len(dref)))()
        for i, (j, k, pi, _, _) in enumerate(dref):
            self.darr[i].op = k
            self.darr[i].param = pi
            self.darr[i].dmatrix = N.dptr(dflat[8 * j:8 * j + 8])
        self.ops = _FrozenOps(buf)
        self.keep = [self.mats, self.dmats]
        return self

    def refill(self, x: Sequence[float]):
        b = {n: float(v) for n, v in zip(self.prog.names, x)}
        nr = len(self.exprs)
        if nr:
            th = np.fromiter((e.eval(b) for e in self.exprs), dtype=np.float64, count=nr)
            c, s = np.cos(th / 2), np.sin(th / 2)
            M, D, kd = self.mats[:nr], self.dmats[:nr], self.kind
            M[:] = 0
            D[:] = 0
            x_, y_, z_ = kd == 0, kd == 1, kd == 2
            # RX = [[c, -i s], [-i s, c]], dRX = [[-s/2, -i c/2], [-i c/2, -s/2]]
            M[x_, 0, 0] = M[x_, 1, 1] = c[x_]
            M[x_, 0, 1] = M[x_, 1, 0] = -1j * s[x_]
            D[x_, 0, 0] = D[x_, 1, 1] = -0.5 * s[x_]
            D[x_, 0, 1] = D[x_, 1, 0] = -0.5j * c[x_]
            # RY = [[c, -s], [s, c]], dRY = [[-s/2, -c/2], [c/2, -s/2]]
            M[y_, 0, 0] = M[y_, 1, 1] = c[y_]
            M[y_, 0, 1] = -s[y_]
            M[y_, 1, 0] = s[y_]
            D[y_, 0, 0] = D[y_, 1, 1] = -0.5 * s[y_]
            D[y_, 0, 1] = -0.5 * c[y_]
            D[y_, 1, 0] = 0.5 * c[y_]
            # RZ = diag(c - i s, c + i s), dRZ = diag(-s/2 - i c/2, -s/2 + i c/2)
            M[z_, 0, 0] = c[z_] - 1j * s[z_]
            M[z_, 1, 1] = c[z_] + 1j * s[z_]
            D[z_, 0, 0] = -0.5 * s[z_] - 0.5j * c[z_]
            D[z_, 1, 1] = -0.5 * s[z_] + 0.5j * c[z_]
        for i, (_, _, _, p, e) in enumerate(self.dref):
            self.darr[i].scale = e.deriv(p, b)
        return self.ops, self.darr, len(self.dref), self.keep


def Adjoint(circuit: Circuit, hamiltonian: PauliSum, param_names: Sequence[str], sim=None,
            precision: int = N.C128):
    """A GradientFunc-compatible constructor beside ParameterShift: x -> dE/dx via one
    forward and one reverse sweep on the device.  The returned function has a
    `.value_and_grad(x)` attribute returning (E, grads)."""
    from . import statevector

    prog = AdjointProgram(circuit, param_names)
    terms = hamiltonian.Terms()
    xm, zm, ny, cf = masks(terms)
    own = sim is None
    state = {"sim": sim}

    def value_and_grad(x):
        s = state["sim"]
        if s is None:
            s = statevector.New(circuit.NumQubits(), precision)
            if not own:
                state["sim"] = s
        buf, darr, nd, keep = prog.build(x)
        grads = np.zeros(len(prog.names))
        e = C.c_double(0)
        try:
            N.check(N.lib().goqu_sv_adjoint(s._h, buf.array(), len(buf), darr, nd, len(prog.names), len(terms),
                                             N.u64ptr(xm), N.u64ptr(zm), N.i32ptr(ny), N.dptr(cf), C.byref(e),
                                             N.dptr(grads)))
        finally:
            if own:
                s.Close()
        return e.value, grads

    def g(x):
        return value_and_grad(x)[1]

    g.value_and_grad = value_and_grad
    return g


def AdjointDensity(circuit: Circuit, hamiltonian: PauliSum, param_names: Sequence[str], noise=None,
                   checkpoint_every: int = 16):
    """Adjoint gradient of Tr(H rho(theta)) on a noisy density matrix (checkpointed)."""
    from . import densitymatrix

    prog = AdjointProgram(circuit, param_names, density=True, noise=noise)
    terms = hamiltonian.Terms()
    xm, zm, ny, cf = masks(terms)

    def value_and_grad(x):
        s = densitymatrix.New(circuit.NumQubits())
        buf, darr, nd, keep = prog.build(x)
        grads = np.zeros(len(prog.names))
        e = C.c_double(0)
        try:
            N.check(N.lib().goqu_dm_adjoint(s._h, buf.array(), len(buf), darr, nd, len(prog.names), len(terms),
                                             N.u64ptr(xm), N.u64ptr(zm), N.i32ptr(ny), N.dptr(cf),
                                             int(checkpoint_every), C.byref(e), N.dptr(grads)))
        finally:
            s.Close()
        return e.value, grads

    def g(x):
        return value_and_grad(x)[1]

    g.value_and_grad = value_and_grad
    return g
