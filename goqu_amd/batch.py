"""Batches of circuits on one register size, run together on the device (goqu_sv_batch_* /
goqu_dm_batch_*): the points of a parameter sweep, the circuits of a mitigation executor, the
parameter sets of a VQE step.  The reference evaluates each on a fresh simulator
(sweep/run.go:75-97, algorithm/mitigation/executor.go:20-41); the results here are the same
per circuit (energies summed over the terms in order, as ExpectPauliSum does, sim.go:355-363)."""
from __future__ import annotations

from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _native as N
from . import gate as G
from .circuit import Bind, Circuit, FreeParameters
from .pauli import PauliSum, masks


def _marshal_sv(circuits: Sequence[Circuit]) -> Tuple["N.OpBuffer", np.ndarray]:
    buf = N.OpBuffer()
    offs = [0]
    for c in circuits:
        for op in c.Ops():
            if op.Gate is not None:
                if isinstance(op.Gate, G.StatePrepGate):
                    raise NotImplementedError("StatePrep in a batch: evolve that circuit on a simulator")
                N.add_sv_gate(buf, op.Gate, op.Qubits)
        offs.append(len(buf))
    return buf, np.asarray(offs, dtype=np.int64)


def _marshal_dm(circuits: Sequence[Circuit], noise) -> Tuple["N.OpBuffer", np.ndarray]:
    from .densitymatrix import _marshal_op

    buf = N.OpBuffer()
    offs = [0]
    for c in circuits:
        for op in c.Ops():
            _marshal_op(buf, op, noise)
        offs.append(len(buf))
    return buf, np.asarray(offs, dtype=np.int64)


def _check(circuits: Sequence[Circuit]) -> int:
    if not circuits:
        return 0
    n = circuits[0].NumQubits()
    for c in circuits:
        if c.NumQubits() != n:
            raise ValueError("batch: every circuit must have the same qubit count")
        if c.IsDynamic():
            raise ValueError("batch: dynamic circuits run through RunDynamic")
    return n


def expect_terms(circuits: Sequence[Circuit], H: PauliSum, precision: int = N.C128,
                 plan: Tuple[int, int] = (0, 0)) -> np.ndarray:
    """coeff_t <psi_i|P_t|psi_i> for every circuit i and term t: complex array (circuits, terms)."""
    n = _check(circuits)
    terms = H.Terms()
    out = np.zeros((len(circuits), len(terms)), dtype=np.complex128)
    if not circuits:
        return out
    if H.NumQubits() != n:
        raise ValueError(f"batch: PauliSum has {H.NumQubits()} qubits, circuits have {n}")
    buf, offs = _marshal_sv(circuits)
    x, z, ny, cf = masks(terms)
    flat = out.view(np.float64).reshape(-1)
    N.check(N.lib().goqu_sv_batch_expect(n, precision, plan[0], plan[1], len(circuits), buf.array(), N.i64ptr(offs),
                                         len(terms), N.u64ptr(x), N.u64ptr(z), N.i32ptr(ny), N.dptr(cf),
                                         N.dptr(flat)))
    return out


def _sum_in_order(vals: np.ndarray) -> np.ndarray:
    e = np.empty(vals.shape[0])
    for i in range(vals.shape[0]):
        total = 0j
        for v in vals[i]:
            total += v
        e[i] = total.real
    return e


def energies(circuits: Sequence[Circuit], H: PauliSum, precision: int = N.C128,
             plan: Tuple[int, int] = (0, 0)) -> np.ndarray:
    """ExpectPauliSum of every circuit (StatevectorExecutor's value, executor.go:20-29)."""
    return _sum_in_order(expect_terms(circuits, H, precision, plan))


def sample(circuits: Sequence[Circuit], shots: int, uniforms: Optional[np.ndarray] = None, rng=None,
           precision: int = N.C128, plan: Tuple[int, int] = (0, 0)) -> np.ndarray:
    """Basis-state indices of `shots` samples per circuit (Run's sampling, sim.go:274-298):
    array (circuits, shots); row i uses uniforms[i]."""
    n = _check(circuits)
    out = np.zeros((len(circuits), max(shots, 0)), dtype=np.uint64)
    if not circuits or shots <= 0:
        return out
    if uniforms is None:
        rng = rng if rng is not None else np.random.default_rng()
        uniforms = rng.random((len(circuits), shots))
    u = np.ascontiguousarray(uniforms, dtype=np.float64).reshape(len(circuits), shots)
    buf, offs = _marshal_sv(circuits)
    N.check(N.lib().goqu_sv_batch_sample(n, precision, plan[0], plan[1], len(circuits), buf.array(), N.i64ptr(offs),
                                         shots, N.dptr(u), N.u64ptr(out)))
    return out


def dm_expect_terms(circuits: Sequence[Circuit], H: PauliSum, noise=None) -> np.ndarray:
    n = _check(circuits)
    terms = H.Terms()
    out = np.zeros((len(circuits), len(terms)), dtype=np.complex128)
    if not circuits:
        return out
    buf, offs = _marshal_dm(circuits, noise)
    x, z, ny, cf = masks(terms)
    flat = out.view(np.float64).reshape(-1)
    N.check(N.lib().goqu_dm_batch_expect(n, len(circuits), buf.array(), N.i64ptr(offs), len(terms), N.u64ptr(x),
                                         N.u64ptr(z), N.i32ptr(ny), N.dptr(cf), N.dptr(flat)))
    return out


def dm_energies(circuits: Sequence[Circuit], H: PauliSum, noise=None) -> np.ndarray:
    """Tr(rho H) of every circuit (DensityMatrixExecutor's value, executor.go:31-41)."""
    return _sum_in_order(dm_expect_terms(circuits, H, noise))


def dm_sample(circuits: Sequence[Circuit], shots: int, noise=None, uniforms: Optional[np.ndarray] = None,
              rng=None) -> np.ndarray:
    """Samples of each circuit's diagonal after the noise model's readout errors (DM Run,
    sim.go:51-67)."""
    n = _check(circuits)
    out = np.zeros((len(circuits), max(shots, 0)), dtype=np.uint64)
    if not circuits or shots <= 0:
        return out
    if uniforms is None:
        rng = rng if rng is not None else np.random.default_rng()
        uniforms = rng.random((len(circuits), shots))
    u = np.ascontiguousarray(uniforms, dtype=np.float64).reshape(len(circuits), shots)
    p01 = p10 = None
    if noise is not None:
        from .densitymatrix import _readout_arrays

        p01, p10 = _readout_arrays(noise, n)
    buf, offs = _marshal_dm(circuits, noise)
    N.check(N.lib().goqu_dm_batch_sample(n, len(circuits), buf.array(), N.i64ptr(offs),
                                         N.dptr(p01) if p01 is not None else None,
                                         N.dptr(p10) if p10 is not None else None, shots, N.dptr(u),
                                         N.u64ptr(out)))
    return out


# Matrices of the parameterised gates at many parameter values at once (same entries as the
# gate constructors in goqu_amd.gate, evaluated with numpy over the batch).
def _e(x):
    return np.exp(1j * x)


def _stack(*cols):
    P = max(np.shape(c)[0] for c in cols if np.ndim(c))
    return np.stack([np.broadcast_to(np.asarray(c, dtype=np.complex128), (P,)) for c in cols], axis=1)


def _rx(t):
    c, s = np.cos(t / 2), np.sin(t / 2)
    return _stack(c, -1j * s, -1j * s, c)


def _ry(t):
    c, s = np.cos(t / 2), np.sin(t / 2)
    return _stack(c, -s, s, c)


def _rz(t):
    return _stack(_e(-t / 2), 0, 0, _e(t / 2))


def _phase(p):
    return _stack(1, 0, 0, _e(p))


def _u3(t, p, l):
    c, s = np.cos(t / 2), np.sin(t / 2)
    return _stack(c, -_e(l) * s, _e(p) * s, _e(p + l) * c)


def _diag4(a, b, c, d):
    z = 0
    return _stack(a, z, z, z, z, b, z, z, z, z, c, z, z, z, z, d)


def _ctrl_block(a, b, c, d):
    z = 0
    return _stack(1, z, z, z, z, 1, z, z, z, z, a, b, z, z, c, d)


def _rxx(t):
    c, s = np.cos(t / 2), np.sin(t / 2)
    i, z = -1j * s, 0
    return _stack(c, z, z, i, z, c, i, z, z, i, c, z, i, z, z, c)


def _ryy(t):
    c, s = np.cos(t / 2), np.sin(t / 2)
    ip, im, z = 1j * s, -1j * s, 0
    return _stack(c, z, z, ip, z, c, im, z, z, im, c, z, ip, z, z, c)


def _rzz(t):
    em, ep = _e(-t / 2), _e(t / 2)
    return _diag4(em, ep, ep, em)


_VECTOR_MATRICES = {
    "RX": _rx, "RY": _ry, "RZ": _rz, "P": _phase, "U1": _phase, "U3": _u3,
    "CP": lambda p: _diag4(1, 1, 1, _e(p)), "CRZ": lambda t: _diag4(1, 1, _e(-t / 2), _e(t / 2)),
    "CRX": lambda t: _ctrl_block(np.cos(t / 2), -1j * np.sin(t / 2), -1j * np.sin(t / 2), np.cos(t / 2)),
    "CRY": lambda t: _ctrl_block(np.cos(t / 2), -np.sin(t / 2), np.sin(t / 2), np.cos(t / 2)),
    "RXX": _rxx, "RYY": _ryy, "RZZ": _rzz, "ZZ": _rzz,
}


def _eval_rows(e, cols: dict, P: int) -> np.ndarray:
    """An expression's value at every row (the parameter columns in cols)."""
    from .circuit import Add, Literal, Mul, Parameter

    if isinstance(e, Literal):
        return np.full(P, float(e.v))
    if isinstance(e, Parameter):
        if e.Name() not in cols:
            raise KeyError(f"param: unbound parameter {e.Name()!r}")
        return cols[e.Name()]
    if isinstance(e, Mul):
        return _eval_rows(e.a, cols, P) * _eval_rows(e.b, cols, P)
    if isinstance(e, Add):
        return _eval_rows(e.a, cols, P) + _eval_rows(e.b, cols, P)
    names = list(cols)
    return np.array([e.eval({k: float(cols[k][p]) for k in names}) for p in range(P)])


class ParamRecords:
    """Op records of one parameterised circuit at every row of a parameter matrix, built once per
    op (its matrices for all rows at once) rather than once per bound circuit: instance p's ops
    are records[p * ops_per_circuit ...].  Keeps the matrix storage alive."""

    def __init__(self, c: Circuit, params: np.ndarray, names: Sequence[str]):
        from .circuit import CONSTRUCTORS, SymbolicGate

        params = np.atleast_2d(np.asarray(params, dtype=np.float64))
        P = params.shape[0]
        cols = {nm: params[:, j] for j, nm in enumerate(names)}
        ops = [op for op in c.Ops() if op.Gate is not None]
        nops = len(ops)
        tmpl = N.OpBuffer()  # kinds, tiers and qubits from the first row (they do not depend on values)
        per_row = []  # (op index, matrices (P, d*d)) of the parameterised ops
        for i, op in enumerate(ops):
            g = op.Gate
            if isinstance(g, SymbolicGate):
                vals = [_eval_rows(e, cols, P) for e in g.exprs]
                g0 = CONSTRUCTORS[g.ctor_name](*[v[0] for v in vals])
                fn = _VECTOR_MATRICES.get(g.ctor_name)
                if fn is not None:
                    M = fn(*vals)
                else:
                    M = np.stack([np.asarray(CONSTRUCTORS[g.ctor_name](*[v[p] for v in vals]).Matrix(),
                                             dtype=np.complex128).reshape(-1) for p in range(P)])
                per_row.append((i, np.ascontiguousarray(M)))
                N.add_sv_gate(tmpl, g0, op.Qubits)
            else:
                if isinstance(g, G.StatePrepGate):
                    raise NotImplementedError("StatePrep in a batch: evolve that circuit on a simulator")
                N.add_sv_gate(tmpl, g, op.Qubits)
        base = tmpl.structured()[:nops]
        rec = np.tile(base, P)
        self._keep = [tmpl]
        for i, M in per_row:
            self._keep.append(M)
            stride = M.shape[1] * 16
            rec["matrix"][i::nops] = M.ctypes.data + stride * np.arange(P, dtype=np.uint64)
        self.records = rec
        self.offsets = np.arange(P + 1, dtype=np.int64) * nops
        self.count = P

    def array(self):
        return N.op_array(self.records)


def param_expect_terms(c: Circuit, H: PauliSum, params: np.ndarray, names: Optional[List[str]] = None,
                       precision: int = N.C128, plan: Tuple[int, int] = (0, 0)) -> np.ndarray:
    """expect_terms of c at every row of params, marshalled through ParamRecords."""
    if c.IsDynamic():
        raise ValueError("batch: dynamic circuits run through RunDynamic")
    names = names if names is not None else FreeParameters(c)
    R = ParamRecords(c, params, names)
    n = c.NumQubits()
    terms = H.Terms()
    if H.NumQubits() != n:
        raise ValueError(f"batch: PauliSum has {H.NumQubits()} qubits, circuits have {n}")
    out = np.zeros((R.count, len(terms)), dtype=np.complex128)
    x, z, ny, cf = masks(terms)
    flat = out.view(np.float64).reshape(-1)
    N.check(N.lib().goqu_sv_batch_expect(n, precision, plan[0], plan[1], R.count, R.array(), N.i64ptr(R.offsets),
                                         len(terms), N.u64ptr(x), N.u64ptr(z), N.i32ptr(ny), N.dptr(cf),
                                         N.dptr(flat)))
    return out


def param_sample(c: Circuit, params: np.ndarray, shots: int, names: Optional[List[str]] = None,
                 uniforms: Optional[np.ndarray] = None, rng=None, precision: int = N.C128,
                 plan: Tuple[int, int] = (0, 0)) -> np.ndarray:
    """sample() of c at every row of params, marshalled through ParamRecords."""
    names = names if names is not None else FreeParameters(c)
    R = ParamRecords(c, params, names)
    out = np.zeros((R.count, max(shots, 0)), dtype=np.uint64)
    if shots <= 0 or not R.count:
        return out
    if uniforms is None:
        rng = rng if rng is not None else np.random.default_rng()
        uniforms = rng.random((R.count, shots))
    u = np.ascontiguousarray(uniforms, dtype=np.float64).reshape(R.count, shots)
    N.check(N.lib().goqu_sv_batch_sample(c.NumQubits(), precision, plan[0], plan[1], R.count, R.array(),
                                         N.i64ptr(R.offsets), shots, N.dptr(u), N.u64ptr(out)))
    return out


def bind_all(c: Circuit, params: np.ndarray, names: Optional[List[str]] = None) -> List[Circuit]:
    """One bound circuit per row of params (columns in ir.FreeParameters order unless names)."""
    names = names if names is not None else FreeParameters(c)
    params = np.atleast_2d(np.asarray(params, dtype=np.float64))
    if params.shape[1] != len(names):
        raise ValueError(f"params: {params.shape[1]} columns for {len(names)} free parameters")
    return [Bind(c, dict(zip(names, row))) for row in params]


def vqe_energies(c: Circuit, H: PauliSum, params: np.ndarray, names: Optional[List[str]] = None,
                 precision: int = N.C128, plan: Tuple[int, int] = (0, 0)) -> np.ndarray:
    """<H> at every parameter set (rows of params) of a parameterised ansatz, in one batch."""
    return _sum_in_order(param_expect_terms(c, H, params, names, precision, plan))
