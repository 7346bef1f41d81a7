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
    return energies(bind_all(c, params, names), H, precision, plan)
