"""ctypes binding of the CPU oracle (oracle/oracle.c).  TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker or the
timed CPU baseline; the goqu_amd package never imports it.  Op records are the C ABI's
(goqu_op_t), so the oracle consumes exactly what the device library consumes.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path
from typing import Optional, Sequence

import numpy as np

from goqu_amd import _native as N
from goqu_amd import gate as G
from goqu_amd.circuit import Circuit

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"
_lib = None

_D = C.POINTER(C.c_double)
_U64 = C.POINTER(C.c_uint64)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
        L = C.CDLL(str(LIB))
        L.orc_set_workers.argtypes = [C.c_int]
        L.orc_sv_apply.argtypes = [_D, C.c_int, C.POINTER(N.GoquOp), C.c_int]
        L.orc_sv_apply.restype = C.c_int
        L.orc_sv_sample.argtypes = [_D, C.c_int, _D, C.c_int64, _U64]
        L.orc_sample_probs.argtypes = [_D, C.c_int64, _D, C.c_int64, _U64]
        L.orc_sv_probabilities.argtypes = [_D, C.c_int, _D]
        L.orc_sv_expect.argtypes = [_D, C.c_int, C.c_uint64, C.c_uint64, C.c_int, C.c_double, C.c_double, _D]
        L.orc_sv_expectation_value.argtypes = [_D, C.c_int, C.c_uint64, C.c_int]
        L.orc_sv_expectation_value.restype = C.c_double
        L.orc_sv_measure_qubit.argtypes = [_D, C.c_int, C.c_int, C.c_double]
        L.orc_sv_measure_qubit.restype = C.c_int
        L.orc_stream_copy.argtypes = [C.c_int64, C.c_int, C.c_int]
        L.orc_stream_copy.restype = C.c_double
        L.orc_dm_measure_qubit.argtypes = [_D, C.c_int, C.c_int, C.c_double]
        L.orc_dm_measure_qubit.restype = C.c_int
        L.orc_dm_apply.argtypes = [_D, C.c_int, C.POINTER(N.GoquOp), C.c_int]
        L.orc_dm_apply.restype = C.c_int
        L.orc_dm_diag_probs.argtypes = [_D, C.c_int, _D]
        L.orc_dm_readout.argtypes = [_D, C.c_int, _D, _D, _D]
        L.orc_dm_expect.argtypes = [_D, C.c_int, C.c_uint64, C.c_uint64, C.c_int, C.c_double, C.c_double, _D]
        L.orc_dm_expectation_value.argtypes = [_D, C.c_int, C.c_uint64, C.c_int]
        L.orc_dm_expectation_value.restype = C.c_double
        L.orc_dm_purity.argtypes = [_D, C.c_int]
        L.orc_dm_purity.restype = C.c_double
        L.orc_dm_fidelity.argtypes = [_D, C.c_int, _D]
        L.orc_dm_fidelity.restype = C.c_double
        _lib = L
    return _lib


def set_workers(w: int) -> None:
    lib().orc_set_workers(int(w))


def _dp(a: np.ndarray):
    return a.view(np.float64).ctypes.data_as(_D)


def zero_state(n: int) -> np.ndarray:
    s = np.zeros(1 << n, dtype=np.complex128)
    s[0] = 1
    return s


def sv_apply_ops(state: np.ndarray, n: int, buf: "N.OpBuffer") -> int:
    if not len(buf):
        return 0
    return lib().orc_sv_apply(_dp(state), n, buf.array(), len(buf))


def sv_marshal(c: Circuit) -> "N.OpBuffer":
    buf = N.OpBuffer()
    for op in c.Ops():
        if op.Gate is None or op.ControlFlow is not None:
            continue
        N.add_sv_gate(buf, op.Gate, op.Qubits)
    return buf


def sv_evolve(c: Circuit, state: Optional[np.ndarray] = None) -> np.ndarray:
    """Evolve (sim/statevector/sim.go:82-85) on the CPU oracle.  StatePrep follows sim.go:105-136:
    a full-register copy on qubits [0..n-1], else the gate's decomposition."""
    n = c.NumQubits()
    s = zero_state(n) if state is None else np.ascontiguousarray(state, dtype=np.complex128).copy()
    buf = N.OpBuffer()

    def flush():
        rc = sv_apply_ops(s, n, buf)
        if rc:
            raise RuntimeError(f"oracle apply failed with status {rc}")

    for op in c.Ops():
        g = op.Gate
        if g is None or op.ControlFlow is not None:
            continue
        if isinstance(g, G.StatePrepGate):
            flush()
            buf = N.OpBuffer()
            if list(op.Qubits) == list(range(n)):
                s[:] = g.Amplitudes()
            else:
                for dg, dq in g.Decompose(op.Qubits):
                    if dg.Matrix() is not None:
                        N.add_sv_gate(buf, dg, dq)
            continue
        N.add_sv_gate(buf, g, op.Qubits)
    flush()
    return s


def sv_sample(state: np.ndarray, uniforms: np.ndarray) -> np.ndarray:
    n = int(np.log2(len(state)))
    u = np.ascontiguousarray(uniforms, dtype=np.float64)
    out = np.empty(len(u), dtype=np.uint64)
    lib().orc_sv_sample(_dp(state), n, u.ctypes.data_as(_D), len(u), out.ctypes.data_as(_U64))
    return out


def sample_probs(probs: np.ndarray, uniforms: np.ndarray) -> np.ndarray:
    p = np.ascontiguousarray(probs, dtype=np.float64)
    u = np.ascontiguousarray(uniforms, dtype=np.float64)
    out = np.empty(len(u), dtype=np.uint64)
    lib().orc_sample_probs(p.ctypes.data_as(_D), len(p), u.ctypes.data_as(_D), len(u), out.ctypes.data_as(_U64))
    return out


def sv_expect(state: np.ndarray, ps) -> complex:
    """pauli.Expect (sim/pauli/expect_sv.go:21-47)."""
    n = int(np.log2(len(state)))
    out = np.zeros(2)
    lib().orc_sv_expect(_dp(state), n, ps.xmask(), ps.zmask(), ps.numY(), ps.coeff.real, ps.coeff.imag,
                        out.ctypes.data_as(_D))
    return complex(out[0], out[1])


def sv_expectation_value(state: np.ndarray, qubits: Sequence[int], round14: bool = True) -> float:
    n = int(np.log2(len(state)))
    mask = sum(1 << q for q in qubits)
    return lib().orc_sv_expectation_value(_dp(state), n, mask, int(round14))


def sv_measure_qubit(state: np.ndarray, q: int, r: float) -> int:
    n = int(np.log2(len(state)))
    return lib().orc_sv_measure_qubit(_dp(state), n, q, r)


def dm_zero(n: int) -> np.ndarray:
    r = np.zeros(1 << (2 * n), dtype=np.complex128)
    r[0] = 1
    return r


def dm_apply_ops(rho: np.ndarray, n: int, buf: "N.OpBuffer") -> int:
    if not len(buf):
        return 0
    return lib().orc_dm_apply(_dp(rho), n, buf.array(), len(buf))


def dm_measure_qubit(rho: np.ndarray, n: int, q: int, r: float) -> int:
    return lib().orc_dm_measure_qubit(_dp(rho), n, q, r)


def dm_probs(rho: np.ndarray, n: int, p01=None, p10=None) -> np.ndarray:
    p = np.empty(1 << n)
    lib().orc_dm_diag_probs(_dp(rho), n, p.ctypes.data_as(_D))
    out = np.empty(1 << n)
    a = None if p01 is None else np.ascontiguousarray(p01, dtype=np.float64)
    b = None if p10 is None else np.ascontiguousarray(p10, dtype=np.float64)
    lib().orc_dm_readout(p.ctypes.data_as(_D), n, a.ctypes.data_as(_D) if a is not None else None,
                         b.ctypes.data_as(_D) if b is not None else None, out.ctypes.data_as(_D))
    return out


def dm_expect(rho: np.ndarray, n: int, ps) -> complex:
    out = np.zeros(2)
    lib().orc_dm_expect(_dp(rho), n, ps.xmask(), ps.zmask(), ps.numY(), ps.coeff.real, ps.coeff.imag,
                        out.ctypes.data_as(_D))
    return complex(out[0], out[1])


def dm_expectation_value(rho: np.ndarray, n: int, qubits: Sequence[int], round14: bool = True) -> float:
    return lib().orc_dm_expectation_value(_dp(rho), n, sum(1 << q for q in qubits), int(round14))


def dm_purity(rho: np.ndarray, n: int) -> float:
    return lib().orc_dm_purity(_dp(rho), n)


def dm_fidelity(rho: np.ndarray, n: int, pure: np.ndarray) -> float:
    p = np.ascontiguousarray(pure, dtype=np.complex128)
    return lib().orc_dm_fidelity(_dp(rho), n, _dp(p))


def stream_copy_gbps(n_doubles: int = 1 << 27, threads: int = 1, reps: int = 3) -> float:
    """Host copy bandwidth (bytes read + written per second, GB/s) with `threads` threads."""
    return lib().orc_stream_copy(n_doubles, threads, reps)
